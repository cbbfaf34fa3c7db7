"""Seeded synthetic river networks in the shape of CaMa-Flood's global maps.

Generators (all deterministic for a given seed):

* elevation: fractal Gaussian field whose 1-D (profile) power spectrum falls as
  k^-2, i.e. an isotropic 2-D spectrum ~ k^-3, on a global lat-lon grid (x wraps);
  land = top ``land_frac`` of the field;
* drainage: D8 directions from priority-flood on the land mask, river mouths at
  the coast (``nextx = -9``);
* channel: river length = cell-centre distance x U(1.1, 1.5); width =
  clip(0.6 A^0.5, 10, 3000) m; bank height = clip(0.1 A^0.4, 1, 20) m with A the
  upstream area in km^2; Manning n = 0.03;
* floodplain: ten layers with exponential(mean 1 m) height increments;
* runoff: spatially correlated gamma(k=0.5, mean 1.5 mm/day) through a Gaussian
  copula, converted to m3/s per cell;
* optional bifurcation paths, levees and dams.

Maps are float32 2-D arrays indexed [iy, ix] (x fastest in memory), exactly the
layout of the reference's binary map files (NX,NY Fortran order).
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Optional

import numpy as np

from . import _native
from ._abi import ptr

R_EARTH = 6371000.0
NLFP = 10


@dataclasses.dataclass
class SynthNetwork:
    nx: int
    ny: int
    west: float
    east: float
    north: float
    south: float
    nextx: np.ndarray      # int32 [ny,nx], 1-based or -9 (mouth) / -9999 (none)
    nexty: np.ndarray
    elevtn: np.ndarray     # float32 [ny,nx]
    nxtdst: np.ndarray
    rivlen: np.ndarray
    rivwth: np.ndarray
    rivhgt: np.ndarray
    rivman: np.ndarray
    ctmare: np.ndarray     # cell area [m2]
    fldhgt: np.ndarray     # float32 [NLFP,ny,nx]
    uparea: np.ndarray     # float64 [ny,nx] upstream area [m2]
    # optional features
    bifprm: Optional[list] = None     # rows (ix,iy,jx,jy,dst,elv,dph,[wth...]) 1-based
    npthlev: int = 0
    levhgt: Optional[np.ndarray] = None
    levfrc: Optional[np.ndarray] = None
    dams: Optional[list] = None       # rows of the dam table
    slpmix_mask: Optional[np.ndarray] = None   # int32 [ny,nx], 1 = kinematic wave (LSLPMIX)

    @property
    def land(self) -> np.ndarray:
        return self.nextx != -9999

    @property
    def ncells(self) -> int:
        return int(self.land.sum())


def _fractal(nx, ny, beta, rng):
    white = rng.standard_normal((ny, nx))
    spec = np.fft.rfft2(white)
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = np.inf
    spec *= k ** (-beta / 2.0)
    z = np.fft.irfft2(spec, s=(ny, nx))
    z -= z.mean()
    z /= z.std()
    return z


def _cell_geometry(nx, ny, west, east, north, south):
    dlon = (east - west) / nx
    dlat = (north - south) / ny
    lat_c = north - (np.arange(ny) + 0.5) * dlat
    lon_c = west + (np.arange(nx) + 0.5) * dlon
    lat_top = np.radians(north - np.arange(ny) * dlat)
    lat_bot = np.radians(north - (np.arange(ny) + 1) * dlat)
    area_row = R_EARTH ** 2 * np.radians(dlon) * (np.sin(lat_top) - np.sin(lat_bot))
    area = np.repeat(area_row[:, None], nx, axis=1)
    return lat_c, lon_c, area


def _haversine(lat1, lon1, lat2, lon2):
    p1, p2 = np.radians(lat1), np.radians(lat2)
    dp = p2 - p1
    dl = np.radians(lon2 - lon1)
    a = np.sin(dp / 2) ** 2 + np.cos(p1) * np.cos(p2) * np.sin(dl / 2) ** 2
    return 2 * R_EARTH * np.arcsin(np.sqrt(np.minimum(a, 1.0)))


def make_network(nx: int, ny: int, seed: int, land_frac: float = 0.2315,
                 relief: float = 3000.0, west=-180.0, east=180.0, north=90.0,
                 south=-90.0, regional: bool = False, max_lat: Optional[float] = None) -> SynthNetwork:
    """River network on an nx x ny grid (see module docstring). regional=True
    makes the outer ring of the grid ocean, so that no river crosses the edges
    of a regional extent. max_lat makes every cell poleward of +-max_lat ocean
    (land_frac stays the fraction of the whole grid)."""
    rng = np.random.default_rng(seed)
    z = _fractal(nx, ny, 3.0, rng)
    if regional:
        z[0, :] = z[-1, :] = z.min()
        z[:, 0] = z[:, -1] = z.min()
    if max_lat is not None:
        lat_row = north - (np.arange(ny) + 0.5) * ((north - south) / ny)
        z[np.abs(lat_row) > max_lat, :] = z.min()
    thr = np.quantile(z, 1.0 - land_frac)
    land = z > thr
    elev = np.where(land, (z - thr) / (z.max() - thr) * relief, 0.0).astype(np.float32)

    lib = _native.host()
    landu8 = np.ascontiguousarray(land.astype(np.uint8))
    filled = np.empty((ny, nx), np.float32)
    nextx = np.empty((ny, nx), np.int32)
    nexty = np.empty((ny, nx), np.int32)
    rc = lib.cmfh_priority_flood(nx, ny, ptr(elev), ptr(landu8), np.float32(0.01),
                                 ptr(filled), ptr(nextx), ptr(nexty))
    if rc != 0:
        raise RuntimeError(f"priority flood failed ({rc})")
    lat_c, lon_c, area = _cell_geometry(nx, ny, west, east, north, south)
    area = np.ascontiguousarray(area)
    uparea = np.empty((ny, nx), np.float64)
    lib.cmfh_upstream_area(nx, ny, ptr(nextx), ptr(nexty), ptr(area), ptr(uparea))

    mouth = nextx == -9
    river = nextx > 0
    elevtn = np.where(land, filled, -9999.0).astype(np.float32)
    elevtn[mouth] = 0.0                            # sea level 0 at the outlets

    # distance to the downstream cell centre (mouths: mean cell size)
    LAT = np.broadcast_to(lat_c[:, None], (ny, nx))
    LON = np.broadcast_to(lon_c[None, :], (ny, nx))
    jx = np.clip(nextx - 1, 0, nx - 1)
    jy = np.clip(nexty - 1, 0, ny - 1)
    dist = _haversine(LAT, LON, lat_c[jy], lon_c[jx])
    cellsize = np.sqrt(area)
    dist = np.where(river, dist, cellsize)
    nxtdst = np.where(land, dist, -9999.0).astype(np.float32)
    rivlen = np.where(land, dist * rng.uniform(1.1, 1.5, (ny, nx)), -9999.0).astype(np.float32)

    a_km2 = uparea * 1e-6
    rivwth = np.where(land, np.clip(0.6 * np.sqrt(a_km2), 10.0, 3000.0), -9999.0).astype(np.float32)
    rivhgt = np.where(land, np.clip(0.1 * a_km2 ** 0.4, 1.0, 20.0), -9999.0).astype(np.float32)
    rivman = np.where(land, 0.03, -9999.0).astype(np.float32)
    ctmare = np.where(land, area, -9999.0).astype(np.float32)
    inc = rng.exponential(1.0, (NLFP, ny, nx))
    fldhgt = np.where(land[None], np.cumsum(inc, axis=0), -9999.0).astype(np.float32)
    return SynthNetwork(nx, ny, west, east, north, south, nextx, nexty, elevtn, nxtdst,
                        rivlen, rivwth, rivhgt, rivman, ctmare, fldhgt, uparea)


def add_bifurcations(net: SynthNetwork, npaths: int, nlev: int, seed: int,
                     wmin=10.0, wmax=500.0) -> None:
    """Bifurcation paths between D8-adjacent land cells that are not already
    linked by the main drainage, widths log-uniform in [wmin, wmax] per level."""
    rng = np.random.default_rng(seed + 101)
    nx, ny = net.nx, net.ny
    land_idx = np.flatnonzero(net.land.ravel())
    rows = []
    seen = set()
    DX = np.array([1, 1, 0, -1, -1, -1, 0, 1])
    DY = np.array([0, 1, 1, 1, 0, -1, -1, -1])
    lat_c, lon_c, _ = _cell_geometry(nx, ny, net.west, net.east, net.north, net.south)
    tries = 0
    while len(rows) < npaths and tries < npaths * 50:
        tries += 1
        k = int(land_idx[rng.integers(len(land_idx))])
        iy, ix = divmod(k, nx)
        d = int(rng.integers(8))
        jy = int(iy + DY[d])
        if jy < 0 or jy >= ny:
            continue
        jx = int((ix + DX[d]) % nx)
        if not net.land[jy, jx]:
            continue
        if (net.nextx[iy, ix] == jx + 1 and net.nexty[iy, ix] == jy + 1) or \
           (net.nextx[jy, jx] == ix + 1 and net.nexty[jy, jx] == iy + 1):
            continue
        key = (min(k, jy * nx + jx), max(k, jy * nx + jx))
        if key in seen:
            continue
        seen.add(key)
        dst = float(_haversine(lat_c[iy], lon_c[ix], lat_c[jy], lon_c[jx]))
        elv = float(min(net.elevtn[iy, ix], net.elevtn[jy, jx]))
        dph = float(net.rivhgt[iy, ix]) * float(rng.uniform(0.3, 1.0))
        wth = np.exp(rng.uniform(math.log(wmin), math.log(wmax), nlev))
        rows.append((ix + 1, iy + 1, jx + 1, jy + 1, dst, elv, dph, [float(w) for w in wth]))
    net.bifprm = rows
    net.npthlev = nlev


def add_levees(net: SynthNetwork, frac: float, seed: int, hmin=2.0, hmax=8.0,
               fmin=0.2, fmax=0.8) -> None:
    rng = np.random.default_rng(seed + 202)
    land = net.land
    pick = land & (rng.random(land.shape) < frac)
    net.levhgt = np.where(pick, rng.uniform(hmin, hmax, land.shape), 0.0).astype(np.float32)
    net.levfrc = np.where(pick, rng.uniform(fmin, fmax, land.shape), 1.0).astype(np.float32)


def add_dams(net: SynthNetwork, ndam: int, seed: int, mu=math.log(1e8), sigma=1.5,
             mean_runoff_mm_day=1.5, years=None) -> None:
    """Dams on river cells with large upstream area; capacity lognormal(mu, sigma) m3.
    Flood-control volume 30%, conservation volume 50% of capacity; normal discharge
    from mean runoff over the upstream area, flood discharge 3x normal. years (a
    sequence, cycled over the dams) gives every dam a construction year, the last
    column of the reference's dam table with LDAMYBY (damout:173-175)."""
    rng = np.random.default_rng(seed + 303)
    river = net.nextx > 0
    idx = np.flatnonzero(river.ravel())
    up = net.uparea.ravel()[idx]
    cand = idx[up >= np.quantile(up, 0.8)] if len(idx) > 50 else idx
    ndam = min(ndam, len(cand))
    sel = np.sort(rng.choice(cand, ndam, replace=False))
    lat_c, lon_c, _ = _cell_geometry(net.nx, net.ny, net.west, net.east, net.north, net.south)
    rows = []
    for d, k in enumerate(sel):
        iy, ix = divmod(int(k), net.nx)
        cap = float(np.exp(rng.normal(mu, sigma)))
        a = float(net.uparea[iy, ix])
        qn = mean_runoff_mm_day * 1e-3 / 86400.0 * a
        rows.append(dict(id=d + 1, name=f"SYN{d + 1:05d}", lat=float(lat_c[iy]), lon=float(lon_c[ix]),
                         upreal=a * 1e-6, ix=ix + 1, iy=iy + 1,
                         fldvol=0.3 * cap * 1e-6, convol=0.5 * cap * 1e-6, totvol=cap * 1e-6,
                         qn=qn, qf=3.0 * qn))
        if years is not None:
            rows[-1]["year"] = int(years[d % len(years)])
    net.dams = rows


def add_slope_mask(net: SynthNetwork, frac: float) -> None:
    """The LSLPMIX mask (mask_slope of the reference's netCDF river map): 1 on the
    fraction `frac` of river cells with the steepest bed slope to their downstream
    cell, where the kinematic wave replaces the local inertial equations."""
    ny, nx = net.ny, net.nx
    iy, ix = np.nonzero(net.nextx > 0)
    jx, jy = net.nextx[iy, ix] - 1, net.nexty[iy, ix] - 1
    slope = (net.elevtn[iy, ix].astype(np.float64) - net.elevtn[jy, jx]) / net.nxtdst[iy, ix]
    mask = np.zeros((ny, nx), np.int32)
    k = int(round(frac * slope.size))
    if k > 0:
        sel = np.argsort(-slope, kind="stable")[:k]
        mask[iy[sel], ix[sel]] = 1
    net.slpmix_mask = mask


def make_runoff(net: SynthNetwork, nrec: int, seed: int, mean_mm_day=1.5, shape=0.5,
                beta=3.0, scale_factors=None) -> np.ndarray:
    """Runoff records [nrec, ny, nx] float64 in m3/s per cell: a smooth Gaussian field
    per record mapped to gamma(shape, mean) marginals (Gaussian copula)."""
    from scipy.special import gammaincinv, ndtr
    rng = np.random.default_rng(seed + 404)
    out = np.zeros((nrec, net.ny, net.nx), np.float64)
    land = net.land
    area = net.ctmare.astype(np.float64)
    theta = mean_mm_day / shape
    for r in range(nrec):
        g = _fractal(net.nx, net.ny, beta, rng)
        u = np.clip(ndtr(g[land]), 1e-12, 1 - 1e-12)
        mm = gammaincinv(shape, u) * theta
        if scale_factors is not None:
            mm = mm * scale_factors[r]
        q = np.zeros((net.ny, net.nx))
        q[land] = mm * 1e-3 / 86400.0 * area[land]
        out[r] = q
    return out


# ---------------------------------------------------------------------------
# Regional domains of the batched runoff ensemble

def ensemble_domain_sizes(ndom: int = 16, smin: float = 2.0e4, smax: float = 4.0e5,
                          power: float = 10.0) -> np.ndarray:
    """Target river-cell counts of ndom regional domains of uneven size, from smin
    to smax with most domains near smax: smin + (smax - smin) (1 - (1 - x)^power),
    x = 0 .. 1 (16 domains: 2e4 .. 4e5 cells, about 5.7e6 in total)."""
    x = np.arange(ndom) / max(ndom - 1, 1)
    return smin + (smax - smin) * (1.0 - (1.0 - x) ** power)


def regional_networks(sizes, seed: int, land_frac: float = 0.6, res: float = 0.1) -> list:
    """One regional network of about sizes[d] river cells per domain, on a res-degree
    grid of aspect 3:2 (outer ring ocean, so no river leaves the domain), placed
    side by side in longitude; seed + 101 d seeds domain d."""
    nets = []
    for d, target in enumerate(sizes):
        side = int(np.sqrt(target / land_frac / 1.5))
        nx, ny = int(side * 1.5), side
        north = 70.0 - 8.0 * (d % 8)
        west = -180.0 + 20.0 * (d % 18)
        nets.append(make_network(nx, ny, seed=seed + 101 * d, land_frac=land_frac, west=west,
                                 east=west + res * nx, north=north, south=north - res * ny, regional=True))
    return nets


# ---------------------------------------------------------------------------
# Runoff on a coarser input grid and its input matrix (the inpmat that the
# reference's map tools write for LINTERP, src/cmf_ctrl_forcing_mod.F90:418-458)

RMIS = float(np.float32(1e20))    # the reference's default RMIS (float) widened to double


def make_inpmat(net: SynthNetwork, cells_seq_xy, res_in: float):
    """Input matrix of the NSEQ cells (seqx, seqy 1-based) on an input grid of
    res_in degrees over the network's extent: for every model cell the input
    cells it overlaps (row-major order, at most 4 when res_in >= the model
    resolution) and the overlap areas [m2] on the sphere. Returns (inpx, inpy,
    inpa as [inpn, n] arrays, nxin, nyin)."""
    seqx, seqy = (np.asarray(a) for a in cells_seq_xy)
    dx = (net.east - net.west) / net.nx
    dy = (net.north - net.south) / net.ny
    if res_in < max(dx, dy) * (1.0 - 1e-9):
        raise ValueError("make_inpmat: the input grid must not be finer than the model grid")
    nxin = int(round((net.east - net.west) / res_in))
    nyin = int(round((net.north - net.south) / res_in))
    lon0 = net.west + (seqx - 1) * dx
    lat0 = net.north - (seqy - 1) * dy                       # north edge
    n = seqx.size
    parts = []
    for oy in (0, 1):
        for ox in (0, 1):
            jx = np.floor((lon0 - net.west) / res_in + 1e-9).astype(np.int64) + ox
            jy = np.floor((net.north - lat0) / res_in + 1e-9).astype(np.int64) + oy
            w = np.clip(np.minimum(lon0 + dx, net.west + (jx + 1) * res_in) - np.maximum(lon0, net.west + jx * res_in),
                        0.0, None)
            top = np.minimum(lat0, net.north - jy * res_in)
            bot = np.maximum(lat0 - dy, net.north - (jy + 1) * res_in)
            h = np.clip(np.sin(np.radians(top)) - np.sin(np.radians(bot)), 0.0, None)
            area = R_EARTH ** 2 * np.radians(w) * h
            ok = (area > 0) & (jx < nxin) & (jy < nyin)
            parts.append((np.where(ok, jx + 1, 0), np.where(ok, jy + 1, 0), np.where(ok, area, 0.0)))
    inpx = np.zeros((4, n), np.int32)
    inpy = np.zeros((4, n), np.int32)
    inpa = np.zeros((4, n))
    cnt = np.zeros(n, np.int64)
    for px, py, pa in parts:                                  # compact the entries of each cell
        has = px > 0
        inpx[cnt[has], np.flatnonzero(has)] = px[has]
        inpy[cnt[has], np.flatnonzero(has)] = py[has]
        inpa[cnt[has], np.flatnonzero(has)] = pa[has]
        cnt += has
    inpn = int(cnt.max())
    return inpx[:inpn].copy(), inpy[:inpn].copy(), inpa[:inpn].copy(), nxin, nyin


def make_runoff_grid(nxin: int, nyin: int, nrec: int, seed: int, mean_mm_day: float = 1.5,
                     missing: float = 0.02) -> np.ndarray:
    """Runoff fields [nrec, nyin, nxin] in mm/day (gamma(0.5) marginals of a smooth
    field) with a fraction `missing` of the cells set to RMIS."""
    from scipy.special import gammaincinv, ndtr
    rng = np.random.default_rng(seed + 505)
    out = np.empty((nrec, nyin, nxin))
    for r in range(nrec):
        g = _fractal(nxin, nyin, 3.0, rng)
        out[r] = gammaincinv(0.5, np.clip(ndtr(g), 1e-12, 1 - 1e-12)) * (mean_mm_day / 0.5)
        out[r][rng.random((nyin, nxin)) < missing] = RMIS
    return out


# ---------------------------------------------------------------------------
# Reference input formats

def write_reference_inputs(net: SynthNetwork, d: str, runoff: np.ndarray, *, dt: float,
                           ifrq_inp: int, ndays: int, lpthout=False,
                           llevee=False, ldamout=False, ladpstp=True, lfldout=True,
                           pmanriv=0.03, pmanfld=0.10, pcadp=0.7, pdstmth=10000.0,
                           restart: bool = False, outputs: Optional[str] = None,
                           lkine: bool = False, restart_in: Optional[str] = None,
                           lstoonly: bool = False, ldamh22: bool = False,
                           slpmix_mask: Optional[np.ndarray] = None,
                           inpmat: Optional[tuple] = None) -> None:
    """Write the map files, tables, runoff records and input_cmf.nam that the
    reference needs (formats of src/cmf_ctrl_maps_mod.F90 READ_MAP_BIN,
    READ_TOPO_BIN, READ_BIFPARAM; src/cmf_ctrl_damout_mod.F90:141-179;
    src/cmf_ctrl_levee_mod.F90:121-133). restart_in names an FP64 binary restart
    file in d to start from (LRESTART, CRESTSTO; LSTOONLY with lstoonly). Dams with a
    "year" write the year column and switch LDAMYBY on; ldamh22 selects the
    Hanazaki-2022 dam rule (LDAMH22). slpmix_mask ([ny, nx], 1 = kinematic wave)
    switches LSLPMIX on and is written to slpmix_mask.bin, which the test driver
    oracle/ref_driver.F90 applies to I2MASK (the reference reads it from its
    netCDF river map only). inpmat = (inpx, inpy, inpa, nxin, nyin, seqx, seqy),
    the input matrix of make_inpmat with the NSEQ cell positions, switches LINTERP
    on: inpmat.bin in the layout CMF_INPMAT_INIT_BIN reads (3*INPN records of NX*NY
    4-byte values: INPX, INPY, INPA as float32; src/cmf_ctrl_forcing_mod.F90:418-458),
    NXIN/NYIN/INPN in diminfo.txt, and runoff records on the input grid in mm/day
    (DROFUNIT 86400*1000)."""
    os.makedirs(d, exist_ok=True)
    nx, ny = net.nx, net.ny

    def wf(name, a):
        np.ascontiguousarray(a, dtype=np.float32).tofile(os.path.join(d, name))

    np.concatenate([net.nextx.ravel(), net.nexty.ravel()]).astype(np.int32).tofile(
        os.path.join(d, "nextxy.bin"))
    wf("ctmare.bin", net.ctmare)
    wf("elevtn.bin", net.elevtn)
    wf("nxtdst.bin", net.nxtdst)
    wf("rivlen.bin", net.rivlen)
    wf("rivwth.bin", net.rivwth)
    wf("rivhgt.bin", net.rivhgt)
    wf("rivman.bin", net.rivman)
    wf("fldhgt.bin", net.fldhgt)
    nxin, nyin, inpn = nx, ny, 1
    if inpmat is not None:
        ix, iy, ia, nxin, nyin, sx, sy = inpmat
        inpn = int(np.asarray(ix).shape[0])
        cell = (np.asarray(sy, np.int64) - 1) * nx + (np.asarray(sx, np.int64) - 1)
        recs = []
        for tab, dt_ in ((ix, np.int32), (iy, np.int32), (ia, np.float32)):
            for k in range(inpn):
                mp_ = np.zeros(nx * ny, dt_)
                mp_[cell] = np.asarray(tab[k]).astype(dt_)
                recs.append(mp_.tobytes())
        with open(os.path.join(d, "inpmat.bin"), "wb") as f:
            f.write(b"".join(recs))
    with open(os.path.join(d, "diminfo.txt"), "w") as f:
        f.write(f"{nx}\n{ny}\n{NLFP}\n{nxin}\n{nyin}\n{inpn}\nNONE\n{net.west}\n{net.east}\n"
                f"{net.north}\n{net.south}\n")
    if lpthout:
        with open(os.path.join(d, "bifprm.txt"), "w") as f:
            f.write(f"{len(net.bifprm)} {net.npthlev}\n")
            for (ix, iy, jx, jy, dst, elv, dph, wth) in net.bifprm:
                f.write(f"{ix} {iy} {jx} {jy} {dst!r} {elv!r} {dph!r} " +
                        " ".join(repr(w) for w in wth) + "\n")
    if llevee:
        wf("levhgt.bin", net.levhgt)
        wf("levfrc.bin", net.levfrc)
    ldamyby = bool(ldamout and net.dams and "year" in net.dams[0])
    if ldamout:
        with open(os.path.join(d, "dam_params.csv"), "w") as f:
            f.write(f"{len(net.dams)}\n")
            f.write("ID,Name,Lat,Lon,Area,IX,IY,FldVol_mcm,ConVol_mcm,TotVol_mcm,Qn,Qf" +
                    (",year\n" if ldamyby else "\n"))
            for r in net.dams:
                f.write(f"{r['id']},{r['name']},{r['lat']!r},{r['lon']!r},{r['upreal']!r},"
                        f"{r['ix']},{r['iy']},{r['fldvol']!r},{r['convol']!r},{r['totvol']!r},"
                        f"{r['qn']!r},{r['qf']!r}" + (f",{r['year']}\n" if ldamyby else "\n"))
    np.ascontiguousarray(runoff, dtype=np.float64).tofile(os.path.join(d, "runoff.bin"))
    if slpmix_mask is not None:
        np.ascontiguousarray(slpmix_mask, dtype=np.int32).tofile(os.path.join(d, "slpmix_mask.bin"))

    import datetime as _dt
    t0 = _dt.datetime(2001, 1, 1)
    t1 = t0 + _dt.timedelta(days=ndays)
    L = lambda b: ".TRUE." if b else ".FALSE."
    # restart at the end of the run (IFRQ_RST = 0), binary FP64, restart<YYYYMMDDHH>.bin
    restart_nml = ('IFRQ_RST = 0\nCRESTDIR = "./"\nCVNREST = "restart"\nLRESTCDF = .FALSE.\n'
                   'LRESTDBL = .TRUE.') if restart else "IFRQ_RST = -1"
    if restart_in is not None:
        restart_nml += f'\nCRESTSTO = "{restart_in}"' + ('' if restart else '\nLRESTCDF = .FALSE.\nLRESTDBL = .TRUE.')
    nml = f"""&NRUNVER
LADPSTP  = {L(ladpstp)}
LFLDOUT  = {L(lfldout)}
LPTHOUT  = {L(lpthout)}
LDAMOUT  = {L(ldamout)}
LLEVEE   = {L(llevee)}
LKINE    = {L(lkine)}
LSLPMIX  = {L(slpmix_mask is not None)}
LRESTART = {L(restart_in is not None)}
LSTOONLY = {L(lstoonly)}
LOUTPUT  = {L(outputs is not None)}
LSPAMAT  = .TRUE.
/
&NDIMTIME
CDIMINFO = "diminfo.txt"
DT       = {dt!r}
IFRQ_INP = {ifrq_inp}
/
&NPARAM
PMANRIV  = {pmanriv!r}
PMANFLD  = {pmanfld!r}
PDSTMTH  = {pdstmth!r}
PCADP    = {pcadp!r}
/
&NSIMTIME
SYEAR = {t0.year}
SMON  = {t0.month}
SDAY  = {t0.day}
SHOUR = 0
EYEAR = {t1.year}
EMON  = {t1.month}
EDAY  = {t1.day}
EHOUR = 0
/
&NMAP
LMAPCDF = .FALSE.
CNEXTXY = "nextxy.bin"
CGRAREA = "ctmare.bin"
CELEVTN = "elevtn.bin"
CNXTDST = "nxtdst.bin"
CRIVLEN = "rivlen.bin"
CFLDHGT = "fldhgt.bin"
CRIVWTH = "rivwth.bin"
CRIVHGT = "rivhgt.bin"
CRIVMAN = "rivman.bin"
CPTHOUT = "bifprm.txt"
/
&NFORCE
LINPCDF = .FALSE.
LINTERP = {L(inpmat is not None)}
CINPMAT = "inpmat.bin"
/
&NRESTART
{restart_nml}
/
&NOUTPUT
COUTDIR  = "./"
CVARSOUT = "{outputs or 'outflw'}"
COUTTAG  = "_cmf"
LOUTCDF  = .FALSE.
LOUTVEC  = .FALSE.
IFRQ_OUT = 24
/
&NDAMOUT
CDAMFILE = "dam_params.csv"
LDAMTXT  = .FALSE.
LDAMH22  = {L(ldamh22)}
LDAMYBY  = {L(ldamyby)}
/
&NLEVEE
CLEVHGT = "levhgt.bin"
CLEVFRC = "levfrc.bin"
/
"""
    with open(os.path.join(d, "input_cmf.nam"), "w") as f:
        f.write(nml)
