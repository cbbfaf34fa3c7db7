"""1-D (NSEQ-order) map tables built from 2-D maps: the host initialisation that
precedes the device physics.

Each step restates the reference's initialisation (cited per function) so that a
host without the Fortran layer builds the same tables the Fortran host would pass
through the C-ABI. All arrays are 0-based in NSEQ order; floodplain and path-level
tables are planar ([layer, cell], [level, path]).
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

from . import _native
from ._abi import ptr

F1E20 = float(np.float32(1e20))     # default-kind 1.E20 widened (maps:580, damout:302)


@dataclasses.dataclass
class Options:
    """Namelist switches and parameters the physics reads (src/cmf_ctrl_nmlist_mod.F90
    defaults :77-115, :227-232)."""
    ladpstp: bool = True
    lfplain: bool = True
    lfldout: bool = True
    lpthout: bool = False
    ldamout: bool = False
    llevee: bool = False
    lkine: bool = False           # kinematic-wave outflow (CMF_CALC_OUTFLW_KINE)
    ldamh22: bool = False         # Hanazaki-2022 dam rule (NDAMOUT LDAMH22)
    lslpmix: bool = False         # slope-mixed outflow (CMF_CALC_OUTFLW_KINEMIX; I2MASK = 1 cells kinematic)
    pmanriv: float = 0.03
    pmanfld: float = 0.10
    pgrv: float = 9.8
    pdstmth: float = 10000.0
    pcadp: float = 0.7
    pminslp: float = float(np.float32(1e-5))


@dataclasses.dataclass
class CMFMap:
    nx: int
    ny: int
    n: int
    nseqriv: int
    nlfp: int
    opts: Options
    seqx: np.ndarray
    seqy: np.ndarray
    next: np.ndarray
    up_ptr: np.ndarray
    up_idx: np.ndarray
    grarea: np.ndarray
    elevtn: np.ndarray
    nxtdst: np.ndarray
    rivlen: np.ndarray
    rivwth: np.ndarray
    rivhgt: np.ndarray
    rivman: np.ndarray
    rivstomax: np.ndarray
    rivelv: np.ndarray
    dwnelv: np.ndarray
    fldhgt: np.ndarray
    fldstomax: np.ndarray
    fldgrd: np.ndarray
    dfrcinc: float
    i2mask: np.ndarray
    # bifurcation
    npth: int = 0
    npthlev: int = 0
    pth_upst: Optional[np.ndarray] = None
    pth_down: Optional[np.ndarray] = None
    pth_dst: Optional[np.ndarray] = None
    pth_elv: Optional[np.ndarray] = None
    pth_wth: Optional[np.ndarray] = None
    pth_man: Optional[np.ndarray] = None
    pout_ptr: Optional[np.ndarray] = None
    pout_idx: Optional[np.ndarray] = None
    pinf_ptr: Optional[np.ndarray] = None
    pinf_idx: Optional[np.ndarray] = None
    # levee
    levhgt: Optional[np.ndarray] = None
    levfrc: Optional[np.ndarray] = None
    bashgt: Optional[np.ndarray] = None
    levdst: Optional[np.ndarray] = None
    levbassto: Optional[np.ndarray] = None
    levtopsto: Optional[np.ndarray] = None
    levfilsto: Optional[np.ndarray] = None
    # dams
    ndam: int = 0
    i1dam: Optional[np.ndarray] = None
    damseq: Optional[np.ndarray] = None
    damstat: Optional[np.ndarray] = None
    dam_convol: Optional[np.ndarray] = None
    dam_emevol: Optional[np.ndarray] = None
    dam_adjvol: Optional[np.ndarray] = None
    dam_qa: Optional[np.ndarray] = None
    dam_qn: Optional[np.ndarray] = None
    dam_qf: Optional[np.ndarray] = None
    dam_norvol: Optional[np.ndarray] = None     # Hanazaki-2022 rule: NorVol
    dam_r_volupa: Optional[np.ndarray] = None   # and FldVol / upstream area

    def vec(self, map2d: np.ndarray) -> np.ndarray:
        """mapR2vecD (src/cmf_utils_mod.F90:282-299): float32 map -> float64 vector."""
        return np.ascontiguousarray(map2d[self.seqy - 1, self.seqx - 1], dtype=np.float64)


def _csr(n, owners, items):
    """CSR lists per owner cell, items in the given (ascending) order."""
    ptr_ = np.zeros(n + 1, np.int32)
    np.add.at(ptr_, np.asarray(owners, np.int64) + 1, 1)
    ptr_ = np.cumsum(ptr_).astype(np.int32)
    idx = np.zeros(len(items), np.int32)
    fill = ptr_[:-1].copy()
    for o, it in zip(owners, items):
        idx[fill[o]] = it
        fill[o] += 1
    return ptr_, idx


def build_map(net, opts: Options) -> CMFMap:
    """Initialisation sequence of CMF_RIVMAP_INIT, CMF_TOPO_INIT and CMF_LEVEE_INIT
    (src/cmf_drv_control_mod.F90:171-180) for a :class:`cmf.synth.SynthNetwork`.
    Dam tables are added by :func:`init_dams` after the cold start, in the
    reference's order (drv_control:252)."""
    lib = _native.host()
    nx, ny = net.nx, net.ny
    ng = nx * ny
    nextx = np.ascontiguousarray(net.nextx, np.int32)
    nexty = np.ascontiguousarray(net.nexty, np.int32)
    seqx = np.zeros(ng, np.int32)
    seqy = np.zeros(ng, np.int32)
    nxt = np.zeros(ng, np.int32)
    nriv = np.zeros(1, np.int32)
    n = lib.cmfh_calc_1d_seq(nx, ny, ptr(nextx), ptr(nexty), ptr(seqx), ptr(seqy), ptr(nxt), ptr(nriv))
    if n <= 0:
        raise RuntimeError("calc_1d_seq failed")
    seqx, seqy, nxt = seqx[:n].copy(), seqy[:n].copy(), nxt[:n].copy()
    nseqriv = int(nriv[0])
    up_ptr = np.zeros(n + 1, np.int32)
    up_idx = np.zeros(max(nseqriv, 1), np.int32)
    if lib.cmfh_build_upstream_csr(n, nseqriv, ptr(nxt), ptr(up_ptr), ptr(up_idx)) != 0:
        raise RuntimeError("upstream CSR failed")

    def v(a):
        return np.ascontiguousarray(a[seqy - 1, seqx - 1], dtype=np.float64)

    nlfp = net.fldhgt.shape[0]
    grarea, elevtn, nxtdst = v(net.ctmare), v(net.elevtn), v(net.nxtdst)
    rivlen, rivwth, rivhgt, rivman = v(net.rivlen), v(net.rivwth), v(net.rivhgt), v(net.rivman)
    fldhgt = np.ascontiguousarray(np.stack([v(net.fldhgt[l]) for l in range(nlfp)]))
    rivstomax = np.zeros(n)
    rivelv = np.zeros(n)
    fldstomax = np.zeros((nlfp, n))
    fldgrd = np.zeros((nlfp, n))
    dfr = np.zeros(1)
    lib.cmfh_set_fldstg(n, nlfp, int(opts.lfplain), ptr(grarea), ptr(rivlen), ptr(rivwth),
                        ptr(rivhgt), ptr(elevtn), ptr(fldhgt), ptr(rivstomax), ptr(rivelv),
                        ptr(fldstomax), ptr(fldgrd), ptr(dfr))
    m = CMFMap(nx, ny, n, nseqriv, nlfp, opts, seqx, seqy, nxt, up_ptr, up_idx,
               grarea, elevtn, nxtdst, rivlen, rivwth, rivhgt, rivman, rivstomax, rivelv,
               elevtn.copy(), fldhgt, fldstomax, fldgrd, float(dfr[0]), np.zeros(n, np.int32))
    if opts.lpthout:
        _read_bifparam(m, net)
    if opts.llevee:
        _levee_init(m, net)
    return m


def _read_bifparam(m: CMFMap, net) -> None:
    """READ_BIFPARAM, src/cmf_ctrl_maps_mod.F90:531-639."""
    vecidx = np.zeros((m.ny, m.nx), np.int32)          # I2VECTOR (1-based, 0 = none)
    vecidx[m.seqy - 1, m.seqx - 1] = np.arange(1, m.n + 1)
    rows = net.bifprm
    npth, nlev = len(rows), net.npthlev
    up = np.zeros(npth, np.int32)
    dn = np.zeros(npth, np.int32)
    dst = np.zeros(npth)
    elv = np.zeros((nlev, npth))
    wth = np.zeros((nlev, npth))
    for p, (ix, iy, jx, jy, d, pelv, pdph, w) in enumerate(rows):
        up[p] = vecidx[iy - 1, ix - 1] - 1
        dn[p] = vecidx[jy - 1, jx - 1] - 1
        dst[p] = d
        for l in range(nlev):
            wth[l, p] = w[l]
            if l == 0:
                elv[l, p] = (pelv - pdph) if w[l] > 0 else F1E20
            else:
                elv[l, p] = (pelv + float(l + 1)) - 2.0 if w[l] > 0 else F1E20
    man = np.array([m.opts.pmanriv if l == 0 else m.opts.pmanfld for l in range(nlev)])
    ok = (up >= 0) & (dn >= 0)
    pids = np.flatnonzero(ok)
    m.pout_ptr, m.pout_idx = _csr(m.n, up[pids], pids)
    m.pinf_ptr, m.pinf_idx = _csr(m.n, dn[pids], pids)
    m.npth, m.npthlev = npth, nlev
    m.pth_upst, m.pth_down, m.pth_dst = up, dn, dst
    m.pth_elv, m.pth_wth, m.pth_man = np.ascontiguousarray(elv), np.ascontiguousarray(wth), man


def _levee_init(m: CMFMap, net) -> None:
    """CMF_LEVEE_INIT, src/cmf_ctrl_levee_mod.F90:93-264."""
    lib = _native.host()
    m.levhgt = m.vec(net.levhgt)
    m.levfrc = m.vec(net.levfrc)
    n = m.n
    m.bashgt, m.levdst = np.zeros(n), np.zeros(n)
    m.levbassto, m.levtopsto, m.levfilsto = np.zeros(n), np.zeros(n), np.zeros(n)
    rc = lib.cmfh_levee_init(n, m.nlfp, ptr(m.grarea), ptr(m.rivlen), ptr(m.rivwth),
                             ptr(m.rivstomax), ptr(m.fldhgt), ptr(m.fldstomax), ptr(m.fldgrd),
                             ptr(m.levhgt), ptr(m.levfrc), ptr(m.bashgt), ptr(m.levdst),
                             ptr(m.levbassto), ptr(m.levtopsto), ptr(m.levfilsto))
    if rc != 0:
        raise RuntimeError("levee init failed")


def cold_start(m: CMFMap) -> dict:
    """CMF_PROG_INIT + STORAGE_SEA_SURFACE (src/cmf_ctrl_vars_mod.F90:28-168):
    prognostic state dict with zero flows and channels filled to the downstream
    water surface."""
    lib = _native.host()
    n = m.n
    st = {k: np.zeros(n) for k in ("rivsto", "fldsto", "rivout", "fldout", "rivout_pre",
                                   "fldout_pre", "rivdph_pre", "fldsto_pre", "runoff")}
    lib.cmfh_storage_sea_surface(n, m.nseqriv, ptr(m.next), ptr(m.dwnelv), ptr(m.rivelv),
                                 ptr(m.rivhgt), ptr(m.rivlen), ptr(m.rivwth), ptr(m.rivstomax),
                                 ptr(st["rivsto"]), ptr(st["rivdph_pre"]))
    if m.opts.llevee:
        st["levsto"] = np.zeros(n)
    if m.opts.ldamout:
        st["damsto"] = np.zeros(n)
        st["daminf"] = np.zeros(n)
    if m.opts.lpthout:
        st["pthflw"] = np.zeros((m.npthlev, m.npth))
        st["pthflw_pre"] = np.zeros((m.npthlev, m.npth))
    return st


def apply_slope_mask(m: CMFMap, mask: np.ndarray) -> None:
    """LSLPMIX: mark the cells of the slope mask ([ny, nx], 1 = kinematic wave;
    SET_SLOPEMIX, src/cmf_ctrl_maps_mod.F90:1005-1040) in I2MASK. Cells already
    coded by the dam set-up (2 at dams, 1 upstream) keep their codes, as when the
    mask is read first and CMF_DAMOUT_INIT writes over it."""
    v = np.asarray(mask)[m.seqy - 1, m.seqx - 1]
    m.i2mask[(m.i2mask == 0) & (v == 1)] = 1


def init_dams(m: CMFMap, net, st: dict, restart: bool = False, syear: int = 2001) -> None:
    """CMF_DAMOUT_INIT (src/cmf_ctrl_damout_mod.F90:123-308): the Yamazaki & Funato
    rule, or with m.opts.ldamh22 the Hanazaki-2022 rule (NorVol and the ratio of
    flood-control volume to upstream area, damout:202-205, stored per dam: the
    reference writes every dam's ratio to R_VolUpa(NDAM)); year-by-year activation
    (LDAMYBY) when the dam rows carry a "year", against the start year syear
    (ISYYYY, damout:216-225). Must run after the initial flood-stage call, as in
    src/cmf_drv_control_mod.F90:243-252; updates st in place. With restart (LRESTART)
    the dam storage read from the restart file is kept, except that a dam new in the
    start year (LDAMYBY) starts from its cell's river and floodplain storage
    (damout:256-285, LiVnorm off)."""
    f32 = lambda x: float(np.float32(x))
    vecidx = np.zeros((m.ny, m.nx), np.int32)
    vecidx[m.seqy - 1, m.seqx - 1] = np.arange(1, m.n + 1)
    rows = net.dams
    nd = len(rows)
    i1dam = np.zeros(m.n, np.int32)
    damseq = np.full(nd, -9999, np.int32)
    damstat = np.full(nd, -9999, np.int32)
    convol, emevol, adjvol = np.zeros(nd), np.zeros(nd), np.zeros(nd)
    qn, qf, qa = np.zeros(nd), np.zeros(nd), np.zeros(nd)
    norvol, r_volupa = np.zeros(nd), np.zeros(nd)
    h22 = bool(m.opts.ldamh22)
    yby = bool(rows) and "year" in rows[0]
    for d, r in enumerate(rows):
        fldvol = r["fldvol"] * 1.0e6
        convol[d] = r["convol"] * 1.0e6
        emevol[d] = convol[d] + fldvol * 0.95
        qn[d], qf[d] = r["qn"], r["qf"]
        ix, iy = r["ix"], r["iy"]
        if ix <= 0 or ix > m.nx or iy <= 0 or iy > m.ny:
            continue
        iseq = int(vecidx[iy - 1, ix - 1])
        if iseq <= 0 or m.next[iseq - 1] == -9999:
            continue
        s = iseq - 1
        damseq[d], damstat[d] = s, 2
        i1dam[s] = 1
        m.i2mask[s] = 2
        if h22:                                             # damout:202-205
            norvol[d] = convol[d] * 0.5
            r_volupa[d] = fldvol * f32(1.e-6) / r["upreal"]
        else:                                               # damout:207-212
            vyr = qn[d] * f32(365. * 24. * 60. * 60.)
            qsto = (convol[d] * f32(0.7) + vyr / 4.0) / f32(180. * 24. * 60. * 60.)
            qn[d] = min(qn[d], qsto) * 1.5
            adjvol[d] = convol[d] + fldvol * f32(0.1)
            qa[d] = (qn[d] + qf[d]) * 0.5
        if yby:                                             # damout:216-225
            if syear == r["year"]:
                damstat[d] = 1
            elif syear < r["year"] and r["year"] > 0:
                damstat[d] = -1
                i1dam[s] = -1
                convol[d] = 0.0
    if not (damstat == 2).any():
        raise RuntimeError("no dam allocated in the domain")
    for i in range(m.n):                                   # damout:238-254, ascending
        j = int(m.next[i])
        if i1dam[i] <= 0 and j >= 0:
            if i1dam[j] in (1, 11):
                i1dam[i] = 10
                m.i2mask[i] = 1
        if i1dam[i] == 1 and j >= 0:
            if i1dam[j] in (1, 11):
                i1dam[i] = 11
                m.i2mask[i] = 2
    if not restart:
        st["damsto"] = np.zeros(m.n)
    if restart and yby:                                    # damout:274-285 (LiVnorm = .FALSE.)
        for d in range(nd):
            if damstat[d] == 1:
                s = damseq[d]
                st["damsto"][s] = st["rivsto"][s] + st["fldsto"][s]
    for d in range(nd):                                    # damout:257-272
        if restart or damstat[d] == -9999:
            continue
        s = damseq[d]
        st["damsto"][s] = st["rivsto"][s] + st["fldsto"][s]
        if damstat[d] != -1 and st["damsto"][s] < convol[d]:
            st["damsto"][s] = convol[d]
            st["rivsto"][s] = convol[d]
            st["fldsto"][s] = 0.0
    st["daminf"] = np.zeros(m.n)
    if m.opts.lpthout:                                      # damout:295-306
        for p in range(m.npth):
            a, b = m.pth_upst[p], m.pth_down[p]
            if a < 0 or b < 0:
                continue
            if i1dam[a] > 0 or i1dam[b] > 0:
                m.pth_elv[:, p] = F1E20
    m.ndam = nd
    m.i1dam, m.damseq, m.damstat = i1dam, damseq, damstat
    m.dam_convol, m.dam_emevol, m.dam_adjvol = convol, emevol, adjvol
    m.dam_qn, m.dam_qf, m.dam_qa = qn, qf, qa
    m.dam_norvol, m.dam_r_volupa = norvol, r_volupa


_CELL = ("grarea", "elevtn", "nxtdst", "rivlen", "rivwth", "rivhgt", "rivman", "rivstomax", "rivelv",
         "dwnelv", "i2mask", "seqx", "seqy")
_LEVEE = ("levhgt", "levfrc", "bashgt", "levdst", "levbassto", "levtopsto", "levfilsto")
_DAM = ("damstat", "dam_convol", "dam_emevol", "dam_adjvol", "dam_qa", "dam_qn", "dam_qf", "dam_norvol",
        "dam_r_volupa")


def concat_maps(ms) -> tuple:
    """Concatenate the maps of independent regional domains into one map for a
    batched handle: domain d owns the next ms[d].n cells (its NSEQ order kept),
    downstream links, upstream lists, paths and dams are shifted into the
    concatenated numbering. Returns (map, dom_cells). All domains must share the
    options, NLFP and NPTHLEV."""
    ms = list(ms)
    o = ms[0].opts
    for m in ms:
        if m.opts != o or m.nlfp != ms[0].nlfp:
            raise ValueError("domains differ in options or NLFP")
    offs = np.cumsum([0] + [m.n for m in ms])
    n = int(offs[-1])
    cat = lambda k: np.concatenate([getattr(m, k) for m in ms])
    nxt = np.concatenate([np.where(m.next >= 0, m.next + off, m.next) for m, off in zip(ms, offs)]).astype(np.int32)
    owners = np.flatnonzero(nxt >= 0)
    up_ptr, up_idx = _csr(n, nxt[owners], owners)
    fl = lambda k: np.ascontiguousarray(np.concatenate([getattr(m, k) for m in ms], axis=1))
    out = CMFMap(0, 0, n, int(sum(m.nseqriv for m in ms)), ms[0].nlfp, o, cat("seqx"), cat("seqy"), nxt,
                 up_ptr, up_idx if len(up_idx) else np.zeros(1, np.int32),
                 *[cat(k) for k in _CELL[:10]], fl("fldhgt"), fl("fldstomax"), fl("fldgrd"), ms[0].dfrcinc,
                 cat("i2mask").astype(np.int32))
    if o.lpthout:
        if len({m.npthlev for m in ms}) != 1:
            raise ValueError("domains differ in NPTHLEV")
        sh = lambda a, off: np.where(a >= 0, a + off, a).astype(np.int32)
        out.npth, out.npthlev = int(sum(m.npth for m in ms)), ms[0].npthlev
        out.pth_upst = np.concatenate([sh(m.pth_upst, off) for m, off in zip(ms, offs)])
        out.pth_down = np.concatenate([sh(m.pth_down, off) for m, off in zip(ms, offs)])
        out.pth_dst = cat("pth_dst")
        out.pth_elv, out.pth_wth = fl("pth_elv"), fl("pth_wth")
        out.pth_man = ms[0].pth_man.copy()
        ok = (out.pth_upst >= 0) & (out.pth_down >= 0)
        pids = np.flatnonzero(ok)
        out.pout_ptr, out.pout_idx = _csr(n, out.pth_upst[pids], pids)
        out.pinf_ptr, out.pinf_idx = _csr(n, out.pth_down[pids], pids)
    if o.llevee:
        for k in _LEVEE:
            setattr(out, k, cat(k))
    if o.ldamout:
        out.ndam = int(sum(m.ndam for m in ms))
        out.i1dam = cat("i1dam").astype(np.int32)
        out.damseq = np.concatenate([np.where(m.damseq >= 0, m.damseq + off, m.damseq)
                                     for m, off in zip(ms, offs)]).astype(np.int32)
        for k in _DAM:
            setattr(out, k, cat(k))
    return out, [m.n for m in ms]


def basin_of(m: CMFMap) -> np.ndarray:
    """Index of the mouth each cell drains to (its river basin)."""
    b = np.full(m.n, -1, np.int64)
    mouths = np.flatnonzero(m.next < 0)
    b[mouths] = mouths
    # NSEQ order puts every cell before its downstream cell: walk it backwards
    order = np.arange(m.n)[::-1]
    for i in order:
        if b[i] < 0:
            b[i] = b[m.next[i]]
    return b


def subset_map(m: CMFMap, cells: np.ndarray) -> CMFMap:
    """The map restricted to `cells` (whole basins, ascending NSEQ order kept):
    downstream links renumbered, upstream lists rebuilt. Options without paths,
    levees or dams (the spin-up basins use the default physics)."""
    cells = np.sort(np.asarray(cells, np.int64))
    new = np.full(m.n, -1, np.int64)
    new[cells] = np.arange(cells.size)
    nxt = m.next[cells]
    nn = np.where(nxt >= 0, new[np.where(nxt >= 0, nxt, 0)], nxt).astype(np.int32)
    if np.any((nxt >= 0) & (nn < 0)):
        raise ValueError("subset_map: cells must be whole basins")
    n = cells.size
    owners = np.flatnonzero(nn >= 0)
    up_ptr, up_idx = _csr(n, nn[owners], owners)
    s = lambda a: np.ascontiguousarray(a[cells])
    sp = lambda a: np.ascontiguousarray(a[:, cells])
    return CMFMap(m.nx, m.ny, n, int((nn >= 0).sum()), m.nlfp, m.opts, s(m.seqx), s(m.seqy), nn, up_ptr,
                  up_idx if len(up_idx) else np.zeros(1, np.int32), s(m.grarea), s(m.elevtn), s(m.nxtdst),
                  s(m.rivlen), s(m.rivwth), s(m.rivhgt), s(m.rivman), s(m.rivstomax), s(m.rivelv), s(m.dwnelv),
                  sp(m.fldhgt), sp(m.fldstomax), sp(m.fldgrd), m.dfrcinc, s(m.i2mask))
