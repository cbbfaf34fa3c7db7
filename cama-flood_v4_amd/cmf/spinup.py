"""Spin-up to a periodic equilibrium, and the runoff climatology that drives it.

The reference spins a model up by repeating one forcing year from the cold start
of STORAGE_SEA_SURFACE (src/cmf_ctrl_vars_mod.F90:129-168) a fixed number of
times (gosh/test1-glb_15min.sh:311-345: NSP loops, no convergence test). Here the
year is repeated until the total storage of a domain changes by less than a
tolerance from one year end to the next (0.1 % by default), and the number of
simulated years needed is the figure of merit that the learned warm start
(:mod:`cmf.surrogate`) is measured by.

Climatology: twelve monthly runoff fields per cell (m3/s), a spatially correlated
gamma field scaled by a seasonal cycle and by random month-to-month factors;
daily runoff is the periodic linear interpolation between month centres.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import List, Optional, Sequence

import numpy as np

from . import _native, synth
from ._abi import ptr
from .maps import CMFMap

DAYS = 365
MONTH_MID = (np.arange(12) + 0.5) * DAYS / 12.0


@dataclasses.dataclass
class Climatology:
    monthly: np.ndarray          # [12, n] runoff per cell, m3/s
    mean_mm_day: float
    amp: float
    phase: float

    def daily(self, day: int) -> np.ndarray:
        """Runoff [m3/s] on day 0..364 of the climatological year."""
        t = (day % DAYS) + 0.5
        k = int(np.searchsorted(MONTH_MID, t)) - 1          # month centre at or before t
        k0, k1 = k % 12, (k + 1) % 12
        t0 = MONTH_MID[k0] if k >= 0 else MONTH_MID[11] - DAYS
        w = (t - t0) / (DAYS / 12.0)
        return (1.0 - w) * self.monthly[k0] + w * self.monthly[k1]


def make_climatology(net, m: CMFMap, seed: int, mean_mm_day: float, amp: float, phase: float,
                     month_sigma: float = 0.2) -> Climatology:
    """Monthly runoff fields of one domain: gamma(k=0.5) spatial field (the
    generator of the daily forcing) x (1 + amp sin(2 pi (month - phase)/12)) x
    lognormal month factors."""
    rng = np.random.default_rng(seed)
    g = synth.make_runoff(net, 1, seed, mean_mm_day=mean_mm_day)[0]       # [ny, nx] m3/s
    base = m.vec(g)
    mon = np.arange(12)
    cyc = 1.0 + amp * np.sin(2.0 * np.pi * (mon - phase) / 12.0)
    fac = np.exp(rng.normal(-0.5 * month_sigma ** 2, month_sigma, 12))
    monthly = np.maximum(cyc * fac, 0.0)[:, None] * base[None, :]
    return Climatology(np.ascontiguousarray(monthly), mean_mm_day, amp, phase)


def accumulate(m: CMFMap, vals: np.ndarray) -> np.ndarray:
    """Upstream accumulation of per-cell values ([n] or [nrec, n])."""
    lib = _native.host()
    v = np.ascontiguousarray(np.atleast_2d(vals), np.float64)
    out = np.empty_like(v)
    nxt = np.ascontiguousarray(np.where(m.next >= 0, m.next, -1), np.int32)
    rc = lib.cmfh_accumulate(m.n, ptr(nxt), v.shape[0], ptr(v), ptr(out))
    if rc != 0:
        raise RuntimeError(f"accumulate failed ({rc})")
    return out.reshape(np.shape(vals)) if np.ndim(vals) == 1 else out


def steady_state(m: CMFMap, q: np.ndarray) -> dict:
    """Steady state of the local-inertial outflow carrying the constant outflow
    q [m3/s] out of every cell (include/cmf_host.h cmfh_steady_state): rivsto,
    fldsto, sfcelv, rivout, fldout."""
    lib = _native.host()
    ch, keep = _channel(m)
    qq = np.ascontiguousarray(q, np.float64)
    out = {k: np.empty(m.n) for k in ("rivsto", "fldsto", "sfcelv", "rivout", "fldout")}
    rc = lib.cmfh_steady_state(C.byref(ch), ptr(qq), *(ptr(out[k]) for k in
                                                        ("rivsto", "fldsto", "sfcelv", "rivout", "fldout")))
    if rc != 0:
        raise RuntimeError(f"steady state failed ({rc})")
    return out


def travel_time(m: CMFMap, q: np.ndarray, tau: np.ndarray) -> np.ndarray:
    """Runoff-weighted mean travel time from the runoff sources upstream of each
    cell to the cell: T_i = sum_u r_u tau(u -> i) / sum_u r_u, where tau(u -> i)
    sums the residence times tau of the cells from u down to i (u included, i
    excluded) and r = the local runoff, q = its upstream accumulation. With
    B_i = T_i q_i, B_i = sum over upstream neighbours c of (B_c + q_c tau_c), so
    B = accumulate(q tau) - q tau."""
    x = q * tau
    b = accumulate(m, x) - x
    return b / np.maximum(q, 1e-30)


def _channel(m: CMFMap):
    """cmfh_channel of a map plus the arrays it points to (kept alive by the caller)."""
    from ._abi import cmfh_channel
    f64 = lambda a: np.ascontiguousarray(a, np.float64)
    keep = {k: f64(getattr(m, k)) for k in ("elevtn", "rivelv", "rivhgt", "rivwth", "rivlen", "rivman",
                                              "nxtdst", "grarea", "dwnelv", "rivstomax", "fldstomax", "fldgrd")}
    keep["next"] = np.ascontiguousarray(np.where(m.next >= 0, m.next, -1), np.int32)
    ch = cmfh_channel(n=m.n, nlfp=m.nlfp, lfldout=int(m.opts.lfldout), dfrcinc=m.dfrcinc,
                      pmanfld=m.opts.pmanfld, pdstmth=m.opts.pdstmth, **{k: ptr(v) for k, v in keep.items()})
    return ch, keep


def runoff_harmonics(clim: Climatology, nh: int) -> np.ndarray:
    """Complex amplitudes r_k (k = 1..nh) of the local daily runoff over the
    climatological year, r(t) = r_0 + 2 Re sum_k r_k exp(i w_k t), w_k = 2 pi k / 365 d."""
    d = np.arange(DAYS)
    R = np.stack([clim.daily(k) for k in d])                               # [365, n]
    w = 2.0 * np.pi * np.arange(1, nh + 1) / DAYS
    return (np.exp(-1j * np.outer(w, d + 0.5)) / DAYS) @ R


def periodic_state(m: CMFMap, clim: Climatology, sfc_mean: np.ndarray, t_day: float = 0.0, nh: int = 6) -> dict:
    """State at day t_day of the climatological year from the network's linear
    periodic response around the steady surface sfc_mean (include/cmf_host.h
    cmfh_periodic_response): surface elevation, river / floodplain storage, and
    the outflow's departure from the steady flow (dq)."""
    lib = _native.host()
    ch, keep = _channel(m)
    rh = runoff_harmonics(clim, nh)
    omega = np.ascontiguousarray(2.0 * np.pi * np.arange(1, nh + 1) / (DAYS * 86400.0))
    rr, ri = np.ascontiguousarray(rh.real), np.ascontiguousarray(rh.imag)
    hr, hi = np.empty_like(rr), np.empty_like(ri)
    sm = np.ascontiguousarray(sfc_mean, np.float64)
    orr, oi = np.empty_like(rr), np.empty_like(ri)
    rc = lib.cmfh_periodic_response(C.byref(ch), ptr(sm), nh, ptr(omega), ptr(rr), ptr(ri), ptr(hr), ptr(hi),
                                    ptr(orr), ptr(oi))
    if rc != 0:
        raise RuntimeError(f"periodic response failed ({rc})")
    w = 2.0 * np.pi * np.arange(1, nh + 1) / DAYS
    ph = np.exp(1j * w * t_day)[:, None]
    eta = sm + 2.0 * np.sum(((hr + 1j * hi) * ph).real, axis=0)
    eta = np.ascontiguousarray(np.maximum(eta, m.rivelv))
    dq = 2.0 * np.sum(((orr + 1j * oi) * ph).real, axis=0)
    out = {"sfcelv": eta, "rivsto": np.empty(m.n), "fldsto": np.empty(m.n), "dq": dq}
    rc = lib.cmfh_storage_at_level(C.byref(ch), ptr(eta), ptr(out["rivsto"]), ptr(out["fldsto"]))
    if rc != 0:
        raise RuntimeError(f"storage at level failed ({rc})")
    return out


def domain_storage(st: dict, offs: Sequence[int]) -> np.ndarray:
    """Total storage (river + floodplain) per domain of a concatenated map."""
    tot = np.asarray(st["rivsto"]) + np.asarray(st["fldsto"])
    if tot.ndim == 2:
        tot = tot[0]
    return np.array([tot[offs[d]:offs[d + 1]].sum() for d in range(len(offs) - 1)])


@dataclasses.dataclass
class SpinupResult:
    years: np.ndarray            # per domain: first year whose annual change is below tol (0 = not reached)
    storage: np.ndarray          # [years + 1, ndom] storage at every year end (row 0: start)
    state: dict                  # final rivsto, fldsto, rivdph, rivout, fldout, outflw_avg (last outer step)
    seconds: float
    daily: Optional[np.ndarray] = None   # single-domain models: total storage at the end of every day
    substeps: int = 0


def spinup(model, clims: List[Climatology], offs: Sequence[int], tol: float = 1e-3, max_years: int = 30,
           stop_tol: Optional[float] = None, dt: float = 86400.0, log=None) -> SpinupResult:
    """Repeat the climatological year on a device model (single domain or a
    concatenated-domain handle with one member) until every domain's annual
    storage change is below stop_tol (default tol) or max_years pass.

    Annual change of year y: |S_y - S_{y-1}| / S_y, S_y = total storage at the end
    of year y, S_0 = the initial storage. years[d] = first y with change < tol."""
    import time
    stop = tol if stop_tol is None else stop_tol
    nd = len(offs) - 1
    st = model.get_state(["rivsto", "fldsto"])
    hist = [domain_storage(st, offs)]
    years = np.zeros(nd, np.int64)
    t0 = time.time()
    steps_per_day = int(round(86400.0 / dt))
    daily = []
    nsub = 0
    for y in range(1, max_years + 1):
        for day in range(DAYS):
            r = np.concatenate([c.daily(day) for c in clims])
            model.set_runoff(r)
            for _ in range(steps_per_day):
                if model.ensemble:
                    nt, _ = model.ensemble_advance(dt)
                    nsub += int(nt.max())
                else:
                    # water-budget sums of the last substep: P0GLBRIVSTO + P0GLBFLDSTO
                    nt, glb = model.physics_advance(dt, budget=True)
                    nsub += nt
            if not model.ensemble:
                daily.append(glb[-1, 7] + glb[-1, 8])
        st = model.get_state(["rivsto", "fldsto", "rivdph", "rivout", "fldout", "outflw_avg"])
        s = domain_storage(st, offs)
        ch = np.abs(s - hist[-1]) / np.maximum(s, 1e-30)
        hist.append(s)
        newly = (years == 0) & (ch < tol)
        years[newly] = y
        if log:
            log(f"year {y}: max annual change {ch.max():.3e}, converged {int((years > 0).sum())}/{nd}, "
                f"{time.time() - t0:.1f}s")
        if (ch < stop).all() and (years > 0).all():
            break
    return SpinupResult(years, np.asarray(hist), st, time.time() - t0,
                        np.asarray(daily) if daily else None, nsub)


def days_to_equilibrium(daily: np.ndarray, eq_year: np.ndarray, tol: float = 1e-3) -> int:
    """First day after which the storage stays within tol (relative) of the
    periodic equilibrium trajectory eq_year[365] for the rest of the record
    (len(daily) + 1 if it never does)."""
    eq = np.resize(eq_year, daily.size)
    bad = np.flatnonzero(np.abs(daily - eq) > tol * np.abs(eq))
    return 0 if bad.size == 0 else int(bad[-1]) + 1
