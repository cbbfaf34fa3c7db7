"""Shared builders for the parity tests: seeded synthetic networks with the
feature switches of the reference (bifurcation, levee, dam), their host map,
the cold-start state and the initial stage, prepared identically for the oracle
and the device model."""
from __future__ import annotations

import numpy as np

from cmf import maps, synth

CASES = {
    #        nx,  ny, seed, lpthout, llevee, ldamout, runoff mean [mm/day], levee fraction
    "plain":  (90, 45, 1, False, False, False, 1.5, 0.0),
    "pth":    (90, 45, 2, True, False, False, 1.5, 0.0),
    "lev":    (90, 45, 3, False, True, False, 1.5, 0.2),
    "levpth": (90, 45, 4, True, True, False, 1.5, 0.2),
    "dam":    (90, 45, 5, False, False, True, 1.5, 0.0),
    "all":    (90, 45, 6, True, True, True, 1.5, 0.2),
    "levwet": (90, 45, 7, False, True, False, 60.0, 0.5),
    "allwet": (90, 45, 8, True, True, True, 60.0, 0.5),
    "damh22": (120, 60, 9, False, False, True, 4.0, 0.0),
    "slpmix": (120, 60, 10, True, True, True, 20.0, 0.3),
}
# cases run with the Hanazaki-2022 reservoir rule (LDAMH22) and one dam per 25 cells
H22_CASES = ("damh22",)
# cases run with LSLPMIX (CMF_CALC_OUTFLW_KINEMIX) and this fraction of the steepest
# river cells in the slope mask
SLPMIX_CASES = {"slpmix": 0.3}


def make_network(name):
    nx, ny, seed, lpth, llev, ldam, mean, levfrac = CASES[name]
    net = synth.make_network(nx, ny, seed=seed)
    if lpth:
        synth.add_bifurcations(net, max(10, net.ncells // 20), 5, seed)
    if llev:
        if mean > 10:
            synth.add_levees(net, levfrac, seed, hmin=0.5, hmax=3.0, fmin=0.05, fmax=0.5)
        else:
            synth.add_levees(net, levfrac, seed, hmin=0.5, hmax=3.0, fmin=0.05, fmax=0.5)
    if ldam:
        synth.add_dams(net, max(3, net.ncells // (25 if name in H22_CASES else 100)), seed)
    if name in SLPMIX_CASES:
        synth.add_slope_mask(net, SLPMIX_CASES[name])
    opts = maps.Options(lpthout=lpth, llevee=llev, ldamout=ldam, ldamh22=name in H22_CASES,
                        lslpmix=name in SLPMIX_CASES)
    return net, opts, mean


def runoff_records(net, nrec, seed, mean):
    return synth.make_runoff(net, nrec, seed, mean_mm_day=mean)


def initial_state(m, net, stage_fn):
    """Cold start, initial CMF_PHYSICS_FLDSTG (stage_fn(state) -> state) and dam
    initialisation, in the order of src/cmf_drv_control_mod.F90:239-252."""
    st = maps.cold_start(m)
    st = stage_fn(st)
    if m.opts.ldamout:
        maps.init_dams(m, net, st)
    if m.opts.lslpmix:
        maps.apply_slope_mask(m, net.slpmix_mask)
    return st


def prepare(name, ndays):
    """(map, network, initial state, runoff vectors per day) for a case, with the
    initial stage taken from the oracle (test infrastructure)."""
    from oracle.oracle import OracleModel
    net, opts, mean = make_network(name)
    m = maps.build_map(net, opts)
    seed = CASES[name][2]
    ro = runoff_records(net, ndays, seed, mean)
    st = maps.cold_start(m)
    om = OracleModel(m, st)
    om.physics_fldstg()
    st = {k: v.copy() for k, v in om.st.items()}
    if opts.ldamout:
        maps.init_dams(m, net, st)
    if opts.lslpmix:
        maps.apply_slope_mask(m, net.slpmix_mask)
    rvec = [m.vec(ro[d]) for d in range(ndays)]
    return m, net, st, rvec


STATE_KEYS = ["rivsto", "fldsto", "rivout", "fldout", "rivdph", "flddph", "fldfrc", "fldare",
              "sfcelv", "storge", "outflw", "rivvel", "rivinf", "fldinf", "pthout", "rivdph_pre",
              "fldsto_pre", "rivout_avg", "fldout_avg", "outflw_avg", "rivvel_avg", "pthout_avg",
              "runoff_avg", "outflw_max", "rivdph_max", "storge_max"]


def upload_keys(st):
    keys = ["rivsto", "fldsto", "rivout", "fldout", "rivout_pre", "fldout_pre", "rivdph_pre",
            "fldsto_pre", "rivdph", "flddph", "fldfrc", "fldare", "sfcelv", "storge", "levsto",
            "levdph", "damsto", "daminf", "pthflw", "pthflw_pre", "pthflwsum"]
    return {k: st[k] for k in keys if k in st}


def rel_err(a, b, floor=1e-30):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(float(np.max(np.abs(b))) if b.size else 0.0, floor)
    return float(np.max(np.abs(a - b))) / scale if a.size else 0.0


def prepare_domain(nx, ny, seed, lpth=True, llev=True, ldam=True, levfrac=0.2, pdstmth=10000.0):
    """(map, network, initial state) of one regional domain for the ensemble
    tests; every domain of one ensemble uses the same switches."""
    from oracle.oracle import OracleModel
    net = synth.make_network(nx, ny, seed=seed)
    if lpth:
        synth.add_bifurcations(net, max(10, net.ncells // 20), 5, seed)
    if llev:
        synth.add_levees(net, levfrac, seed, hmin=0.5, hmax=3.0, fmin=0.05, fmax=0.5)
    if ldam:
        synth.add_dams(net, max(3, net.ncells // 100), seed)
    opts = maps.Options(lpthout=lpth, llevee=llev, ldamout=ldam, pdstmth=pdstmth)
    m = maps.build_map(net, opts)
    st = maps.cold_start(m)
    om = OracleModel(m, st)
    om.physics_fldstg()
    st = {k: v.copy() for k, v in om.st.items()}
    if ldam:
        maps.init_dams(m, net, st)
    return m, net, st
