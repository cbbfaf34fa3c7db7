"""Loader of the golden fixtures in tests/golden (written by
tests/golden/make_golden.py from the reference Fortran) and a runner that replays
a fixture's run with the oracle or the device model."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

from cmf import maps

HERE = os.path.dirname(os.path.abspath(__file__))
GDIR = os.path.join(HERE, "golden")
NAMES = ["g_plain", "g_pth", "g_levwet", "g_dam", "g_allwet", "g_hourly", "c1_1deg", "g_kine", "g_dam_yby",
         "g_dam_h22", "g_slpmix"]


def sha(a):
    a = np.ascontiguousarray(a)
    if a.dtype.kind == "f":
        a = a.astype("<f8")
    elif a.dtype.kind in "iu":
        a = a.astype("<i4")
    return hashlib.sha1(a.tobytes()).hexdigest()


def load(name):
    import sys
    sys.path.insert(0, GDIR)
    from make_golden import load_network
    net, ro = load_network(os.path.join(GDIR, f"{name}_inputs.npz"))
    with open(os.path.join(GDIR, f"{name}.json")) as f:
        ref = json.load(f)
    fin = dict(np.load(os.path.join(GDIR, f"{name}_final.npz")))
    return net, ro, ref, fin


def setup(name):
    """Map, initial state (oracle initial stage) and runoff vectors of a fixture."""
    from oracle.oracle import OracleModel
    net, ro, ref, fin = load(name)
    nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac = ref["spec"]
    opts = maps.Options(lpthout=lpth, llevee=llev, ldamout=ldam, lkine=bool(ref.get("lkine", False)),
                        ldamh22=bool(ref.get("ldamh22", False)), lslpmix=bool(ref.get("lslpmix", False)))
    m = maps.build_map(net, opts)
    st = maps.cold_start(m)
    om = OracleModel(m, st)
    om.physics_fldstg()
    st = {k: v.copy() for k, v in om.st.items()}
    if ldam:
        maps.init_dams(m, net, st)
    if opts.lslpmix:
        maps.apply_slope_mask(m, net.slpmix_mask)
    rvec = [m.vec(ro[r]) for r in range(ro.shape[0])]
    nsteps = int(round(days * 86400.0 / dt))
    spr = max(1, int(round(max(dt, 3600.0 * int(dt // 3600)) / dt)))   # steps per forcing record
    return m, net, st, rvec, ref, fin, dt, nsteps


def runoff_for_step(rvec, step, dt):
    """Record read by oracle/ref_driver.F90 at outer step `step` (1-based): one record
    per forcing interval DTIN = IFRQ_INP h, cycled."""
    ifrq = int(dt // 3600)
    spr = int(round(ifrq * 3600.0 / dt))
    return rvec[((step - 1) // spr) % len(rvec)]


RESTART_IN_NAMES = ["g_allwet_sto", "g_allwet_rst"]


def restart_init(m, st, binb, pthb, lstoonly):
    """CMF_RESTART_INIT (src/cmf_ctrl_restart_mod.F90:95-133) on host arrays: zero
    the storages, flows and t-1 values, read the FP64 records of READ_REST_BIN
    (:140-213) placed by I1SEQX/I1SEQY, and with LSTOONLY set D2FLDSTO_PRE =
    P2FLDSTO and D1PTHFLW_PRE = 0. Updates st in place."""
    n = m.n
    nxy = m.nx * m.ny
    g = (m.seqy - 1) * m.nx + (m.seqx - 1)
    recs = np.frombuffer(binb, np.float64)
    fields = ["rivsto", "fldsto"]
    if not lstoonly:
        fields += ["rivout_pre", "fldout_pre", "rivdph_pre", "fldsto_pre"]
    if m.opts.ldamout:
        fields.append("damsto")
    if m.opts.llevee:
        fields.append("levsto")
    for k in ("rivsto", "fldsto", "rivout", "fldout", "rivout_pre", "fldout_pre", "rivdph_pre", "fldsto_pre",
              "damsto", "levsto"):
        if k in st:
            st[k] = np.zeros(n)
    for r, k in enumerate(fields):
        st[k] = recs[r * nxy:(r + 1) * nxy][g].copy()
    if m.opts.lpthout:
        st["pthflw"] = np.zeros((m.npthlev, m.npth))
        if lstoonly:
            st["pthflw_pre"] = np.zeros((m.npthlev, m.npth))
        else:
            st["pthflw_pre"] = np.frombuffer(pthb, np.float64).reshape(m.npthlev, m.npth).copy()
    if lstoonly:
        st["fldsto_pre"] = st["fldsto"].copy()
    return st


def restart_setup(name):
    """Inputs of a run that starts from a restart file (tests/golden make_golden
    RESTART_IN): map with dam tables, the state before CMF_RESTART_INIT (cold
    start and its stage), the restart file bytes, runoff vectors and the
    reference's outputs."""
    import sys
    from oracle.oracle import OracleModel
    sys.path.insert(0, GDIR)
    from make_golden import load_network, restart_input
    with open(os.path.join(GDIR, f"{name}.json")) as f:
        ref = json.load(f)
    base, lstoonly = ref["base"], ref["lstoonly"]
    net, ro = load_network(os.path.join(GDIR, f"{base}_inputs.npz"))
    nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac = ref["spec"]
    m = maps.build_map(net, maps.Options(lpthout=lpth, llevee=llev, ldamout=ldam))
    om = OracleModel(m, maps.cold_start(m))
    om.physics_fldstg()                               # drv_control:240 (before the restart)
    st = {k: v.copy() for k, v in om.st.items()}
    if ldam:
        maps.init_dams(m, net, {"rivsto": st["rivsto"].copy(), "fldsto": st["fldsto"].copy()}, restart=True)
        st["daminf"] = np.zeros(m.n)
    binb, pthb = restart_input(base, lstoonly)
    step0 = dict(np.load(os.path.join(GDIR, f"{name}_step0.npz")))
    fin = dict(np.load(os.path.join(GDIR, f"{name}_final.npz")))
    rvec = [m.vec(ro[r]) for r in range(ro.shape[0])]
    nsteps = int(round(days * 86400.0 / dt))
    return m, net, st, binb, pthb, lstoonly, rvec, ref, step0, fin, dt, nsteps
