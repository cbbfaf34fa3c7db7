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
NAMES = ["g_plain", "g_pth", "g_levwet", "g_dam", "g_allwet", "g_hourly", "c1_1deg", "g_kine"]


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
    opts = maps.Options(lpthout=lpth, llevee=llev, ldamout=ldam, lkine=bool(ref.get("lkine", False)))
    m = maps.build_map(net, opts)
    st = maps.cold_start(m)
    om = OracleModel(m, st)
    om.physics_fldstg()
    st = {k: v.copy() for k, v in om.st.items()}
    if ldam:
        maps.init_dams(m, net, st)
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
