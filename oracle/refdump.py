"""Reader for the stream dumps written by oracle/ref_driver.F90 (the reference run).

TEST INFRASTRUCTURE ONLY.
"""
from __future__ import annotations

import numpy as np

MAP_D = "grarea elevtn nxtdst rivlen rivwth rivhgt rivman rivstomax rivelv dwnelv".split()
LEV_D = "levhgt levfrc bashgt levdst levbassto levtopsto levfilsto".split()
STATE_P = ["rivsto", "fldsto"]
STATE_D = ("rivout fldout rivout_pre fldout_pre rivdph_pre fldsto_pre rivdph flddph fldfrc "
           "fldare sfcelv storge outflw rivvel rivinf fldinf pthout rivout_avg fldout_avg "
           "outflw_avg rivvel_avg pthout_avg runoff_avg outflw_max rivdph_max storge_max").split()
GLB = "stopre stonxt stonew rivinf rivout stopre2 stonew2 rivsto fldsto fldare".split()


class _Cursor:
    def __init__(self, buf):
        self.b = buf
        self.o = 0

    def take(self, dtype, count):
        a = np.frombuffer(self.b, dtype=dtype, count=count, offset=self.o)
        self.o += a.nbytes
        return a.copy()

    @property
    def done(self):
        return self.o >= len(self.b)


def read_dump(path: str, lpthout=False, llevee=False, ldamout=False):
    """Returns (map dict, list of state dicts). Index arrays stay 1-based as written."""
    with open(path, "rb") as f:
        c = _Cursor(f.read())
    nx, ny, nlfp, nseqriv, n, npth, nlev = (int(v) for v in c.take(np.int32, 7))
    mp = dict(nx=nx, ny=ny, nlfp=nlfp, nseqriv=nseqriv, n=n, npth=npth, npthlev=nlev)
    mp["seqx"] = c.take(np.int32, n)
    mp["seqy"] = c.take(np.int32, n)
    mp["next"] = c.take(np.int32, n)
    mp["i2mask"] = c.take(np.int32, n)
    for k in MAP_D:
        mp[k] = c.take(np.float64, n)
    for k in ("fldhgt", "fldstomax", "fldgrd"):
        mp[k] = c.take(np.float64, nlfp * n).reshape(nlfp, n)
    mp["dfrcinc"] = float(c.take(np.float64, 1)[0])
    if lpthout:
        mp["pth_upst"] = c.take(np.int32, npth)
        mp["pth_down"] = c.take(np.int32, npth)
        mp["pth_dst"] = c.take(np.float64, npth)
        mp["pth_elv"] = c.take(np.float64, nlev * npth).reshape(nlev, npth)
        mp["pth_wth"] = c.take(np.float64, nlev * npth).reshape(nlev, npth)
        mp["pth_man"] = c.take(np.float64, nlev)
    if llevee:
        for k in LEV_D:
            mp[k] = c.take(np.float64, n)
    if ldamout:
        mp["i1dam"] = c.take(np.int32, n)
    states = []
    while not c.done:
        s = {"step": int(c.take(np.int32, 1)[0])}
        for k in STATE_P + STATE_D:
            s[k] = c.take(np.float64, n)
        s["nadd"] = float(c.take(np.float64, 1)[0])
        s["glb"] = dict(zip(GLB, c.take(np.float64, 10)))
        if lpthout:
            s["pthflw"] = c.take(np.float64, nlev * npth).reshape(nlev, npth)
            s["pthflw_pre"] = c.take(np.float64, nlev * npth).reshape(nlev, npth)
            s["pthflwsum"] = c.take(np.float64, npth)
            s["pthflw_avg"] = c.take(np.float64, nlev * npth).reshape(nlev, npth)
            s["pthflwsum_avg"] = c.take(np.float64, npth)
        if llevee:
            s["levsto"] = c.take(np.float64, n)
            s["levdph"] = c.take(np.float64, n)
        if ldamout:
            s["damsto"] = c.take(np.float64, n)
            s["daminf"] = c.take(np.float64, n)
            s["daminf_avg"] = c.take(np.float64, n)
        states.append(s)
    return mp, states


def count_substeps(log_path: str) -> list:
    """NT per outer step from the reference log: one water-budget line per substep
    (src/cmf_ctrl_physics_mod.F90:241-246), grouped by the per-step end line
    (src/cmf_drv_advance_mod.F90:165)."""
    nts, cur = [], 0
    with open(log_path, errors="replace") as f:
        for line in f:
            if " flx: " in line:
                cur += 1
            elif "DRV_ADVANCE END" in line:
                nts.append(cur)
                cur = 0
    return nts
