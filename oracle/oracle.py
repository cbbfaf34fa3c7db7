"""Python binding of the C oracle (cmf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker of the HIP physics.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_P = C.c_void_p
_FIELDS = [
    ("n", C.c_int32), ("nseqriv", C.c_int32), ("nlfp", C.c_int32), ("npth", C.c_int32),
    ("npthlev", C.c_int32), ("ndam", C.c_int32),
    ("ladpstp", C.c_int32), ("lfldout", C.c_int32), ("lpthout", C.c_int32),
    ("ldamout", C.c_int32), ("llevee", C.c_int32),
    ("pgrv", C.c_double), ("pmanfld", C.c_double), ("pdstmth", C.c_double),
    ("pcadp", C.c_double), ("pminslp", C.c_double), ("dfrcinc", C.c_double),
]
_PTRS = ("next up_ptr up_idx grarea elevtn nxtdst rivlen rivwth rivhgt rivman rivstomax rivelv "
         "dwnelv fldstomax fldgrd fldhgt i2mask pth_upst pth_down pth_dst pth_elv pth_wth pth_man "
         "pout_ptr pout_idx pinf_ptr pinf_idx levhgt levfrc bashgt levdst levbassto levtopsto "
         "levfilsto i1dam damseq damstat dam_convol dam_emevol dam_adjvol dam_qa dam_qn dam_qf "
         "rivsto fldsto levsto damsto daminf rivout fldout rivout_pre fldout_pre rivdph_pre "
         "fldsto_pre pthflw pthflw_pre runoff rivdph flddph fldfrc fldare sfcelv storge outflw "
         "rivvel rivinf fldinf pthout levdph pthflwsum rivout_avg fldout_avg outflw_avg rivvel_avg "
         "pthout_avg runoff_avg outflw_max rivdph_max storge_max pthflw_avg pthflwsum_avg").split()
_TAIL = [("nadd", C.c_double), ("glb", C.c_double * 11)]
_WORK = "w_sfcelv_pre w_dwnelv_pre w_flddph_pre w_stoout w_rate w_rivinf w_fldinf w_pthout".split()


class _Model(C.Structure):
    _fields_ = _FIELDS + [(p, _P) for p in _PTRS] + _TAIL + [(w, _P) for w in _WORK] + [
        ("upsinf", _P), ("lkine", C.c_int32), ("glb_hist", _P), ("glb_cap", C.c_int32),
        ("ldamh22", C.c_int32), ("dam_norvol", _P), ("dam_r_volupa", _P), ("lslpmix", C.c_int32),
        ("daminf_avg", _P)]


MAP_KEYS = ("next up_ptr up_idx grarea elevtn nxtdst rivlen rivwth rivhgt rivman rivstomax "
            "rivelv dwnelv fldstomax fldgrd fldhgt i2mask").split()
PTH_KEYS = "pth_upst pth_down pth_dst pth_elv pth_wth pth_man pout_ptr pout_idx pinf_ptr pinf_idx".split()
LEV_KEYS = "levhgt levfrc bashgt levdst levbassto levtopsto levfilsto".split()
DAM_KEYS = "i1dam damseq damstat dam_convol dam_emevol dam_adjvol dam_qa dam_qn dam_qf".split()
CELL_STATE = ("rivsto fldsto rivout fldout rivout_pre fldout_pre rivdph_pre fldsto_pre runoff "
              "rivdph flddph fldfrc fldare sfcelv storge outflw rivvel rivinf fldinf pthout "
              "rivout_avg fldout_avg outflw_avg rivvel_avg pthout_avg runoff_avg outflw_max "
              "rivdph_max storge_max").split()
GLB_NAMES = ("stopre stonxt stonew rivinf rivout stopre2 stonew2 rivsto fldsto fldare levsto").split()


def _lib():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} missing: run make -C oracle")
    lib = C.CDLL(LIB)
    lib.cmfo_physics_advance.restype = C.c_int32
    lib.cmfo_physics_advance.argtypes = [C.POINTER(_Model), C.POINTER(C.c_double), C.c_int32]
    lib.cmfo_dam_rule_h22.restype = C.c_double
    lib.cmfo_dam_rule_h22.argtypes = [C.c_double] * 8
    lib.cmfo_calc_adpstp.restype = C.c_int32
    lib.cmfo_calc_adpstp.argtypes = [C.POINTER(_Model), C.c_double, C.POINTER(C.c_double)]
    for f in ("cmfo_physics_fldstg", "cmfo_calc_fldstg_def", "cmfo_levee_fldstg", "cmfo_calc_outpre",
              "cmfo_calc_vars_pre", "cmfo_diag_reset_adpstp", "cmfo_diag_getave_adpstp"):
        getattr(lib, f).restype = None
        getattr(lib, f).argtypes = [C.POINTER(_Model)]
    for f in ("cmfo_calc_outflw", "cmfo_calc_pthout", "cmfo_levee_opt_pthout", "cmfo_damout_calc",
              "cmfo_calc_inflow_lspamat", "cmfo_damout_watbal", "cmfo_calc_stonxt",
              "cmfo_diag_avemax_adpstp"):
        getattr(lib, f).restype = None
        getattr(lib, f).argtypes = [C.POINTER(_Model), C.c_double]
    P = C.c_void_p
    for f in ("cmfo_diag_reset_output", "cmfo_diag_getave_output"):
        getattr(lib, f).restype = None
        getattr(lib, f).argtypes = [P, C.c_int32, C.c_int32, C.c_int32]
    lib.cmfo_diag_avemax_output.restype = None
    lib.cmfo_diag_avemax_output.argtypes = [P, P, C.c_int32, C.c_int32, C.c_int32, C.c_double]
    lib.cmfo_roff_interp.restype = None
    lib.cmfo_roff_interp.argtypes = [C.c_int32] * 4 + [P, P, P, P, C.c_double, C.c_double, P]
    return lib


OUT_FIELDS = ("rivout", "fldout", "outflw", "rivvel", "pthout", "runoff", "daminf",
              "outflw_max", "rivdph_max", "storge_max", "pthflw")


class _Output(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in OUT_FIELDS] + [("nadd", C.c_double)]


class OracleOutput:
    """Output-period accumulators of the oracle (CMF_DIAG_*_OUTPUT,
    src/cmf_calc_diag_mod.F90:222-356) over n cells and npthl path levels."""

    def __init__(self, n: int, npthl: int, ldamout: bool):
        self.n, self.npthl, self.ldamout = n, npthl, int(ldamout)
        self.arr = {k: np.zeros(npthl if k == "pthflw" else n) for k in OUT_FIELDS}
        self.s = _Output(**{k: a.ctypes.data for k, a in self.arr.items()}, nadd=0.0)
        self.lib = _lib()
        self.lib.cmfo_diag_reset_output(C.byref(self.s), n, npthl, self.ldamout)

    def avemax(self, a: dict, dt: float) -> None:
        """a: substep averages / maxima of one outer step (keys of OUT_FIELDS)."""
        keep = {k: np.ascontiguousarray(a[k], np.float64) for k in OUT_FIELDS}
        src = _Output(**{k: v.ctypes.data for k, v in keep.items()}, nadd=0.0)
        self.lib.cmfo_diag_avemax_output(C.byref(self.s), C.byref(src), self.n, self.npthl, self.ldamout,
                                         float(dt))

    def getave(self) -> float:
        self.lib.cmfo_diag_getave_output(C.byref(self.s), self.n, self.npthl, self.ldamout)
        return self.s.nadd


def dam_rule_h22(vol, inflow, norvol, convol, emevol, qn, qf, r_volupa):
    """The oracle's Hanazaki-2022 dam rule (src/cmf_ctrl_damout_mod.F90:354-378)."""
    return float(_lib().cmfo_dam_rule_h22(vol, inflow, norvol, convol, emevol, qn, qf, r_volupa))


def roff_interp(inpx, inpy, inpa, nxin, nyin, grid, drofunit, rmis):
    """ROFF_INTERP of the C oracle (src/cmf_ctrl_forcing_mod.F90:757-790):
    inpx/inpy/inpa [inpn, n], grid [nyin, nxin] -> runoff [n]."""
    ix = np.ascontiguousarray(inpx, np.int32)
    iy = np.ascontiguousarray(inpy, np.int32)
    a = np.ascontiguousarray(inpa, np.float64)
    g = np.ascontiguousarray(grid, np.float64)
    inpn, n = ix.shape
    out = np.empty(n)
    _lib().cmfo_roff_interp(n, inpn, nxin, nyin, ix.ctypes.data, iy.ctypes.data, a.ctypes.data,
                            g.ctypes.data, float(drofunit), float(rmis), out.ctypes.data)
    return out


class OracleModel:
    """FP64 CPU oracle over a cmf.maps.CMFMap and a prognostic state dict."""

    def __init__(self, m, state: dict):
        self.lib = _lib()
        self.map = m
        o = m.opts
        self.s = _Model()
        s = self.s
        s.n, s.nseqriv, s.nlfp = m.n, m.nseqriv, m.nlfp
        s.npth, s.npthlev, s.ndam = m.npth, m.npthlev, m.ndam
        s.ladpstp, s.lfldout, s.lpthout = int(o.ladpstp), int(o.lfldout), int(o.lpthout)
        s.ldamout, s.llevee = int(o.ldamout), int(o.llevee)
        s.lkine = int(getattr(o, "lkine", False))
        s.lslpmix = int(getattr(o, "lslpmix", False))
        s.pgrv, s.pmanfld, s.pdstmth, s.pcadp, s.pminslp = o.pgrv, o.pmanfld, o.pdstmth, o.pcadp, o.pminslp
        s.dfrcinc = m.dfrcinc
        self._keep = {}
        n = m.n
        for k in MAP_KEYS:
            self._bind(k, getattr(m, k))
        if o.lpthout:
            for k in PTH_KEYS:
                self._bind(k, getattr(m, k))
        if o.llevee:
            for k in LEV_KEYS:
                self._bind(k, getattr(m, k))
        if o.ldamout:
            for k in DAM_KEYS:
                self._bind(k, getattr(m, k))
            if getattr(o, "ldamh22", False):
                s.ldamh22 = 1
                self._bind("dam_norvol", m.dam_norvol)
                self._bind("dam_r_volupa", m.dam_r_volupa)
        self.st = {}
        for k in CELL_STATE:
            self.st[k] = np.ascontiguousarray(state.get(k, np.zeros(n)), dtype=np.float64).copy()
        if o.llevee:
            self.st["levsto"] = np.array(state.get("levsto", np.zeros(n)), np.float64)
            self.st["levdph"] = np.array(state.get("levdph", np.zeros(n)), np.float64)
        if o.ldamout:
            self.st["damsto"] = np.array(state["damsto"], np.float64)
            self.st["daminf"] = np.array(state.get("daminf", np.zeros(n)), np.float64)
            self.st["daminf_avg"] = np.array(state.get("daminf_avg", np.zeros(n)), np.float64)
        if o.lpthout:
            shp = (m.npthlev, m.npth)
            for k in ("pthflw", "pthflw_pre", "pthflw_avg"):
                self.st[k] = np.ascontiguousarray(state.get(k, np.zeros(shp)), np.float64).copy()
            for k in ("pthflwsum", "pthflwsum_avg"):
                self.st[k] = np.ascontiguousarray(state.get(k, np.zeros(m.npth)), np.float64).copy()
        for k, a in self.st.items():
            setattr(s, k, a.ctypes.data)
        if state.get("upsinf") is not None:                 # LUPSINF (stonxt:81)
            self._bind("upsinf", np.ascontiguousarray(state["upsinf"], np.float64))
        self.work = {w: np.zeros(n) for w in _WORK}
        for w, a in self.work.items():
            setattr(s, w, a.ctypes.data)

    def _bind(self, k, a):
        if a is None:
            return
        a = np.ascontiguousarray(a)
        self._keep[k] = a
        setattr(self.s, k, a.ctypes.data)

    # reference entry points ------------------------------------------------
    def physics_fldstg(self):
        self.lib.cmfo_physics_fldstg(C.byref(self.s))

    def calc_outpre(self):
        """CMF_CALC_OUTPRE (src/cmf_opt_outflw_mod.F90:308-421)."""
        self.lib.cmfo_calc_outpre(C.byref(self.s))

    def physics_advance(self, dt: float, nt_cap: int = 0, budget: bool = False):
        """One outer step; returns NT, or (NT, budget sums [NT, 11] of every
        substep) when budget=True."""
        d = C.c_double(dt)
        if not budget:
            return int(self.lib.cmfo_physics_advance(C.byref(self.s), C.byref(d), nt_cap))
        cap = 8192
        hist = np.zeros((cap, 11))
        self.s.glb_hist, self.s.glb_cap = hist.ctypes.data, cap
        try:
            nt = int(self.lib.cmfo_physics_advance(C.byref(self.s), C.byref(d), nt_cap))
        finally:
            self.s.glb_hist, self.s.glb_cap = None, 0
        nrun = min(nt if nt_cap <= 0 else min(nt, nt_cap), cap)
        return nt, hist[:nrun].copy()

    @property
    def glb(self) -> dict:
        return dict(zip(GLB_NAMES, list(self.s.glb)))
