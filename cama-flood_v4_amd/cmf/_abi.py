"""ctypes prototypes for include/cmf_host.h and include/cmf_hip.h."""
from __future__ import annotations

import ctypes as C

P = C.c_void_p
I = C.c_int
I32P = C.POINTER(C.c_int32)
DP = C.POINTER(C.c_double)

HOST_FUNCS = {
    "cmfh_calc_1d_seq": (I, [I, I, P, P, P, P, P, P]),
    "cmfh_build_upstream_csr": (I, [I, I, P, P, P]),
    "cmfh_set_fldstg": (I, [I, I, I, P, P, P, P, P, P, P, P, P, P, P]),
    "cmfh_levee_init": (I, [I, I, P, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "cmfh_storage_sea_surface": (I, [I, I, P, P, P, P, P, P, P, P, P]),
    "cmfh_priority_flood": (I, [I, I, P, P, C.c_float, P, P, P]),
    "cmfh_upstream_area": (I, [I, I, P, P, P, P]),
}

HIP_FUNCS: dict = {}


def _declare(lib, table):
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def declare_host(lib):
    _declare(lib, HOST_FUNCS)


def declare_hip(lib):
    _declare(lib, HIP_FUNCS)


def ptr(a):
    """Raw data pointer of a C-contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data
