"""The C-ABI libraries load and export every function their headers declare
(no device calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cmf[h]?_[a-z0-9_]+)\s*\(", txt)))


@pytest.mark.parametrize("header,lib", [("cmf_hip.h", "libcmf_hip.so"), ("cmf_mlp.h", "libcmf_hip.so"),
                                        ("cmf_host.h", "libcmf_host.so")])
def test_exports(header, lib):
    path = os.path.join(ROOT, "cama-flood_v4_amd", "cmf", lib)
    assert os.path.exists(path), f"{lib} not built"
    so = ctypes.CDLL(path)
    names = declared(header)
    assert names, "no declarations parsed"
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_python_bindings_match_headers():
    from cmf import _abi
    assert set(declared("cmf_hip.h")) == set(_abi.HIP_FUNCS)
    assert set(declared("cmf_mlp.h")) == set(_abi.MLP_FUNCS)
    assert set(declared("cmf_host.h")) == set(_abi.HOST_FUNCS)


def test_fortran_bindings_name_declared_functions():
    """Every BIND(C, NAME='...') interface of the Fortran drop-in modules names a
    function that include/cmf_hip.h declares and libcmf_hip.so exports."""
    fdir = os.path.join(ROOT, "cama-flood_v4_amd", "fortran")
    names = set()
    for f in os.listdir(fdir):
        if f.endswith(".F90"):
            txt = open(os.path.join(fdir, f)).read()
            names |= set(re.findall(r"BIND\(C,\s*NAME='([A-Za-z0-9_]+)'\)", txt))
    assert names, "no BIND(C) interfaces found"
    declared_hip = set(declared("cmf_hip.h"))
    assert not (names - declared_hip), f"undeclared: {sorted(names - declared_hip)}"
    so = ctypes.CDLL(os.path.join(ROOT, "cama-flood_v4_amd", "cmf", "libcmf_hip.so"))
    assert all(hasattr(so, n) for n in names)
