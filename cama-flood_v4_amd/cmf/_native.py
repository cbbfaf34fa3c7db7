"""Locate and load the native libraries of the package.

The HIP library (libcmf_hip.so) carries the device physics; the host library
(libcmf_host.so) carries CPU-side map initialisation. Both are built in-tree by
``cmf.build`` (or ``__graft_entry__.build()``). Loading fails loudly: there is no
CPU fallback for the device physics.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(os.path.dirname(PKG_DIR))
HOST_LIB = os.path.join(PKG_DIR, "libcmf_host.so")
HIP_LIB = os.path.join(PKG_DIR, "libcmf_hip.so")

_host = None
_hip = None


class NativeLibraryMissing(RuntimeError):
    pass


def _load(path: str) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{os.path.basename(path)} not found at {path}; build it first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def host() -> ctypes.CDLL:
    global _host
    if _host is None:
        _host = _load(HOST_LIB)
        from . import _abi
        _abi.declare_host(_host)
    return _host


def hip() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        _hip = _load(HIP_LIB)
        from . import _abi
        _abi.declare_hip(_hip)
    return _hip
