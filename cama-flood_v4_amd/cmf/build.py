"""In-tree build of the native libraries.

* libcmf_host.so -- host map initialisation (gcc)
* libcmf_hip.so  -- HIP physics for gfx950 behind the C-ABI of include/cmf_hip.h

FP64 physics is compiled without floating-point contraction so that every
multiply and add rounds exactly as the reference's statements do.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

from ._native import PKG_DIR, ROOT_DIR, HOST_LIB, HIP_LIB

CSRC = os.path.join(os.path.dirname(PKG_DIR), "csrc")
INCLUDE = os.path.join(ROOT_DIR, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CMF_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    print("[cmf.build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_host() -> str:
    src = os.path.join(CSRC, "host", "cmf_host.c")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-std=c11",
          f"-I{INCLUDE}", src, "-o", HOST_LIB, "-lm"])
    return HOST_LIB


def build_hip() -> str:
    srcs = sorted(glob.glob(os.path.join(CSRC, "hip", "*.hip")))
    if not srcs:
        raise RuntimeError("no HIP sources found")
    _run([HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", *srcs, "-o", HIP_LIB])
    return HIP_LIB


def build_all(hip: bool = True) -> None:
    build_host()
    if hip:
        build_hip()


if __name__ == "__main__":
    build_all(hip="--host-only" not in sys.argv)
