"""The drop-in configuration: the reference's own Fortran host (initialisation,
time loop, forcing, dam and levee set-up) linked with this package's
CMF_CTRL_PHYSICS_MOD replacement, which drives the HIP physics through
include/cmf_hip.h. Run on the golden fixtures, it must reproduce the reference's
substep counts and final state to 1e-10 relative.

The binary is built by cama-flood_v4_amd/fortran/build_dropin.sh where the
reference sources exist; the test is skipped when it has not been built."""
import os
import subprocess

import numpy as np
import pytest

from cases import rel_err
from golden_cases import load

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cama-flood_v4_amd", "fortran", "_build", "cmf_dropin")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["g_allwet", "g_hourly", "c1_1deg"])
def test_dropin_matches_reference(name, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("drop-in binary not built")
    from cmf import synth
    from oracle import refdump
    print(f"drop-in case {name}", flush=True)
    net, ro, ref, fin = load(name)
    nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac = ref["spec"]
    d = str(tmp_path)
    synth.write_reference_inputs(net, d, ro, dt=dt, ifrq_inp=int(dt // 3600), ndays=days,
                                 lpthout=lpth, llevee=llev, ldamout=ldam)
    r = subprocess.run([EXE, "runoff.bin", "dump.bin", "0"], cwd=d, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    mp, states = refdump.read_dump(os.path.join(d, "dump.bin"), lpth, llev, ldam)
    nts = refdump.count_substeps(os.path.join(d, "log_CaMa.txt"))
    assert nts == ref["nt"]
    # dump_every=0 writes the initial state only; the final state is read back from
    # a second dump written at the end (dump_every = number of steps)
    nsteps = len(ref["nt"])
    r = subprocess.run([EXE, "runoff.bin", "dump2.bin", str(nsteps)], cwd=d, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    mp, states = refdump.read_dump(os.path.join(d, "dump2.bin"), lpth, llev, ldam)
    last = states[-1]
    assert last["step"] == nsteps
    for k, v in fin.items():
        e = rel_err(last[k], v)
        print(f"  {k}: rel err {e:.3e}", flush=True)
        assert e <= 1e-10, f"{name}: {k} rel err {e:.3e}"
