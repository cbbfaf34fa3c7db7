"""HIP physics (FP64, through the C-ABI) against the reference Fortran's golden
fixtures: identical substep counts NT at every outer step and final fields within
1e-10 relative of each field's magnitude (the required bound on storage per
simulated day, applied here to the whole run)."""
import numpy as np
import pytest

from cases import rel_err, upload_keys
from golden_cases import NAMES, runoff_for_step, setup

TOL = 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_device_matches_reference(name):
    from cmf.model import CMFModel
    print(f"golden case {name}", flush=True)
    m, net, st, rvec, ref, fin, dt, nsteps = setup(name)
    gm = CMFModel(m, "fp64")
    gm.set_state(upload_keys(st))
    nts = []
    for step in range(1, nsteps + 1):
        gm.set_runoff(runoff_for_step(rvec, step, dt))
        nts.append(gm.physics_advance(dt))
    g = gm.get_state(list(fin))
    gm.close()
    assert nts == ref["nt"]
    for k, v in fin.items():
        e = rel_err(g[k], v)
        print(f"  {k}: rel err {e:.3e}", flush=True)
        assert e <= TOL, f"{name}: {k} rel err {e:.3e}"
