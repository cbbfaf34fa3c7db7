"""The device physics is deterministic: two runs of the same inputs give
bit-identical state (all neighbour sums are ordered gathers, all reductions use
a fixed tree)."""
import numpy as np
import pytest

from cases import upload_keys
from golden_cases import runoff_for_step, setup


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_two_runs_identical(precision):
    from cmf.model import CMFModel
    m, net, st, rvec, ref, fin, dt, nsteps = setup("g_allwet")
    outs = []
    for rep in range(2):
        gm = CMFModel(m, precision)
        gm.set_state(upload_keys(st))
        nts = []
        for step in range(1, nsteps + 1):
            gm.set_runoff(runoff_for_step(rvec, step, dt))
            nts.append(gm.physics_advance(dt))
        outs.append((nts, gm.get_state()))
        gm.close()
    assert outs[0][0] == outs[1][0]
    for k, a in outs[0][1].items():
        assert np.array_equal(a, outs[1][1][k]), f"{precision}: {k} differs between identical runs"
