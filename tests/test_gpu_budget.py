"""Water-budget sums of the device (CALC_WATBAL, src/cmf_ctrl_physics_mod.F90:216-248):
the P0GLB* values the reference logs after every substep.

Two pins:
* against the reference's own values: the golden fixtures hold the sums of the
  last substep of every outer step, as the reference Fortran computed them;
* against the oracle, substep by substep (the oracle's sums are bit-identical to
  the reference's, tests/test_oracle_golden.py).

Tolerance: 1e-12 relative, with a floor of 1e-12 of the total storage for sums
that are zero or nearly so. The device adds the same per-cell FP64 terms but in
another order (per-tile sums over the depth-first cell order, then a fixed tree
over tiles and across the wavefront) than the reference's serial loop over NSEQ,
so the sums agree to FP64 rounding of a sum (at most ~n*2^-53 relative for
sums of nonnegative terms, in practice ~sqrt(n)*2^-53 ~ 1e-14), not bit for bit.
A dropped or double-counted cell moves a sum by at least one cell's share, far
above this bound.
"""
import numpy as np
import pytest

from cases import prepare, upload_keys
from golden_cases import NAMES, runoff_for_step, setup

GLB10 = ("stopre", "stonxt", "stonew", "rivinf", "rivout", "stopre2", "stonew2", "rivsto", "fldsto", "fldare")
TOL = 1e-12


def budget_err(dev, ref):
    """Largest error over the entries, relative to each entry with a floor of
    1e-12 of the total storage (entry 0, STOPRE)."""
    dev = np.asarray(dev, np.float64)
    ref = np.asarray(ref, np.float64)
    floor = abs(ref[..., :1]) if ref.ndim == 1 else np.abs(ref[:, :1])
    scale = np.maximum(np.abs(ref), floor)
    scale = np.where(scale > 0, scale, 1.0)
    return float(np.max(np.abs(dev - ref) / scale))


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_budget_matches_reference(name):
    """The last substep's sums of every outer step against the reference's."""
    from cmf.model import CMFModel
    print(f"golden case {name}", flush=True)
    m, net, st, rvec, ref, fin, dt, nsteps = setup(name)
    gm = CMFModel(m, "fp64")
    gm.set_state(upload_keys(st))
    worst = 0.0
    for step in range(1, nsteps + 1):
        gm.set_runoff(runoff_for_step(rvec, step, dt))
        nt, glb = gm.physics_advance(dt, budget=True)
        assert nt == ref["nt"][step - 1]
        want = [ref["steps"][step]["glb"][k] for k in GLB10]
        e = budget_err(glb[-1, :10], want)
        worst = max(worst, e)
        assert e <= TOL, f"{name} step {step}: budget rel err {e:.3e}\n dev {glb[-1, :10]}\n ref {want}"
    gm.close()
    print(f"  worst budget rel err {worst:.3e}", flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["pth", "levwet", "dam", "allwet"])
def test_budget_matches_oracle_every_substep(name):
    from cmf.model import CMFModel
    from oracle.oracle import OracleModel
    print(f"case {name}", flush=True)
    m, net, st, rvec = prepare(name, 2)
    om = OracleModel(m, st)
    gm = CMFModel(m, "fp64")
    gm.set_state(upload_keys(st))
    for d in range(2):
        om.st["runoff"][:] = rvec[d]
        gm.set_runoff(rvec[d])
        nt_o, g_o = om.physics_advance(86400.0, budget=True)
        nt_g, g_g = gm.physics_advance(86400.0, budget=True)
        assert nt_o == nt_g and g_o.shape == g_g.shape
        e = budget_err(g_g, g_o)
        print(f"  day {d + 1}: NT {nt_g}, worst budget rel err over substeps {e:.3e}", flush=True)
        assert e <= TOL
    gm.close()
