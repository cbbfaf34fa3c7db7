"""Long-run drift of the FP64 device path against the oracle on the 30-day 1-degree
configuration (15001 cells, ~5400 substeps): at every simulated day the substep
count must agree and every field stay within 1e-10 relative of the oracle. The
oracle is bit-identical to the reference on this run (test_oracle_golden)."""
import numpy as np
import pytest

from cases import STATE_KEYS, rel_err, upload_keys
from golden_cases import runoff_for_step, setup


@pytest.mark.gpu
def test_c1_daily_drift():
    from cmf.model import CMFModel
    from oracle.oracle import OracleModel
    m, net, st, rvec, ref, fin, dt, nsteps = setup("c1_1deg")
    om = OracleModel(m, st)
    gm = CMFModel(m, "fp64")
    gm.set_state(upload_keys(st))
    worst_all = 0.0
    for step in range(1, nsteps + 1):
        r = runoff_for_step(rvec, step, dt)
        om.st["runoff"][:] = r
        gm.set_runoff(r)
        nt_o = om.physics_advance(dt)
        nt_g = gm.physics_advance(dt)
        g = gm.get_state()
        e = {k: rel_err(g[k], om.st[k]) for k in STATE_KEYS}
        k = max(e, key=e.get)
        d = np.abs(g[k] - om.st[k])
        i = int(np.argmax(d))
        print(f"day {step}: NT {nt_o}/{nt_g} worst {k} {e[k]:.3e} at cell {i} "
              f"(oracle {om.st[k][i]:.17g}, device {g[k][i]:.17g}); rivsto {e['rivsto']:.2e}", flush=True)
        worst_all = max(worst_all, e["rivsto"])
        assert nt_o == nt_g, f"day {step}: NT {nt_o} vs {nt_g}"
    gm.close()
    assert worst_all <= 1e-10
