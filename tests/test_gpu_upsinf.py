"""LUPSINF on the device (cmf_hip_set_upsinf, src/cmf_calc_stonxt_mod.F90:81)
against the oracle with the same D2UPSINF (itself pinned by the one-substep
difference in tests/test_steady_state.py): every state field within the FP64
parity bound (1e-10 relative) and identical NT after every day, with the inflow
changed between days; and a mixed-precision 2-member ensemble whose members both
see the cell's inflow matches two single-model runs bit for bit."""
import numpy as np
import pytest

from cases import STATE_KEYS, prepare, rel_err, upload_keys


def _upsinf(m, day):
    u = np.zeros(m.n)
    rng = np.random.default_rng(40 + day)
    k = rng.choice(m.n, max(5, m.n // 50), replace=False)
    u[k] = rng.uniform(10.0, 500.0, k.size)
    return u


@pytest.mark.gpu
def test_upsinf_matches_oracle():
    from cmf.model import CMFModel
    from oracle.oracle import OracleModel
    m, net, st, rvec = prepare("all", 3)
    u0 = _upsinf(m, 0)
    om = OracleModel(m, dict(st, upsinf=u0))
    gm = CMFModel(m, "fp64")
    gm.set_state(upload_keys(st))
    for d in range(3):
        u = _upsinf(m, d)
        om._keep["upsinf"][:] = u                  # the oracle reads D2UPSINF through its bound array
        gm.set_upsinf(u)
        om.st["runoff"][:] = rvec[d]
        gm.set_runoff(rvec[d])
        nt_o = om.physics_advance(86400.0)
        nt_g = gm.physics_advance(86400.0)
        g = gm.get_state()
        print(f"day {d + 1}: NT oracle {nt_o} device {nt_g}", flush=True)
        assert nt_o == nt_g
        for k in STATE_KEYS:
            e = rel_err(g[k], om.st[k])
            assert e <= 1e-10, f"day {d + 1}: {k} rel err {e:.3e}"
    gm.close()


@pytest.mark.gpu
def test_upsinf_ensemble_members():
    from cmf.model import CMFModel
    m, net, st, rvec = prepare("plain", 1)
    u = _upsinf(m, 0)
    up = upload_keys(st)
    e = CMFModel(m, "mixed", members=2, dom_cells=[m.n])
    e.set_state({k: np.stack([v, v]) for k, v in up.items()})
    e.set_upsinf(np.stack([u, 2.0 * u]))
    e.set_runoff(np.stack([rvec[0], rvec[0]]))
    e.ensemble_advance(86400.0)
    se = e.get_state()
    e.close()
    for mem, scale in enumerate((1.0, 2.0)):
        s = CMFModel(m, "mixed")
        s.set_state(up)
        s.set_upsinf(scale * u)
        s.set_runoff(rvec[0])
        s.physics_advance(86400.0)
        ss = s.get_state()
        s.close()
        for k in ss:
            assert np.array_equal(se[k][mem], ss[k]), f"member {mem}: {k}"
