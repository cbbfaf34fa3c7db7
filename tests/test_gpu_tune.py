"""Launch configurations change no state value: block sizes, XCD-contiguous tile
order, the fused persistent substep kernel and graph replay must all give
bit-identical state (only the rounding of the logged budget sums may move, so
those are compared at 1e-12). Also checks that the autotuner restores the state
it times on."""
import numpy as np
import pytest

from cases import STATE_KEYS, prepare, upload_keys

CONFIGS = [
    dict(),
    dict(bs_outflow=64, bs_rate=128, bs_update=512),
    dict(xcd=1),
    dict(fused=1, bs_fused=256, grid_fused=0),
    dict(fused=1, bs_fused=512, grid_fused=0, xcd=1),
    dict(graph=1),
    dict(graph=1, xcd=1, bs_update=128),
]


def _run(name, cfg, ndays, precision="fp64"):
    from cmf.model import CMFModel
    m, net, st, rvec = prepare(name, ndays)
    gm = CMFModel(m, precision)
    gm.set_tune(**cfg)
    gm.set_state(upload_keys(st))
    out = []
    for d in range(ndays):
        gm.set_runoff(rvec[d])
        nt, glb = gm.physics_advance(86400.0, budget=True)
        out.append((nt, glb, gm.get_state()))
    gm.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["all", "allwet"])
def test_configs_identical(name):
    ref = _run(name, CONFIGS[0], 2)
    for cfg in CONFIGS[1:]:
        print(f"{name}: {cfg}", flush=True)
        got = _run(name, cfg, 2)
        for (nt0, g0, s0), (nt1, g1, s1) in zip(ref, got):
            assert nt0 == nt1
            for k in s0:
                assert np.array_equal(s0[k], s1[k]), f"{cfg}: {k} differs"
            assert np.allclose(g0, g1, rtol=1e-12, atol=0.0), f"{cfg}: budget sums differ"


@pytest.mark.gpu
def test_autotune_restores_state():
    from cmf.model import CMFModel
    m, net, st, rvec = prepare("all", 1)
    a = CMFModel(m, "fp64")
    a.set_state(upload_keys(st))
    a.set_runoff(rvec[0])
    best, rep = a.autotune(86400.0, nsub=4)
    print(f"autotune best {best}; {len(rep)} trials", flush=True)
    assert len(rep) >= 8
    nt_a = a.physics_advance(86400.0)
    sa = a.get_state()
    b = CMFModel(m, "fp64")
    b.set_state(upload_keys(st))
    b.set_runoff(rvec[0])
    nt_b = b.physics_advance(86400.0)
    sb = b.get_state()
    assert nt_a == nt_b
    for k in STATE_KEYS:
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()


@pytest.mark.gpu
def test_graph_replay_after_budget_buffer_growth():
    """Graph replay stays correct when the per-substep budget buffers grow: the
    substep count is capped 10, 10, 20, 10. The second call captures the
    10-substep graph, the third needs more budget rows (the buffers are
    reallocated and every cached graph is dropped), the fourth replays a
    10-substep graph again. State and budget sums equal the run without
    graphs."""
    from cmf.model import CMFModel
    m, net, st, rvec = prepare("all", 1)
    runs = []
    for graph in (0, 1):
        gm = CMFModel(m, "fp64")
        gm.set_tune(graph=graph)
        gm.set_state(upload_keys(st))
        gm.set_runoff(rvec[0])
        out = []
        for cap in (10, 10, 20, 10):
            nt, glb = gm.physics_advance(86400.0, nt_cap=cap, budget=True)
            out.append((glb, gm.get_state()))
        gm.close()
        runs.append(out)
    for k, ((g0, s0), (g1, s1)) in enumerate(zip(*runs)):
        assert g0.shape == g1.shape
        for key in s0:
            assert np.array_equal(s0[key], s1[key]), f"call {k + 1}: {key} differs"
        assert np.allclose(g0, g1, rtol=1e-12, atol=0.0), f"call {k + 1}: budget sums differ"
