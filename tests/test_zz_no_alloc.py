"""The allocation-free advance (cmf_hip_physics_advance_noalloc) against the
regular entry point (cmf_hip_physics_advance), through the C-ABI: two calls in
a row on two models, the second smaller than the first, each compared with a
fresh model stepped by the regular entry point on the same inputs. Both run the
same kernels in the same order, so the state is bit-identical and the budget
sums of every substep equal (the allocation-free variant reduces each substep's
tile sums with the same fixed tree, right after the substep). The second call
must not see the first call's scratch contents."""
import numpy as np
import pytest

from cases import STATE_KEYS, prepare, upload_keys


def _pair(name, ndays):
    m, net, st, rvec = prepare(name, ndays)
    return m, st, rvec


def _small():
    """A 48 x 24 network without options: fewer cells than every 90 x 45 case."""
    from cmf import maps, synth
    from oracle.oracle import OracleModel
    net = synth.make_network(48, 24, seed=21)
    m = maps.build_map(net, maps.Options())
    om = OracleModel(m, maps.cold_start(m))
    om.physics_fldstg()
    st = {k: v.copy() for k, v in om.st.items()}
    return m, st, [m.vec(synth.make_runoff(net, 1, 21, mean_mm_day=5.0)[0])]


@pytest.mark.gpu
def test_noalloc_matches_regular_entry_point():
    from cmf.model import CMFModel
    big = _pair("allwet", 1)          # 90 x 45, paths, levees, dams, flooded
    small = _small()                  # fewer cells, no options
    assert small[0].n < big[0].n
    for (m, st, rvec), label in ((big, "first, larger"), (small, "second, smaller")):
        a = CMFModel(m, "fp64")
        b = CMFModel(m, "fp64")
        for g in (a, b):
            g.set_state(upload_keys(st))
            g.set_runoff(rvec[0])
        nt_a, glb_a = a.physics_advance_noalloc(86400.0)
        nt_b, glb_b = b.physics_advance(86400.0, budget=True)
        sa, sb = a.get_state(), b.get_state()
        print(f"  {label}: n={m.n} NT {nt_a}/{nt_b}", flush=True)
        assert nt_a == nt_b and glb_a.shape == glb_b.shape
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), f"{label}: {k} differs"
        assert np.array_equal(glb_a, glb_b), f"{label}: budget sums differ"
        a.close()
        b.close()


@pytest.mark.gpu
def test_noalloc_rejects_ensembles():
    from cmf import maps
    from cmf.model import CMFError, CMFModel
    m, st, rvec = _pair("plain", 1)
    cm, dc = maps.concat_maps([m, m])
    e = CMFModel(cm, "fp64", members=2, dom_cells=dc)
    with pytest.raises(CMFError, match="single-model"):
        e.physics_advance_noalloc(86400.0)
    e.close()
