"""LKINE (kinematic-wave outflow, CMF_CALC_OUTFLW_KINE,
src/cmf_opt_outflw_mod.F90:27-127) in the oracle, beside its bit-exact golden
fixture (tests/golden/g_kine*, test_oracle_golden.py): the closed form on single
cells and the mass balance of a whole network (CPU)."""
import numpy as np
import pytest

from cmf import maps, synth


def _basin(seed=11, mean=20.0):
    net = synth.make_network(90, 45, seed=seed)
    m = maps.build_map(net, maps.Options(lkine=True))
    r = m.vec(synth.make_runoff(net, 1, seed, mean_mm_day=mean)[0])
    return m, r


@pytest.mark.parametrize("river", [True, False])
def test_kinematic_closed_form_single_cell(river):
    """Q = W h (1/n) S^(1/2) h^(2/3), S = max((ELEVTN(i) - ELEVTN(next)) / NXTDST,
    PMINSLP) for a river cell and PMINSLP at a mouth, for an in-bank cell whose
    storage covers the substep's outflow (no cap, rate 1)."""
    from oracle.oracle import OracleModel
    m, _ = _basin()
    if river:
        i = int(np.flatnonzero((m.next >= 0) & (m.rivhgt > 2.0))[0])
        sub = maps.subset_map(m, np.array([i, int(m.next[i])]))
        k = 0
    else:
        i = int(np.flatnonzero(m.next < 0)[0])
        sub = maps.subset_map(m, np.array([i]))
        k = 0
    st = maps.cold_start(sub)
    h = 0.5 * float(sub.rivhgt[k])
    st["rivsto"][:] = 0.5 * sub.rivhgt * sub.rivlen * sub.rivwth
    st["fldsto"][:] = 0.0
    om = OracleModel(sub, st)
    om.physics_fldstg()
    assert abs(om.st["rivdph"][k] - h) < 1e-9 * h
    dt = 60.0
    om.st["runoff"][:] = 0.0
    nt = om.physics_advance(dt, nt_cap=1)
    if river and sub.next[k] >= 0:
        s = max((float(sub.elevtn[k]) - float(sub.elevtn[1])) / float(sub.nxtdst[k]), sub.opts.pminslp)
    else:
        s = sub.opts.pminslp
    w, n = float(sub.rivwth[k]), float(sub.rivman[k])
    q = w * h * (1.0 / n) * np.sqrt(s) * h ** (2.0 / 3.0)
    assert nt >= 1
    assert abs(om.st["rivout"][k] - q) <= 1e-6 * q, (om.st["rivout"][k], q)
    assert om.st["fldout"][k] == 0.0


def test_kinematic_network_mass_balance():
    """One hourly outer step of a whole network under LKINE: the storage change
    equals runoff in minus mouth outflow out (to rounding)."""
    from oracle.oracle import OracleModel
    m, r = _basin(seed=12, mean=40.0)
    st = maps.cold_start(m)
    om = OracleModel(m, st)
    om.physics_fldstg()
    om.st["runoff"][:] = r
    s0 = float((om.st["rivsto"] + om.st["fldsto"]).sum())
    dt = 3600.0
    nt = om.physics_advance(dt)
    s1 = float((om.st["rivsto"] + om.st["fldsto"]).sum())
    mouths = m.next < 0
    # the substep average of the outflow (RIVOUT + FLDOUT) over the outer step
    out = float(om.st["outflw_avg"][mouths].sum()) * dt
    inflow = float(r.sum()) * dt
    assert nt >= 1
    assert abs((s1 - s0) - (inflow - out)) <= 1e-9 * max(s0, inflow)
