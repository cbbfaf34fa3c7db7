"""CMF_CALC_OUTPRE (src/cmf_opt_outflw_mod.F90:308-421) and the restart
initialisation of the oracle against the reference's own runs from a restart
file (tests/golden make_golden RESTART_IN): a storage-only restart (LSTOONLY,
where CMF_PHYSICS_FLDSTG and CMF_CALC_OUTPRE rebuild the previous-step flows,
drv_control:273-278) and a full restart, both of the g_allwet case (paths,
levees and dams). The state after initialisation and the state of every outer
step after it are bit-identical to the reference's (SHA-1 of every array).

A closed form pins OUTPRE where no fixture reaches: a single river cell's
previous-step outflow is the Manning discharge W h (h^(2/3) S^(1/2) / n) with
the sign of the slope.
"""
import numpy as np
import pytest

from golden_cases import RESTART_IN_NAMES, restart_init, restart_setup, runoff_for_step, sha

STEP0_KEYS = ["rivsto", "fldsto", "rivout", "fldout", "rivout_pre", "fldout_pre", "rivdph_pre", "fldsto_pre",
              "sfcelv", "rivdph"]


@pytest.mark.parametrize("name", RESTART_IN_NAMES)
def test_oracle_restart_matches_reference(name):
    from oracle.oracle import OracleModel
    m, net, st, binb, pthb, lstoonly, rvec, ref, step0, fin, dt, nsteps = restart_setup(name)
    restart_init(m, st, binb, pthb, lstoonly)
    om = OracleModel(m, st)
    if lstoonly:
        om.physics_fldstg()                           # drv_control:276
        om.calc_outpre()                              # drv_control:277
    for k in STEP0_KEYS + [k for k in ("damsto", "levsto") if k in step0]:
        assert np.array_equal(om.st[k], step0[k]), f"{name} after init: {k}"
    if m.opts.lpthout:
        assert np.array_equal(om.st["pthflw_pre"], step0["pthflw_pre"])
    for step in range(1, nsteps + 1):
        om.st["runoff"][:] = runoff_for_step(rvec, step, dt)
        nt = om.physics_advance(dt)
        assert nt == ref["nt"][step - 1], f"step {step}: NT {nt} vs {ref['nt'][step - 1]}"
        want = ref["steps"][step]["sha"]
        for k in ("rivsto", "fldsto", "rivout", "fldout", "rivdph", "flddph", "outflw", "rivvel", "storge"):
            assert sha(om.st[k]) == want[k], f"{name} step {step}: {k}"


def test_outpre_single_cell_closed_form():
    """One river cell draining to a mouth and the mouth itself: OUTPRE equals
    the Manning discharge of each cell with the slope's sign."""
    from cmf import maps
    from oracle.oracle import OracleModel
    from cases import make_network
    net, opts, mean = make_network("plain")
    m = maps.build_map(net, opts)
    st = maps.cold_start(m)
    rng = np.random.default_rng(4)
    st["rivsto"] = st["rivsto"] * (1.0 + rng.uniform(0.0, 2.0, m.n))
    om = OracleModel(m, st)
    om.physics_fldstg()
    om.calc_outpre()
    s = om.st
    sfc = m.rivelv + s["rivdph"]
    i = int(np.argmax(s["rivdph"][:m.nseqriv]))
    j = int(m.next[i])
    slope = (sfc[i] - sfc[j]) / m.nxtdst[i]
    h = max(sfc[i], sfc[j]) - m.rivelv[i]
    q = m.rivwth[i] * h * h ** (2.0 / 3.0) * abs(slope) ** 0.5 / m.rivman[i] * np.sign(slope)
    assert s["rivout_pre"][i] == pytest.approx(q, rel=1e-6)
    k = m.nseqriv                                      # first mouth
    slope_k = (sfc[k] - m.dwnelv[k]) / m.opts.pdstmth
    hk = s["rivdph"][k]
    qk = m.rivwth[k] * hk * hk ** (2.0 / 3.0) * abs(slope_k) ** 0.5 / m.rivman[k] * np.sign(slope_k)
    assert s["rivout_pre"][k] == pytest.approx(qk, rel=1e-6)
    assert np.array_equal(s["rivdph_pre"], s["rivdph"])
