"""Output-period diagnostics on the device (cmf_hip_diag_*_output) against the
oracle's restatement, itself bit-identical to the reference's routines
(tests/test_output_diag.py): the oracle is fed the device's own substep averages
of every outer step, so the comparison isolates the output accumulation; FP64
and mixed, single model and a 2-member ensemble, with bifurcation paths, levees
and dams, across a reset in the middle of the run. Bit-identical."""
import numpy as np
import pytest

from cases import prepare, upload_keys

A_KEYS = {"rivout": "rivout_avg", "fldout": "fldout_avg", "outflw": "outflw_avg", "rivvel": "rivvel_avg",
          "pthout": "pthout_avg", "runoff": "runoff_avg", "daminf": "daminf_avg", "outflw_max": "outflw_max",
          "rivdph_max": "rivdph_max", "storge_max": "storge_max", "pthflw": "pthflw_avg"}


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("members", [1, 2])
def test_output_diag_matches_oracle(precision, members):
    from cmf.model import CMFModel
    from oracle.oracle import OracleOutput
    m, net, st, rvec = prepare("all", 4)
    kw = {} if members == 1 else dict(members=members, dom_cells=[m.n])
    up = upload_keys(st)
    if members > 1:
        up = {k: np.stack([v] * members) for k, v in up.items()}
    g = CMFModel(m, precision, **kw)
    g.set_state(up)
    npthl = m.npth * m.npthlev

    def fresh():
        return [OracleOutput(m.n, npthl, True) for _ in range(members)]

    def check(orc, nadd_dev):
        got = g.get_output()
        for mem, o in enumerate(orc):
            assert o.s.nadd == nadd_dev
            for k, v in got.items():
                d = v[mem] if members > 1 else v
                assert np.array_equal(d.reshape(-1), o.arr[k]), f"{precision} member {mem}: {k}"

    g.diag_reset_output()
    orc = fresh()
    for day in range(4):
        print(f"{precision} members={members} day {day}", flush=True)
        g.set_runoff(rvec[day] if members == 1 else np.stack([rvec[day] * (1.0 + 0.3 * k) for k in range(members)]))
        if members == 1:
            g.physics_advance(86400.0)
        else:
            g.ensemble_advance(86400.0)
        g.diag_avemax_output(86400.0)
        a = g.get_state(list(A_KEYS.values()))
        for mem, o in enumerate(orc):
            o.avemax({k: (a[v][mem] if members > 1 else a[v]).reshape(-1) for k, v in A_KEYS.items()}, 86400.0)
        if day in (1, 3):                                   # output time: average, compare, reset
            nadd = g.diag_getave_output()
            for o in orc:
                o.getave()
            check(orc, nadd)
            g.diag_reset_output()
            orc = fresh()
    g.close()


@pytest.mark.gpu
def test_output_diag_errors():
    from cmf.model import CMFError, CMFModel
    m, net, st, rvec = prepare("plain", 1)
    g = CMFModel(m, "fp64")
    with pytest.raises(CMFError):
        g.diag_getave_output()                               # nothing accumulated yet
    with pytest.raises(CMFError):
        g.get_output()
    g.close()
