"""HIP physics (through the C-ABI) against the FP64 oracle on the same seeded
inputs: every feature branch of the reference's substep (default physics,
bifurcation, levee, dam, their combination, and flood-heavy levee cases).

Tolerance (FP64 path): relative error 1e-10 of each field's magnitude after every
simulated day, and identical substep counts NT. The only operations that can
differ from the oracle are the fractional powers (device pow vs host libm pow,
both within 1 ulp); every add, multiply, divide and square root is evaluated in
the reference's order and rounds identically.
"""
import numpy as np
import pytest

from cases import CASES, STATE_KEYS, prepare, rel_err, upload_keys

TOL_FP64 = 1e-10
NDAYS = 3


def _run(name, ndays=NDAYS, precision="fp64"):
    from cmf.model import CMFModel
    from oracle.oracle import OracleModel
    m, net, st, rvec = prepare(name, ndays)
    om = OracleModel(m, st)
    gm = CMFModel(m, precision)
    gm.set_state(upload_keys(st))
    errs = []
    for d in range(ndays):
        om.st["runoff"][:] = rvec[d]
        gm.set_runoff(rvec[d])
        nt_o = om.physics_advance(86400.0)
        nt_g = gm.physics_advance(86400.0)
        g = gm.get_state()
        e = {k: rel_err(g[k], om.st[k]) for k in STATE_KEYS}
        if m.opts.llevee:
            e["levsto"] = rel_err(g["levsto"], om.st["levsto"])
            e["levdph"] = rel_err(g["levdph"], om.st["levdph"])
        if m.opts.ldamout:
            e["damsto"] = rel_err(g["damsto"], om.st["damsto"])
        if m.opts.lpthout:
            e["pthflw"] = rel_err(g["pthflw"], om.st["pthflw"])
            e["pthflwsum"] = rel_err(g["pthflwsum"], om.st["pthflwsum"])
        errs.append((nt_o, nt_g, e))
    gm.close()
    return errs


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_fp64_matches_oracle(name):
    print(f"case {name}", flush=True)
    for day, (nt_o, nt_g, e) in enumerate(_run(name)):
        worst = max(e, key=e.get)
        print(f"  day {day + 1}: NT oracle {nt_o} device {nt_g}; worst rel err {worst} {e[worst]:.3e}", flush=True)
        assert nt_o == nt_g
        for k, v in e.items():
            assert v <= TOL_FP64, f"{name} day {day + 1}: {k} rel err {v:.3e}"
