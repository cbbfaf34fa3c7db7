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
            for k in ("damsto", "daminf", "daminf_avg"):
                e[k] = rel_err(g[k], om.st[k])
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


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["plain", "allwet"])
def test_dt_min_closed_form(name):
    """DT_MIN of CALC_ADPSTP (physics:178-197) returned by cmf_hip_get_dt_min
    equals the closed form min(DT_DEF, min over I2MASK == 0 of
    PCADP * dx / sqrt(PGRV * max(RIVDPH, 0.01))), dx = NXTDST for river cells
    and PDSTMTH for mouths, on the stage at the start of the outer step; and NT
    follows from it as physics:205 writes it."""
    from cmf.model import CMFModel
    from oracle.oracle import OracleModel
    m, net, st, rvec = prepare(name, 2)
    om = OracleModel(m, st)
    gm = CMFModel(m, "fp64")
    gm.set_state(upload_keys(st))
    o = m.opts
    for d in range(2):
        om.physics_fldstg()                       # the stage CALC_ADPSTP sees (physics:54)
        ddph = np.maximum(om.st["rivdph"], 0.01)
        dx = np.where(m.next >= 0, m.nxtdst, o.pdstmth)
        c = (o.pcadp * dx) * (1.0 / np.sqrt(o.pgrv * ddph))
        want = min(86400.0, float(c[m.i2mask == 0].min()))
        om.st["runoff"][:] = rvec[d]
        gm.set_runoff(rvec[d])
        nt_o = om.physics_advance(86400.0)
        nt_g = gm.physics_advance(86400.0)
        got = float(gm.dt_min()[0])
        print(f"  day {d + 1}: DT_MIN device {got!r} closed form {want!r}, NT {nt_g}", flush=True)
        assert got == want
        assert nt_g == nt_o == int(86400.0 * (1.0 / got) - float(np.float32(0.01))) + 1
    gm.close()


@pytest.mark.gpu
def test_daminf_set_outside_dams():
    """DAMINF changes only in cells with I1DAM > 0 (damout:442-465); every other
    cell keeps the value the host gave it, and its D2DAMINF_aAVG (diag:130-135)
    accumulates that value every substep. The device skips those cells while all
    of them hold zero (the reference's initial value); a host state with nonzero
    values there takes the full path. Both against the oracle (1e-10, identical
    NT), and back to zero on the same handle."""
    from cmf.model import CMFModel
    from oracle.oracle import OracleModel
    m, net, st, rvec = prepare("dam", 2)
    rng = np.random.default_rng(5)
    dense = dict(st)
    dense["daminf"] = np.where(m.i1dam == 0, rng.uniform(0.5, 50.0, m.n), 0.0)
    gm = CMFModel(m, "fp64")
    for label, s0 in (("nonzero outside dams", dense), ("zero outside dams", st)):
        om = OracleModel(m, s0)
        gm.set_state(upload_keys(s0))
        for d in range(2):
            om.st["runoff"][:] = rvec[d]
            gm.set_runoff(rvec[d])
            nt_o = om.physics_advance(86400.0)
            nt_g = gm.physics_advance(86400.0)
            g = gm.get_state(["damsto", "daminf", "daminf_avg", "rivsto", "fldsto"])
            e = {k: rel_err(g[k], om.st[k]) for k in g}
            print(f"  {label}, day {d + 1}: NT {nt_o}/{nt_g}, {e}", flush=True)
            assert nt_o == nt_g
            for k, v in e.items():
                assert v <= TOL_FP64, f"{label} day {d + 1}: {k} rel err {v:.3e}"
        avg = gm.get_state(["daminf_avg"])["daminf_avg"]
        outside = m.i1dam == 0
        if s0 is dense:
            assert np.all(avg[outside] > 0.0)
        else:
            assert np.all(avg[outside] == 0.0)
    gm.close()
