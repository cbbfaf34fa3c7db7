"""The Hanazaki-2022 reservoir rule of the oracle (LDAMH22,
src/cmf_ctrl_damout_mod.F90:354-378), pinned case by case.

The golden fixture g_dam_h22 (the reference's own run) pins the normal, water
supply and emergency zones through the whole physics; its dams have no
flood-control volume, because the reference reads R_VolUpa(IDAM) in the flood
control zone while damout:204 writes every dam's ratio to R_VolUpa(NDAM). That
zone is pinned here instead, with values worked out by hand for one dam
(NorVol 50, ConVol 100, EmeVol 195, Qn 10, Qf 40), and with the single-precision
constant 0.2 of `R_VolUpa/0.2` (0.2 as a default real is 0.20000000298...), which
decides whether a ratio of exactly 0.2 still lets a little inflow through."""
import numpy as np
import pytest

from oracle.oracle import dam_rule_h22

NV, CV, EV, QN, QF = 50.0, 100.0, 195.0, 10.0, 40.0


def rule(vol, inflow, r=0.0):
    return dam_rule_h22(vol, inflow, NV, CV, EV, QN, QF, r)


@pytest.mark.parametrize("vol,inflow,want", [
    (0.0, 100.0, 0.0),                 # empty reservoir releases nothing
    (25.0, 5.0, 2.5),                  # normal zone: Qn * V / ConVol
    (50.0, 100.0, 5.0),                # V = NorVol still belongs to the normal zone
    (75.0, 40.0, 20.0),                # supply, inflow >= Qf: Qn/2 + (25/50) * 30
    (100.0, 60.0, 35.0),               # V = ConVol, inflow >= Qf: Qn/2 + 30
    (75.0, 30.0, 5.0 + 30.0 * (25.0 / 145.0) ** 2),    # supply, inflow < Qf: quadratic in V
    (195.0, 70.0, 70.0),               # V = EmeVol: emergency, max(inflow, Qf)
    (300.0, 20.0, 40.0),               # emergency with inflow below Qf
])
def test_zones_by_hand(vol, inflow, want):
    assert rule(vol, inflow) == pytest.approx(want, rel=1e-15, abs=1e-15)


def test_flood_control_zone():
    # r = 0: Qf + (V - ConVol)/(EmeVol - ConVol) * (inflow - Qf) = 40 + (50/95) * 60
    assert rule(150.0, 100.0, 0.0) == pytest.approx(40.0 + 60.0 * 50.0 / 95.0, rel=1e-15)
    # r = 0.1: half of that excess passes (1 - 0.1/0.2), up to 0.2 held in single precision
    want = 40.0 + 0.5 * 60.0 * 50.0 / 95.0
    assert rule(150.0, 100.0, 0.1) == pytest.approx(want, rel=1e-8)
    assert rule(150.0, 100.0, 0.1) != want
    # r = 0.2 exactly: 1 - 0.2/0.20000000298 = 1.49e-8 > 0, a trace passes
    got = rule(150.0, 100.0, 0.2)
    assert 40.0 < got < 40.0 + 1e-6
    assert got - 40.0 == pytest.approx((1.0 - 0.2 / float(np.float32(0.2))) * 60.0 * 50.0 / 95.0, rel=1e-6)
    # r >= 0.2 (beyond the rounding): everything above Qf is held back
    assert rule(150.0, 100.0, 0.3) == 40.0
    # inflow below Qf: the supply zone's quadratic continues into the flood zone
    assert rule(150.0, 30.0, 0.0) == pytest.approx(5.0 + 30.0 * (100.0 / 145.0) ** 2, rel=1e-15)


def test_continuity_and_order():
    eps = 1e-9
    # the normal and supply zones meet at NorVol: Qn/2 from both sides
    for inflow in (10.0, 100.0):
        assert abs(rule(NV + eps, inflow) - rule(NV, inflow)) < 1e-6
    # inflow < Qf: one quadratic from NorVol to EmeVol, continuous at ConVol
    assert abs(rule(CV + eps, 30.0) - rule(CV, 30.0)) < 1e-6
    # inflow >= Qf: a step at ConVol from Qf - Qn/2 to Qf (the rule's own design)
    assert rule(CV, 100.0) == pytest.approx(QF - QN / 2)
    assert rule(CV + eps, 100.0) == pytest.approx(QF, abs=1e-6)
    # release grows with storage inside each zone
    v = np.linspace(0.0, 194.0, 389)
    for inflow in (20.0, 100.0):
        q = [rule(x, inflow) for x in v]
        zone = np.digitize(v, [NV + 1e-12, CV + 1e-12, EV])
        for z in range(3):
            qz = np.asarray(q)[zone == z]
            assert (np.diff(qz) >= 0).all()
