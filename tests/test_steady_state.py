"""Steady state of the local-inertial outflow (cmfh_steady_state), checked against
the oracle's restatement of the reference physics (CPU):

* storage <-> depth: the storages it returns, passed through the reference's
  stage inversion (CMF_PHYSICS_FLDSTG), give back its surface elevations;
* fixed point: one hourly outer step of the reference physics from the steady
  state under the runoff that produces its flows changes no storage, no outflow;
* Manning closed form on a single in-bank mouth cell.

and the linear periodic response around it (cmfh_periodic_response): the closed
form of a single linear reservoir, and at zero frequency the sensitivity of the
steady state itself (a finite difference of two steady solves).
"""
import numpy as np
import pytest

from cmf import maps, spinup, synth


def _basin(seed, mean):
    net = synth.make_network(90, 45, seed=seed)
    m = maps.build_map(net, maps.Options())
    r = m.vec(synth.make_runoff(net, 1, seed, mean_mm_day=mean)[0])
    return m, r


@pytest.mark.parametrize("seed,mean", [(11, 1.5), (12, 40.0)])
def test_steady_state_is_a_fixed_point_of_the_reference_physics(seed, mean):
    from oracle.oracle import OracleModel
    m, r = _basin(seed, mean)
    q = spinup.accumulate(m, r)
    ss = spinup.steady_state(m, q)
    assert np.all(ss["rivsto"] >= 0) and np.all(ss["fldsto"] >= 0)
    flooded = int((ss["fldsto"] > 0).sum())
    if mean > 10:
        assert flooded > 50                      # the wet case exercises the floodplain tables
    st = maps.cold_start(m)
    st["rivsto"], st["fldsto"] = ss["rivsto"].copy(), ss["fldsto"].copy()
    om = OracleModel(m, st)
    om.physics_fldstg()
    # storage -> stage of the reference returns the solver's surface elevation
    depth = ss["sfcelv"] - m.rivelv
    assert np.max(np.abs(om.st["rivdph"] - depth) / np.maximum(depth, 1e-3)) < 1e-9
    # previous-step values equal to the current ones: a steady flow
    om.st["rivout_pre"][:] = ss["rivout"]
    om.st["fldout_pre"][:] = ss["fldout"]
    om.st["rivdph_pre"][:] = om.st["rivdph"]
    om.st["fldsto_pre"][:] = om.st["fldsto"]
    om.st["runoff"][:] = r
    sto0 = om.st["rivsto"] + om.st["fldsto"]
    om.physics_advance(3600.0)
    sto1 = om.st["rivsto"] + om.st["fldsto"]
    # per-cell storage change relative to the hour's throughput (q * 3600 s). Not
    # 0: at sea-level mouths the surface slope over PDSTMTH is ~1e-9 m, so the
    # rounding of RIVDPH = RIVSTO/RIVLEN/RIVWTH moves the outflow by ~1e-6
    ch = np.abs(sto1 - sto0) / np.maximum(q * 3600.0, 1e-3)
    assert ch.max() < 1e-4, ch.max()
    assert np.median(ch) < 1e-9
    out = om.st["rivout"] + om.st["fldout"]
    assert np.max(np.abs(out - q) / np.maximum(q, 1e-9)) < 1e-5


def test_manning_closed_form_single_cell():
    """One in-bank mouth cell: Q = W h^(5/3) ((h + RIVELV - DWNELV) / PDSTMTH)^(1/2) / n."""
    m, _ = _basin(11, 1.5)
    i = int(np.flatnonzero(m.next < 0)[0])
    sub = maps.subset_map(m, np.array([i]))
    assert sub.n == 1 and sub.next[0] < 0
    sub.dwnelv = sub.rivelv - 1.0                  # downstream water surface 1 m below the bed
    h = 0.3 * float(sub.rivhgt[0])
    s = float(sub.rivelv[0]) + h
    q = float(sub.rivwth[0]) * h ** (5 / 3) * np.sqrt((s - float(sub.dwnelv[0])) / sub.opts.pdstmth) / float(sub.rivman[0])
    ss = spinup.steady_state(sub, np.array([q]))
    assert abs(ss["sfcelv"][0] - s) < 1e-9
    assert abs(ss["rivsto"][0] - h * float(sub.rivlen[0] * sub.rivwth[0])) < 1e-9 * ss["rivsto"][0]
    assert ss["fldsto"][0] == 0.0 and ss["fldout"][0] == 0.0


def _response(m, sfc, omega, r):
    """cmfh_periodic_response for one harmonic (complex local runoff r)."""
    import ctypes as C
    from cmf import _native
    from cmf._abi import ptr
    lib = _native.host()
    ch, keep = spinup._channel(m)
    rr, ri = np.ascontiguousarray(r.real[None]), np.ascontiguousarray(r.imag[None])
    hr, hi = np.empty_like(rr), np.empty_like(ri)
    w = np.array([omega])
    s = np.ascontiguousarray(sfc, np.float64)
    orr, oi = np.empty_like(rr), np.empty_like(ri)
    assert lib.cmfh_periodic_response(C.byref(ch), ptr(s), 1, ptr(w), ptr(rr), ptr(ri), ptr(hr), ptr(hi),
                                      ptr(orr), ptr(oi)) == 0
    _response.outflow = orr[0] + 1j * oi[0]
    return hr[0] + 1j * hi[0]


def test_periodic_response_single_reservoir_closed_form():
    """One in-bank mouth cell is a reservoir of area RIVLEN*RIVWTH with outflow
    Q(s) = W h^(5/3) ((s - DWNELV) / PDSTMTH)^(1/2) / n: the response to runoff
    amplitude r at frequency w is r / (i w A + dQ/ds)."""
    m, _ = _basin(11, 1.5)
    sub = maps.subset_map(m, np.array([int(np.flatnonzero(m.next < 0)[0])]))
    sub.dwnelv = sub.rivelv - 1.0
    W, L, n, A = float(sub.rivwth[0]), float(sub.opts.pdstmth), float(sub.rivman[0]), float(sub.rivlen[0] * sub.rivwth[0])
    h = 0.3 * float(sub.rivhgt[0])
    s = float(sub.rivelv[0]) + h
    q = W * h ** (5 / 3) * np.sqrt((s - float(sub.dwnelv[0])) / L) / n
    dq = q * (5.0 / (3.0 * h) + 0.5 / (s - float(sub.dwnelv[0])))
    for w in (0.0, 2 * np.pi / (365 * 86400.0), 2 * np.pi / 86400.0):
        r = np.array([0.3 * q * np.exp(0.7j)])
        got = _response(sub, np.array([s]), w, r)[0]
        want = r[0] / (1j * w * A + dq)
        assert abs(got - want) < 1e-5 * abs(want), (w, got, want)
        # outflow amplitude dQ/ds * h, and the balance r = i w A h + o
        o = _response.outflow[0]
        assert abs(o - dq * want) < 1e-5 * abs(dq * want)
        assert abs(r[0] - (1j * w * A * got + o)) < 1e-6 * abs(r[0])


def test_periodic_response_zero_frequency_is_the_steady_sensitivity():
    """At w = 0 the response to a small runoff change dr is the change of the
    steady state: steady(q + acc(dr)) - steady(q), for the whole coupled network
    (downstream backwater included)."""
    m, r = _basin(12, 40.0)
    q = spinup.accumulate(m, r)
    s0 = spinup.steady_state(m, q)
    dr = 1e-3 * r
    s1 = spinup.steady_state(m, q + spinup.accumulate(m, dr))
    got = _response(m, s0["sfcelv"], 0.0, dr.astype(complex))
    want = s1["sfcelv"] - s0["sfcelv"]
    # at zero frequency the outflow change is the accumulated runoff change
    dq = spinup.accumulate(m, dr)
    assert np.max(np.abs(_response.outflow.real - dq)) < 1e-6 * np.abs(dq).max()
    assert np.abs(got.imag).max() == 0.0
    scale = np.abs(want).max()
    assert scale > 0
    assert np.max(np.abs(got.real - want)) < 2e-2 * scale
    assert np.median(np.abs(got.real - want) / np.maximum(np.abs(want), 1e-3 * scale)) < 1e-2


def test_storage_at_level_inverts_the_stage():
    """cmfh_storage_at_level on the steady surface returns the steady storages."""
    import ctypes as C
    from cmf import _native
    from cmf._abi import ptr
    m, r = _basin(12, 40.0)
    ss = spinup.steady_state(m, spinup.accumulate(m, r))
    ch, keep = spinup._channel(m)
    riv, fld = np.empty(m.n), np.empty(m.n)
    assert _native.host().cmfh_storage_at_level(C.byref(ch), ptr(ss["sfcelv"]), ptr(riv), ptr(fld)) == 0
    assert np.allclose(riv, ss["rivsto"], rtol=1e-12, atol=1e-6)
    assert np.allclose(fld, ss["fldsto"], rtol=1e-12, atol=1e-6)


def test_oracle_upstream_inflow_adds_to_the_river():
    """LUPSINF (src/cmf_calc_stonxt_mod.F90:81): over one substep the inflow
    D2UPSINF*DT lands in the river storage and nothing else moves: both runs see
    the same outflows (computed before STONXT), so the total storage after the
    substep differs by exactly the added volume (to rounding)."""
    from oracle.oracle import OracleModel
    m, r = _basin(12, 40.0)
    ss = spinup.steady_state(m, spinup.accumulate(m, r))
    st = maps.cold_start(m)
    st["rivsto"], st["fldsto"] = ss["rivsto"].copy(), ss["fldsto"].copy()
    base = OracleModel(m, st)
    base.physics_fldstg()
    st1 = {k: v.copy() for k, v in base.st.items()}
    st1["runoff"] = r.copy()
    ups = np.zeros(m.n)
    k = np.random.default_rng(0).choice(m.n, 20, replace=False)
    ups[k] = 50.0
    a = OracleModel(m, dict(st1))
    b = OracleModel(m, dict(st1, upsinf=ups))
    dt = 600.0
    na = a.physics_advance(dt, nt_cap=1)
    nb = b.physics_advance(dt, nt_cap=1)
    assert na == nb
    sa = a.st["rivsto"] + a.st["fldsto"]
    sb = b.st["rivsto"] + b.st["fldsto"]
    dts = dt / na
    added = ups * dts
    assert np.allclose(sb - sa, added, rtol=0, atol=1e-4)          # m3, against 3e4 m3 added
    assert np.all(np.abs(sb - sa)[ups == 0] <= 1e-4)
