"""Host-side logic (no GPU): upstream accumulation, the climatological year, the
spin-up convergence measure, domain concatenation and the surrogate features."""
import numpy as np

from cmf import maps, spinup, synth


class _Map:
    def __init__(self, nxt):
        self.next = np.asarray(nxt, np.int32)
        self.n = len(nxt)


def test_accumulate_closed_form():
    # 0 -> 2, 1 -> 2, 2 -> 4, 3 -> 4, 4 mouth; 5 isolated mouth
    m = _Map([2, 2, 4, 4, -9, -10])
    v = np.array([1.0, 2.0, 4.0, 8.0, 16.0, 32.0])
    out = spinup.accumulate(m, v)
    assert out.tolist() == [1.0, 2.0, 7.0, 8.0, 31.0, 32.0]
    two = spinup.accumulate(m, np.stack([v, 2 * v]))
    assert np.array_equal(two[1], 2 * out)


def test_climatology_daily_is_periodic_and_hits_month_centres():
    mon = np.arange(12, dtype=float)[:, None] * np.ones((1, 3))
    c = spinup.Climatology(mon, 1.0, 0.0, 0.0)
    assert np.allclose(c.daily(0), c.daily(365))
    # day d covers [d, d+1): its midpoint d+0.5 equals a month centre for d = 15 (15.208 ~ 15.5)
    mids = (np.arange(12) + 0.5) * 365 / 12
    for k in (3, 7):
        d = int(np.floor(mids[k] - 0.5))
        w = (d + 0.5 - mids[k - 1]) / (365 / 12)
        assert np.allclose(c.daily(d), (1 - w) * mon[k - 1] + w * mon[k])
    # continuity across the year end
    assert np.abs(c.daily(364) - c.daily(0)).max() < np.abs(mon[11] - mon[0]).max()


def test_days_to_equilibrium():
    eq = np.ones(365)
    run = np.concatenate([np.linspace(0.5, 0.9, 100), np.ones(630)])   # days 0..99 outside the band
    assert spinup.days_to_equilibrium(run, eq, 1e-3) == 100
    assert spinup.days_to_equilibrium(np.ones(730), eq, 1e-3) == 0


def test_concat_maps_shifts_indices():
    ms = []
    for seed in (1, 2):
        net = synth.make_network(40, 20, seed=seed)
        synth.add_bifurcations(net, 10, 3, seed)
        ms.append(maps.build_map(net, maps.Options(lpthout=True)))
    cm, dom = maps.concat_maps(ms)
    n0 = ms[0].n
    assert dom == [ms[0].n, ms[1].n] and cm.n == sum(dom)
    assert np.array_equal(cm.next[n0:][ms[1].next >= 0], ms[1].next[ms[1].next >= 0] + n0)
    assert cm.npth == ms[0].npth + ms[1].npth
    p0 = ms[0].npth
    ok = ms[1].pth_upst >= 0
    assert np.array_equal(cm.pth_upst[p0:][ok], ms[1].pth_upst[ok] + n0)
    assert np.array_equal(cm.fldhgt[:, n0:], ms[1].fldhgt)


def test_surrogate_features_finite():
    from cmf import surrogate
    net = synth.make_network(40, 20, seed=3)
    m = maps.build_map(net, maps.Options())
    clim = spinup.make_climatology(net, m, 3, 1.5, 0.5, 2.0)
    pr = surrogate.prior(m, clim)
    X = surrogate.features(m, clim, pr)
    assert X.shape == (m.n, surrogate.N_IN) and np.isfinite(X).all()
    # storage-ratio targets invert to storages; the flow-ratio floor applies
    s_eq = 1.5 * pr.storage
    y = surrogate.targets(pr, s_eq, 1e-9 * pr.qm)
    fast = surrogate.fast_cells(X)
    assert 0 < fast.sum() < m.n
    got = surrogate.storage(pr, y, X)
    assert np.allclose(got[fast], s_eq[fast], rtol=1e-6)
    assert np.allclose(got[~fast], pr.storage[~fast], rtol=1e-12)     # slow cells keep the prior
    Xs, Ys, ws = surrogate.training_set(X, y, s_eq)
    assert Xs.shape[0] == fast.sum() == ws.size and ws.min() >= 0.05 and ws.max() <= 50.0
    assert np.allclose(y[:, 1], np.log(surrogate.Q_FLOOR))
    # window ratios of a climatology without seasonality are 1 (log 0), and its
    # periodic prior is the steady state
    flat = spinup.Climatology(np.repeat(clim.monthly.mean(axis=0, keepdims=True), 12, axis=0), 1.5, 0.0, 0.0)
    pf = surrogate.prior(m, flat)
    Xf = surrogate.features(m, flat, pf)
    assert np.abs(Xf[:, 7:11]).max() < 1e-5 and np.abs(Xf[:, 16:18]).max() < 1e-5
    assert np.abs(Xf[:, 11]).max() < 1e-4 and np.abs(Xf[:, 12]).max() < 1e-4


def test_travel_time_closed_form():
    # chain 0 -> 1 -> 2 (mouth) plus 3 -> 2; local runoff r, residence times tau
    m = _Map([1, 2, -9, 2])
    r = np.array([1.0, 2.0, 4.0, 8.0])
    tau = np.array([10.0, 20.0, 30.0, 40.0])
    q = spinup.accumulate(m, r)
    T = spinup.travel_time(m, q, tau)
    # cell 1: runoff of 0 travels tau0 = 10, its own 0 -> (1*10 + 2*0) / 3
    assert np.isclose(T[1], 10.0 / 3.0)
    # cell 2: 0 travels 10+20, 1 travels 20, 3 travels 40, own 0
    assert np.isclose(T[2], (1 * 30 + 2 * 20 + 8 * 40) / 15.0)
    assert T[0] == 0.0 and T[3] == 0.0

