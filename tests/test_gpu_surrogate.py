"""What the learned warm start gains over the plain cold start.

Equilibrium storages of 24 small glb_15min-shaped basins with varied runoff
climatologies come from the FP64 device physics (batched spin-up); the network
is trained on them by the project's HIP kernels and applied to three held-out
basins. Each held-out basin is spun up from the cold start (STORAGE_SEA_SURFACE)
and from the warm start; the cold run's last year is the equilibrium the two are
judged against (the cold start stays the fallback and the judge of drift).
Checked: the warm start's initial storage is closer to the equilibrium (median
gain > 3x) and it reaches the 0.1 % band of the equilibrium trajectory no later
than the cold start. On these networks the cold start itself settles within one
simulated year, so on the annual-change criterion both need two years."""
import numpy as np
import pytest


@pytest.mark.gpu
def test_warm_start_beats_cold_start():
    from cmf import surrogate as sg
    train_b = sg.make_basins(24, seed=11)
    test_b = sg.make_basins(3, seed=12)
    X, Y, res = sg.equilibrium_labels(train_b, tol=1e-6, max_years=6,
                                      log=lambda s: print("labels:", s, flush=True))
    print(f"training cells {X.shape[0]}; cold-start years {res.years.tolist()}", flush=True)
    sur, hist = sg.train(X, Y, epochs=60, batch=512, lr=2e-3, seed=0)
    print(f"loss {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    assert hist[-1] < 0.2 * hist[0]
    gains = []
    for b in test_b:
        r = sg.compare_spinup(sur, b, years=3)
        print(f"basin {b.m.n} cells: {r}", flush=True)
        assert r["warm"]["init_rel_err"] < r["cold"]["init_rel_err"]
        assert r["warm"]["days"] <= r["cold"]["days"]
        assert r["warm"]["years"] <= r["cold"]["years"]
        gains.append(r["cold"]["init_rel_err"] / r["warm"]["init_rel_err"])
    assert np.median(gains) > 3.0
