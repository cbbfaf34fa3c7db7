"""What the learned warm start gains over the plain cold start.

Equilibria of river basins cut from a small glb_15min-shaped network (one runoff
climatology per basin) come from the FP64 device physics (batched spin-up from
the physical prior, the network's linear periodic response); the network learns
the residual of their year-start storage over that prior with the project's HIP
kernels and is applied to a held-out network. The held-out network is spun up
from the cold start (STORAGE_SEA_SURFACE), from the steady state of the
annual-mean flow, from the prior and from the learned warm start as four members
of one handle.
Checked: the prior and the warm start begin within 5% of the cold start's
distance from the equilibrium, the warm start changes less than 5% as much in
its first year, and neither needs more years than the cold start to reach an
annual storage change below 0.1%. What the trained network adds to the method
without it (the physical prior it corrects): the warm start begins closer to the
equilibrium than the prior, changes less in its first year, and needs no more
years."""
import numpy as np
import pytest


@pytest.mark.gpu
def test_warm_start_beats_cold_start():
    from cmf import maps, synth
    from cmf import surrogate as sg
    from cmf.spinup import make_climatology
    basins = sg.basins_from_networks(1, 60, seed=11, nx=360, ny=180, min_cells=10)
    X, Y, res = sg.equilibrium_labels(basins, tol=1e-5, max_years=30,
                                      log=lambda s: print("labels:", s, flush=True))
    print(f"training cells {X.shape[0]}; label years {np.bincount(res.years).tolist()}", flush=True)
    Xs, Ys, ws = sg.training_set(X, Y, res.state["rivsto"] + res.state["fldsto"])
    sur, hist = sg.train(Xs, Ys, epochs=60, batch=512, lr=2e-3, seed=0, weights=ws)
    print(f"loss {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    assert hist[-1] < 0.2 * hist[0]
    net = synth.make_network(360, 180, seed=12)
    m = maps.build_map(net, maps.Options())
    clim = make_climatology(net, m, 12, mean_mm_day=1.5, amp=0.5, phase=3.0)
    r = sg.compare_spinup(sur, sg.Basin(net, m, clim), years=12, log=lambda s: print("spin:", s, flush=True))
    c, p, w = r["cold"], r["prior"], r["warm"]
    print(f"cold {c['years']} y, init err {c['init_rel_err']:.3e}; prior {p['years']} y, "
          f"init err {p['init_rel_err']:.3e}; warm {w['years']} y, init err {w['init_rel_err']:.3e}", flush=True)
    assert p["init_rel_err"] < 0.05 * c["init_rel_err"]
    assert w["init_rel_err"] < 0.05 * c["init_rel_err"]
    assert w["annual_change"][0] < 0.05 * c["annual_change"][0]
    assert w["years"] > 0 and p["years"] > 0
    if c["years"]:
        assert p["years"] <= c["years"] and w["years"] <= c["years"]
