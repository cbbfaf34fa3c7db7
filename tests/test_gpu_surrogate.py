"""What the learned warm start gains over the plain cold start.

Equilibrium storages of river basins cut from a small glb_15min-shaped network
(one runoff climatology per basin) come from the FP64 device physics (batched
spin-up); the network is trained on them by the project's HIP kernels and applied
to a held-out network. The held-out network is spun up from the cold start
(STORAGE_SEA_SURFACE) and from the warm start as two members of one handle; the
cold run's last year is the equilibrium both are judged against (the cold start
stays the fallback and the judge of drift). Checked: the warm start begins closer
to the equilibrium, changes less in its first year, and needs no more years to
reach an annual storage change below 0.1%."""
import numpy as np
import pytest


@pytest.mark.gpu
def test_warm_start_beats_cold_start():
    from cmf import maps, synth
    from cmf import surrogate as sg
    from cmf.spinup import make_climatology
    basins = sg.basins_from_networks(1, 60, seed=11, nx=360, ny=180, min_cells=10)
    X, Y, res = sg.equilibrium_labels(basins, tol=1e-4, max_years=30,
                                      log=lambda s: print("labels:", s, flush=True))
    print(f"training cells {X.shape[0]}; cold-start years {np.bincount(res.years).tolist()}", flush=True)
    sur, hist = sg.train(X, Y, epochs=60, batch=512, lr=2e-3, seed=0, w_over=16.0)
    print(f"loss {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    assert hist[-1] < 0.2 * hist[0]
    net = synth.make_network(360, 180, seed=12)
    m = maps.build_map(net, maps.Options())
    clim = make_climatology(net, m, 12, mean_mm_day=1.5, amp=0.5, phase=3.0)
    r = sg.compare_spinup(sur, sg.Basin(net, m, clim), years=12, log=lambda s: print("spin:", s, flush=True))
    c, w = r["cold"], r["warm"]
    print(f"cold {c['years']} y, init err {c['init_rel_err']:.3f}; warm {w['years']} y, "
          f"init err {w['init_rel_err']:.3f}", flush=True)
    assert w["init_rel_err"] < 0.5 * c["init_rel_err"]
    assert w["annual_change"][0] < 0.5 * c["annual_change"][0]
    if c["years"]:
        assert 0 < w["years"] <= c["years"]
