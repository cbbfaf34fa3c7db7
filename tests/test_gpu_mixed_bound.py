"""What bounds the mixed path's error with bifurcations, and why.

The glb_15min-shaped 1440x720 network with 6e4 bifurcation pathways (5 levels
each), daily DT with adaptive substeps, 30 days from the initial stage, run three
ways on the device: FP64, FP64 with the initial river storage multiplied by
(1 + 1e-12), and mixed precision (FP32 flux/stage arithmetic, FP64 storages and
sums) with its per-substep water-budget sums.

Checked:
* the mixed path conserves water: the sum over every substep of the reference's
  budget error DERROR = -(STOPRE - STONXT + RIVINF - RIVOUT)
  (src/cmf_ctrl_physics_mod.F90:238) stays below 1e-12 of the total storage, so
  the storage update loses nothing that compensated summation could recover;
* the model itself amplifies small differences: the FP64 run perturbed by 1e-12
  differs from the unperturbed one by more than 1e-8 of the largest cell storage
  within the first days (switches such as DFLW_im > 1.E-5 of the path and
  floodplain flows, src/cmf_calc_pthout_mod.F90:60-63, flip on different
  substeps and the cells downstream inherit the difference);
* under that amplification the mixed path's FP32 rounding (~6e-8) gives its
  error: total storage within 1e-4 on every day, within 1e-5 after 30 days, the
  cells within 1e-2 of the largest cell storage.
Measured on MI355X (365 days, same network): total storage error 3.5e-5 at day 3
(the cold start's first wetting), 2e-6 at day 30, 1.8e-7 at day 365; the
perturbed FP64 run reaches 9e-7 of the largest cell on day 1.
"""
import numpy as np
import pytest

SEED = 2026
DAYS = 30


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_mixed_error_is_the_models_amplification():
    import sys
    import os
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    from cmf import maps, synth
    from cmf.model import CMFModel
    net = synth.make_network(1440, 720, seed=SEED)
    synth.add_bifurcations(net, 60000, 5, SEED)
    ro = synth.make_runoff(net, 32, SEED)
    m = maps.build_map(net, maps.Options(lpthout=True))
    st = bench.initial_state(net, m)
    print(f"C2 network: {m.n} cells, {m.npth} paths", flush=True)
    runs = {"fp64": CMFModel(m, "fp64"), "fp64p": CMFModel(m, "fp64"), "mixed": CMFModel(m, "mixed")}
    for name, g in runs.items():
        s = dict(st)
        if name == "fp64p":
            s["rivsto"] = st["rivsto"] * (1.0 + 1e-12)
        g.set_state(s)
    miss = 0.0
    worst_total = 0.0
    amp = 0.0
    for d in range(DAYS):
        r = m.vec(ro[d % ro.shape[0]])
        for name, g in runs.items():
            g.set_runoff(r)
            if name == "mixed":
                nt, glb = g.physics_advance(86400.0, budget=True)
                miss += float(np.sum(-(glb[:, 0] - glb[:, 1] + glb[:, 3] - glb[:, 4])))
            else:
                g.physics_advance(86400.0)
        a = runs["fp64"].get_state(["rivsto", "fldsto"])
        sa = a["rivsto"] + a["fldsto"]
        b = runs["fp64p"].get_state(["rivsto", "fldsto"])
        amp = max(amp, float(np.abs(b["rivsto"] + b["fldsto"] - sa).max() / np.abs(sa).max()))
        c = runs["mixed"].get_state(["rivsto", "fldsto"])
        sc = c["rivsto"] + c["fldsto"]
        total = abs(sc.sum() - sa.sum()) / sa.sum()
        cell = float(np.abs(sc - sa).max() / np.abs(sa).max())
        worst_total = max(worst_total, total)
        if d < 3 or (d + 1) % 10 == 0:
            print(f"  day {d + 1}: mixed total {total:.2e} cell {cell:.2e}; perturbed FP64 cell {amp:.2e}; "
                  f"mixed mass residual {miss / sa.sum():.2e}", flush=True)
        assert cell <= 1e-2
    for g in runs.values():
        g.close()
    assert abs(miss) / sa.sum() <= 1e-12
    assert amp >= 1e-8
    assert worst_total <= 1e-4
    assert total <= 1e-5
