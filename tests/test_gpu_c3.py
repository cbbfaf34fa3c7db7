"""Parity at the benchmark's own size: the glb_06min-shaped 3600x1800 network
(1,500,120 cells, levees on 3% of cells, 2500 dams under the Yamazaki-Funato
rule, hourly gamma runoff, DT = 3600 s with adaptive substeps), the workload whose
cell-substeps/s bench.py reports.

The device (FP64, through the C-ABI, with the XCD-contiguous tile order and
mixed block sizes the autotuner picks among) against the oracle on the same
inputs for three hourly outer steps from the initial stage: identical NT at
every step, every state field within 1e-10 of its magnitude, and the budget
sums of every substep within 1e-12 (the bound and its reason are those of
test_gpu_budget.py). At this size every tile edge, every XCD range boundary,
the depth-first permutation of 1.5e6 cells and the compacted levee and dam
tables are exercised, which the small parity cases cannot show.
"""
import time

import numpy as np
import pytest

from cases import STATE_KEYS, rel_err, upload_keys
from test_gpu_budget import budget_err

NX, NY, SEED = 3600, 1800, 2026
STEPS = 3
TOL = 1e-10
TOL_GLB = 1e-12


def _workload():
    from cmf import maps, synth
    from oracle.oracle import OracleModel
    net = synth.make_network(NX, NY, seed=SEED)
    synth.add_levees(net, 0.03, SEED)
    synth.add_dams(net, 2500, SEED)
    ro = synth.make_runoff(net, STEPS, SEED)
    m = maps.build_map(net, maps.Options(llevee=True, ldamout=True))
    om = OracleModel(m, maps.cold_start(m))
    om.physics_fldstg()
    st = {k: v.copy() for k, v in om.st.items()}
    maps.init_dams(m, net, st)
    return net, m, st, [m.vec(ro[k]) for k in range(STEPS)]


@pytest.mark.gpu
@pytest.mark.timeout(300)
def test_c3_fp64_matches_oracle():
    from cmf.model import CMFModel
    from oracle.oracle import OracleModel
    t0 = time.time()
    net, m, st, rvec = _workload()
    print(f"C3 workload: {m.n} cells, {int(np.count_nonzero(m.levhgt > 0))} levee cells, {m.ndam} dams, "
          f"built in {time.time() - t0:.1f}s", flush=True)
    om = OracleModel(m, st)
    gm = CMFModel(m, "fp64")
    gm.set_tune(bs_outflow=128, bs_rate=256, bs_update=256, xcd=1)
    gm.set_state(upload_keys(st))
    extra = ["levsto", "levdph", "damsto", "daminf"]
    for step in range(STEPS):
        om.st["runoff"][:] = rvec[step]
        gm.set_runoff(rvec[step])
        t1 = time.time()
        nt_o, g_o = om.physics_advance(3600.0, budget=True)
        t2 = time.time()
        nt_g, g_g = gm.physics_advance(3600.0, budget=True)
        g = gm.get_state(STATE_KEYS + extra)
        e = {k: rel_err(g[k], om.st[k]) for k in STATE_KEYS + extra}
        worst = max(e, key=e.get)
        eg = budget_err(g_g, g_o) if g_o.shape == g_g.shape else float("inf")
        flooded = int(np.count_nonzero(om.st["fldsto"] > 0.0))
        print(f"  step {step + 1}: NT oracle {nt_o} device {nt_g} (oracle {t2 - t1:.1f}s); worst rel err "
              f"{worst} {e[worst]:.3e}; budget {eg:.3e}; flooded cells {flooded}", flush=True)
        assert nt_o == nt_g
        for k, v in e.items():
            assert v <= TOL, f"step {step + 1}: {k} rel err {v:.3e}"
        assert eg <= TOL_GLB, f"step {step + 1}: budget rel err {eg:.3e}"
    gm.close()
