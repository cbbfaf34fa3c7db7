"""The batched ensemble at its benchmark size: 16 regional domains of 2e4 .. 4e5
river cells (5.66e6 in total) x 64 runoff members = 3.6e8 member-cells on one
handle (FP64), one daily outer step from the cold start, every instance with its
own adaptive step.

* Mass balance of every one of the 1024 instances: the change of its total
  storage equals its runoff minus its outflow at the river mouths, both taken
  from the instance's substep averages times its NADD (the sum of its substep
  lengths), within 1e-12 of the instance's storage plus runoff. Water crossing
  between members or domains, or a slot updated twice or not at all, breaks this
  by a cell's share of the volume, orders of magnitude above the bound (FP64
  rounding of the storage updates and sums).
* Two members of the smallest and of the largest domain against the oracle run
  on that domain alone (identical NT, fields within 1e-10).
"""
import time

import numpy as np
import pytest

MEMBERS = 64
SEED = 2026
SIGMA = 0.3


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_c4_full_size_mass_balance_and_oracle():
    from cmf import maps, synth
    from cmf.model import CMFModel
    from oracle.oracle import OracleModel
    t0 = time.time()
    nets = synth.regional_networks(synth.ensemble_domain_sizes(16), SEED)
    ms = [maps.build_map(net, maps.Options()) for net in nets]
    cm, dom_cells = maps.concat_maps(ms)
    nd = len(ms)
    print(f"C4: {nd} domains, {cm.n} cells x {MEMBERS} members = {cm.n * MEMBERS} member-cells "
          f"({time.time() - t0:.1f}s)", flush=True)
    st0 = [maps.cold_start(m) for m in ms]
    init = {k: np.concatenate([s[k] for s in st0])
            for k in ("rivsto", "fldsto", "rivout", "fldout", "rivdph_pre", "fldsto_pre")}
    em = CMFModel(cm, "fp64", members=MEMBERS, dom_cells=dom_cells)
    em.set_state(init, shared=True)
    em.physics_fldstg()
    ro = [synth.make_runoff(net, 1, SEED + 7 * d)[0] for d, net in enumerate(nets)]
    base = np.concatenate([m.vec(r) for m, r in zip(ms, ro)])
    fac = np.exp(np.random.default_rng(SEED).normal(-0.5 * SIGMA ** 2, SIGMA, size=nd * MEMBERS))
    em.set_runoff_scaled(base, fac)
    s0 = em.get_state(["rivsto", "fldsto"])
    sto0 = s0["rivsto"] + s0["fldsto"]
    del s0
    t1 = time.time()
    nt, dt = em.ensemble_advance(86400.0)
    em.synchronize()
    print(f"  one day: NT {nt.min()}..{nt.max()} ({time.time() - t1:.1f}s)", flush=True)
    g = em.get_state(["rivsto", "fldsto", "runoff_avg", "outflw_avg"])
    em.close()
    sto1 = g["rivsto"] + g["fldsto"]
    off = np.cumsum([0] + dom_cells)
    worst = 0.0
    for d in range(nd):
        a, b = off[d], off[d + 1]
        mouth = np.flatnonzero(ms[d].next < 0) + a
        for mi in range(MEMBERS):
            k = d * MEMBERS + mi
            nadd = 0.0
            for _ in range(int(nt[k])):                 # diag:108, in substep order
                nadd = nadd + float(dt[k])
            ds = sto1[mi, a:b].sum() - sto0[mi, a:b].sum()
            rin = g["runoff_avg"][mi, a:b].sum() * nadd
            out = g["outflw_avg"][mi, mouth].sum() * nadd
            e = abs(ds - (rin - out)) / (sto0[mi, a:b].sum() + rin)
            worst = max(worst, e)
            assert e <= 1e-12, f"domain {d} member {mi}: mass balance rel err {e:.3e}"
    print(f"  mass balance of {nd * MEMBERS} instances: worst rel err {worst:.3e}", flush=True)
    # the oracle on the smallest and the largest domain, two members each
    for d in (0, nd - 1):
        m = ms[d]
        a, b = off[d], off[d + 1]
        for mi in (0, MEMBERS - 1):
            k = d * MEMBERS + mi
            om = OracleModel(m, {kk: v.copy() for kk, v in st0[d].items()})
            om.physics_fldstg()
            om.st["runoff"][:] = base[a:b] * fac[k]
            nt_o = om.physics_advance(86400.0)
            assert nt_o == nt[k], f"domain {d} member {mi}: NT {nt[k]} vs oracle {nt_o}"
            for key in ("rivsto", "fldsto", "runoff_avg", "outflw_avg"):
                got = g[key][mi, a:b]
                e = float(np.max(np.abs(got - om.st[key])) / max(float(np.max(np.abs(om.st[key]))), 1e-30))
                assert e <= 1e-10, f"domain {d} member {mi}: {key} rel err {e:.3e}"
            print(f"  domain {d} ({m.n} cells) member {mi}: NT {nt_o}, matches the oracle", flush=True)
