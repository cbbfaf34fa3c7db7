"""Batched ensembles: every instance (domain, member) of one ensemble handle
must reproduce, bit for bit, a single-model handle run on that domain with that
member's runoff. Both share the per-slot arithmetic; the ensemble adds the
member-minor layout, domain concatenation, per-instance adaptive time steps
(instances with fewer substeps idle while others continue) and the scaled
forcing. The single-model handle itself is pinned against the oracle by
test_gpu_parity.py."""
import numpy as np
import pytest

from cases import STATE_KEYS, prepare_domain, upload_keys

DOMAINS = [(60, 30, 21), (90, 45, 22), (40, 20, 23)]   # uneven sizes
MEMBERS = 4
NDAYS = 3
# a long mouth distance (PDSTMTH) keeps the bank-full river mouths, whose depth
# is held by sea level, from setting every instance's CFL bound: interior cells
# then set it, and members with more runoff need more substeps
PDSTMTH = 4.0e5
EXTRA = ["levsto", "levdph", "damsto", "daminf", "pthflw", "pthflwsum", "pthflw_avg",
         "pthflwsum_avg", "daminf_avg"]


def _factors(nd, nm, seed=5):
    rng = np.random.default_rng(seed)
    return np.exp(rng.normal(0.0, 1.0, size=(nd, nm)))


def _domains():
    from cmf import synth
    out = []
    for nx, ny, seed in DOMAINS:
        m, net, st = prepare_domain(nx, ny, seed, pdstmth=PDSTMTH)
        ro = synth.make_runoff(net, NDAYS, seed, mean_mm_day=20.0)
        out.append((m, st, [m.vec(ro[d]) for d in range(NDAYS)]))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ensemble_matches_single_runs(precision):
    from cmf import maps
    from cmf.model import CMFModel
    doms = _domains()
    fac = _factors(len(doms), MEMBERS)
    cm, dom_cells = maps.concat_maps([d[0] for d in doms])
    print(f"ensemble {precision}: domains {dom_cells}, members {MEMBERS}", flush=True)
    em = CMFModel(cm, precision, members=MEMBERS, dom_cells=dom_cells)
    init = {}
    for k in upload_keys(doms[0][1]):
        parts = [d[1][k] for d in doms]
        if parts[0].ndim == 2 and k.startswith("pthflw"):
            cat = np.concatenate(parts, axis=1)                 # [lev, paths]
        else:
            cat = np.concatenate(parts)
        init[k] = np.stack([cat] * MEMBERS)
    em.set_state(init)
    singles = []
    for di, (m, st, rv) in enumerate(doms):
        row = []
        for mi in range(MEMBERS):
            gm = CMFModel(m, precision)
            gm.set_state(upload_keys(st))
            row.append(gm)
        singles.append(row)
    nts = set()
    off = np.cumsum([0] + dom_cells)
    poff = np.cumsum([0] + [d[0].npth for d in doms])
    for day in range(NDAYS):
        base = np.concatenate([d[2][day] for d in doms])
        em.set_runoff_scaled(base, fac.reshape(-1))
        nt_e, dt_e = em.ensemble_advance(86400.0)
        ge = em.get_state()
        for di, (m, st, rv) in enumerate(doms):
            for mi in range(MEMBERS):
                gm = singles[di][mi]
                gm.set_runoff(rv[day] * fac[di, mi])
                nt_s = gm.physics_advance(86400.0)
                k = di * MEMBERS + mi
                assert nt_e[k] == nt_s, f"instance {k}: NT {nt_e[k]} vs {nt_s}"
                gs = gm.get_state()
                for key in STATE_KEYS + [x for x in EXTRA if x in gs]:
                    a = ge[key][mi]
                    if key.startswith("pthflw"):
                        a = a[..., poff[di]:poff[di + 1]]
                    else:
                        a = a[off[di]:off[di + 1]]
                    assert np.array_equal(a, gs[key]), \
                        f"day {day + 1} dom {di} member {mi}: {key} differs (max {np.max(np.abs(a - gs[key])):.3e})"
        print(f"  day {day + 1}: NT per instance {nt_e.tolist()}", flush=True)
        nts.update(nt_e.tolist())
    assert len(nts) > 1, "instances should need different substep counts"
    em.close()
    for row in singles:
        for gm in row:
            gm.close()


SAMPLED = [(0, 0), (0, 63), (1, 17), (1, 40), (2, 5), (2, 62)]   # (domain, member)


@pytest.mark.gpu
def test_64_member_instances_match_oracle():
    """A 64-member ensemble (3 uneven domains with paths, levees and dams; every
    member from the same initial state, uploaded once for all members; lognormal
    member factors): sampled instances against the oracle run on that domain with
    that member's runoff, at the single-model bound (identical NT, every field
    within 1e-10 of its magnitude) after each of three days."""
    from cmf import maps
    from cmf.model import CMFModel
    from oracle.oracle import OracleModel
    members = 64
    doms = _domains()
    rng = np.random.default_rng(11)
    fac = np.exp(rng.normal(0.0, 0.6, size=(NDAYS, len(doms), members)))
    cm, dom_cells = maps.concat_maps([d[0] for d in doms])
    em = CMFModel(cm, "fp64", members=members, dom_cells=dom_cells)
    init = {}
    for k in upload_keys(doms[0][1]):
        parts = [d[1][k] for d in doms]
        init[k] = np.concatenate(parts, axis=1) if (parts[0].ndim == 2 and k.startswith("pthflw")) else \
            np.concatenate(parts)
    em.set_state(init, shared=True)
    oracles = {}
    for di, mi in SAMPLED:
        m, st, rv = doms[di]
        oracles[(di, mi)] = OracleModel(m, {k: v.copy() for k, v in st.items()})
    off = np.cumsum([0] + dom_cells)
    poff = np.cumsum([0] + [d[0].npth for d in doms])
    for day in range(NDAYS):
        base = np.concatenate([d[2][day] for d in doms])
        em.set_runoff_scaled(base, fac[day].reshape(-1))
        nt_e, _ = em.ensemble_advance(86400.0)
        ge = em.get_state()
        worst = 0.0
        for (di, mi), om in oracles.items():
            m, st, rv = doms[di]
            om.st["runoff"][:] = rv[day] * fac[day, di, mi]
            nt_o = om.physics_advance(86400.0)
            k = di * members + mi
            assert nt_e[k] == nt_o, f"day {day + 1} instance {k}: NT {nt_e[k]} vs oracle {nt_o}"
            for key in STATE_KEYS + ["levsto", "levdph", "damsto", "pthflw", "pthflwsum"]:
                a = ge[key][mi]
                a = a[..., poff[di]:poff[di + 1]] if key.startswith("pthflw") else a[off[di]:off[di + 1]]
                b = om.st[key]
                e = float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-30))
                worst = max(worst, e)
                assert e <= 1e-10, f"day {day + 1} dom {di} member {mi}: {key} rel err {e:.3e}"
        print(f"  day {day + 1}: NT range {nt_e.min()}..{nt_e.max()}, sampled worst rel err {worst:.3e}", flush=True)
    em.close()
