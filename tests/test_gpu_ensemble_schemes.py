"""Batched ensembles with the kinematic (LKINE) and slope-mixed (LSLPMIX) outflow
schemes: every instance (domain, member) of one ensemble handle must reproduce,
bit for bit, a single-model handle run on that domain with that member's runoff,
as tests/test_gpu_ensemble.py checks for the default scheme. The single-model
kinematic and slope-mixed paths are pinned against the reference by the golden
tests (g_kine, g_slpmix). Paths, levees and dams on both domains; FP64 and mixed."""
import numpy as np
import pytest

from cases import STATE_KEYS, prepare_domain, upload_keys

DOMAINS = [(60, 30, 31), (90, 45, 32)]
MEMBERS = 4
NDAYS = 2
PDSTMTH = 4.0e5
EXTRA = ["levsto", "levdph", "damsto", "daminf", "pthflw", "pthflwsum", "pthflw_avg", "pthflwsum_avg",
         "daminf_avg"]


def _domains(slpmix):
    from cmf import maps, synth
    out = []
    for nx, ny, seed in DOMAINS:
        m, net, st = prepare_domain(nx, ny, seed, pdstmth=PDSTMTH)
        if slpmix:
            synth.add_slope_mask(net, 0.3)
            maps.apply_slope_mask(m, net.slpmix_mask)
        ro = synth.make_runoff(net, NDAYS, seed, mean_mm_day=20.0)
        out.append((m, st, [m.vec(ro[d]) for d in range(NDAYS)]))
    return out


def _scheme(g, scheme):
    if scheme == "kine":
        g.set_kinematic(True)
    else:
        g.set_slopemix(True)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("scheme", ["kine", "slpmix"])
def test_ensemble_scheme_matches_single_runs(scheme, precision):
    from cmf import maps
    from cmf.model import CMFModel
    doms = _domains(scheme == "slpmix")
    rng = np.random.default_rng(7)
    fac = np.exp(rng.normal(0.0, 1.0, size=(len(doms), MEMBERS)))
    cm, dom_cells = maps.concat_maps([d[0] for d in doms])
    print(f"ensemble {scheme} {precision}: domains {dom_cells}, members {MEMBERS}", flush=True)
    em = CMFModel(cm, precision, members=MEMBERS, dom_cells=dom_cells)
    _scheme(em, scheme)
    init = {}
    for k in upload_keys(doms[0][1]):
        parts = [d[1][k] for d in doms]
        cat = np.concatenate(parts, axis=1) if (parts[0].ndim == 2 and k.startswith("pthflw")) else np.concatenate(parts)
        init[k] = np.stack([cat] * MEMBERS)
    em.set_state(init)
    singles = []
    for m, st, rv in doms:
        row = []
        for _ in range(MEMBERS):
            gm = CMFModel(m, precision)
            _scheme(gm, scheme)
            gm.set_state(upload_keys(st))
            row.append(gm)
        singles.append(row)
    off = np.cumsum([0] + dom_cells)
    poff = np.cumsum([0] + [d[0].npth for d in doms])
    for day in range(NDAYS):
        em.set_runoff_scaled(np.concatenate([d[2][day] for d in doms]), fac.reshape(-1))
        nt_e, _ = em.ensemble_advance(86400.0)
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
                    a = a[..., poff[di]:poff[di + 1]] if key.startswith("pthflw") else a[off[di]:off[di + 1]]
                    assert np.array_equal(a, gs[key]), f"day {day + 1} dom {di} member {mi}: {key} differs"
        print(f"  day {day + 1}: NT per instance {nt_e.tolist()}", flush=True)
    em.close()
    for row in singles:
        for gm in row:
            gm.close()
