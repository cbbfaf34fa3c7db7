#!/usr/bin/env python3
"""Benchmark of the MI355X CaMa-Flood substep physics.

Workload (one GPU): glb_06min-shaped synthetic network, 3600x1800 grid, ~1.5e6
river cells, levees on 3% of cells (crown 2-8 m, distance fraction U(0.2,0.8)),
2500 dams (capacity lognormal(ln 1e8 m3, 1.5), Yamazaki-Funato rule), hourly
gamma(k=0.5, mean 1.5 mm/day) runoff, DT = 3600 s with adaptive substeps.

A step is one outer step CMF_PHYSICS_ADVANCE (stage, CFL, NT substeps) over all
cells with inputs resident in HBM. The metric is cell-substeps/s = N * sum(NT) / t.
Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "cama-flood_v4_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per cell-substep by kernel: the 508 B (FP64) and 276 B (mixed)
# of the per-item byte count of one substep, each item charged to the kernel that moves it
BYTES_FP64 = {"outflow": 180.0, "rate": 44.0, "update": 284.0}
BYTES_MIXED = {"outflow": 96.0, "rate": 24.0, "update": 156.0}
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "traffic_c3.json")
# launch configuration of the single-model C3 runs: the one profiles/traffic_c3.* was
# measured with (the autotuner's usual choice on this workload), so roofline.traffic
# describes the launches that are timed
PINNED_TUNE = '{"bs_outflow":64,"bs_rate":256,"bs_update":128,"xcd":1}'


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_workload(nx, ny, seed, nrec):
    from cmf import maps, synth
    t0 = time.time()
    net = synth.make_network(nx, ny, seed=seed)
    synth.add_levees(net, 0.03, seed)
    synth.add_dams(net, 2500, seed)
    ro = synth.make_runoff(net, nrec, seed)
    log(f"network {nx}x{ny}: {net.ncells} cells, {time.time() - t0:.1f}s")
    opts = maps.Options(llevee=True, ldamout=True)
    m = maps.build_map(net, opts)
    log(f"map init {time.time() - t0:.1f}s")
    return net, m, ro


def device_model(net, m, precision):
    import dataclasses
    from cmf import maps
    from cmf.model import CMFModel
    st = maps.cold_start(m)
    # the initial stage precedes CMF_DAMOUT_INIT (drv_control:243-252): no dam tables yet
    gm0 = CMFModel(m, precision, opts=dataclasses.replace(m.opts, ldamout=False))
    gm0.set_state(st)
    gm0.physics_fldstg()                       # initial CMF_PHYSICS_FLDSTG on the device
    st.update(gm0.get_state(["rivsto", "fldsto", "rivdph", "flddph", "fldfrc", "fldare",
                             "sfcelv", "storge", "levsto", "levdph"]))
    gm0.close()
    maps.init_dams(m, net, st)                 # CMF_DAMOUT_INIT (host, after the stage)
    gm = CMFModel(m, precision)
    gm.set_state(st)
    return gm


def run_gpu(net, m, rvec, precision, steps, warmup, dt, tune="auto"):
    gm = device_model(net, m, precision)
    rec = 0
    for _ in range(warmup):
        gm.set_runoff(rvec[rec % len(rvec)]); rec += 1
        gm.physics_advance(dt)
    gm.synchronize()
    if tune == "auto":
        gm.set_runoff(rvec[rec % len(rvec)])
        best, trials = gm.autotune(dt, nsub=12)       # restores the state it times on
        for r in trials:
            if r[0] >= 2:
                log(f"  trial graph={int(r[0]) - 2}: outer step of 12 substeps {r[7]:.3f} ms wall")
                continue
            log(f"  trial fused={int(r[0])} bs={int(r[1])} xcd={int(r[2])} grid={int(r[3])}: "
                f"outflow {r[4]:.3f} rate {r[5]:.3f} update {r[6]:.3f} total {r[7]:.3f} ms / 12 substeps")
        log(f"{precision}: autotuned launch configuration {best}")
    elif tune:
        gm.set_tune(**tune)
    tune_used = gm.get_tune()
    # timed region: runoff for each step is uploaded before the clock starts
    sum_nt = 0
    t_total = 0.0
    for s in range(steps):
        gm.set_runoff(rvec[rec % len(rvec)]); rec += 1
        gm.synchronize()
        t0 = time.perf_counter()
        nt = gm.physics_advance(dt)
        gm.synchronize()
        t_total += time.perf_counter() - t0
        sum_nt += nt
    # kernel timing pass (HIP events on the model's stream)
    gm.kernel_times(True)
    gm.set_runoff(rvec[rec % len(rvec)])
    gm.physics_advance(dt)
    kt = gm.kernel_times(False)
    gm.close()
    return sum_nt, t_total, kt, tune_used


def _reference_case(d, net, ro, dt, steps, llevee=True, ldamout=True, inpmat=None):
    """Reference inputs of the workload (levees, dams unless switched off) in
    directory d, the run shortened to `steps` outer steps; with inpmat, LINTERP
    forcing on an input grid (ro then holds its records)."""
    from cmf import synth
    days = max(1, int(np.ceil(steps * dt / 86400.0)))
    synth.write_reference_inputs(net, d, ro, dt=dt, ifrq_inp=int(dt // 3600), ndays=days,
                                 llevee=llevee, ldamout=ldamout, inpmat=inpmat)
    # end time = start + steps*dt
    nml = open(os.path.join(d, "input_cmf.nam")).read()
    import datetime as _dt
    t1 = _dt.datetime(2001, 1, 1) + _dt.timedelta(seconds=steps * dt)
    nml = nml.replace(nml[nml.index("EYEAR"):nml.index("EHOUR")],
                      f"EYEAR = {t1.year}\nEMON  = {t1.month}\nEDAY  = {t1.day}\n")
    nml = nml.replace("EHOUR = 0", f"EHOUR = {t1.hour}")
    open(os.path.join(d, "input_cmf.nam"), "w").write(nml)


def _big_stack():
    """the reference keeps NX*NY work arrays on the stack (its run scripts: ulimit -s unlimited)"""
    import resource
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def _substeps(log_path):
    """NT per outer step from a CaMa-Flood log (LOGNAM): one water-budget line
    (' flx: ') per substep, steps closed by the 'DRV_ADVANCE END' line
    (src/cmf_ctrl_physics_mod.F90:241-246, src/cmf_drv_advance_mod.F90:165)."""
    nts, cur = [], 0
    with open(log_path, errors="replace") as f:
        for line in f:
            if " flx: " in line:
                cur += 1
            elif "DRV_ADVANCE END" in line:
                nts.append(cur)
                cur = 0
    return nts


def _run_host(exe, d, env):
    """Run a reference-host binary (reference build or drop-in) in d as a child
    process; returns (stdout, per-step seconds, substeps per step) or None."""
    r = subprocess.run([exe, "runoff.bin", "NONE", "0"], cwd=d, env=env, capture_output=True,
                       text=True, timeout=900, preexec_fn=_big_stack)
    if r.returncode != 0:
        log(f"{os.path.basename(exe)} run failed (rc={r.returncode}): {r.stdout[-800:]} {r.stderr[-800:]}")
        try:
            log(open(os.path.join(d, "log_CaMa.txt")).read()[-1500:])
        except OSError:
            pass
        return None
    steps = [float(l.split()[2]) for l in r.stdout.splitlines() if l.startswith("STEPTIME")]
    return r.stdout, steps, _substeps(os.path.join(d, "log_CaMa.txt"))


def cpu_baseline(net, ro, n_cells, dt, sample_steps, variant="omp64"):
    """Reference OpenMP build (omp64: FP64; omp32: the reference's default
    single-precision mode) timed on this host over a bounded sample."""
    exe = os.path.join(ROOT, "oracle", "_ref", variant, "cmf_ref")
    if not os.path.exists(exe):
        return None
    ncores = len(os.sched_getaffinity(0))
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        _reference_case(d, net, ro, dt, sample_steps)
        env = dict(os.environ, OMP_NUM_THREADS=str(ncores), OMP_PROC_BIND="spread", OMP_PLACES="cores",
                   OMP_STACKSIZE="512M")
        res = _run_host(exe, d, env)
        if res is None:
            return None
        out, _steps, nts = res
        line = [l for l in out.splitlines() if l.startswith("REFTIME")][0].split()
        t, nsteps, n = float(line[1]), int(line[2]), int(line[3])
        cs = n * sum(nts)
        return dict(value=cs / t, unit="cell-substeps/s", cores=ncores, kind="reference",
                    sample=f"reference {'FP64' if variant == 'omp64' else 'single-precision'} OpenMP "
                           f"(flang -O3 -fopenmp), {nsteps} outer steps of "
                           f"{dt:.0f} s, {sum(nts)} substeps over {n} cells, {t:.1f} s")


def ensemble_cpu_baseline(nets, seed, steps=2):
    """The reference's OpenMP build on the largest ensemble domain for `steps`
    daily steps. The reference has no batched mode: an ensemble runs its
    instances one after another, so its member-cell-substeps/s is its
    cell-substeps/s on a domain."""
    exe = os.path.join(ROOT, "oracle", "_ref", "omp64", "cmf_ref")
    if not os.path.exists(exe):
        return None
    from cmf import synth
    net = max(nets, key=lambda x: x.ncells)
    ro = synth.make_runoff(net, steps, seed)
    ncores = len(os.sched_getaffinity(0))
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        _reference_case(d, net, ro, 86400.0, steps, llevee=False, ldamout=False)
        env = dict(os.environ, OMP_NUM_THREADS=str(ncores), OMP_PROC_BIND="spread", OMP_PLACES="cores",
                   OMP_STACKSIZE="512M")
        res = _run_host(exe, d, env)
        if res is None:
            return None
        _out, times, nts = res
    cs = net.ncells * sum(nts)
    return dict(value=cs / sum(times), unit="member-cell-substeps/s", cores=ncores, kind="reference",
                sample=f"reference FP64 OpenMP (flang -O3 -fopenmp) on the largest domain, {net.ncells} cells, "
                       f"{len(times)} daily steps, {sum(nts)} substeps, {sum(times):.1f} s; instances run one "
                       f"after another in the reference")


def spinup_cpu_baseline(net, m, clim, days=30):
    """The reference's OpenMP build on the spin-up network under the first
    month's climatological runoff for `days` daily steps, from its cold start;
    reported as seconds per simulated year (the sample scaled by 365 / days)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "omp64", "cmf_ref")
    if not os.path.exists(exe):
        return None
    grid = np.zeros((1, net.ny, net.nx))
    grid[0, m.seqy - 1, m.seqx - 1] = clim.monthly[0]
    ncores = len(os.sched_getaffinity(0))
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        _reference_case(d, net, grid, 86400.0, days, llevee=False, ldamout=False)
        env = dict(os.environ, OMP_NUM_THREADS=str(ncores), OMP_PROC_BIND="spread", OMP_PLACES="cores",
                   OMP_STACKSIZE="512M")
        res = _run_host(exe, d, env)
        if res is None:
            return None
        _out, times, nts = res
    return dict(value=sum(times) * 365.0 / len(times), unit="s per simulated year", cores=ncores,
                kind="reference",
                sample=f"reference FP64 OpenMP (flang -O3 -fopenmp), {m.n} cells, {len(times)} daily steps "
                       f"({sum(nts)} substeps) from the cold start under the first month's runoff, "
                       f"{sum(times):.1f} s, scaled to 365 days")


def dropin_section(net, ro, n_cells, dt, steps, m=None):
    """The Fortran drop-in end to end: the reference's host (serial flang build:
    time loop, forcing, diagnostics, log) with the device physics and the
    device-side output diagnostics (cmf_dropin_devdiag), on the same inputs as the
    CPU baseline, without output or restart files, driven by the package's
    host program (fortran/cmf_dropin_main.F90). Timed per outer step inside
    CMF_DRV_ADVANCE; the first step also uploads the map and the state. Once as
    built (downloads only when the host reads the state) and once with
    CMF_HIP_DOWNLOAD=every (the whole state after every step over PCIe). With the
    map m, also with LINTERP forcing from a 1440x720 (0.25 deg) input grid, the
    forcing put inside the timed step: the reference's CMF_FORCING_PUT on the host
    (then D2RUNOFF goes over PCIe) against CMF_HIP_FORCING_PUT (the input field
    goes over PCIe, the interpolation runs on the device) and against the fields
    kept resident in HBM (CMF_HIP_FORCING=resident: nothing crosses PCIe per step)."""
    exe = os.path.join(ROOT, "cama-flood_v4_amd", "fortran", "_build", "cmf_dropin_devdiag")
    if not os.path.exists(exe):
        return None
    res = {"binary": "cmf_dropin_devdiag", "cells": n_cells, "steps": steps, "dt": dt}
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        _reference_case(d, net, ro, dt, steps)
        for mode in ("due", "every"):
            env = dict(os.environ)
            if mode == "every":
                env["CMF_HIP_DOWNLOAD"] = "every"
            log(f"drop-in run, downloads {mode}")
            r = _run_host(exe, d, env)
            if r is None:
                return None
            _out, st, nts = r
            if len(st) < 3:
                return None
            rest = st[1:]
            res[mode] = dict(first_step_s=round(st[0], 3), step_ms_median=round(1e3 * float(np.median(rest)), 2),
                             cell_substeps_per_s=n_cells * sum(nts[1:]) / sum(rest),
                             substeps=int(sum(nts[1:])))
            log(f"drop-in ({mode}): first step {st[0]:.2f} s, then {1e3 * np.median(rest):.1f} ms per "
                f"{dt:.0f} s step, {res[mode]['cell_substeps_per_s']:.3g} cell-substeps/s")
    if m is None:
        return res
    from cmf import synth
    inpx, inpy, inpa, nxin, nyin = synth.make_inpmat(net, (m.seqx, m.seqy), 0.25)
    grids = synth.make_runoff_grid(nxin, nyin, 4, 2026, mean_mm_day=3.0, missing=0.0)
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        _reference_case(d, net, grids, dt, steps, inpmat=(inpx, inpy, inpa, nxin, nyin, m.seqx, m.seqy))
        for put in ("host", "device", "resident"):
            env = dict(os.environ, CMF_HIP_FORCING=put)
            log(f"drop-in run, LINTERP forcing put on the {put}")
            r = _run_host(exe, d, env)
            if r is None or len(r[1]) < 3:
                return res
            _out, st, nts = r
            rest = st[1:]
            key = f"forcing_{put}"
            res[key] = dict(input_grid=f"{nxin}x{nyin}", inpn=int(inpx.shape[0]),
                            step_ms_median=round(1e3 * float(np.median(rest)), 2),
                            cell_substeps_per_s=n_cells * sum(nts[1:]) / sum(rest), substeps=int(sum(nts[1:])))
            log(f"drop-in (LINTERP, put on the {put}): {1e3 * np.median(rest):.1f} ms per {dt:.0f} s step "
                f"with the put, {res[key]['cell_substeps_per_s']:.3g} cell-substeps/s")
    return res


def initial_state(net, m):
    """Cold start, initial CMF_PHYSICS_FLDSTG on the device (FP64) and dam
    initialisation (drv_control:239-252); host arrays shared by every run."""
    import dataclasses
    from cmf import maps
    from cmf.model import CMFModel
    st = maps.cold_start(m)
    gm0 = CMFModel(m, "fp64", opts=dataclasses.replace(m.opts, ldamout=False))
    gm0.set_state(st)
    gm0.physics_fldstg()
    st.update(gm0.get_state(["rivsto", "fldsto", "rivdph", "flddph", "fldfrc", "fldare", "sfcelv", "storge",
                             "levsto", "levdph"]))
    gm0.close()
    if m.opts.ldamout:
        maps.init_dams(m, net, st)
    return st


def precision_report(seed):
    """Mixed (FP32 flux/stage, FP64 storage) against FP64 on identical inputs for
    the three single-model configurations; relative errors are max|mixed - fp64|
    over max|fp64| per field (cells) and |sum difference| / sum for the total
    storage, the worst over the run."""
    from cmf import maps, synth
    from cmf.model import CMFModel
    cfgs = [
        ("C1 1-degree 360x180, default physics, daily DT, 30 days", 360, 180, 30, 86400.0, dict(), 0),
        ("C2 glb_15min-shaped 1440x720 + 6e4 bifurcation paths x 5 levels, daily DT, 365 days", 1440, 720, 365,
         86400.0, dict(lpthout=True), 60000),
        ("C3 glb_06min-shaped 3600x1800, 3% levees, 2500 dams, hourly DT, 24 steps", 3600, 1800, 24, 3600.0,
         dict(llevee=True, ldamout=True), 0),
    ]
    out = []
    for name, nx, ny, nstep, dt, flags, npth in cfgs:
        t0 = time.time()
        net = synth.make_network(nx, ny, seed=seed)
        if npth:
            synth.add_bifurcations(net, npth, 5, seed)
        if flags.get("llevee"):
            synth.add_levees(net, 0.03, seed)
        if flags.get("ldamout"):
            synth.add_dams(net, 2500, seed)
        nrec = min(nstep, 32)
        ro = synth.make_runoff(net, nrec, seed)
        m = maps.build_map(net, maps.Options(**flags))
        st = initial_state(net, m)
        runs = {p: CMFModel(m, p) for p in ("fp64", "mixed")}
        for g in runs.values():
            g.set_state(st)
        worst = {"storage_cell": 0.0, "storage_total": 0.0, "outflow": 0.0, "depth": 0.0}
        nts = {"fp64": 0, "mixed": 0}
        for k in range(nstep):
            r = m.vec(ro[k % nrec])
            for p, g in runs.items():
                g.set_runoff(r)
                nts[p] += g.physics_advance(dt)
            a = runs["fp64"].get_state(["rivsto", "fldsto", "levsto", "outflw", "rivdph"])
            b = runs["mixed"].get_state(["rivsto", "fldsto", "levsto", "outflw", "rivdph"])
            sa = a["rivsto"] + a["fldsto"] + (a["levsto"] if "levsto" in a else 0.0)
            sb = b["rivsto"] + b["fldsto"] + (b["levsto"] if "levsto" in b else 0.0)
            worst["storage_cell"] = max(worst["storage_cell"], float(np.abs(sb - sa).max() / np.abs(sa).max()))
            worst["storage_total"] = max(worst["storage_total"], float(abs(sb.sum() - sa.sum()) / sa.sum()))
            worst["outflow"] = max(worst["outflow"], float(np.abs(b["outflw"] - a["outflw"]).max()
                                                          / np.abs(a["outflw"]).max()))
            worst["depth"] = max(worst["depth"], float(np.abs(b["rivdph"] - a["rivdph"]).max()
                                                      / np.abs(a["rivdph"]).max()))
        for g in runs.values():
            g.close()
        out.append(dict(config=name, cells=m.n, steps=nstep, dt=dt, substeps=nts, rel_err=worst))
        log(f"{name}: {worst} ({time.time() - t0:.1f}s)")
    return out


ENS_DOMAINS = 16
ENS_MEMBERS = 64          # runoff realisations of every domain: 16 x 64 = 1024 instances
ENS_SIGMA = 0.3           # multiplicative lognormal runoff perturbation
ENS_RECORDS = 3           # daily runoff records per domain, cycled


def ensemble_domains(seed, ndom=ENS_DOMAINS):
    """Regional 0.1-degree domains of uneven size, 2e4 .. 4e5 river cells, most near
    4e5 (about 5.7e6 cells in total)."""
    from cmf import maps, synth
    nets = synth.regional_networks(synth.ensemble_domain_sizes(ndom), seed)
    return nets, [maps.build_map(net, maps.Options()) for net in nets]


def run_ensemble(nets, ms, precision, days, warmup, seed, tune=None, timing_days=2, members=ENS_MEMBERS):
    """Batched ensemble on one handle: every domain x `members` members, daily DT
    with a per-instance adaptive substep, forcing = domain runoff x member factor
    (lognormal, sigma ENS_SIGMA, redrawn each day); every member starts from the
    domain's cold start."""
    from cmf import maps, synth
    from cmf.model import CMFModel
    cm, dom_cells = maps.concat_maps(ms)
    nm = members
    ninst = len(ms) * nm
    st0 = [maps.cold_start(m) for m in ms]
    init = {k: np.concatenate([s_[k] for s_ in st0])
            for k in ("rivsto", "fldsto", "rivout", "fldout", "rivdph_pre", "fldsto_pre")}
    em = CMFModel(cm, precision, members=nm, dom_cells=dom_cells)
    em.set_state(init, shared=True)
    em.physics_fldstg()
    ndays = warmup + days + timing_days
    ro = [synth.make_runoff(net, ENS_RECORDS, seed + 7 * d) for d, net in enumerate(nets)]
    base = [np.concatenate([m.vec(ro[d][k]) for d, m in enumerate(ms)]) for k in range(ENS_RECORDS)]
    rng = np.random.default_rng(seed)
    fac = np.exp(rng.normal(-0.5 * ENS_SIGMA ** 2, ENS_SIGMA, size=(ndays, ninst)))
    cells_inst = np.repeat(np.asarray(dom_cells, np.float64), nm)
    if tune == "auto":
        em.set_runoff_scaled(base[0], fac[0])
        best, rep = em.autotune(86400.0, nsub=8)
        tune_used, trials = best, rep.tolist()
    else:
        if tune:
            em.set_tune(**tune)
        tune_used, trials = em.get_tune(), []
    day = 0
    for _ in range(warmup):
        em.set_runoff_scaled(base[day % ENS_RECORDS], fac[day]); day += 1
        em.ensemble_advance(86400.0)
    em.synchronize()
    work, t_total = 0.0, 0.0
    for _ in range(days):
        em.set_runoff_scaled(base[day % ENS_RECORDS], fac[day]); day += 1
        em.synchronize()
        t0 = time.perf_counter()
        nt, _dt = em.ensemble_advance(86400.0)
        em.synchronize()
        t_total += time.perf_counter() - t0
        work += float(np.dot(cells_inst, nt))
    # latency pass: per-instance completion time of its simulated day
    lat = []
    em.kernel_times(True)
    for _ in range(timing_days):
        em.set_runoff_scaled(base[day % ENS_RECORDS], fac[day]); day += 1
        em.ensemble_advance(86400.0)
        lat.extend(em.instance_times().tolist())
    kt = em.kernel_times(False)
    g = em.get_state(["rivsto", "fldsto"])
    em.close()
    return dict(work=work, t=t_total, lat=np.asarray(lat), kt=kt, tune=tune_used, trials=trials,
                storage=g["rivsto"] + g["fldsto"], cells=int(sum(dom_cells)), ninst=ninst, dom_cells=dom_cells)


def ensemble_section(a):
    t0 = time.time()
    nets, ms = ensemble_domains(a.seed)
    log(f"ensemble: {len(ms)} domains, {sum(m.n for m in ms)} cells x {ENS_MEMBERS} members "
        f"({time.time() - t0:.1f}s)")
    res = {}
    for prec in ("mixed", "fp64"):
        r = run_ensemble(nets, ms, prec, a.ens_days, 1, a.seed,
                         tune="auto" if a.ens_tune == "auto" else json.loads(a.ens_tune))
        res[prec] = r
        steps = max(r["kt"][4], 1.0)
        r["kernels_ms"] = {k: round(r["kt"][i] / steps, 4) for i, k in enumerate(("outflow", "rate", "update"))}
        log(f"ensemble {prec}: kernels ms per substep {r['kernels_ms']}")
        log(f"ensemble {prec}: {r['work'] / r['t']:.4g} member-cell-substeps/s; tune {r['tune']}; "
            f"latency p50 {np.percentile(r['lat'], 50):.1f} ms p99 {np.percentile(r['lat'], 99):.1f} ms")
    # mixed vs FP64 storage error per instance after the same days
    tot_m, tot_d = res["mixed"].pop("storage"), res["fp64"].pop("storage")
    offs = np.cumsum([0] + res["fp64"]["dom_cells"])
    rel_tot, rel_max = [], []
    for mi in range(ENS_MEMBERS):
        for d in range(len(ms)):
            a_ = tot_m[mi, offs[d]:offs[d + 1]]
            b_ = tot_d[mi, offs[d]:offs[d + 1]]
            rel_tot.append(abs(a_.sum() - b_.sum()) / b_.sum())
            rel_max.append(float(np.max(np.abs(a_ - b_)) / np.max(np.abs(b_))))
    r = res["mixed"]
    rd = res["fp64"]
    cpu = None
    if not a.no_cpu:
        try:
            cpu = ensemble_cpu_baseline(nets, a.seed)
            if cpu:
                log(f"ensemble CPU reference: {cpu['value']:.4g} member-cell-substeps/s on {cpu['cores']} cores; "
                    f"mixed ensemble / CPU = {r['work'] / r['t'] / cpu['value']:.0f}x")
        except Exception as e:          # a reported baseline only: never fails the run
            log(f"ensemble CPU baseline failed: {e!r}")
    return dict(
        metric="member-cell-substeps/s", value=r["work"] / r["t"], unit="member-cell-substeps/s",
        precision="mixed", domains=len(ms), members_per_domain=ENS_MEMBERS, instances=r["ninst"],
        cells=r["cells"], member_cells=r["cells"] * ENS_MEMBERS, days=a.ens_days, dt=86400.0,
        ms_per_day=1e3 * r["t"] / a.ens_days,
        latency_ms={"p50": float(np.percentile(r["lat"], 50)), "p99": float(np.percentile(r["lat"], 99)),
                    "samples": int(r["lat"].size)},
        tune=r["tune"], kernels_ms=r["kernels_ms"],
        fp64={"value": rd["work"] / rd["t"], "ms_per_day": 1e3 * rd["t"] / a.ens_days,
              "kernels_ms": rd["kernels_ms"],
              "latency_ms": {"p50": float(np.percentile(rd["lat"], 50)),
                             "p99": float(np.percentile(rd["lat"], 99))}, "tune": rd["tune"]},
        mixed_vs_fp64_ensemble=(r["work"] / r["t"]) / (rd["work"] / rd["t"]),
        mixed_vs_fp64_storage={"total_rel_max": float(max(rel_tot)), "cell_rel_max": float(max(rel_max)),
                               "instances": len(rel_tot)},
        cpu_baseline=cpu,
    )


def spinup_section(a):
    """Learned spin-up warm start at 30 arc-minutes: labels from a batched FP64
    spin-up of 600 river basins cut from six 720x360 (0.5-degree) networks (own
    climatology per basin), training by the project's BF16 matrix-core kernels,
    then FP64 spin-ups of a held-out 720x360 network from the cold start, the
    mean-flow steady state, the physical prior and the learned warm start (four
    members of one handle). The glb_03min-sized run is spinup_full_section."""
    from cmf import maps, synth
    from cmf import surrogate as sg
    from cmf.spinup import make_climatology
    nx, ny = 720, 360
    nets, per = 6, 100
    t0 = time.time()
    basins = sg.basins_from_networks(nets, per, seed=a.seed + 7, nx=nx, ny=ny)
    X, Y, res = sg.equilibrium_labels(basins, tol=1e-5, max_years=40, log=lambda s_: log("labels " + s_))
    t_labels = time.time() - t0
    log(f"spin-up labels: {len(basins)} basins from {nets} {nx}x{ny} networks, {X.shape[0]} cells, "
        f"years to < 0.1% annual change from the prior: median {int(np.median(res.years))} "
        f"max {int(res.years.max())}, {t_labels:.1f}s")
    t1 = time.time()
    Xs, Ys, ws = sg.training_set(X, Y, res.state["rivsto"] + res.state["fldsto"])
    sur, hist = sg.train(Xs, Ys, epochs=a.spin_epochs, batch=1024, lr=2e-3, seed=0, weights=ws,
                         log=lambda s_: log("train " + s_))
    t_train = time.time() - t1
    net = synth.make_network(nx, ny, seed=a.seed + 991)
    m = maps.build_map(net, maps.Options())
    clim = make_climatology(net, m, a.seed + 991, mean_mm_day=1.5, amp=0.5, phase=3.0)
    # inference speed of the device network at this size and at 6e6 cells
    net_dev = sg.MLP(max_batch=1 << 18)
    net_dev.set_params(sur.params)
    xs = (sg.features(m, clim) - sur.x_mu) / sur.x_sd
    big = np.resize(xs, (6_000_000, xs.shape[1])).astype(np.float32)
    net_dev.forward(big[: 1 << 18])                      # warm-up
    ti = time.perf_counter()
    net_dev.forward(big)
    t_inf = time.perf_counter() - ti
    net_dev.close()
    flops = 2.0 * big.shape[0] * (24 * 256 + 2 * 256 * 256 + 256 * 2)
    log(f"inference on 6e6 cells: {1e3 * t_inf:.1f} ms ({flops / t_inf / 1e12:.2f} TFLOP/s incl. host transfers)")
    cmp_ = sg.compare_spinup(sur, sg.Basin(net, m, clim), years=a.spin_years, log=lambda s_: log("spin-up " + s_))
    c, w, pr, sd = cmp_["cold"], cmp_["warm"], cmp_["prior"], cmp_["steady"]
    out = dict(
        application=f"30-arcmin {nx}x{ny} held-out network, {m.n} cells, monthly gamma climatology "
                    f"repeated, DT=86400, criterion: annual storage change < 0.1%",
        cold_years=c["years"] or f">{cmp_['years_run']}", warm_years=w["years"] or f">{cmp_['years_run']}",
        physics_prior_years=pr["years"] or f">{cmp_['years_run']}",
        mean_flow_steady_years=sd["years"] or f">{cmp_['years_run']}",
        years_ratio=(c["years"] / w["years"]) if (c["years"] and w["years"]) else None,
        days_to_0p1pct={k: v["days"] for k, v in (("cold", c), ("steady", sd), ("prior", pr), ("warm", w))},
        init_rel_err={k: v["init_rel_err"] for k, v in (("cold", c), ("steady", sd), ("prior", pr), ("warm", w))},
        annual_change={k: v["annual_change"] for k, v in (("cold", c), ("steady", sd), ("prior", pr),
                                                           ("warm", w))},
        equilibrium_reference=cmp_["reference_member"],
        training={"basins": len(basins), "cells": int(X.shape[0]), "epochs": a.spin_epochs,
                  "corrected_cells": int(Xs.shape[0]),
                  "loss_first": hist[0], "loss_last": hist[-1], "label_seconds": round(t_labels, 1),
                  "train_seconds": round(t_train, 1)},
        inference={"cells": big.shape[0], "ms": round(1e3 * t_inf, 2),
                   "tflops_incl_transfers": round(flops / t_inf / 1e12, 3)},
        spinup_seconds=cmp_["spinup_seconds"],
        member_years=cmp_["member_years"],
    )
    if not a.no_cpu:
        try:
            cpu = spinup_cpu_baseline(net, m, clim)
            out["cpu_baseline"] = cpu
            if cpu:
                gpu_y = cmp_["spinup_seconds"] / cmp_["member_years"]
                log(f"spin-up CPU reference: {cpu['value']:.1f} s per simulated year on {cpu['cores']} cores; "
                    f"device {gpu_y:.2f} s per member-year in the 4-member run")
        except Exception as e:          # a reported baseline only: never fails the run
            log(f"spin-up CPU baseline failed: {e!r}")
    log(f"spin-up: cold {c['years']} y, mean-flow steady {sd['years']} y, physics prior {pr['years']} y, "
        f"warm {w['years']} y; first-year change cold {c['annual_change'][0]:.3e} steady "
        f"{sd['annual_change'][0]:.3e} prior {pr['annual_change'][0]:.3e} warm {w['annual_change'][0]:.3e}")
    return out


C5_NX, C5_NY = 7200, 3600        # glb_03min-shaped: 0.05-degree global grid
C5_MAX_LAT = 70.0                # no land poleward of +-70 degrees (cells there are metres wide)
# training set: 2000 basins of twenty 1440x720 (glb_15min-shaped) networks, labelled by the
# FP64 device spin-up (tests/golden/make_surrogate_set.py); 200 basins are too few (the
# network then fits its basins and is worse than the prior on unseen ones)
C5_TRAIN_SET = os.path.join(ROOT, "tests", "golden", "surrogate_train.npz")


def c5_network(seed):
    from cmf import maps, synth
    from cmf.spinup import make_climatology
    net = synth.make_network(C5_NX, C5_NY, seed=seed + 991, max_lat=C5_MAX_LAT)
    m = maps.build_map(net, maps.Options())
    clim = make_climatology(net, m, seed + 991, mean_mm_day=1.5, amp=0.5, phase=3.0)
    return net, m, clim


def mixed_drift(seed, days, npth=60000):
    """Mixed vs FP64 error growth on the glb_15min-shaped 1440x720 network with and
    without bifurcation paths: the mixed path against FP64, an FP64 run whose
    initial river storage is scaled by (1 + 1e-12) against FP64 (the model's own
    amplification of a small difference), and the mixed path's own mass residual
    (sum over substeps of DERROR = -(STOPRE - STONXT + RIVINF - RIVOUT),
    src/cmf_ctrl_physics_mod.F90:238). Rows on days 1-5 and every 5th day."""
    from cmf import maps, synth
    from cmf.model import CMFModel
    out = {}
    for label, np_ in (("paths", npth), ("nopaths", 0)):
        net = synth.make_network(1440, 720, seed=seed)
        if np_:
            synth.add_bifurcations(net, np_, 5, seed)
        ro = synth.make_runoff(net, 32, seed)
        m = maps.build_map(net, maps.Options(lpthout=bool(np_)))
        st = initial_state(net, m)
        runs = {p: CMFModel(m, "fp64" if p.startswith("fp64") else "mixed") for p in ("fp64", "fp64p", "mixed")}
        for p, g in runs.items():
            s = dict(st)
            if p == "fp64p":
                s["rivsto"] = st["rivsto"] * (1.0 + 1e-12)
            g.set_state(s)
        rows, miss = [], 0.0
        for k in range(days):
            r = m.vec(ro[k % ro.shape[0]])
            for p, g in runs.items():
                g.set_runoff(r)
                if p == "mixed":
                    _, glb = g.physics_advance(86400.0, budget=True)
                    miss += float(np.sum(-(glb[:, 0] - glb[:, 1] + glb[:, 3] - glb[:, 4])))
                else:
                    g.physics_advance(86400.0)
            if (k + 1) % 5 == 0 or k < 5:
                a = runs["fp64"].get_state(["rivsto", "fldsto", "outflw"])
                sa = a["rivsto"] + a["fldsto"]
                row = dict(day=k + 1)
                for p in ("fp64p", "mixed"):
                    b = runs[p].get_state(["rivsto", "fldsto", "outflw"])
                    sb = b["rivsto"] + b["fldsto"]
                    row[p] = dict(cell=float(np.abs(sb - sa).max() / np.abs(sa).max()),
                                  total=float(abs(sb.sum() - sa.sum()) / sa.sum()),
                                  outflw=float(np.abs(b["outflw"] - a["outflw"]).max() / np.abs(a["outflw"]).max()))
                row["mixed_mass_residual"] = miss / float(sa.sum())
                rows.append(row)
                log(f"{label} day {k + 1}: mixed total {row['mixed']['total']:.2e} cell {row['mixed']['cell']:.2e}; "
                    f"perturbed FP64 cell {row['fp64p']['cell']:.2e}")
        out[label] = rows
        for g in runs.values():
            g.close()
    return out


def spinup_full_section(a):
    """The learned warm start at its specified size, in phases that exchange small
    files: each phase writes to a.spin_dir and reads what earlier phases wrote from
    a.spin_in (default a.spin_dir); a.spinup_full may list several phases, run in
    order. With --spin-keep-state a spin-up phase also saves its full state
    (hundreds of MB at this size) and resumes from it when run again.
      train  training on the device from the stored FP64 equilibria of 2000 basins
             cut from twenty glb_15min-shaped 1440x720 networks
      cold   FP64 spin-up of the held-out glb_03min-shaped 7200x3600 network
             (~6e6 cells) from the cold start (STORAGE_SEA_SURFACE)
      warm   the same from the physical prior and from the learned warm start
             (two members of one handle)
      report years and days to an annual storage change below 0.1% of every
             start, against the equilibrium of the start that changes least"""
    phases = a.spinup_full.split(",")
    if len(phases) > 1:
        return [spinup_full_section(argparse.Namespace(**{**vars(a), "spinup_full": p_})) for p_ in phases]
    from cmf import surrogate as sg
    d = a.spin_dir
    src = a.spin_in or d
    os.makedirs(d, exist_ok=True)
    ph = a.spinup_full
    t0 = time.time()
    if ph == "train":
        Xs, Ys, ws, bid = sg.load_training_set(C5_TRAIN_SET)
        t1 = time.time()
        sur, hist = sg.train(Xs, Ys, epochs=a.spin_epochs, batch=1024, lr=2e-3, seed=0, weights=ws,
                             log=lambda s_: log("train " + s_))
        sur.save(os.path.join(d, "surrogate.npz"))
        info = {"basins": int(bid.max()) + 1, "networks": "20 x 1440x720 (glb_15min-shaped, land within +-70 deg)",
                "training_set": os.path.relpath(C5_TRAIN_SET, ROOT), "corrected_cells": int(Xs.shape[0]),
                "epochs": a.spin_epochs, "loss_first": hist[0], "loss_last": hist[-1],
                "train_seconds": round(time.time() - t1, 1)}
        with open(os.path.join(d, "train.json"), "w") as f:
            json.dump(info, f)
        return {"phase": "train", **info}
    if ph in ("cold", "warm"):
        net, m, clim = c5_network(a.seed)
        log(f"glb_03min-shaped network: {m.n} cells ({time.time() - t0:.1f}s)")
        state_in = os.path.join(src, f"{ph}_state.npz")
        carry_in = os.path.join(src, f"{ph}_carry.npz")
        resume = None
        if a.spin_keep_state and os.path.exists(state_in):
            z = np.load(state_in, allow_pickle=False)
            resume = {k: z[k] for k in z.files}
            starts = None
            log(f"{ph}: resuming after {resume['hist'].shape[0] - 1} years")
        elif a.spin_carry and os.path.exists(carry_in):
            # continue from the float32 storages of an earlier run, entered as a
            # storage-only restart (stage, then CMF_CALC_OUTPRE flows): the float32
            # rounding (6e-8) is far below the 0.1% annual-change criterion
            z = np.load(carry_in, allow_pickle=False)
            sts = [sg.warm_state(m, z["rivsto"][k].astype(np.float64), z["fldsto"][k].astype(np.float64))
                   for k in range(z["rivsto"].shape[0])]
            resume = {k: np.stack([s_[k] for s_ in sts]) for k in sg.RESUME_KEYS}
            resume.update({k: z[k] for k in ("hist", "daily", "conv", "seconds", "member_substeps", "cells")})
            starts = None
            log(f"{ph}: continuing after {resume['hist'].shape[0] - 1} years from float32 storages")
        elif ph == "cold":
            starts = {"cold": sg.cold_state(m)}
        else:
            sur = sg.Surrogate.load(os.path.join(src, "surrogate.npz"))
            tp = time.time()
            pr = sg.prior(m, clim)
            t_prior = time.time() - tp
            ti = time.time()
            ws_ = sur.warm_start(m, clim, pr)
            t_inf = time.time() - ti
            starts = {"prior": sg.prior_state(m, clim, pr), "warm": sg.warm_state(m, ws_, np.zeros(m.n))}
            with open(os.path.join(d, "warm_setup.json"), "w") as f:
                json.dump({"prior_seconds": round(t_prior, 1), "features_and_inference_seconds": round(t_inf, 1)}, f)
        names = ["cold"] if ph == "cold" else ["prior", "warm"]
        r = sg.spinup_from_starts(m, clim, starts, years=a.spin_years, names=names, resume=resume,
                                  time_budget=a.spin_seconds, log=lambda s_: log(f"{ph} " + s_), progress_days=30)
        if a.spin_keep_state:
            np.savez(os.path.join(d, f"{ph}_state.npz"), **r["resume"])
        small = {k: r["resume"][k] for k in ("hist", "daily", "conv", "seconds", "member_substeps", "cells")}
        np.savez(os.path.join(d, f"{ph}_summary.npz"), **small)
        done = all(r["conv"])
        if a.spin_carry and not done:
            np.savez_compressed(os.path.join(d, f"{ph}_carry.npz"), **small,
                                rivsto=r["resume"]["rivsto"].astype(np.float32),
                                fldsto=r["resume"]["fldsto"].astype(np.float32))
        log(f"{ph}: years run {r['hist'].shape[0] - 1}, converged {r['conv']}, {r['seconds']} s this run")
        return {"phase": ph, "cells": m.n, "years_run": int(r["hist"].shape[0] - 1), "converged": done,
                "conv": r["conv"], "seconds": r["seconds"]}
    if ph == "report":
        runs = []
        for nm in ("cold", "warm"):
            z = np.load(os.path.join(src, f"{nm}_summary.npz"), allow_pickle=False)
            runs.append(dict(names=["cold"] if nm == "cold" else ["prior", "warm"], hist=z["hist"], daily=z["daily"],
                             conv=[int(c) for c in z["conv"]], seconds=float(z["seconds"]),
                             member_substeps=int(z["member_substeps"]), cells=int(z["cells"])))
        res = sg.summarize_spinups(runs)
        with open(os.path.join(src, "train.json")) as f:
            train = json.load(f)
        with open(os.path.join(src, "warm_setup.json")) as f:
            setup = json.load(f)
        c, p_, w = res["cold"], res["prior"], res["warm"]
        ys = lambda v: v["years"] or f">{v['years_run']}"
        # a start that did not converge in the run: its annual change has decayed
        # geometrically over the last years; the year it would pass the tolerance
        # at that rate (an extrapolation, reported as such)
        extra = None
        ch = c["annual_change"]
        if not c["years"] and len(ch) >= 3 and ch[-1] < ch[-3]:
            rate = (ch[-1] / ch[-3]) ** 0.5
            extra = dict(decay_per_year=rate, years=len(ch) + float(np.log(1e-3 / ch[-1]) / np.log(rate)))
        member_years = sum((r["hist"].shape[0] - 1) * len(r["names"]) for r in runs)
        secs = sum(r["seconds"] for r in runs)
        out = dict(
            application=f"glb_03min-shaped {C5_NX}x{C5_NY} held-out network (land within +-{C5_MAX_LAT:g} deg), "
                        f"{runs[0]['cells']} cells, monthly gamma climatology repeated, DT=86400, FP64, "
                        f"criterion: annual storage change < 0.1%",
            cold_years=ys(c), physics_prior_years=ys(p_), warm_years=ys(w),
            cold_years_extrapolated=extra,
            years_ratio=(c["years"] / w["years"]) if (c["years"] and w["years"]) else None,
            days_to_0p1pct={k: res[k]["days"] for k in ("cold", "prior", "warm")},
            init_rel_err={k: res[k]["init_rel_err"] for k in ("cold", "prior", "warm")},
            annual_change={k: res[k]["annual_change"] for k in ("cold", "prior", "warm")},
            equilibrium_reference=res["reference_member"], training=train, warm_setup=setup,
            member_years=member_years, spinup_seconds=round(secs, 1),
            member_cell_substeps_per_s=sum(r["member_substeps"] for r in runs) * runs[0]["cells"] / secs)
        return out
    raise SystemExit(f"unknown spin-up phase {ph!r}")


def forcing_section(net, m, seed, res_in=0.25, calls=20):
    """Runoff interpolation on the device (ROFF_INTERP with the input matrix) for
    the benchmark map: input grid of res_in degrees, 2 x 2 overlaps per cell.
    Timed with the field resident in device memory (kernel + stream sync per
    call) and with the host-to-device copy of the field included; and the
    output-period accumulation (CMF_DIAG_AVEMAX_OUTPUT) of the same map."""
    import torch
    from cmf import synth
    from cmf.model import CMFModel
    inpx, inpy, inpa, nxin, nyin = synth.make_inpmat(net, (m.seqx, m.seqy), res_in)
    grids = synth.make_runoff_grid(nxin, nyin, 2, seed)
    g = CMFModel(m, "fp64")
    g.set_inpmat(inpx, inpy, inpa, nxin, nyin)
    dev = torch.from_numpy(grids[0]).to("cuda")
    torch.cuda.synchronize()
    for _ in range(3):
        g.interp_runoff(dev.data_ptr())
    t0 = time.perf_counter()
    for _ in range(calls):
        g.interp_runoff(dev.data_ptr())
    t_dev = (time.perf_counter() - t0) / calls
    g.set_runoff_grid(grids[1])
    t0 = time.perf_counter()
    for k in range(calls):
        g.set_runoff_grid(grids[k % 2])
    t_h2d = (time.perf_counter() - t0) / calls
    # output-period accumulation (CMF_DIAG_AVEMAX_OUTPUT) on the same handle:
    # 10 substep averages/maxima read, 10 accumulators read and written per cell
    g.diag_reset_output()
    for _ in range(3):
        g.diag_avemax_output(3600.0)
    t0 = time.perf_counter()
    for _ in range(calls):
        g.diag_avemax_output(3600.0)
    t_out = (time.perf_counter() - t0) / calls
    g.close()
    inpn = inpx.shape[0]
    nbytes = m.n * (inpn * (4 + 4 + 8) + 8) + 8 * nxin * nyin
    obytes = m.n * (10 * 8 + 10 * 16)
    out = dict(input_grid=f"{nxin}x{nyin} ({res_in} deg)", inpn=inpn, cells=m.n,
               us_per_call_resident=round(1e6 * t_dev, 1), us_per_call_with_h2d=round(1e6 * t_h2d, 1),
               bytes_per_call=nbytes, gbs_resident=round(nbytes / t_dev / 1e9, 1),
               output_avemax_us_per_call=round(1e6 * t_out, 1), output_avemax_bytes=obytes,
               output_avemax_gbs=round(obytes / t_out / 1e9, 1))
    log(f"runoff interpolation and output accumulation (host-timed calls incl. stream sync): {out}")
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs of the FP64 C3 workload (scaled by the copy-kernel
    calibration), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            return json.load(f)["kernels"][kernel]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def roofline(kt, n_cells, bytes_tab):
    steps = max(kt[4], 1.0)
    if kt[0] == 0.0 and kt[1] == 0.0:
        # fused persistent substep: one kernel carries all three phases
        per = {"substep": kt[2] / steps}
        bytes_tab = dict(bytes_tab, substep=sum(bytes_tab.values()))
    else:
        per = {"outflow": kt[0] / steps, "rate": kt[1] / steps, "update": kt[2] / steps}
    dom = max(per, key=per.get)
    ms = per[dom]
    achieved = bytes_tab[dom] * n_cells / (ms * 1e-3) / 1e9
    return dom, per, dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                          frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                          kernel=f"k_{dom}", avg_ms=round(ms, 4),
                          bytes_per_cell_substep=bytes_tab[dom])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=3600)
    ap.add_argument("--ny", type=int, default=1800)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--tune", type=str, default=PINNED_TUNE,
                    help='launch configuration of the single-model runs: a JSON object of cmf_tune fields '
                         '(default: the configuration the committed traffic counters were taken with) or "auto" '
                         '(the autotuner)')
    ap.add_argument("--no-spinup", action="store_true")
    ap.add_argument("--no-forcing", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--dropin-steps", type=int, default=12)
    ap.add_argument("--spinup-only", action="store_true")
    ap.add_argument("--spin-epochs", type=int, default=150)
    ap.add_argument("--spin-years", type=int, default=20)
    ap.add_argument("--spinup-full", default=None,
                    help="phases of the glb_03min-sized learned warm start, comma-separated, from train, cold, "
                         "warm, report (JSON of the phases on stdout)")
    ap.add_argument("--spin-dir", default=os.path.join(ROOT, "gpurun_out", "spinup_full"),
                    help="directory the --spinup-full phases write to")
    ap.add_argument("--spin-in", default=None,
                    help="directory the --spinup-full phases read earlier phases' files from (default: --spin-dir)")
    ap.add_argument("--spin-keep-state", action="store_true",
                    help="save (and resume from) the full spin-up state of the cold/warm phases")
    ap.add_argument("--spin-carry", action="store_true",
                    help="an unfinished cold/warm phase saves its storages in float32 (<phase>_carry.npz); when "
                         "--spin-in holds one, the phase continues from it as a storage-only restart")
    ap.add_argument("--spin-seconds", type=float, default=900.0,
                    help="a spin-up phase stops after the simulated year that passes this many seconds")
    ap.add_argument("--mixed-drift", type=int, default=0, metavar="DAYS",
                    help="only the mixed-vs-FP64 drift study with and without bifurcation paths over DAYS days")
    ap.add_argument("--precision-report", action="store_true",
                    help="only report mixed-vs-FP64 errors of the C1/C2/C3 configurations (JSON)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 2 GiB device copy three times (HBM counter calibration)")
    ap.add_argument("--ens-days", type=int, default=3)
    ap.add_argument("--ens-tune", type=str, default="auto",
                    help='launch configuration of both ensemble runs: "auto" (autotuner) or a JSON object '
                         'of cmf_tune fields')
    ap.add_argument("--compare-order", action="store_true",
                    help="also time the FP64 path with cells in NSEQ order")
    a = ap.parse_args()
    if a.spinup_full:
        bad = set(a.spinup_full.split(",")) - {"train", "cold", "warm", "report"}
        if bad:
            ap.error(f"unknown --spinup-full phase(s): {sorted(bad)}")
        print(json.dumps({"spinup_full": spinup_full_section(a)}), flush=True)
        return
    if a.spinup_only:
        print(json.dumps({"spinup": spinup_section(a)}), flush=True)
        return
    if a.precision_report:
        print(json.dumps({"mixed_vs_fp64": precision_report(a.seed)}), flush=True)
        return
    if a.mixed_drift:
        print(json.dumps({"mixed_drift": mixed_drift(a.seed, a.mixed_drift)}), flush=True)
        return
    if a.calibrate_copy:
        import torch
        # 2 GiB, far larger than the 256 MB Infinity Cache, so every byte crosses HBM
        src = torch.ones(512 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copied 3 x 2 GiB")
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); multi-GPU sharding is not "
              "part of this build", file=sys.stderr)
        sys.exit(2)
    dt = 3600.0
    net, m, ro = build_workload(a.nx, a.ny, a.seed, 4)
    rvec = [m.vec(ro[r]) for r in range(ro.shape[0])]
    n = m.n
    log("FP64 run")
    tune_arg = "auto" if a.tune == "auto" else json.loads(a.tune)
    sum_nt, t, kt, tune64 = run_gpu(net, m, rvec, "fp64", a.steps, a.warmup, dt, tune=tune_arg)
    value = n * sum_nt / t
    dom, per, roof = roofline(kt, n, BYTES_FP64)
    log(f"FP64: {value:.4g} cell-substeps/s, sum NT {sum_nt}, {t:.3f}s; kernels ms {per}")
    order_cmp = None
    if a.compare_order:
        from cmf import maps
        os.environ["CMF_HIP_ORDER"] = "nseq"
        m3 = maps.build_map(net, maps.Options(llevee=True, ldamout=True))
        s3, t3, kt3, _ = run_gpu(net, m3, rvec, "fp64", a.steps, a.warmup, dt, tune=tune64)
        del os.environ["CMF_HIP_ORDER"]
        order_cmp = dict(value=n * s3 / t3, kernels_ms=roofline(kt3, n, BYTES_FP64)[1])
        log(f"FP64 NSEQ order: {order_cmp['value']:.4g} cell-substeps/s; kernels ms {order_cmp['kernels_ms']}")
    mixed = None
    if not a.no_mixed:
        net2, m2, _ = net, None, None
        from cmf import maps
        m2 = maps.build_map(net, maps.Options(llevee=True, ldamout=True))
        sum_nt2, t2, kt2, tune32 = run_gpu(net, m2, rvec, "mixed", a.steps, a.warmup, dt, tune=tune_arg)
        dom2, per2, roof2 = roofline(kt2, n, BYTES_MIXED)
        mixed = dict(value=n * sum_nt2 / t2, unit="cell-substeps/s", sum_nt=sum_nt2,
                     ms_per_step=1e3 * t2 / a.steps, roofline=roof2, tune=tune32)
        log(f"mixed: {mixed['value']:.4g} cell-substeps/s; kernels ms {per2}")
    ens = None
    if not a.no_ensemble:
        try:
            ens = ensemble_section(a)
            ens["vs_fp64_single_gpu"] = ens["value"] / value
            log(f"ensemble (mixed) / FP64 single model: {ens['vs_fp64_single_gpu']:.2f}x")
        except Exception as e:              # the headline line is still printed
            log(f"ensemble section failed: {e!r}")
            ens = {"error": repr(e)}
    forcing = None
    if not a.no_forcing:
        try:
            forcing = forcing_section(net, m, a.seed)
        except Exception as e:
            log(f"forcing section failed: {e!r}")
            forcing = {"error": repr(e)}
    spin = None
    if not a.no_spinup:
        try:
            spin = spinup_section(a)
        except Exception as e:
            log(f"spin-up section failed: {e!r}")
            spin = {"error": repr(e)}
    cpu = None
    if not a.no_cpu:
        log("CPU baseline (reference OpenMP)")
        cpu = cpu_baseline(net, ro, n, dt, a.cpu_steps)
        if cpu:
            log(f"CPU reference FP64: {cpu['value']:.4g} cell-substeps/s on {cpu['cores']} cores")
            if mixed is not None:
                c32 = cpu_baseline(net, ro, n, dt, a.cpu_steps, "omp32")
                mixed["cpu_baseline"] = c32
                if c32:
                    log(f"CPU reference single precision: {c32['value']:.4g} cell-substeps/s")
    dropin = None
    if not a.no_dropin:
        try:
            dropin = dropin_section(net, ro, n, dt, a.dropin_steps, m)
        except Exception as e:
            log(f"drop-in section failed: {e!r}")
            dropin = {"error": repr(e)}
    roof["traffic"] = measured_traffic(roof["kernel"]) if a.nx == 3600 and a.ny == 1800 else None
    sub_ms = sum(per.values())
    roof_sub = dict(bound="hbm", kernels="k_outflow + k_rate + k_update", avg_ms=round(sub_ms, 4),
                    bytes_per_cell_substep=508.0,
                    achieved=round(508.0 * n / (sub_ms * 1e-3) / 1e9, 1), peak=HBM_PEAK_GBS, unit="GB/s")
    roof_sub["frac"] = round(roof_sub["achieved"] / HBM_PEAK_GBS, 4)
    out = {
        "metric": "cell-substeps/s (1 GPU, 1.5M cells) FP64 vs mixed+batched; spin-up speedup",
        "value": value, "unit": "cell-substeps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * t / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"glb_06min-shaped synthetic network {a.nx}x{a.ny}, {n} cells, "
                               f"3% levees, 2500 dams, hourly runoff, DT=3600 s adaptive",
                   "cells": n, "sum_nt": sum_nt, "precision": "fp64", "tune": tune64},
        "roofline": roof,
        "roofline_substep": roof_sub,
        "cpu_baseline": cpu,
        "mixed": mixed,
        "fp64_nseq_order": order_cmp,
        "ensemble": ens,
        "forcing_interpolation": forcing,
        "spinup": spin,
        # the glb_03min-sized spin-up (6e6 cells, ~90 GPU-minutes) is not re-run here;
        # its result from the --spinup-full phases is kept in this file
        "spinup_glb03min_profile": "profiles/spinup_glb03min_7200x3600.json",
        "dropin": dropin,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
