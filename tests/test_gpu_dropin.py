"""The drop-in configuration: the reference's own Fortran host (initialisation,
time loop, forcing, dam and levee set-up) linked with this package's
CMF_CTRL_PHYSICS_MOD replacement, which drives the HIP physics through
include/cmf_hip.h. Run on the golden fixtures, it must reproduce the reference's
substep counts and final state to 1e-10 relative.

Without CMF_HIP_DOWNLOAD=every the module downloads the state only on output and
restart steps; the output-file and restart-file tests below run that way, so a
step that the host reads but the module skipped would show as stale records.

The output files that the reference's own writer (CMF_OUTPUT_WRITE) produces in
the drop-in must match the reference's (tests/golden/g_pth_output.npz) to one
float32 ulp, both with the reference's host-side CMF_CALC_DIAG_MOD (cmf_dropin)
and with this package's device-side one (cmf_dropin_devdiag: output-period sums
on the device, no per-step download of the substep averages).

The binaries are built by cama-flood_v4_amd/fortran/build_dropin.sh where the
reference sources exist; the tests are skipped when they have not been built."""
import os
import subprocess

import numpy as np
import pytest

from cases import rel_err
from golden_cases import load

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "cama-flood_v4_amd", "fortran", "_build")
EXE = os.path.join(BUILD, "cmf_dropin")
OUTPUT_VARS = ("rivout,fldout,outflw,rivvel,pthout,runoff,maxflw,maxdph,maxsto,rivsto,fldsto,rivdph,"
               "flddph,fldfrc,fldare,sfcelv,storge,pthflw")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["g_allwet", "g_hourly", "c1_1deg", "g_kine", "g_dam_yby", "g_dam_h22",
                                  "g_slpmix"])
def test_dropin_matches_reference(name, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("drop-in binary not built")
    from cmf import synth
    from oracle import refdump
    print(f"drop-in case {name}", flush=True)
    net, ro, ref, fin = load(name)
    nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac = ref["spec"]
    d = str(tmp_path)
    synth.write_reference_inputs(net, d, ro, dt=dt, ifrq_inp=int(dt // 3600), ndays=days,
                                 lpthout=lpth, llevee=llev, ldamout=ldam, lkine=bool(ref.get("lkine", False)),
                                 ldamh22=bool(ref.get("ldamh22", False)), slpmix_mask=net.slpmix_mask)
    # the test driver reads the state after arbitrary steps (its dumps): whole
    # state downloaded after every step
    env = dict(os.environ, CMF_HIP_DOWNLOAD="every")
    r = subprocess.run([EXE, "runoff.bin", "dump.bin", "0"], cwd=d, capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    mp, states = refdump.read_dump(os.path.join(d, "dump.bin"), lpth, llev, ldam)
    nts = refdump.count_substeps(os.path.join(d, "log_CaMa.txt"))
    assert nts == ref["nt"]
    # dump_every=0 writes the initial state only; the final state is read back from
    # a second dump written at the end (dump_every = number of steps)
    nsteps = len(ref["nt"])
    r = subprocess.run([EXE, "runoff.bin", "dump2.bin", str(nsteps)], cwd=d, capture_output=True,
                       text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    mp, states = refdump.read_dump(os.path.join(d, "dump2.bin"), lpth, llev, ldam)
    last = states[-1]
    assert last["step"] == nsteps
    for k, v in fin.items():
        e = rel_err(last[k], v)
        print(f"  {k}: rel err {e:.3e}", flush=True)
        assert e <= 1e-10, f"{name}: {k} rel err {e:.3e}"


@pytest.mark.gpu
@pytest.mark.parametrize("exe", ["cmf_dropin", "cmf_dropin_devdiag"])
def test_dropin_output_files_match_reference(exe, tmp_path):
    """g_pth (bifurcations, 4 daily steps, daily output of 18 variables): the files
    CMF_OUTPUT_WRITE writes in the drop-in against the reference's own files."""
    path = os.path.join(BUILD, exe)
    if not os.path.exists(path):
        pytest.skip(f"{exe} not built")
    from cmf import synth
    from golden_cases import setup
    print(f"drop-in output files, {exe}", flush=True)
    net, ro, ref, fin = load("g_pth")
    nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac = ref["spec"]
    m = setup("g_pth")[0]
    nsteps = int(round(days * 86400.0 / dt))
    d = str(tmp_path)
    synth.write_reference_inputs(net, d, ro, dt=dt, ifrq_inp=int(dt // 3600), ndays=days,
                                 lpthout=lpth, llevee=llev, ldamout=ldam, outputs=OUTPUT_VARS)
    r = subprocess.run([path, "runoff.bin", "NONE", "0"], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    z = np.load(os.path.join(ROOT, "tests", "golden", "g_pth_output.npz"), allow_pickle=False)
    cell = (m.seqy.astype(np.int64) - 1) * m.nx + (m.seqx.astype(np.int64) - 1)
    off = np.ones(m.nx * m.ny, bool)
    off[cell] = False
    for v in OUTPUT_VARS.split(","):
        files = [f for f in os.listdir(d) if f.startswith(v + "_cmf")]
        assert len(files) == 1, (v, files)
        got = np.fromfile(os.path.join(d, files[0]), np.float32)
        want = z[v].view(np.float32)
        assert got.size == want.size, v
        if v != "pthflw":
            got, want = got.reshape(nsteps, -1), want.reshape(nsteps, -1)
            assert np.all(got[:, off] == want[:, off]), v
            got, want = got[:, cell], want[:, cell]
        ulp = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
        same_sign = np.sign(got) == np.sign(want)
        ok = (got == want) | (same_sign & (ulp <= 1))
        print(f"  {v}: {int((got != want).sum())} of {got.size} values differ", flush=True)
        assert ok.all(), v


@pytest.mark.gpu
@pytest.mark.parametrize("exe", ["cmf_dropin", "cmf_dropin_devdiag"])
@pytest.mark.parametrize("name", ["g_plain", "g_allwet"])
def test_dropin_restart_matches_reference(name, exe, tmp_path):
    """The end-of-run restart that the reference's CMF_RESTART_WRITE writes in the
    drop-in (the state downloaded because restart_is_write_time holds on the last
    step) against the reference's own file: 1e-10 of each record's magnitude, the
    bound of the device physics; g_allwet adds the path, levee and dam records."""
    path = os.path.join(BUILD, exe)
    if not os.path.exists(path):
        pytest.skip(f"{exe} not built")
    from cmf import synth
    from golden_cases import setup
    print(f"drop-in restart {name}, {exe}", flush=True)
    net, ro, ref, fin = load(name)
    nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac = ref["spec"]
    m = setup(name)[0]
    d = str(tmp_path)
    synth.write_reference_inputs(net, d, ro, dt=dt, ifrq_inp=int(dt // 3600), ndays=days,
                                 lpthout=lpth, llevee=llev, ldamout=ldam, restart=True)
    r = subprocess.run([path, "runoff.bin", "NONE", "0"], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    files = sorted(f for f in os.listdir(d) if f.startswith("restart") and f.endswith(".bin"))
    assert len(files) == 1, files
    z = np.load(os.path.join(ROOT, "tests", "golden", f"{name}_restart.npz"), allow_pickle=False)
    assert files[0] == bytes(z["name"]).decode()
    keys = ["rivsto", "fldsto", "rivout_pre", "fldout_pre", "rivdph_pre", "fldsto_pre"]
    keys += (["damsto"] if m.opts.ldamout else []) + (["levsto"] if m.opts.llevee else [])
    ng = m.nx * m.ny
    want = z["bin"].view(np.float64).reshape(len(keys), ng)
    got = np.fromfile(os.path.join(d, files[0]), np.float64)
    assert got.size == want.size
    got = got.reshape(len(keys), ng)
    for k, (g, w) in enumerate(zip(got, want)):
        e = np.abs(g - w).max() / max(np.abs(w[w < 1e19]).max(), 1e-30)
        print(f"  {keys[k]}: rel err {e:.3e}", flush=True)
        assert e <= 1e-10, keys[k]
    if lpth:
        gp = np.fromfile(os.path.join(d, files[0] + ".pth"), np.float64)
        wp = z["pth"].view(np.float64)
        assert gp.size == wp.size
        assert np.abs(gp - wp).max() <= 1e-10 * max(np.abs(wp).max(), 1e-30)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["g_allwet_sp", "g_hourly_sp"])
def test_single_precision_dropin_matches_reference(name, tmp_path):
    """The reference built in its default single precision (-DSinglePrec_CMF,
    JPRB = 4) linked with this package's module: the JPRB arrays cross the C-ABI
    as float, the device runs its mixed path (FP32 flux arithmetic and flux/stage
    arrays, FP64 storage and stage inversion). Against the reference's own single-precision run of the same
    inputs (tests/golden/*_sp: paths, levees, dams; daily and hourly DT):
    identical NT at every step, final storage within 2e-5 of the largest cell
    storage and 1e-6 in total, outflow within 1e-3 of the largest. The two single-
    precision paths round differently (the reference's FP32 flux arithmetic
    against the device's difference-form FP32 arithmetic), so they agree to the
    size of either's distance from the FP64 reference (measured on these cases:
    1e-6 cell, 2e-8 total, 3e-5 outflow), not bit for bit."""
    import json
    path = os.path.join(BUILD, "cmf_dropin32")
    if not os.path.exists(path):
        pytest.skip("cmf_dropin32 not built")
    from cmf import synth
    from oracle import refdump
    gdir = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(gdir, f"{name}.json")) as f:
        sp = json.load(f)
    fin = dict(np.load(os.path.join(gdir, f"{name}_final.npz")))
    net, ro, ref, _ = load(sp["base"])
    nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac = ref["spec"]
    d = str(tmp_path)
    synth.write_reference_inputs(net, d, ro, dt=dt, ifrq_inp=int(dt // 3600), ndays=days,
                                 lpthout=lpth, llevee=llev, ldamout=ldam)
    nsteps = len(sp["nt"])
    env = dict(os.environ, CMF_HIP_DOWNLOAD="every")
    r = subprocess.run([path, "runoff.bin", "dump.bin", str(nsteps)], cwd=d, capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    nts = refdump.count_substeps(os.path.join(d, "log_CaMa.txt"))
    log = open(os.path.join(d, "log_CaMa.txt")).read()
    assert "ADPSTP: NT=" in log
    mp, states = refdump.read_dump(os.path.join(d, "dump.bin"), lpth, llev, ldam)
    last = states[-1]
    assert last["step"] == nsteps
    print(f"  NT device {nts[:6]}... reference {sp['nt'][:6]}...", flush=True)
    assert nts == sp["nt"]
    tot_d = last["rivsto"] + last["fldsto"]
    tot_r = fin["rivsto"] + fin["fldsto"]
    e_cell = float(np.abs(tot_d - tot_r).max() / np.abs(tot_r).max())
    e_tot = abs(tot_d.sum() - tot_r.sum()) / tot_r.sum()
    e_out = rel_err(last["outflw"], fin["outflw"])
    print(f"  storage cell {e_cell:.3e} total {e_tot:.3e} outflow {e_out:.3e}", flush=True)
    assert e_cell <= 2e-5 and e_tot <= 1e-6 and e_out <= 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["g_allwet", "g_hourly"])
def test_dropin_device_forcing_put(name, tmp_path):
    """LINTERP forcing (an input matrix onto a 5-degree input grid, runoff fields
    with missing values) in cmf_dropin_devdiag: CMF_HIP_FORCING_PUT, which runs
    ROFF_INTERP on the device, and the same from fields kept resident in HBM
    (CMF_HIP_FORCING=resident), against the reference's CMF_FORCING_PUT on the host
    (CMF_HIP_FORCING=host). The interpolation is bit-identical and the physics the
    same, so every output file of the three runs is byte-identical; the runoff
    output shows that the forcing arrived."""
    path = os.path.join(BUILD, "cmf_dropin_devdiag")
    if not os.path.exists(path):
        pytest.skip("cmf_dropin_devdiag not built")
    from cmf import synth
    from golden_cases import setup
    print(f"drop-in device forcing put, {name}", flush=True)
    net, ro, ref, fin = load(name)
    nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac = ref["spec"]
    m = setup(name)[0]
    inpx, inpy, inpa, nxin, nyin = synth.make_inpmat(net, (m.seqx, m.seqy), 5.0)
    grids = synth.make_runoff_grid(nxin, nyin, 4, seed, mean_mm_day=3.0 * mean, missing=0.05)
    variables = OUTPUT_VARS if lpth else OUTPUT_VARS.replace(",pthflw", "")
    got = {}
    for put in ("host", "device", "resident"):
        d = os.path.join(str(tmp_path), put)
        synth.write_reference_inputs(net, d, grids, dt=dt, ifrq_inp=int(dt // 3600), ndays=days,
                                     lpthout=lpth, llevee=llev, ldamout=ldam, outputs=variables,
                                     inpmat=(inpx, inpy, inpa, nxin, nyin, m.seqx, m.seqy))
        env = dict(os.environ, CMF_HIP_FORCING=put)
        r = subprocess.run([path, "runoff.bin"], cwd=d, capture_output=True, text=True, timeout=300, env=env)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        got[put] = {}
        for v in variables.split(","):
            files = [f for f in os.listdir(d) if f.startswith(v + "_cmf")]
            assert len(files) == 1, (put, v, files)
            got[put][v] = np.fromfile(os.path.join(d, files[0]), np.uint8)
    for v in variables.split(","):
        same = [np.array_equal(got["host"][v], got[put][v]) for put in ("device", "resident")]
        print(f"  {v}: {got['host'][v].size} bytes, device / resident identical {same}", flush=True)
        assert all(same), v
    roff = got["device"]["runoff"].view(np.float32)
    assert np.isfinite(roff).all() and roff.max() > 0.0
