#!/usr/bin/env python3
"""Writes the golden fixtures of the reference path (run in the build container,
where /root/reference and oracle/_ref exist).

For each case: a seeded synthetic network (inputs stored in the fixture), the
reference Fortran (oracle/_ref/serial64/cmf_ref: FP64, flang -O2
-ffp-contract=off, serial) run on it through oracle/ref_driver.F90, and its
outputs reduced to:
  * map tables after initialisation: SHA-1 of each array (float64/int32 bytes);
  * per outer step: NT, SHA-1 of every state array, the 10 budget sums;
  * the full final state of the main prognostic/diagnostic fields.
and, for the cases of RESTART_CASES, the reference's end-of-run binary restart file
and, for OUTPUT_CASES, its daily binary output files.
Usage: python tests/golden/make_golden.py [case ... | case_restart ... | case_output ...]
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "cama-flood_v4_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

from cmf import synth  # noqa: E402
from oracle import refdump  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "serial64", "cmf_ref")

# name: nx, ny, seed, days, dt, lpthout, llevee, ldamout, runoff mean, levee frac
GOLDEN = {
    "g_plain":  (90, 45, 11, 4, 86400.0, False, False, False, 1.5, 0.0),
    "g_pth":    (90, 45, 12, 4, 86400.0, True, False, False, 1.5, 0.0),
    "g_levwet": (90, 45, 13, 4, 86400.0, False, True, False, 60.0, 0.5),
    "g_dam":    (90, 45, 14, 4, 86400.0, False, False, True, 1.5, 0.0),
    "g_allwet": (90, 45, 15, 4, 86400.0, True, True, True, 60.0, 0.5),
    "g_hourly": (90, 45, 16, 1, 3600.0, True, True, True, 20.0, 0.3),
    "c1_1deg":  (360, 180, 2026, 30, 86400.0, False, False, False, 1.5, 0.0),
    "g_kine":   (90, 45, 17, 4, 86400.0, False, False, True, 20.0, 0.0),
    "g_dam_yby": (90, 45, 18, 4, 86400.0, False, False, True, 20.0, 0.0),
    "g_dam_h22": (90, 45, 19, 4, 86400.0, False, False, True, 4.0, 0.0),
    "g_slpmix": (90, 45, 20, 4, 86400.0, True, True, False, 20.0, 0.3),
}
# cases run with LSLPMIX (CMF_CALC_OUTFLW_KINEMIX, src/cmf_opt_outflw_mod.F90:118-300)
# and this fraction of the steepest river cells in the slope mask. Without dams: on
# this network with dams, one reservoir enters a substep-to-substep oscillation of
# the Yamazaki-Funato rule around ConVol (d(outflow)/d(storage) * DT ~ 8), which
# multiplies a one-ulp difference of a fractional power several times per substep;
# the oracle still matches the reference there bit for bit, a device whose pow
# rounds differently by one ulp cannot.
SLPMIX_CASES = {"g_slpmix": 0.3}
# cases whose dams carry construction years (LDAMYBY, damout:216-225): cycled over
# the dams, against the start year 2001 of the reference inputs: 0 (no year, active),
# 1990 (active), 2001 (new this year), 2005 (not yet built)
YBY_YEARS = {"g_dam_yby": (0, 1990, 2001, 2005)}
# cases run with LKINE = .TRUE. (kinematic-wave outflow, src/cmf_opt_outflw_mod.F90:27-127)
KINE_CASES = ("g_kine",)
# cases run with LDAMH22 = .TRUE. (Hanazaki-2022 rule, damout:354-378). Their dams
# have no flood-control volume (FldVol = 0, so EmeVol = ConVol): the flood-control
# zone, the only one that reads R_VolUpa, is empty, and the reference's R_VolUpa,
# written only at index NDAM (damout:204) and otherwise uninitialised, is never read.
H22_CASES = ("g_dam_h22",)
FINAL_KEYS = ["rivsto", "fldsto", "rivout", "fldout", "outflw", "rivvel"]


def final_keys(ldam):
    """Fields of the final-state fixture: D2DAMINF_aAVG too where dams operate."""
    return FINAL_KEYS + (["daminf_avg"] if ldam else [])


def sha(a):
    a = np.ascontiguousarray(a)
    if a.dtype.kind == "f":
        a = a.astype("<f8")
    elif a.dtype.kind in "iu":
        a = a.astype("<i4")
    return hashlib.sha1(a.tobytes()).hexdigest()


def network_for(name):
    nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac = GOLDEN[name]
    net = synth.make_network(nx, ny, seed=seed)
    if lpth:
        synth.add_bifurcations(net, max(10, net.ncells // 20), 5, seed)
    if llev:
        synth.add_levees(net, levfrac, seed, hmin=0.5, hmax=3.0, fmin=0.05, fmax=0.5)
    if ldam:
        synth.add_dams(net, max(3, net.ncells // 100), seed, years=YBY_YEARS.get(name))
        if name in H22_CASES:
            for r in net.dams:
                r["fldvol"] = 0.0
    if name in SLPMIX_CASES:
        synth.add_slope_mask(net, SLPMIX_CASES[name])
    nrec = min(days, 5) if dt >= 86400 else 4   # records are cycled by the driver
    ro = synth.make_runoff(net, nrec, seed, mean_mm_day=mean)
    return net, ro


def save_network(path, net, ro):
    arrays = {k: getattr(net, k) for k in ("nextx", "nexty", "elevtn", "nxtdst", "rivlen", "rivwth",
                                           "rivhgt", "rivman", "ctmare", "fldhgt")}
    if net.levhgt is not None:
        arrays["levhgt"] = net.levhgt
        arrays["levfrc"] = net.levfrc
    if net.slpmix_mask is not None:
        arrays["slpmix_mask"] = net.slpmix_mask
    arrays["runoff"] = ro
    meta = dict(nx=net.nx, ny=net.ny, west=net.west, east=net.east, north=net.north, south=net.south,
                bifprm=net.bifprm, npthlev=net.npthlev, dams=net.dams)
    np.savez_compressed(path, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8), **arrays)


def load_network(path):
    z = np.load(path, allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    up = np.zeros((meta["ny"], meta["nx"]))
    net = synth.SynthNetwork(meta["nx"], meta["ny"], meta["west"], meta["east"], meta["north"],
                             meta["south"], z["nextx"], z["nexty"], z["elevtn"], z["nxtdst"], z["rivlen"],
                             z["rivwth"], z["rivhgt"], z["rivman"], z["ctmare"], z["fldhgt"], up)
    if meta["bifprm"]:
        net.bifprm = [tuple(r[:7]) + (list(r[7]),) for r in meta["bifprm"]]
        net.npthlev = meta["npthlev"]
    if "levhgt" in z:
        net.levhgt = z["levhgt"]
        net.levfrc = z["levfrc"]
    net.dams = meta["dams"]
    if "slpmix_mask" in z:
        net.slpmix_mask = z["slpmix_mask"]
    return net, z["runoff"]


def make(name):
    nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac = GOLDEN[name]
    net, ro = network_for(name)
    inp = os.path.join(HERE, f"{name}_inputs.npz")
    save_network(inp, net, ro)
    net, ro = load_network(inp)            # the fixture, not the generator, is the source
    with tempfile.TemporaryDirectory() as d:
        ifrq = int(dt // 3600)
        synth.write_reference_inputs(net, d, ro, dt=dt, ifrq_inp=ifrq, ndays=days,
                                     lpthout=lpth, llevee=llev, ldamout=ldam, lkine=name in KINE_CASES,
                                     ldamh22=name in H22_CASES, slpmix_mask=net.slpmix_mask)
        r = subprocess.run([REF, "runoff.bin", "dump.bin", "1"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stderr)
        mp, states = refdump.read_dump(os.path.join(d, "dump.bin"), lpth, llev, ldam)
        nts = refdump.count_substeps(os.path.join(d, "log_CaMa.txt"))
    keys_map = ["seqx", "seqy", "next", "i2mask", "rivstomax", "rivelv", "fldstomax", "fldgrd"]
    if lpth:
        keys_map += ["pth_elv", "pth_dst", "pth_wth"]
    if llev:
        keys_map += list(refdump.LEV_D)
    if ldam:
        keys_map += ["i1dam"]
    out = dict(case=name, spec=GOLDEN[name], lkine=name in KINE_CASES, ldamh22=name in H22_CASES,
               lslpmix=name in SLPMIX_CASES, n=mp["n"], nseqriv=mp["nseqriv"], nt=nts,
               map_sha={k: sha(mp[k]) for k in keys_map}, steps=[])
    skeys = refdump.STATE_P + refdump.STATE_D + (["levsto", "levdph"] if llev else []) + \
        (["damsto", "daminf", "daminf_avg"] if ldam else []) + (["pthflw", "pthflwsum"] if lpth else [])
    for s in states:
        out["steps"].append(dict(step=s["step"], sha={k: sha(s[k]) for k in skeys},
                                 glb={k: float(v) for k, v in s["glb"].items()}))
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(out, f, indent=0)
    fin = states[-1]
    np.savez_compressed(os.path.join(HERE, f"{name}_final.npz"), **{k: fin[k] for k in final_keys(ldam)})
    print(f"{name}: n={mp['n']} steps={len(states) - 1} sum NT={sum(nts)}")


RESTART_CASES = ("g_plain", "g_allwet")


def make_restart(name):
    """The reference's own end-of-run binary restart (CMF_RESTART_WRITE, FP64
    records) of a golden case, as raw bytes."""
    nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac = GOLDEN[name]
    net, ro = load_network(os.path.join(HERE, f"{name}_inputs.npz"))
    with tempfile.TemporaryDirectory() as d:
        synth.write_reference_inputs(net, d, ro, dt=dt, ifrq_inp=int(dt // 3600), ndays=days,
                                     lpthout=lpth, llevee=llev, ldamout=ldam, restart=True)
        r = subprocess.run([REF, "runoff.bin", "NONE", "0"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stderr)
        files = sorted(f for f in os.listdir(d) if f.startswith("restart") and f.endswith(".bin"))
        if len(files) != 1:
            raise RuntimeError(f"expected one restart file, found {files}")
        rec = dict(name=np.frombuffer(files[0].encode(), np.uint8),
                   bin=np.fromfile(os.path.join(d, files[0]), np.uint8))
        if lpth:
            rec["pth"] = np.fromfile(os.path.join(d, files[0] + ".pth"), np.uint8)
    np.savez_compressed(os.path.join(HERE, f"{name}_restart.npz"), **rec)
    print(f"{name}: restart {files[0]} ({rec['bin'].size} bytes)")


# runs that start from the end-of-run restart file of a golden case: storage only
# (LSTOONLY: CMF_RESTART_INIT reads P2RIVSTO/P2FLDSTO, then CMF_PHYSICS_FLDSTG and
# CMF_CALC_OUTPRE rebuild the previous-step flows, drv_control:273-278) and the
# full restart (the t-1 records read from the file). The dump after the
# initialisation holds OUTPRE's result.
RESTART_IN = {"g_allwet_sto": ("g_allwet", True, 2), "g_allwet_rst": ("g_allwet", False, 2)}


def restart_input(base, lstoonly):
    """(file bytes, .pth bytes or None) of the restart file a RESTART_IN run
    reads: the base case's end-of-run restart; for LSTOONLY its records in the
    storage-only layout of WRTE_REST_BIN (restart:418-447): P2RIVSTO, P2FLDSTO,
    then P2DAMSTO (LDAMOUT) and P2LEVSTO (LLEVEE), the four t-1 records left out."""
    nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac = GOLDEN[base]
    z = np.load(os.path.join(HERE, f"{base}_restart.npz"), allow_pickle=False)
    raw = z["bin"].tobytes()
    rec = 8 * nx * ny
    nfull = 6 + int(ldam) + int(llev)
    if len(raw) != nfull * rec:
        raise RuntimeError(f"{base}: restart file has {len(raw)} bytes, expected {nfull} records of {rec}")
    recs = [raw[k * rec:(k + 1) * rec] for k in range(nfull)]
    if lstoonly:
        return b"".join(recs[:2] + recs[6:]), None
    return raw, (z["pth"].tobytes() if lpth else None)


def make_restart_in(name):
    base, lstoonly, days = RESTART_IN[name]
    nx, ny, seed, _days, dt, lpth, llev, ldam, mean, levfrac = GOLDEN[base]
    net, ro = load_network(os.path.join(HERE, f"{base}_inputs.npz"))
    binb, pthb = restart_input(base, lstoonly)
    with tempfile.TemporaryDirectory() as d:
        synth.write_reference_inputs(net, d, ro, dt=dt, ifrq_inp=int(dt // 3600), ndays=days,
                                     lpthout=lpth, llevee=llev, ldamout=ldam, restart_in="restart_in.bin",
                                     lstoonly=lstoonly)
        with open(os.path.join(d, "restart_in.bin"), "wb") as f:
            f.write(binb)
        if pthb is not None:
            with open(os.path.join(d, "restart_in.bin.pth"), "wb") as f:
                f.write(pthb)
        r = subprocess.run([REF, "runoff.bin", "dump.bin", "1"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr)
        mp, states = refdump.read_dump(os.path.join(d, "dump.bin"), lpth, llev, ldam)
        nts = refdump.count_substeps(os.path.join(d, "log_CaMa.txt"))
    spec = [nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac]
    out = dict(case=name, base=base, lstoonly=lstoonly, spec=spec, n=mp["n"], nseqriv=mp["nseqriv"], nt=nts,
               steps=[])
    skeys = refdump.STATE_P + refdump.STATE_D + (["levsto", "levdph"] if llev else []) + \
        (["damsto", "daminf", "daminf_avg"] if ldam else []) + (["pthflw", "pthflwsum"] if lpth else [])
    for s_ in states:
        out["steps"].append(dict(step=s_["step"], sha={k: sha(s_[k]) for k in skeys},
                                 glb={k: float(v) for k, v in s_["glb"].items()}))
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(out, f, indent=0)
    s0 = states[0]
    k0 = ["rivsto", "fldsto", "rivout", "fldout", "rivout_pre", "fldout_pre", "rivdph_pre", "fldsto_pre", "sfcelv",
          "rivdph"] + (["pthflw", "pthflw_pre"] if lpth else []) + (["damsto"] if ldam else []) + \
        (["levsto"] if llev else [])
    np.savez_compressed(os.path.join(HERE, f"{name}_step0.npz"), **{k: s0[k] for k in k0})
    np.savez_compressed(os.path.join(HERE, f"{name}_final.npz"),
                        **{k: states[-1][k] for k in final_keys(ldam)})
    print(f"{name}: n={mp['n']} steps={len(states) - 1} sum NT={sum(nts)} (from {base}, lstoonly={lstoonly})")


# the reference in its default single-precision mode (-DSinglePrec_CMF, serial,
# no FMA contraction) on the inputs of golden cases: the counterpart of the
# single-precision drop-in (a JPRB = 4 host over the mixed device path)
REF32 = os.path.join(ROOT, "oracle", "_ref", "serial32", "cmf_ref")
SP_CASES = {"g_allwet_sp": "g_allwet", "g_hourly_sp": "g_hourly"}


def make_sp(name):
    base = SP_CASES[name]
    nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac = GOLDEN[base]
    net, ro = load_network(os.path.join(HERE, f"{base}_inputs.npz"))
    with tempfile.TemporaryDirectory() as d:
        synth.write_reference_inputs(net, d, ro, dt=dt, ifrq_inp=int(dt // 3600), ndays=days,
                                     lpthout=lpth, llevee=llev, ldamout=ldam)
        r = subprocess.run([REF32, "runoff.bin", "dump.bin", "1"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr)
        mp, states = refdump.read_dump(os.path.join(d, "dump.bin"), lpth, llev, ldam)
        nts = refdump.count_substeps(os.path.join(d, "log_CaMa.txt"))
    out = dict(case=name, base=base, spec=GOLDEN[base], n=mp["n"], nt=nts,
               glb=[{k: float(v) for k, v in s_["glb"].items()} for s_ in states])
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(out, f, indent=0)
    np.savez_compressed(os.path.join(HERE, f"{name}_final.npz"), **{k: states[-1][k] for k in FINAL_KEYS})
    print(f"{name}: n={mp['n']} steps={len(states) - 1} sum NT={sum(nts)} (single precision, from {base})")


OUTPUT_CASES = ("g_pth",)
OUTPUT_VARS = ("rivout,fldout,outflw,rivvel,pthout,runoff,maxflw,maxdph,maxsto,rivsto,fldsto,rivdph,"
               "flddph,fldfrc,fldare,sfcelv,storge,pthflw")


def make_output(name):
    """The reference's own daily binary output files (CMF_OUTPUT_WRITE: float32
    NX*NY records, pthflw as NPTHOUT*NPTHLEV records) of a golden case."""
    nx, ny, seed, days, dt, lpth, llev, ldam, mean, levfrac = GOLDEN[name]
    net, ro = load_network(os.path.join(HERE, f"{name}_inputs.npz"))
    rec = {}
    with tempfile.TemporaryDirectory() as d:
        synth.write_reference_inputs(net, d, ro, dt=dt, ifrq_inp=int(dt // 3600), ndays=days,
                                     lpthout=lpth, llevee=llev, ldamout=ldam, outputs=OUTPUT_VARS)
        r = subprocess.run([REF, "runoff.bin", "NONE", "0"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stderr)
        for v in OUTPUT_VARS.split(","):
            files = [f for f in os.listdir(d) if f.startswith(v + "_cmf")]
            if len(files) != 1:
                raise RuntimeError(f"{v}: expected one output file, found {files}")
            rec[v] = np.fromfile(os.path.join(d, files[0]), np.uint8)
    np.savez_compressed(os.path.join(HERE, f"{name}_output.npz"), **rec)
    print(f"{name}: output files of {len(rec)} variables")


if __name__ == "__main__":
    names = sys.argv[1:] or list(GOLDEN)
    for name in names:
        if name in RESTART_IN:
            make_restart_in(name)
        elif name in SP_CASES:
            make_sp(name)
        elif name.endswith("_restart"):
            make_restart(name[:-len("_restart")])
        elif name.endswith("_output"):
            make_output(name[:-len("_output")])
        else:
            make(name)
    if not sys.argv[1:]:
        for name in RESTART_CASES:
            make_restart(name)
        for name in OUTPUT_CASES:
            make_output(name)
        for name in RESTART_IN:
            make_restart_in(name)
        for name in SP_CASES:
            make_sp(name)
