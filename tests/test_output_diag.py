"""Output-period diagnostics (CMF_DIAG_RESET/AVEMAX/GETAVE_OUTPUT,
src/cmf_calc_diag_mod.F90:222-356), CPU side: the C oracle against the
reference's own routines (oracle/_ref serial64 build of the unmodified Fortran
plus oracle/ref_diag.F90) over several outer steps of random substep averages
with paths and dams: bit-identical, NADD_out included."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "serial64",
                   "cmf_ref_diag")
A_ORDER = ("rivout", "fldout", "rivvel", "outflw", "pthout", "runoff", "daminf", "outflw_max", "rivdph_max",
           "storge_max")
O_ORDER = ("rivout", "fldout", "outflw", "rivvel", "pthout", "runoff", "daminf", "outflw_max", "rivdph_max",
           "storge_max")


def _steps(n, npth, nlev, nstep, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nstep):
        a = {k: rng.lognormal(0.0, 2.0, n) * rng.choice([-1.0, 1.0], n) for k in A_ORDER}
        for k in ("rivdph_max", "storge_max"):
            a[k] = np.abs(a[k])
        a["pthflw"] = rng.normal(0.0, 10.0, nlev * npth)
        out.append((float(rng.choice([3600.0, 86400.0, 1800.0])), a))
    return out


def test_oracle_reset_is_zero_and_maxima_keep_largest():
    from oracle.oracle import OracleOutput
    o = OracleOutput(5, 2, True)
    a = {k: np.full(5 if k != "pthflw" else 2, 2.0) for k in O_ORDER + ("pthflw",)}
    a["outflw_max"][:] = -7.0
    o.avemax(a, 10.0)
    assert np.allclose(o.arr["rivout"], 20.0) and np.allclose(o.arr["outflw_max"], 7.0)
    assert o.getave() == 10.0 and np.allclose(o.arr["rivout"], 2.0) and np.allclose(o.arr["pthflw"], 2.0)


@pytest.mark.skipif(not os.path.exists(REF), reason="reference diag driver not built (oracle/build_ref.sh)")
@pytest.mark.parametrize("ldamout", [0, 1])
def test_oracle_matches_reference_output_diag(ldamout):
    from oracle.oracle import OracleOutput
    n, npth, nlev, nstep = 257, 13, 3, 5
    steps = _steps(n, npth, nlev, nstep, 3 + ldamout)
    o = OracleOutput(n, npth * nlev, bool(ldamout))
    for dt, a in steps:
        o.avemax(a, dt)
    nadd = o.getave()
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([n, npth, nlev, nstep, ldamout], np.int32).tofile(f)
            for dt, a in steps:
                np.array([dt]).tofile(f)
                for k in A_ORDER:
                    a[k].astype(np.float64).tofile(f)
                a["pthflw"].astype(np.float64).tofile(f)      # D1PTHFLW(NPTHOUT,NPTHLEV): path fastest
        r = subprocess.run([REF, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        raw = np.fromfile(fout, np.float64)
    assert raw[0] == nadd
    ref = dict(zip(O_ORDER, raw[1:1 + 10 * n].reshape(10, n)))
    ref["pthflw"] = raw[1 + 10 * n:]
    for k in O_ORDER + ("pthflw",):
        if k == "daminf" and not ldamout:
            continue
        assert np.array_equal(o.arr[k].view(np.uint64), ref[k].view(np.uint64)), k
