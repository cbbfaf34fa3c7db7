"""Runoff interpolation with the input matrix (ROFF_INTERP inside CMF_FORCING_PUT,
src/cmf_ctrl_forcing_mod.F90:757-790), CPU side.

* The C oracle against the reference's own CMF_FORCING_PUT (oracle/_ref
  serial64 build of the unmodified Fortran plus oracle/ref_forcing.F90) on an
  input matrix with empty entries, entries outside the grid (the reference logs
  and skips them) and RMIS cells: bit-identical.
* Mass conservation: the interpolated total equals the area-weighted sum of the
  field over the entries in the grid that are not RMIS.
* The synthetic input matrix covers every cell with its own area.
* The input matrix file written for the reference has the layout its binary
  reader expects, and the reference's initialisation reads it.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from cmf import maps, synth

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "serial64",
                   "cmf_ref_forcing")


def _case(seed=5, res_in=5.0):
    net = synth.make_network(120, 60, seed=seed)          # 3 degree cells, 5 degree input grid
    m = maps.build_map(net, maps.Options())
    inpx, inpy, inpa, nxin, nyin = synth.make_inpmat(net, (m.seqx, m.seqy), res_in)
    grid = synth.make_runoff_grid(nxin, nyin, 1, seed, missing=0.1)[0]
    return net, m, inpx, inpy, inpa, nxin, nyin, grid


def _edge_cases(inpx, inpy, nxin, nyin, rng):
    """Entries past the grid edge (skipped by the reference) and emptied entries."""
    inpx, inpy = inpx.copy(), inpy.copy()
    n = inpx.shape[1]
    k = rng.choice(n, 5, replace=False)
    inpx[0, k[:3]] = nxin + 1 + np.arange(3)            # IXIN > NXIN
    inpy[0, k[3:]] = nyin + 2                            # IYIN > NYIN
    e = rng.choice(n, 7, replace=False)
    inpx[0, e] = 0                                       # no entry
    return inpx, inpy


def test_synthetic_inpmat_covers_cell_areas():
    net, m, inpx, inpy, inpa, nxin, nyin, grid = _case()
    area = m.vec(net.ctmare.astype(np.float64))
    assert inpx.shape[0] <= 4 and (inpx > 0).any(axis=0).all()
    assert np.allclose(inpa.sum(axis=0), area, rtol=2e-3)
    assert inpx.max() <= nxin and inpy.max() <= nyin and inpy[inpx > 0].min() >= 1


def test_oracle_conserves_mass():
    from oracle.oracle import roff_interp
    net, m, inpx, inpy, inpa, nxin, nyin, grid = _case()
    inpx, inpy = _edge_cases(inpx, inpy, nxin, nyin, np.random.default_rng(1))
    dru = 86400.0 * 1000.0
    out = roff_interp(inpx, inpy, inpa, nxin, nyin, grid, dru, synth.RMIS)
    ok = (inpx > 0) & (inpx <= nxin) & (inpy <= nyin)
    v = np.where(ok, grid[np.clip(inpy - 1, 0, nyin - 1), np.clip(inpx - 1, 0, nxin - 1)], synth.RMIS)
    ok &= v != synth.RMIS
    want = np.where(ok, v * inpa, 0.0).sum() / dru
    assert np.isclose(out.sum(), want, rtol=1e-12)
    assert np.all(out >= 0.0)


@pytest.mark.skipif(not os.path.exists(REF), reason="reference forcing driver not built (oracle/build_ref.sh)")
def test_oracle_matches_reference_forcing_put():
    from oracle.oracle import roff_interp
    net, m, inpx, inpy, inpa, nxin, nyin, grid = _case()
    inpx, inpy = _edge_cases(inpx, inpy, nxin, nyin, np.random.default_rng(2))
    inpn, n = inpx.shape
    dru = 86400.0 * 1000.0
    rmis32 = np.float32(1e20)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([n, inpn, nxin, nyin], np.int32).tofile(f)
            np.array([dru], np.float64).tofile(f)
            np.array([rmis32], np.float32).tofile(f)
            # Fortran INPX(NSEQMAX, INPN): cell fastest = our [inpn, n] row-major
            inpx.astype(np.int32).tofile(f)
            inpy.astype(np.int32).tofile(f)
            inpa.astype(np.float64).tofile(f)
            grid.astype(np.float64).tofile(f)                # PBUFFIN(NXIN, NYIN): x fastest
        r = subprocess.run([REF, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        ref = np.fromfile(fout, np.float64)
    got = roff_interp(inpx, inpy, inpa, nxin, nyin, grid, dru, float(rmis32))
    assert ref.shape == (n,)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    assert "XXX" in r.stdout                                  # the reference logged the skipped entries


def test_inpmat_file_in_the_reference_layout():
    """write_reference_inputs with an input matrix: inpmat.bin holds 3*INPN records
    of NX*NY 4-byte values in the order CMF_INPMAT_INIT_BIN reads them (INPX, INPY,
    then INPA as float32, one record per entry, zero off the river network;
    src/cmf_ctrl_forcing_mod.F90:418-458), diminfo.txt carries NXIN/NYIN/INPN and the
    namelist switches LINTERP on. Where the reference build is present, its own
    initialisation (CMF_DRV_INIT -> CMF_INPMAT_INIT_BIN) reads the file."""
    net, m, inpx, inpy, inpa, nxin, nyin, grid = _case(seed=6)
    inpn = inpx.shape[0]
    with tempfile.TemporaryDirectory() as d:
        synth.write_reference_inputs(net, d, grid[None], dt=86400.0, ifrq_inp=24, ndays=1,
                                     inpmat=(inpx, inpy, inpa, nxin, nyin, m.seqx, m.seqy))
        raw = np.fromfile(os.path.join(d, "inpmat.bin"), np.uint8)
        assert raw.size == 3 * inpn * net.nx * net.ny * 4
        recs = raw.reshape(3 * inpn, net.ny * net.nx, 4)
        cell = (m.seqy.astype(np.int64) - 1) * net.nx + (m.seqx.astype(np.int64) - 1)
        off = np.ones(net.nx * net.ny, bool)
        off[cell] = False
        for k in range(inpn):
            x = recs[k].copy().view(np.int32).ravel()
            y = recs[inpn + k].copy().view(np.int32).ravel()
            a = recs[2 * inpn + k].copy().view(np.float32).ravel()
            assert np.array_equal(x[cell], inpx[k]) and np.array_equal(y[cell], inpy[k])
            assert np.array_equal(a[cell], inpa[k].astype(np.float32))
            assert not x[off].any() and not y[off].any() and not a[off].any()
        dim = open(os.path.join(d, "diminfo.txt")).read().split()
        assert [int(v) for v in dim[:6]] == [net.nx, net.ny, 10, nxin, nyin, inpn]
        nml = open(os.path.join(d, "input_cmf.nam")).read()
        assert "LINTERP = .TRUE." in nml and 'CINPMAT = "inpmat.bin"' in nml
        exe = os.path.join(os.path.dirname(REF), "cmf_ref")
        if os.path.exists(exe):
            # the test driver assigns D2RUNOFF from NX*NY records of its own; the
            # initialisation reads the input matrix all the same
            np.zeros(net.nx * net.ny).tofile(os.path.join(d, "runoff_map.bin"))
            r = subprocess.run([exe, "runoff_map.bin", "NONE", "0"], cwd=d, capture_output=True, text=True,
                               timeout=120)
            log = open(os.path.join(d, "log_CaMa.txt"), errors="replace").read()
            assert r.returncode == 0, r.stdout[-1500:] + r.stderr[-1500:]
            assert "INPUT MATRIX done" in log
            assert [int(v) for v in log.split("NXIN,NYIN,INPN")[1].split()[:3]] == [nxin, nyin, inpn]
