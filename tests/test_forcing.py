"""Runoff interpolation with the input matrix (ROFF_INTERP inside CMF_FORCING_PUT,
src/cmf_ctrl_forcing_mod.F90:757-790), CPU side.

* The C oracle against the reference's own CMF_FORCING_PUT (oracle/_ref
  serial64 build of the unmodified Fortran plus oracle/ref_forcing.F90) on an
  input matrix with empty entries, entries outside the grid (the reference logs
  and skips them) and RMIS cells: bit-identical.
* Mass conservation: the interpolated total equals the area-weighted sum of the
  field over the entries in the grid that are not RMIS.
* The synthetic input matrix covers every cell with its own area.
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
