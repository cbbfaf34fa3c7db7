"""Runoff interpolation on the device (cmf_hip_set_inpmat / cmf_hip_set_runoff_grid)
against the oracle's ROFF_INTERP, itself bit-identical to the reference's
CMF_FORCING_PUT (tests/test_forcing.py): a model fed through the device
interpolation and a model fed the oracle's D2RUNOFF through cmf_hip_set_runoff
end two outer steps in bit-identical state (FP64 and mixed, single model and a
2-member ensemble, whose members both receive the cell's value)."""
import numpy as np
import pytest

from cases import prepare, upload_keys


def _inputs(m, net, seed):
    from cmf import synth
    from oracle.oracle import roff_interp
    inpx, inpy, inpa, nxin, nyin = synth.make_inpmat(net, (m.seqx, m.seqy), 5.0)
    n = m.n
    rng = np.random.default_rng(seed)
    k = rng.choice(n, 4, replace=False)
    inpx[0, k[:2]] = nxin + 3                               # outside the grid: skipped
    inpx[0, k[2:]] = 0                                      # empty entry
    grids = synth.make_runoff_grid(nxin, nyin, 2, seed, mean_mm_day=3.0, missing=0.1)
    dru = 86400.0 * 1000.0
    ref = [roff_interp(inpx, inpy, inpa, nxin, nyin, g, dru, synth.RMIS) for g in grids]
    return (inpx, inpy, inpa, nxin, nyin, dru), grids, ref


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("members", [1, 2])
def test_device_interpolation_matches_oracle(precision, members):
    from cmf.model import CMFModel
    m, net, st, _ = prepare("plain", 1)
    mat, grids, ref = _inputs(m, net, 7)
    kw = {} if members == 1 else dict(members=members, dom_cells=[m.n])
    up = upload_keys(st)
    if members > 1:
        up = {k: np.stack([v] * members) for k, v in up.items()}
    a = CMFModel(m, precision, **kw)
    b = CMFModel(m, precision, **kw)
    a.set_state(up)
    b.set_state(up)
    a.set_inpmat(*mat[:5], drofunit=mat[5])
    for g, r in zip(grids, ref):
        print(f"{precision} members={members}: runoff total {r.sum():.6e}", flush=True)
        a.set_runoff_grid(g)
        b.set_runoff(r if members == 1 else np.stack([r] * members))
        if members == 1:
            na, nb = a.physics_advance(86400.0), b.physics_advance(86400.0)
        else:
            na, nb = a.ensemble_advance(86400.0)[0], b.ensemble_advance(86400.0)[0]
        assert np.array_equal(np.asarray(na), np.asarray(nb))
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()


@pytest.mark.gpu
def test_inpmat_errors():
    from cmf.model import CMFError, CMFModel
    m, net, st, _ = prepare("plain", 1)
    mat, grids, ref = _inputs(m, net, 8)
    g = CMFModel(m, "fp64")
    with pytest.raises(CMFError):
        g.nxin, g.nyin = mat[3], mat[4]
        g.set_runoff_grid(grids[0])                          # no input matrix yet
    bad_y = mat[1].copy()
    bad_y[0, np.flatnonzero(mat[0][0] > 0)[0]] = 0          # IXIN > 0 with IYIN = 0
    with pytest.raises(CMFError):
        g.set_inpmat(mat[0], bad_y, mat[2], mat[3], mat[4])
    g.set_inpmat(*mat[:5])
    with pytest.raises(CMFError):
        g.set_inpmat(*mat[:5])                               # only once
    g.close()


@pytest.mark.gpu
def test_resident_records_match_single_fields():
    """Fields kept resident (cmf_hip_set_runoff_records, used out of order and
    more than once through cmf_hip_use_runoff_record) give the state that one
    cmf_hip_set_runoff_grid per step gives, bit for bit; a record index outside
    the set is an error, and so is a record before any set."""
    from cmf.model import CMFError, CMFModel
    m, net, st, _ = prepare("plain", 1)
    mat, grids, ref = _inputs(m, net, 9)
    order = [1, 0, 1]
    a = CMFModel(m, "fp64")
    b = CMFModel(m, "fp64")
    for g in (a, b):
        g.set_state(upload_keys(st))
        g.set_inpmat(*mat[:5], drofunit=mat[5])
    with pytest.raises(CMFError):
        a.use_runoff_record(0)                               # no records yet
    a.set_runoff_records(np.stack(grids))
    for k in order:
        print(f"record {k}", flush=True)
        a.use_runoff_record(k)
        b.set_runoff_grid(grids[k])
        assert a.physics_advance(86400.0) == b.physics_advance(86400.0)
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    with pytest.raises(CMFError):
        a.use_runoff_record(len(grids))
    a.close()
    b.close()
