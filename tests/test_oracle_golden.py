"""The oracle against the reference Fortran: golden fixtures written by running
the reference (FP64, flang, serial; see tests/golden/make_golden.py).

Every map table built by the host initialisation and every state array of every
outer step must be bit-identical (SHA-1 of the float64 bytes), the substep counts
NT identical, and the ten water-budget sums identical."""
import numpy as np
import pytest

from golden_cases import NAMES, runoff_for_step, setup, sha


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_reference(name):
    from oracle import refdump
    from oracle.oracle import OracleModel
    m, net, st, rvec, ref, fin, dt, nsteps = setup(name)
    assert m.n == ref["n"] and m.nseqriv == ref["nseqriv"]
    mine = {"seqx": m.seqx, "seqy": m.seqy, "next": np.where(m.next >= 0, m.next + 1, m.next),
            "i2mask": m.i2mask, "rivstomax": m.rivstomax, "rivelv": m.rivelv,
            "fldstomax": m.fldstomax, "fldgrd": m.fldgrd}
    if m.opts.lpthout:
        mine.update(pth_elv=m.pth_elv, pth_dst=m.pth_dst, pth_wth=m.pth_wth)
    if m.opts.llevee:
        mine.update({k: getattr(m, k) for k in refdump.LEV_D})
    if m.opts.ldamout:
        mine["i1dam"] = m.i1dam
    for k, h in ref["map_sha"].items():
        assert sha(mine[k]) == h, f"map table {k} differs from the reference"
    om = OracleModel(m, st)
    keys = list(ref["steps"][0]["sha"])
    for k in keys:
        assert sha(om.st[k]) == ref["steps"][0]["sha"][k], f"initial {k} differs"
    for step in range(1, nsteps + 1):
        om.st["runoff"][:] = runoff_for_step(rvec, step, dt)
        nt = om.physics_advance(dt)
        assert nt == ref["nt"][step - 1], f"step {step}: NT {nt} != {ref['nt'][step - 1]}"
        rs = ref["steps"][step]
        bad = [k for k in keys if sha(om.st[k]) != rs["sha"][k]]
        assert not bad, f"step {step}: arrays differ from the reference: {bad}"
        g = om.glb
        for k, v in rs["glb"].items():
            assert g[k] == v, f"step {step}: budget sum {k} {g[k]!r} != {v!r}"
