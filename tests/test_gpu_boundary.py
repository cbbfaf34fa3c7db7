"""Sea-level boundary on the device (CMF_BOUNDARY_UPDATE,
src/cmf_ctrl_boundary_mod.F90:267-305). A model whose D2DWNELV is replaced through
cmf_hip_set_dwnelv, or formed on the device by cmf_hip_update_boundary from
D2MEANSL and D2SEALEV, ends an outer step in bit-identical state to a model
whose map carried that D2DWNELV from the start (FP64), and the device sum
follows the reference's order (ELEVTN + MEANSL) + SEALEV."""
import dataclasses

import numpy as np
import pytest

from cases import prepare, upload_keys


@pytest.mark.gpu
def test_boundary_update_matches_map_upload():
    from cmf.model import CMFModel
    m, net, st, rvec = prepare("plain", 1)
    rng = np.random.default_rng(3)
    meansl = rng.normal(0.0, 0.3, m.n)
    sealev = rng.normal(0.0, 0.5, m.n)
    want = (m.elevtn.astype(np.float64) + meansl) + sealev
    m2 = dataclasses.replace(m, dwnelv=want)
    ref = CMFModel(m2, "fp64")
    a = CMFModel(m, "fp64")
    b = CMFModel(m, "fp64")
    orig = CMFModel(m, "fp64")
    for g in (ref, a, b, orig):
        g.set_state(upload_keys(st))
        g.set_runoff(rvec[0])
    a.set_dwnelv(want)
    b.update_boundary(meansl, sealev)
    for g in (ref, a, b, orig):
        g.physics_advance(86400.0)
    s0, sa, sb, so = ref.get_state(), a.get_state(), b.get_state(), orig.get_state()
    # the boundary matters: the mouths' outflow moved with it
    mouths = m.next < 0
    assert not np.array_equal(s0["rivout"][mouths], so["rivout"][mouths])
    for k in s0:
        assert np.array_equal(s0[k], sa[k]), k
        assert np.array_equal(s0[k], sb[k]), k
    for g in (ref, a, b, orig):
        g.close()
