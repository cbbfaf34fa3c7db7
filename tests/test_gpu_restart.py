"""Binary restart straight from / into device state (cmf_hip_restart_write /
cmf_hip_restart_read, WRTE_REST_BIN / READ_REST_BIN of
src/cmf_ctrl_restart_mod.F90) against the reference's own end-of-run restart
file of two golden cases (tests/golden/*_restart.npz, written by the reference
Fortran; the second has paths, levees and dams, so every optional record and the
.pth file):
* write: same file layout and size, the reference's DMIS fill on every grid cell
  without a river cell, every record within the golden bound of the device
  physics (1e-10 relative of the record's magnitude);
* read: the reference's file loaded into a fresh handle gives back its values
  bit for bit;
* single precision: a float32 file round-trips to the float32 rounding of the
  device state."""
import os
import tempfile

import numpy as np
import pytest

from cases import upload_keys
from golden_cases import runoff_for_step, setup

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-10


def _records(m):
    k = ["rivsto", "fldsto", "rivout_pre", "fldout_pre", "rivdph_pre", "fldsto_pre"]
    if m.opts.ldamout:
        k.append("damsto")
    if m.opts.llevee:
        k.append("levsto")
    return k


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["g_plain", "g_allwet"])
def test_restart_matches_reference(name):
    from cmf.model import CMFModel
    print(f"restart case {name}", flush=True)
    m, net, st, rvec, ref, fin, dt, nsteps = setup(name)
    z = np.load(os.path.join(HERE, "golden", f"{name}_restart.npz"), allow_pickle=False)
    keys = _records(m)
    ng = m.nx * m.ny
    want = z["bin"].view(np.float64).reshape(len(keys), ng)
    gm = CMFModel(m, "fp64")
    gm.set_state(upload_keys(st))
    for step in range(1, nsteps + 1):
        gm.set_runoff(runoff_for_step(rvec, step, dt))
        gm.physics_advance(dt)
    cell = (m.seqy.astype(np.int64) - 1) * m.nx + (m.seqx.astype(np.int64) - 1)
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "restart.bin")
        gm.restart_write(f)
        got = np.fromfile(f, np.float64)
        assert got.size == want.size
        got = got.reshape(len(keys), ng)
        empty = np.ones(ng, bool)
        empty[cell] = False
        assert np.all(got[:, empty] == 1e20) and np.all(want[:, empty] == 1e20)
        for r, k in enumerate(keys):
            scale = max(np.abs(want[r, cell]).max(), 1e-30)
            e = np.abs(got[r, cell] - want[r, cell]).max() / scale
            print(f"  {k}: rel err {e:.3e}", flush=True)
            assert e <= TOL, k
        if m.opts.lpthout:
            gp = np.fromfile(f + ".pth", np.float64)
            wp = z["pth"].view(np.float64)
            assert gp.size == wp.size == m.npth * m.npthlev
            assert np.abs(gp - wp).max() <= TOL * max(np.abs(wp).max(), 1e-30)
        # single precision file: float32 rounding of the device state
        f4 = os.path.join(d, "restart4.bin")
        gm.restart_write(f4, lrestdbl=False)
        g4 = np.fromfile(f4, np.float32).reshape(len(keys), ng)
        assert np.array_equal(g4[:, cell], got[:, cell].astype(np.float32))
        assert np.all(g4[:, empty] == np.float32(1e20))
        # read: the reference's own file into a fresh handle, values back bit for bit
        fr = os.path.join(d, "reference.bin")
        z["bin"].tofile(fr)
        if m.opts.lpthout:
            z["pth"].tofile(fr + ".pth")
        g2 = CMFModel(m, "fp64")
        g2.restart_read(fr)
        back = g2.get_state(keys + (["pthflw_pre"] if m.opts.lpthout else []))
        for r, k in enumerate(keys):
            assert np.array_equal(back[k], want[r, cell]), k
        if m.opts.lpthout:
            assert np.array_equal(back["pthflw_pre"].reshape(-1), z["pth"].view(np.float64))
        g2.restart_read(f4, lrestdbl=False)
        back4 = g2.get_state(keys)
        for r, k in enumerate(keys):
            assert np.array_equal(back4[k], got[r, cell].astype(np.float32).astype(np.float64)), k
        g2.close()
    gm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["g_allwet_sto", "g_allwet_rst"])
def test_run_from_restart_matches_reference(name):
    """A run started from the reference's restart file, storage only (LSTOONLY:
    cmf_hip_restart_read, cmf_hip_physics_fldstg, cmf_hip_calc_outpre, as
    CMF_DRV_INIT does, drv_control:273-278) or full, against the reference's own
    run from the same file (paths, levees, dams): the state after the
    initialisation (OUTPRE's previous-step flows included) within 1e-12, then
    identical NT at every step and the final fields within 1e-10. The first
    substep's RIVVEL reads D2RIVOUT = 0 (CMF_RESTART_INIT zeroes it), which the
    final RIVVEL average carries."""
    from cases import rel_err
    from cmf.model import CMFModel
    from golden_cases import restart_setup
    print(f"run from restart {name}", flush=True)
    m, net, st, binb, pthb, lstoonly, rvec, ref, step0, fin, dt, nsteps = restart_setup(name)
    gm = CMFModel(m, "fp64")
    gm.set_state(upload_keys(st))
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "restart_in.bin")
        with open(f, "wb") as fh:
            fh.write(binb)
        if pthb is not None:
            with open(f + ".pth", "wb") as fh:
                fh.write(pthb)
        gm.restart_read(f, lstoonly=lstoonly)
    if lstoonly:
        gm.physics_fldstg()
        gm.calc_outpre()
    g = gm.get_state(list(step0))
    for k, v in step0.items():
        e = rel_err(g[k], v)
        print(f"  after init {k}: rel err {e:.3e}", flush=True)
        assert e <= 1e-12, f"{name} after init: {k} rel err {e:.3e}"
    nts = []
    for step in range(1, nsteps + 1):
        gm.set_runoff(runoff_for_step(rvec, step, dt))
        nts.append(gm.physics_advance(dt))
    g = gm.get_state(list(fin) + ["rivvel_avg"])
    gm.close()
    assert nts == ref["nt"]
    for k, v in fin.items():
        e = rel_err(g[k], v)
        print(f"  final {k}: rel err {e:.3e}", flush=True)
        assert e <= TOL, f"{name}: {k} rel err {e:.3e}"
