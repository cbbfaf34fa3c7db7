"""Binary output files written from device state (cmf_hip_output_write, the
binary branch of CMF_OUTPUT_WRITE, src/cmf_ctrl_output_mod.F90:474-664) against
the reference's own daily output files of the bifurcation golden case
(tests/golden/g_pth_output.npz: 18 variables, averages, maxima, instantaneous
state and the path flows, 4 daily records each). The device run uses the same
output-period sequence as the reference's CMF_DRV_ADVANCE (avemax after every
step, getave / write / reset at output time). Same file sizes, RMIS off the
river cells, and every value within one float32 ulp of the reference (the FP64
device physics agrees with the reference to ~1e-12; the float32 rounding of the
record can then differ by one ulp)."""
import os
import tempfile

import numpy as np
import pytest

from cases import upload_keys
from golden_cases import runoff_for_step, setup

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.gpu
def test_output_files_match_reference():
    from cmf.model import CMFModel
    m, net, st, rvec, ref, fin, dt, nsteps = setup("g_pth")
    z = np.load(os.path.join(HERE, "golden", "g_pth_output.npz"), allow_pickle=False)
    names = list(z.files)
    g = CMFModel(m, "fp64")
    g.set_state(upload_keys(st))
    g.diag_reset_output()
    with tempfile.TemporaryDirectory() as d:
        for step in range(1, nsteps + 1):
            g.set_runoff(runoff_for_step(rvec, step, dt))
            g.physics_advance(dt)
            g.diag_avemax_output(dt)
            g.diag_getave_output()                       # daily output (IFRQ_OUT = 24, DT = 1 day)
            for v in names:
                g.output_write(v, os.path.join(d, v + ".bin"))
            g.diag_reset_output()
        cell = (m.seqy.astype(np.int64) - 1) * m.nx + (m.seqx.astype(np.int64) - 1)
        for v in names:
            got = np.fromfile(os.path.join(d, v + ".bin"), np.float32)
            want = z[v].view(np.float32)
            assert got.size == want.size, v
            if v != "pthflw":
                got = got.reshape(nsteps, -1)
                want = want.reshape(nsteps, -1)
                off = np.ones(m.nx * m.ny, bool)
                off[cell] = False
                assert np.all(got[:, off] == want[:, off]), v
                got, want = got[:, cell], want[:, cell]
            gi, wi = got.view(np.int32).astype(np.int64), want.view(np.int32).astype(np.int64)
            ulp = np.abs(gi - wi)
            same_sign = np.sign(got) == np.sign(want)
            ok = (got == want) | (same_sign & (ulp <= 1))
            print(f"  {v}: {int((got != want).sum())} of {got.size} values differ, max {int(ulp[same_sign].max())} ulp",
                  flush=True)
            assert ok.all(), v
    g.close()
