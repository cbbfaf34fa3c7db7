"""Mixed precision (FP32 flux/stage, FP64 storage) against the FP64 device path on
identical inputs. The FP64 path is pinned to the oracle at 1e-10
(test_gpu_parity.py), so this bounds the mixed path's error against the oracle.

Bound: 1e-6 relative error of the total storage and of every cell's storage (as a
fraction of the largest storage) over 10 simulated days of the default physics
with levees, and over 3 days with bifurcations, levees and dams. The storage
update, the inflow sums and the stage inversion are FP64, and the FP32 surface
slopes are formed from depth differences and exact elevation differences, so the
FP32 rounding of the fluxes is the only error source."""
import numpy as np
import pytest

from cases import prepare, upload_keys

BOUND = 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("name,ndays", [("levwet", 10), ("allwet", 3), ("plain", 10)])
def test_mixed_storage_error(name, ndays):
    from cmf.model import CMFModel
    m, net, st, rvec = prepare(name, ndays)
    runs = {p: CMFModel(m, p) for p in ("fp64", "mixed")}
    for g in runs.values():
        g.set_state(upload_keys(st))
    for d in range(ndays):
        for g in runs.values():
            g.set_runoff(rvec[d])
            g.physics_advance(86400.0)
        a = runs["fp64"].get_state(["rivsto", "fldsto", "levsto"] if m.opts.llevee else ["rivsto", "fldsto"])
        b = runs["mixed"].get_state(list(a))
        sa = sum(a[k] for k in a)
        sb = sum(b[k] for k in b)
        cell = float(np.abs(sb - sa).max() / np.abs(sa).max())
        total = float(abs(sb.sum() - sa.sum()) / sa.sum())
        print(f"{name} day {d + 1}: storage error cell {cell:.2e} total {total:.2e}", flush=True)
        assert cell <= BOUND and total <= BOUND
    for g in runs.values():
        g.close()
