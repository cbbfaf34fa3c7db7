"""Host-side parts of bench.py that need no GPU: the reference CPU baseline run
(the reference's OpenMP build, oracle/_ref/omp64, on a small 1-degree network
with levees and dams) and the parsing of the test driver's per-step times and
substep counts that the CPU baseline and the drop-in timing share; the HBM
traffic figure read from the committed rocprofv3 summary."""
import os
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "omp64", "cmf_ref")


def _bench():
    sys.path.insert(0, ROOT)
    import bench
    return bench


def _net():
    from cmf import synth
    net = synth.make_network(360, 180, seed=5)
    synth.add_levees(net, 0.03, 5, hmin=0.5, hmax=3.0, fmin=0.05, fmax=0.5)
    synth.add_dams(net, 20, 5)
    return net, synth.make_runoff(net, 2, 5, mean_mm_day=1.5)


@pytest.mark.skipif(not os.path.exists(REF), reason="reference OpenMP build not built")
def test_host_run_parsing():
    bench = _bench()
    net, ro = _net()
    steps = 5
    with tempfile.TemporaryDirectory() as d:
        bench._reference_case(d, net, ro, 3600.0, steps)
        out, times, nts = bench._run_host(REF, d, dict(os.environ, OMP_NUM_THREADS="2"))
    assert len(times) == steps and all(t >= 0.0 for t in times)
    assert len(nts) == steps and all(k >= 1 for k in nts)
    line = [l for l in out.splitlines() if l.startswith("REFTIME")][0].split()
    assert int(line[2]) == steps
    assert abs(float(line[1]) - sum(times)) <= 1e-3 + 1e-6 * sum(times)


@pytest.mark.skipif(not os.path.exists(REF), reason="reference OpenMP build not built")
def test_cpu_baseline_fields():
    bench = _bench()
    net, ro = _net()
    r = bench.cpu_baseline(net, ro, net.ncells, 3600.0, 4)
    assert r is not None and r["value"] > 0 and r["kind"] == "reference"
    assert r["unit"] == "cell-substeps/s" and r["cores"] >= 1
    assert "4 outer steps of 3600 s" in r["sample"]


def test_traffic_profile_is_read():
    bench = _bench()
    t = bench.measured_traffic("k_update")
    assert t is not None and t > 284 * 1500120 * 0.5          # bytes per launch of the 1.5M-cell network
    assert np.isfinite(t)


def test_more_than_one_gpu_is_refused():
    """bench.py measures one GPU; --gpus N > 1 ends with a message and status 2."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


@pytest.mark.skipif(not os.path.exists(REF), reason="reference OpenMP build not built")
def test_ensemble_cpu_baseline_on_regional_domains():
    """The ensemble's CPU baseline runs the reference on the largest regional
    domain (the form bench.py's ensemble domains take) and reports its rate."""
    bench = _bench()
    from cmf import synth
    nets = []
    for d, (nx, ny) in enumerate([(60, 40), (120, 80)]):
        north = 70.0 - 8.0 * d
        nets.append(synth.make_network(nx, ny, seed=2026 + 101 * d, land_frac=0.6, west=-180.0 + 20.0 * d,
                                       east=-180.0 + 20.0 * d + 0.1 * nx, north=north, south=north - 0.1 * ny,
                                       regional=True))
    r = bench.ensemble_cpu_baseline(nets, 2026)
    assert r is not None and r["value"] > 0 and r["unit"] == "member-cell-substeps/s"
    assert f"{max(n.ncells for n in nets)} cells" in r["sample"]


@pytest.mark.skipif(not os.path.exists(REF), reason="reference OpenMP build not built")
def test_spinup_cpu_baseline():
    """The spin-up form's CPU baseline: the reference under the first month's
    climatological runoff, reported per simulated year."""
    bench = _bench()
    from cmf import maps, synth
    from cmf.spinup import make_climatology
    net = synth.make_network(180, 90, seed=991)
    m = maps.build_map(net, maps.Options())
    clim = make_climatology(net, m, 991, mean_mm_day=1.5, amp=0.5, phase=3.0)
    r = bench.spinup_cpu_baseline(net, m, clim, days=5)
    assert r is not None and r["value"] > 0 and r["unit"] == "s per simulated year"
    assert "5 daily steps" in r["sample"]
