"""bench.py's contract on the CPU: one GPU only (N > 1 exits with a
message, before any GPU work), the roofline arithmetic, and the HBM traffic
read from the committed rocprofv3 summary."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "single" in r.stderr and r.stdout == ""


def test_roofline_fraction_is_achieved_over_peak():
    bt = {"work": 2.0e12, "ms": 100.0, "bytes": 4.0e9}
    rf = bench.roofline(bt, bench.PEAK_VALU_TOPS, "k")
    assert abs(rf["achieved"] - 20.0) < 1e-9          # 2e12 ops in 0.1 s = 20 T ops/s
    assert abs(rf["frac"] - 20.0 / bench.PEAK_VALU_TOPS) < 1e-12
    assert abs(rf["hbm_gbs"] - 40.0) < 1e-9
    assert abs(bench.PEAK_VALU_TOPS - 39.3216) < 1e-9


def test_traffic_comes_from_the_committed_profile():
    t = bench.measured_traffic("batched_int32")
    # 2^20 nodes of 62 algorithmic bytes each (uint8 permutation of 50 jobs,
    # two limits, one bound); the counters may not read less than that
    assert t is not None and t >= (1 << 20) * 62
    assert bench.measured_traffic("no such workload") is None
