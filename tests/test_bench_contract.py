"""bench.py's command-line contract on the CPU: it runs on one GPU only and
says so for --gpus N > 1 (one JSON line, exit status 2), and it reads the
calibrated HBM traffic of the elimination class from the committed
rocprofv3 summary."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert out.returncode == 2
    line = json.loads(out.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_traffic_summary_is_read():
    sys.path.insert(0, ROOT)
    try:
        import bench
    finally:
        sys.path.pop(0)
    traffic, src = bench.measured_traffic("factor")
    assert src == "rskelf_circle_N1048576_fp32_traffic.md"
    # the elimination writes the factor (over a GB per factorization over 12
    # launches): the measured bytes per launch are of that order
    assert 5e7 < traffic < 1e9
    # the Gram ID and the sweeps have their own entries
    for cls in ("id", "sweep"):
        t, s = bench.measured_traffic(cls)
        assert s == "rskelf_circle_N1048576_fp32_traffic.md" and t > 1e6
    assert bench.measured_traffic("no-such-kernel") == (None, None)
    # keyed by workload: another workload's class of the same name gets none
    assert bench.measured_traffic("factor", "ellipses_fp32") == (None, None)
