"""bench.py's command-line contract on the host: --gpus above 1 exits with
a message (one GPU only), and the HBM traffic of the dominant kernel is
read from the committed rocprofv3 counter summary."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_several_gpus_exit_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in r.stderr and r.stdout == ""


def test_traffic_read_from_profile_summary():
    sys.path.insert(0, ROOT)
    import bench
    t = bench.measured_traffic("rhs_fused")
    # the fused RHS moves 13 words per cell algorithmically (17.7 GB at
    # 768x384x576 FP64); the counters read somewhat more
    assert t is not None and 1.7e10 < t < 3.0e10
    assert bench.measured_traffic("no_such_stage") is None
