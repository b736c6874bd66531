"""Host checks of bench.py's record: the committed HBM-traffic profile is
found and parsed, and --gpus other than 1 exits with a message."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_traffic_profile_is_read():
    t = bench.measured_traffic("k_limber_gg", 512, "fp64")
    assert t is not None and 1e7 < t < 1e10
    assert bench.measured_traffic("no_such_kernel", 512, "fp64") is None


def test_algorithmic_work_counts():
    from cosmolike_amd import synth
    w = bench.algorithmic_work(synth.full_3x2pt_problem())
    rows = 500 + 19
    assert w["k_limber_gg"] == 20.0 * rows * 96 * 5 + 40.0 * rows * 96 * 5
    assert w["k_legendre"] == 2.0 * 26 * rows * (2 * 55 + 50 + 5)


def test_multi_gpu_request_exits():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)
