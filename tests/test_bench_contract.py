"""bench.py's command-line contract on the CPU: one GPU only (a request for more exits with a
message), and the defaults that name the measured workload."""
import json
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], cwd=REPO,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]


def test_defaults_name_the_measured_workload():
    sys.path.insert(0, REPO)
    import bench
    argv = sys.argv
    try:
        sys.argv = ["bench.py"]
        a = bench.parse()
    finally:
        sys.argv = argv
    assert a.gpus == 1 and a.workload == "sphere3d" and a.points == 1_000_000
    assert a.point_order == "kd" and a.walks_per_lane == 1 and a.bvh_width == 4
