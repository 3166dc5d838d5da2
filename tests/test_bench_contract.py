"""bench.py's contract on the CPU: more than one GPU is refused with a JSON
message and exit status 2 (before any device call), the HBM traffic summary
is read for the default workload only, and every workload's options parse."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"] and line["n_gpus"] == 2


def test_traffic_summary_only_for_the_default_workload():
    import bench

    import re

    sys.argv = ["bench.py"]
    args = bench.parse()
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    with open(os.path.join(here, "profiles", "c2_fp64_traffic.md")) as f:
        names = re.findall(r"`(spmv_\w+<[^`]*>)`", f.read())
    assert names
    t = bench.measured_traffic(names[0], args)
    assert t is not None and 4e8 < t < 2e9
    args.contacts = 1000
    assert bench.measured_traffic(names[0], args) is None


def test_every_workload_parses():
    import bench

    for w in ("single", "batched", "learned", "coupled"):
        sys.argv = ["bench.py", "--workload", w]
        assert bench.parse().workload == w
