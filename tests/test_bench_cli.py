"""bench.py contract pieces that need no GPU."""
import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_before_any_device_work():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU only" in p.stdout + p.stderr
    assert '"metric"' not in p.stdout


def test_committed_traffic_summary_is_read_for_the_roofline():
    path = os.path.join(ROOT, "profiles", "k_pcg_spmv_traffic_bar_twist_fold.txt")
    t = bench.read_traffic(path, "pcg_spmv")
    assert t is not None and 2e7 < t < 4e7  # bytes per SpMV dispatch on the 64x16x16 bar
    assert bench.read_traffic(path, "no_such_kernel") is None
    assert bench.read_traffic(os.path.join(ROOT, "profiles", "absent.txt"), "pcg_spmv") is None


def test_committed_tuning_file_is_valid():
    with open(os.path.join(ROOT, "idp_amd", "tuning.json")) as f:
        d = json.load(f)
    assert 0 <= d["spmv_shape"] <= 7 and 1 <= d["upd_waves"] <= 8
    assert d["measurements"], "the tuning file keeps the measurements it was chosen from"
