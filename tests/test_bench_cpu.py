"""CPU checks of bench.py's host logic: the one-GPU rule and the committed traffic figures."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "--gpus 2 is not supported" in r.stderr
    assert r.stdout == ""  # no JSON line


def test_profiled_traffic_reads_both_passes():
    t = bench.profiled_traffic()
    assert t is not None and set(t) == {"pass_a", "pass_b"}
    # the measured HBM bytes per launch lie between the algorithmic bytes (76 and 36 B per interior
    # voxel of 182x218x182) and twice them
    n_int = 180 * 216 * 180
    assert 76 * n_int <= t["pass_a"] <= 2 * 76 * n_int
    assert 36 * n_int <= t["pass_b"] <= 2 * 36 * n_int
