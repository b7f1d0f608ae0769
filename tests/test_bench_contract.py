"""bench.py's contract on the CPU: more than one GPU is refused with a
message, the committed HBM-traffic summary parses for both precisions, and
the algorithmic bytes per particle follow the documented per-kernel model
(own streams once, k neighbour entries of 4 bytes)."""
import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_committed_traffic_summary_parses():
    for prec in ("fp32c", "fp64"):
        walk = bench.profiled_traffic(prec, "knn_walk")
        # the walk writes at least its neighbour lists (k = 64, 4 bytes) for
        # 4M queries, and reads their positions
        assert walk is not None and walk > 4194304 * 64 * 4
    assert bench.profiled_traffic("fp32c", "no_such_kernel") is None


def test_algorithmic_bytes_model():
    for k in (32, 64):
        assert bench.algo_bytes("knn_walk", k) == 28 + 4 * k
        assert bench.algo_bytes("force", k) == 4 * k + 144
