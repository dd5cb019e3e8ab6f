"""bench.py's host-side contract: more than one GPU is refused with a
message, and the roofline byte models count what their docstrings say on a
corpus small enough to count by hand."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_sweep_bytes_by_hand():
    # two documents: [1, 1, 2] and [5, 3, 3, 3, 9]; V = 8 (word 9 removed), K = 10
    ws = np.array([1, 1, 2, 5, 3, 3, 3, 9], np.uint32)
    offs = np.array([0, 3, 8], np.uint64)
    K, V = 10, 8
    # every column int32 (hot = V): 7 valid tokens x (40 + 8) + 2 documents x 80
    b = bench.sweep_bytes(ws, offs, V, K, V)
    assert b["tokens"] == 7 and b["per_token"] == 7 * 48 + 2 * 80 == b["dense_int32"]
    # runs of equal words within a document: [1 1] [2] | [5] [3 3 3] -> 4 runs
    assert b["runs"] == 4 and b["per_run"] == 4 * 40 + 7 * 8 + 2 * 80
    # words >= 3 narrow (u16 columns, 2K bytes): columns of 1, 1, 2 at 40, of 5, 3, 3, 3 at 20
    n = bench.sweep_bytes(ws, offs, V, K, 3)
    assert n["per_token"] == 3 * 40 + 4 * 20 + 7 * 8 + 2 * 80
    assert n["per_run"] == 2 * 40 + 2 * 20 + 7 * 8 + 2 * 80
    # a run does not continue across a document boundary
    ws2 = np.array([4, 4, 4, 4], np.uint32)
    b2 = bench.sweep_bytes(ws2, np.array([0, 2, 4], np.uint64), V, K, V)
    assert b2["runs"] == 2
