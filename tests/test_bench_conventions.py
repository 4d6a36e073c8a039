"""CPU checks of bench.py's conventions: frame lengths, the algorithmic
byte counts per kernel group, and the refusal of more than one GPU."""
from __future__ import annotations

import subprocess
import sys

from tests.conftest import ROOT


def test_frame_length_divides_the_run():
    sys.path.insert(0, ROOT)
    import bench
    assert bench.frame_length(500, 10) == 10
    assert bench.frame_length(300, 10) == 10
    assert bench.frame_length(12, 10) == 6
    assert bench.frame_length(7, 10) == 7
    assert bench.frame_length(5, 1) == 1
    for steps in range(1, 40):
        f = bench.frame_length(steps, 10)
        assert 1 <= f <= 10 and steps % f == 0


def test_group_bytes_sum_to_the_step_convention():
    """stress+P2G, G2P and binning particle bytes sum to the step's 289 B per
    particle plus binning's 20 B, and the node bytes of P2G's writes and G2P's
    reads to the step's 32 + 16 B per node pair."""
    sys.path.insert(0, ROOT)
    import bench
    n, q, c = 1000, 300, 400
    parts = sum(bench.algorithmic_bytes(k, n, q, c) for k in ("stress_p2g", "g2p"))
    assert parts == bench.step_bytes(n, q, c)
    assert bench.algorithmic_bytes("binning", n, q, c) == 20 * n
    assert bench.algorithmic_bytes("grid_update", n, q, c) == 32 * q + 16 * c
    assert bench.pcg_apply_bytes(n, c) == 48 * c + 16 * n


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], cwd=ROOT, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
