"""Device scene builders (peel_mpm_amd.builders) against the reference's
fixtures and the oracle (needs a GPU).  Every case is bit-exact: the
builders' jitter is NumPy's default_rng stream and their arithmetic is
NumPy's, in NumPy's order and types.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import scene_builders as O
from tests.golden_cases import load_npz
from tests.test_oracle_builders import cloud

pytestmark = pytest.mark.gpu


def test_ball_matches_the_reference_peel_runs():
    from peel_mpm_amd.builders import ball_particles
    z = load_npz("peel_ball.npz")
    print("ball: reference peel runs", z["x_start"].shape, flush=True)
    got = ball_particles(z["center"], float(z["R"]), float(z["spacing"]), float(z["jitter"]), int(z["seed"]))
    assert got.dtype == np.float32 and np.array_equal(got, z["x_start"])


@pytest.mark.parametrize("center,R,spacing,jitter,seed", [
    ((0.3, 0.6, 0.45), 0.1, 0.004, 0.5, 7),
    ((0.5, 0.5, 0.5), 0.28, 1 / 128, 0.3, 0),
    (np.array([0.5, 0.5, 0.5], np.float32), 0.45, 1 / 200, 0.3, 11),   # 3 M points
])
def test_ball_matches_the_oracle(center, R, spacing, jitter, seed):
    from peel_mpm_amd.builders import ball_particles
    print(f"ball R={R} spacing={spacing} seed={seed}", flush=True)
    got = ball_particles(center, R, spacing, jitter, seed)
    want = O.ball_particles(center, R, spacing, jitter, seed)
    print("points", len(want), flush=True)
    assert got.shape == want.shape and np.array_equal(got, want)


@pytest.mark.parametrize("lo,hi,n", [((0.3, 0.4, 0.3), (0.7, 0.5, 0.7), (40, 10, 40)),
                                     ((0.1, 0.2, 0.3), (0.5, 0.4, 0.9), (4, 3, 5)),
                                     ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (128, 96, 64))])
def test_grid_block_matches_the_oracle(lo, hi, n):
    from peel_mpm_amd.builders import grid_block
    print(f"grid_block {n}", flush=True)
    got = grid_block(lo, hi, n)
    want = O.grid_block(lo, hi, n)
    assert got.dtype == np.float64 and np.array_equal(got, want)


def orange_like(n, seed=0):
    """A thick shell of fp32 points (a peel) around a sparser core."""
    g = np.random.default_rng(seed)
    v = g.normal(size=(n, 3))
    r = np.where(g.random(n) < 0.7, 0.9 + 0.1 * g.random(n), g.random(n) ** (1 / 3) * 0.9)
    return (0.05 * r[:, None] * v / np.linalg.norm(v, axis=1, keepdims=True) + [0.01, 0.2, -0.03]).astype(np.float32)


CASES = [
    # the orange pipeline's types: fp32 points, fp32 centre, fp32 scale (fp32 arithmetic)
    ("orange fp32", lambda: orange_like(200_000), "f32", np.float32(12.4), 1 / 256, True, 0),
    # a fp64 centre promotes the sim frame to fp64
    ("orange fp64 centre", lambda: orange_like(200_000, 1), "f64", 12.4, 1 / 256, True, 3),
    ("blob no jitter", lambda: cloud("blob", 30_000, 2), "f32", np.float32(1.0), 0.02, False, 0),
    ("shell + anchor", lambda: np.concatenate([cloud("shell", 50_000, 4), np.full((1, 3), -0.5, np.float32)]), "f32",
     np.float32(1.0), 0.02, True, 9),
    ("orange 2M", lambda: orange_like(2_000_000, 5), "f32", np.float32(12.4), 1 / 512, True, 1),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fill_interior_matches_the_oracle(case):
    from peel_mpm_amd.builders import fill_interior
    name, make, cdt, scale, vox, jitter, seed = case
    pts = make()
    lo, hi = pts.min(0), pts.max(0)
    center = (np.float32(0.5) * (lo + hi)) if cdt == "f32" else (0.5 * (lo.astype(np.float64) + hi))
    print(f"fill_interior {name}: {len(pts)} points", flush=True)
    got = fill_interior(pts, center, scale, vox, jitter=jitter, seed=seed)
    want = O.fill_interior(pts, center, scale, vox, jitter=jitter, seed=seed)
    print("particles", len(want[0]), flush=True)
    assert got[2] == want[2]
    assert got[0].shape == want[0].shape and np.array_equal(got[0], want[0])
    assert got[1].dtype == np.float32 and np.array_equal(got[1], want[1])


def test_builders_return_device_tensors():
    import torch
    from peel_mpm_amd.builders import ball_particles, fill_interior
    t = ball_particles((0.5, 0.5, 0.5), 0.1, 0.01, as_tensor=True)
    assert isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32
    x, v, dx = fill_interior(torch.from_numpy(cloud("blob", 5000)).cuda(), np.zeros(3, np.float32), np.float32(1.0),
                             0.05, as_tensor=True)
    assert x.is_cuda and v.is_cuda and len(x) == len(v) > 0


@pytest.mark.parametrize("seed", range(10))
def test_random_builder_parameters(seed):
    """Random centres, radii, pitches, jitters and seeds for the ball, and
    random clouds, frames, voxel sizes and input types for fill_interior:
    bit-exact against the oracle."""
    from peel_mpm_amd.builders import ball_particles, fill_interior
    g = np.random.default_rng(100 + seed)
    R = float(g.uniform(0.02, 0.3))
    spacing = float(R / g.uniform(4.0, 40.0))
    center = g.uniform(0.2, 0.8, 3)
    if seed % 2:
        center = center.astype(np.float32)
    jitter, rs = float(g.uniform(0.0, 1.0)), int(g.integers(0, 2 ** 31))
    print(f"random ball R={R:.4f} spacing={spacing:.5f} jitter={jitter:.2f} seed={rs}", flush=True)
    assert np.array_equal(ball_particles(center, R, spacing, jitter, rs), O.ball_particles(center, R, spacing, jitter, rs))
    n = int(g.integers(1, 60_000))
    pts = (g.uniform(-1, 1) + g.uniform(0.01, 0.5) * g.normal(size=(n, 3)))
    pts = pts.astype(np.float32 if seed % 3 else np.float64)
    lo, hi = pts.min(0), pts.max(0)
    ext = float((hi - lo).max()) or 1.0
    if seed % 2:
        center, scale = (pts.dtype.type(0.5) * (lo + hi)), np.float32(0.6) / np.float32(ext)
    else:
        center, scale = 0.5 * (lo.astype(np.float64) + hi), 0.6 / ext
    vox = float(g.uniform(1 / 256, 1 / 32))
    jit = bool(seed % 4 != 0)
    print(f"random fill n={n} dtype={pts.dtype} vox={vox:.5f} jitter={jit}", flush=True)
    got = fill_interior(pts, center, scale, vox, jitter=jit, seed=rs)
    want = O.fill_interior(pts, center, scale, vox, jitter=jit, seed=rs)
    assert got[2] == want[2] and np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_degenerate_builder_inputs():
    """A ball smaller than the lattice pitch (no point, or only the centre
    point), a one-point cloud (a single voxel does not survive the closing),
    a one-cell block: as the oracle."""
    from peel_mpm_amd.builders import ball_particles, fill_interior, grid_block
    for R in (0.0, 0.001, 0.005):
        got, want = ball_particles((0.5, 0.5, 0.5), R, 0.01, 0.3, 0), O.ball_particles((0.5, 0.5, 0.5), R, 0.01, 0.3, 0)
        print(f"ball R={R}: {len(want)} points", flush=True)
        assert got.shape == want.shape and np.array_equal(got, want)
    one = np.array([[0.1, 0.2, 0.3]], np.float32)
    for jit in (True, False):
        got = fill_interior(one, np.zeros(3, np.float32), np.float32(1.0), 0.05, jitter=jit)
        want = O.fill_interior(one, np.zeros(3, np.float32), np.float32(1.0), 0.05, jitter=jit)
        assert got[2] == want[2] and got[0].shape == want[0].shape == (0, 3) and got[1].shape == (0,)
    got, want = grid_block((0.1, 0.2, 0.3), (0.2, 0.4, 0.5), (1, 1, 1)), O.grid_block((0.1, 0.2, 0.3), (0.2, 0.4, 0.5), (1, 1, 1))
    assert np.array_equal(got, want)
