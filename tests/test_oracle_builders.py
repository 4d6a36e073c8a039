"""CPU pins of the scene-builder oracle (oracle/scene_builders.py), of the
host-side generator seeding the device builders use, and of the builders'
argument checks (no GPU work).

* ball_particles: the reference's own peel runs hold its output
  (``x_start``), bit-exact;
* grid_block: cell centres of a uniform split of the box, in "ij" order;
* fill_interior: SciPy's closing and hole fill against brute-force voxel
  loops; the low-face erosion (erosion sees 0 outside the array, and the
  grid starts at the lowest voxel) on a solid cube; a closed shell filled, an
  open cup not; points at voxel centres within the jitter bound;
* the PCG64 state the device starts from (peel_mpm_amd.builders.pcg64_state)
  reproduces numpy.random.default_rng(seed) when stepped, and the
  jump-ahead the kernels use lands on the sequentially stepped state.
"""
from __future__ import annotations

import ctypes as C
from collections import deque

import numpy as np
import pytest

from oracle import scene_builders as O
from tests.golden_cases import load_npz

MULT = 0x2360ED051FC65DA44385DF649FCCF645
M128, M64 = (1 << 128) - 1, (1 << 64) - 1


# ---------------------------------------------------------------- ball_particles
def test_ball_matches_the_reference_peel_runs():
    z = load_npz("peel_ball.npz")
    got = O.ball_particles(z["center"], float(z["R"]), float(z["spacing"]), float(z["jitter"]), int(z["seed"]))
    assert got.dtype == np.float32 and got.shape == z["x_start"].shape
    assert np.array_equal(got, z["x_start"])


def test_ball_lattice_and_jitter_bounds():
    c, R, s = np.array([0.3, 0.6, 0.45]), 0.1, 0.01
    plain = O.ball_particles(c, R, s, jitter=0.0, seed=3).astype(np.float64) - c
    assert np.all(np.linalg.norm(plain, axis=1) < R + 1e-6)
    rel = plain / s - 0.5
    assert np.allclose(rel, np.round(rel), atol=1e-3)                # on the centred lattice
    j = O.ball_particles(c, R, s, jitter=0.3, seed=3).astype(np.float64) - c
    assert j.shape == plain.shape and np.abs(j - plain).max() <= 0.15 * s + 1e-6


# ---------------------------------------------------------------- grid_block
def test_grid_block_cell_centres_in_ij_order():
    lo, hi, n = (0.1, 0.2, 0.3), (0.5, 0.4, 0.9), (4, 3, 5)
    p = O.grid_block(lo, hi, n)
    assert p.shape == (60, 3) and p.dtype == np.float64
    for d in range(3):
        h = (hi[d] - lo[d]) / n[d]
        centres = lo[d] + (np.arange(n[d]) + 0.5) * h
        vals = np.unique(p[:, d])
        assert np.allclose(vals, centres, rtol=0, atol=1e-15)
    assert np.all(p[:15, 0] == p[0, 0]) and p[15, 0] > p[0, 0]        # x slowest
    assert np.all(np.diff(p[:5, 2]) > 0)                              # z fastest


# ---------------------------------------------------------------- fill_interior
def brute_morph(src, dilate):
    out = np.zeros_like(src)
    D = src.shape
    for i in range(D[0]):
        for j in range(D[1]):
            for k in range(D[2]):
                vals = []
                for a, b, c in ((i, j, k), (i - 1, j, k), (i + 1, j, k), (i, j - 1, k), (i, j + 1, k), (i, j, k - 1),
                                (i, j, k + 1)):
                    inside = 0 <= a < D[0] and 0 <= b < D[1] and 0 <= c < D[2]
                    vals.append(bool(src[a, b, c]) if inside else False)
                out[i, j, k] = any(vals) if dilate else all(vals)
    return out


def brute_fill(solid):
    D = solid.shape
    reach = np.zeros(D, bool)
    q = deque()
    for idx in np.ndindex(*D):
        if not solid[idx] and any(v == 0 or v == D[a] - 1 for a, v in enumerate(idx)):
            reach[idx] = True
            q.append(idx)
    while q:
        i, j, k = q.popleft()
        for a, b, c in ((i - 1, j, k), (i + 1, j, k), (i, j - 1, k), (i, j + 1, k), (i, j, k - 1), (i, j, k + 1)):
            if 0 <= a < D[0] and 0 <= b < D[1] and 0 <= c < D[2] and not solid[a, b, c] and not reach[a, b, c]:
                reach[a, b, c] = True
                q.append((a, b, c))
    return ~reach


def cloud(kind, n=4000, seed=0):
    g = np.random.default_rng(seed)
    if kind == "shell":
        v = g.normal(size=(n, 3))
        return (0.3 * v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    if kind == "cup":                      # a shell with its top cap cut away
        p = cloud("shell", n, seed)
        return p[p[:, 1] < 0.2]
    if kind == "blob":
        return (0.2 * g.normal(size=(n, 3))).astype(np.float32)
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["shell", "cup", "blob"])
def test_closing_and_fill_match_brute_force(kind):
    pts = cloud(kind)
    c = np.zeros(3, np.float32)
    lo, dims, marked, solid = O.fill_voxels(pts, c, np.float32(1.0), 0.05)
    closed = brute_morph(brute_morph(marked, True), False)
    assert np.array_equal(solid, brute_fill(closed))


def test_solid_cube_loses_its_low_faces():
    """A dense cube of points marks an m^3 block at the grid's low corner;
    the closing's erosion sees 0 outside the array, so the block's faces on
    the low sides go: (m-1)^3 solid voxels."""
    g = (np.arange(40) + 0.5) / 40
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32) * np.float32(0.5)
    lo, dims, marked, solid = O.fill_voxels(pts, np.zeros(3, np.float32), np.float32(1.0), 0.05)
    m = int(marked.sum() ** (1 / 3) + 0.5)
    assert marked.sum() == m ** 3 and solid.sum() == (m - 1) ** 3
    assert not solid[0].any() and solid[1:m, 1:m, 1:m].all()


def test_shell_is_filled_and_cup_is_not():
    """With one extra point at the low corner the shell lies off the grid's
    low faces: its interior is filled, an open cup's is not.  Without it the
    erosion opens the shell's low caps and the interior stays empty."""
    c = np.zeros(3, np.float32)
    anchor = np.full((1, 3), -0.5, np.float32)
    shell = cloud("shell", 20000)
    _, _, m, s = O.fill_voxels(np.concatenate([shell, anchor]), c, np.float32(1.0), 0.05)
    assert s.sum() > 1.8 * m.sum()                         # the interior joined
    _, _, m, s = O.fill_voxels(np.concatenate([cloud("cup", 20000), anchor]), c, np.float32(1.0), 0.05)
    assert s.sum() < 1.1 * m.sum()                         # the open cavity stays empty
    _, _, m, s = O.fill_voxels(shell, c, np.float32(1.0), 0.05)
    assert s.sum() < m.sum()                               # low caps eroded: the shell leaks


def test_points_sit_at_voxel_centres_within_the_jitter():
    pts = cloud("blob")
    c, s = np.array([0.01, -0.02, 0.03], np.float32), np.float32(1.1)
    dx = 0.04
    lo, dims, marked, solid = O.fill_voxels(pts, c, s, dx)
    x0, vol, dxo = O.fill_interior(pts, c, s, dx, jitter=False)
    assert dxo == dx and vol.dtype == np.float32 and len(vol) == solid.sum() == len(x0)
    want = (lo[None] + (np.argwhere(solid) + 0.5) * dx).astype(np.float32)
    assert np.array_equal(x0, want)
    xj, _, _ = O.fill_interior(pts, c, s, dx, jitter=True, seed=5)
    assert np.abs(xj.astype(np.float64) - x0).max() <= 0.3 * dx + 1e-6


# ---------------------------------------------------------------- generator
def step_outputs(state, inc, k):
    out = []
    for _ in range(k):
        state = (state * MULT + inc) & M128
        v = (state >> 64) ^ (state & M64)
        rot = state >> 122
        out.append(((v >> rot) | (v << ((64 - rot) & 63))) & M64)
    return state, out


def advance(state, inc, delta):
    """Jump-ahead by squaring the step map, as the kernels do."""
    am, ap, cm, cp = 1, 0, MULT, inc
    while delta:
        if delta & 1:
            am = (am * cm) & M128
            ap = (ap * cm + cp) & M128
        cp = ((cm + 1) * cp) & M128
        cm = (cm * cm) & M128
        delta >>= 1
    return (am * state + ap) & M128


@pytest.mark.parametrize("seed", [0, 1, 2 ** 40 + 7])
def test_pcg64_seed_state_reproduces_default_rng(seed):
    from peel_mpm_amd.builders import pcg64_state
    s = pcg64_state(seed)
    state, inc = (s.state_hi << 64) | s.state_lo, (s.inc_hi << 64) | s.inc_lo
    _, outs = step_outputs(state, inc, 64)
    got = np.array([(o >> 11) * (1.0 / 9007199254740992.0) for o in outs])
    assert np.array_equal(got, np.random.default_rng(seed).random(64))
    for delta in (0, 1, 2, 5, 63, 1000):
        st, _ = step_outputs(state, inc, delta)
        assert advance(state, inc, delta) == st


# ---------------------------------------------------------------- ABI checks
def test_builder_entry_points_reject_bad_arguments_before_any_gpu_work():
    from peel_mpm_amd import _lib
    L = _lib.lib()
    rng = _lib.Pcg64(0, 1, 0, 1)
    cnt = C.c_int64(0)
    c3 = (C.c_double * 3)(0.5, 0.5, 0.5)
    ws = C.c_void_p(16)                  # never dereferenced: every call below fails validation
    assert L.pmpm_compact_workspace_bytes(1000) > 0
    assert L.pmpm_ball_particles(None, 10, 0.3, 0.01, c3, C.byref(rng), C.c_void_p(16), 1000, C.byref(cnt), ws, 1 << 20,
                                 None) != 0
    assert b"null" in L.pmpm_last_error()
    assert L.pmpm_ball_particles(C.c_void_p(16), 0, 0.3, 0.01, c3, C.byref(rng), C.c_void_p(16), 1000, C.byref(cnt), ws,
                                 1 << 20, None) != 0
    assert L.pmpm_ball_particles(C.c_void_p(16), 10, 0.3, 0.01, c3, C.byref(rng), C.c_void_p(16), 1000, C.byref(cnt), ws,
                                 1, None) != 0
    assert L.pmpm_ball_particles(C.c_void_p(16), 1291, 0.3, 0.01, c3, C.byref(rng), C.c_void_p(16), 1000, C.byref(cnt),
                                 ws, 1 << 40, None) != 0
    assert L.pmpm_grid_block(None, None, None, 2, 2, 2, None, None) != 0
    lo = (C.c_double * 3)()
    dims = (C.c_int64 * 3)()
    nb = L.pmpm_fill_workspace_bytes(None)
    assert nb > 0
    assert L.pmpm_fill_bounds(7, 0, C.c_void_p(16), 10, c3, 1.0, 0.1, lo, dims, ws, nb, None) != 0
    assert L.pmpm_fill_bounds(_lib.PMPM_F64, _lib.PMPM_F32, C.c_void_p(16), 10, c3, 1.0, 0.1, lo, dims, ws, nb,
                              None) != 0
    assert L.pmpm_fill_bounds(_lib.PMPM_F32, _lib.PMPM_F32, C.c_void_p(16), 10, c3, 1.0, 0.0, lo, dims, ws, nb,
                              None) != 0
    huge = (C.c_int64 * 3)(1 << 11, 1 << 10, 1 << 10)
    assert L.pmpm_fill_interior(_lib.PMPM_F32, _lib.PMPM_F32, C.c_void_p(16), 10, c3, 1.0, 0.1, lo, huge, 0.06, None,
                                C.c_void_p(16), 100, C.byref(cnt), ws, 1 << 40, None) != 0
    bad = (C.c_int64 * 3)(4, 0, 4)
    assert L.pmpm_fill_workspace_bytes(bad) == 0
    assert L.pmpm_fill_interior(_lib.PMPM_F32, _lib.PMPM_F32, C.c_void_p(16), 10, c3, 1.0, 0.1, lo, bad, 0.06, None,
                                C.c_void_p(16), 100, C.byref(cnt), ws, nb, None) != 0
    assert b"voxel grid" in L.pmpm_last_error()


def test_empty_cloud_is_refused_as_the_reference_refuses_it():
    """orange_setup.py:69 takes min and max of the cloud; NumPy raises
    ValueError on no points, and so does fill_interior, before any GPU work;
    the C entry point refuses n = 0 too."""
    from peel_mpm_amd import _lib
    from peel_mpm_amd.builders import fill_interior
    empty = np.zeros((0, 3), np.float32)
    with pytest.raises(ValueError):
        O.fill_interior(empty, np.zeros(3, np.float32), np.float32(1.0), 0.05)
    with pytest.raises(ValueError, match="no points"):
        fill_interior(empty, np.zeros(3, np.float32), np.float32(1.0), 0.05)
    L = _lib.lib()
    lo, dims = (C.c_double * 3)(), (C.c_int64 * 3)()
    c3 = (C.c_double * 3)(0.5, 0.5, 0.5)
    assert L.pmpm_fill_bounds(_lib.PMPM_F32, _lib.PMPM_F32, C.c_void_p(16), 0, c3, 1.0, 0.1, lo, dims, C.c_void_p(16),
                              L.pmpm_fill_workspace_bytes(None), None) != 0
    assert b"no points" in L.pmpm_last_error()
