"""Oracle: CPU restatements of the reference's scene builders, for checking
only.

TEST INFRASTRUCTURE.  Only ``tests/`` may import this module, and only as
the checker; ``peel_mpm_amd`` never imports it.

* ``ball_particles`` follows ``anisompm_py/src/peel_test.py:128-137``;
* ``grid_block`` follows ``anisompm_py/src/test_block.py:41-47``;
* ``fill_interior`` follows ``anisompm_py/src/orange_setup.py:65-92``
  (without its progress print), with SciPy's ``ndimage.binary_closing`` and
  ``binary_fill_holes`` as the reference calls them.

The reference runs them with NumPy and SciPy; so do these, in the
reference's operation order, so their results are the reference's bits
under this NumPy (2.x, whose promotion keeps fp32 inputs in fp32).

Pinned by ``tests/test_oracle_builders.py``: ``ball_particles`` against the
reference's own peel-run fixtures (``x_start`` of
``out/{first_run,figs_spread}/run_{correct,wrong}.npz``, bit-exact);
``grid_block`` against cell-centre closed forms; ``fill_interior``'s closing
and hole fill against brute-force voxel loops (cross-shaped dilation and
erosion with 0 outside, breadth-first flood from the faces), and its
points against the voxel-centre closed form and the jitter bound.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage


def ball_particles(center, R, spacing, jitter=0.3, seed=0):
    """peel_test.py:128-137."""
    rng = np.random.default_rng(seed)
    n = int(np.ceil(2.0 * R / spacing)) + 2
    g = (np.arange(n) + 0.5) * spacing
    g -= g.mean()
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    pts = pts[np.linalg.norm(pts, axis=1) < R]
    pts += (rng.random(pts.shape) - 0.5) * (jitter * spacing)
    return (pts + np.asarray(center)[None]).astype(np.float32)


def grid_block(lo, hi, n_per_dim):
    """test_block.py:41-47."""
    xs = [np.linspace(lo[d] + 0.5 / n_per_dim[d] * (hi[d] - lo[d]),
                      hi[d] - 0.5 / n_per_dim[d] * (hi[d] - lo[d]), n_per_dim[d]) for d in range(3)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)


def fill_voxels(world_xyz, grid_center, scale_to_sim, vox_size):
    """The voxel stage of orange_setup.py:74-84: (lo, dims, marked, solid)."""
    sim_xyz = (world_xyz - grid_center[None]) * scale_to_sim + 0.5
    lo = sim_xyz.min(0)
    dx = float(vox_size)
    ijk = np.floor((sim_xyz - lo) / dx).astype(np.int64)
    dims = ijk.max(0) + 3
    occ = np.zeros(dims, bool)
    occ[ijk[:, 0], ijk[:, 1], ijk[:, 2]] = True
    marked = occ.copy()
    occ = ndimage.binary_closing(occ, iterations=1)
    solid = ndimage.binary_fill_holes(occ)
    return lo, dims, marked, solid


def fill_interior(world_xyz, grid_center, scale_to_sim, vox_size, jitter=True, seed=0):
    """orange_setup.py:65-92: (points (M,3) fp32, volumes (M,) fp32, dx)."""
    rng = np.random.default_rng(seed)
    lo, dims, marked, solid = fill_voxels(world_xyz, grid_center, scale_to_sim, vox_size)
    dx = float(vox_size)
    vox_idx = np.argwhere(solid)
    centers = lo[None] + (vox_idx + 0.5) * dx
    if jitter:
        centers = centers + (rng.random(centers.shape) - 0.5) * (0.6 * dx)
    vol = np.full(len(centers), dx ** 3, np.float32)
    return centers.astype(np.float32), vol, dx
