"""Scene builders on the GPU, bit-exact with the reference's NumPy / SciPy
constructors.

* ``ball_particles`` -- jittered lattice points inside a ball
  (``anisompm_py/src/peel_test.py:128-137``, the peel scene);
* ``grid_block`` -- the cell-centred lattice of a box
  (``anisompm_py/src/test_block.py:41-47``);
* ``fill_interior`` -- one jittered particle per solid voxel of a point
  cloud, after a binary closing and a hole fill
  (``anisompm_py/src/orange_setup.py:65-92``, the orange scene).

Each takes the reference's arguments and returns what it returns (NumPy
arrays), or device tensors with ``as_tensor=True``.  The jitter is NumPy's
``default_rng(seed)`` stream: the generator is seeded here on the host with
NumPy's ``SeedSequence`` and the device draws the stream at any offset by
jump-ahead (``pmpm_ball_particles``, ``pmpm_fill_interior``).  The per-axis
lattices (a few hundred values) are formed here with the reference's NumPy
expressions; everything per point or per voxel runs in the kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

_MULT = 0x2360ED051FC65DA44385DF649FCCF645     # PCG64's LCG multiplier
_M128 = (1 << 128) - 1
_M64 = (1 << 64) - 1


def pcg64_state(seed) -> "_lib.Pcg64":
    """The state of ``numpy.random.default_rng(seed)``'s PCG64 after seeding:
    SeedSequence's four words give the initial state and stream; the
    generator then runs PCG's set-sequence seeding (state 0, one step, add
    the initial state, one step)."""
    ss = seed if isinstance(seed, np.random.SeedSequence) else np.random.SeedSequence(seed)
    v = [int(a) for a in ss.generate_state(4, np.uint64)]
    init, seq = (v[0] << 64) | v[1], (v[2] << 64) | v[3]
    inc = ((seq << 1) | 1) & _M128
    state = inc                                   # 0 * mult + inc
    state = (state + init) & _M128
    state = (state * _MULT + inc) & _M128
    return _lib.Pcg64(state >> 64, state & _M64, inc >> 64, inc & _M64)


def _out(t: torch.Tensor, as_tensor: bool):
    return t if as_tensor else t.cpu().numpy()


def _stream(device):
    return _lib.stream_ptr(torch.device(device))


def ball_particles(center, R, spacing, jitter=0.3, seed=0, *, as_tensor=False, device="cuda:0"):
    """Points of the lattice of pitch ``spacing`` centred on the origin that
    lie inside radius ``R``, jittered by up to ``jitter * spacing / 2`` per
    axis and moved to ``center``; (M, 3) fp32."""
    dev = torch.device(device)
    if dev.type != "cuda":
        raise ValueError("the scene builders run on the GPU only")
    n = int(np.ceil(2.0 * R / spacing)) + 2
    g = (np.arange(n) + 0.5) * spacing
    g -= g.mean()
    c = np.asarray(center, dtype=np.float64).reshape(3)
    L = _lib.lib()
    gd = torch.from_numpy(g).to(dev)
    cap = n ** 3
    out = torch.empty((cap, 3), dtype=torch.float32, device=dev)
    ws = torch.empty(L.pmpm_compact_workspace_bytes(cap), dtype=torch.uint8, device=dev)
    count = C.c_int64(0)
    rng = pcg64_state(seed)
    _lib.check(L.pmpm_ball_particles(gd.data_ptr(), n, float(R), float(jitter * spacing),
                                     (C.c_double * 3)(*c.tolist()), C.byref(rng), out.data_ptr(), cap, C.byref(count),
                                     ws.data_ptr(), ws.numel(), _stream(dev)), "pmpm_ball_particles")
    return _out(out[:count.value], as_tensor)


def grid_block(lo, hi, n_per_dim, *, as_tensor=False, device="cuda:0"):
    """Cell centres of the box [lo, hi] split into n_per_dim cells per axis,
    meshgrid "ij" order; (nx ny nz, 3) fp64."""
    dev = torch.device(device)
    if dev.type != "cuda":
        raise ValueError("the scene builders run on the GPU only")
    xs = [np.linspace(lo[d] + 0.5 / n_per_dim[d] * (hi[d] - lo[d]),
                      hi[d] - 0.5 / n_per_dim[d] * (hi[d] - lo[d]), n_per_dim[d]) for d in range(3)]
    ax = [torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(dev) for x in xs]
    nx, ny, nz = (len(x) for x in xs)
    out = torch.empty((nx * ny * nz, 3), dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().pmpm_grid_block(ax[0].data_ptr(), ax[1].data_ptr(), ax[2].data_ptr(), nx, ny, nz,
                                          out.data_ptr(), _stream(dev)), "pmpm_grid_block")
    return _out(out, as_tensor)


def fill_interior(world_xyz, grid_center, scale_to_sim, vox_size, jitter=True, seed=0, *, as_tensor=False,
                  device="cuda:0"):
    """One particle per solid voxel of the point cloud ``world_xyz`` mapped to
    the sim frame; returns (points (M,3) fp32, volumes (M,) fp32, dx).

    The sim-frame arithmetic runs in the type NumPy's promotion gives the
    inputs (fp32 for fp32 points, centre and scale, as the orange pipeline
    passes them), so the voxels are NumPy's."""
    dev = torch.device(device)
    if dev.type != "cuda":
        raise ValueError("the scene builders run on the GPU only")
    w = torch.as_tensor(world_xyz)
    if w.dtype not in (torch.float32, torch.float64) or w.dim() != 2 or w.shape[1] != 3:
        raise ValueError("world_xyz must be (N, 3) fp32 or fp64")
    if w.shape[0] == 0:
        # orange_setup.py:69 takes the cloud's min and max, which NumPy refuses on no points
        raise ValueError("world_xyz holds no points: an empty cloud has no bounds")
    center_np = np.asarray(grid_center)
    ct = np.result_type(np.float32 if w.dtype == torch.float32 else np.float64, center_np.dtype, scale_to_sim, 0.5)
    if ct not in (np.float32, np.float64):
        raise ValueError(f"unsupported arithmetic type {ct}")
    w = w.to(dev).contiguous()
    n = w.shape[0]
    c = (C.c_double * 3)(*np.asarray(center_np, dtype=np.float64).reshape(3).tolist())
    w_code = _lib.PMPM_F32 if w.dtype == torch.float32 else _lib.PMPM_F64
    ct_code = _lib.PMPM_F32 if ct == np.float32 else _lib.PMPM_F64
    dx = float(vox_size)
    L = _lib.lib()
    st = _stream(dev)
    ws = torch.empty(L.pmpm_fill_workspace_bytes(None), dtype=torch.uint8, device=dev)
    lo = (C.c_double * 3)()
    dims = (C.c_int64 * 3)()
    _lib.check(L.pmpm_fill_bounds(w_code, ct_code, w.data_ptr(), n, c, float(scale_to_sim), dx, lo, dims, ws.data_ptr(),
                                  ws.numel(), st), "pmpm_fill_bounds")
    cap = dims[0] * dims[1] * dims[2]
    ws = torch.empty(L.pmpm_fill_workspace_bytes(dims), dtype=torch.uint8, device=dev)
    out = torch.empty((cap, 3), dtype=torch.float32, device=dev)
    count = C.c_int64(0)
    rng = pcg64_state(seed) if jitter else None
    _lib.check(L.pmpm_fill_interior(w_code, ct_code, w.data_ptr(), n, c, float(scale_to_sim), dx, lo, dims, 0.6 * dx,
                                    C.byref(rng) if rng is not None else None, out.data_ptr(), cap, C.byref(count),
                                    ws.data_ptr(), ws.numel(), st), "pmpm_fill_interior")
    pts = out[:count.value]
    vol = torch.full((count.value,), dx ** 3, dtype=torch.float32, device=dev)
    return _out(pts, as_tensor), _out(vol, as_tensor), dx
