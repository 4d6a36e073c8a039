"""Function-level entry points of libpmpm for parity tests and tools.

Each wraps one C entry point of include/pmpm.h on torch device tensors and
runs on the current stream.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib


def _ptr(t):
    return None if t is None else t.data_ptr()


def sym_eig(A: torch.Tensor):
    """Batched register Jacobi (anisompm.py:48-95): (N,3,3) -> evals (N,3), V (N,3,3)."""
    A = A.contiguous()
    n = A.shape[0]
    ev = torch.empty(n, 3, device=A.device, dtype=A.dtype)
    V = torch.empty(n, 3, 3, device=A.device, dtype=A.dtype)
    _lib.check(_lib.lib().pmpm_sym_eig(A.data_ptr(), _lib.dtype_code(A.dtype), n, ev.data_ptr(), V.data_ptr(),
                                       _lib.stream_ptr(A.device)), "pmpm_sym_eig")
    return ev, V


def stencil(x: torch.Tensor, n_grid: int, grid_lim: float = 1.0):
    """Base cells, clamped node indices and weights as the step computes them
    (anisompm.py:544-590).  Returns a dict of device tensors."""
    x = x.contiguous()
    n, dev = x.shape[0], x.device
    i32 = dict(device=dev, dtype=torch.int32)
    out = {"base_q": torch.empty(n, 3, **i32), "idx27": torch.empty(n, 27, **i32),
           "base_c": torch.empty(n, 3, **i32), "idx64": torch.empty(n, 64, **i32),
           "w": torch.empty(n, 3, 3, device=dev, dtype=x.dtype), "wc": torch.empty(n, 4, 3, device=dev, dtype=x.dtype),
           "wc2": torch.empty(n, 4, 3, device=dev, dtype=x.dtype)}
    inv_dx = 1.0 / (grid_lim / n_grid)
    _lib.check(_lib.lib().pmpm_stencil(x.data_ptr(), _lib.dtype_code(x.dtype), n, n_grid, inv_dx,
                                       *(out[k].data_ptr() for k in ("base_q", "idx27", "base_c", "idx64", "w", "wc",
                                                                      "wc2")),
                                       _lib.stream_ptr(dev)), "pmpm_stencil")
    return out


def n_cells(n_grid: int) -> int:
    """Cells of the binning key space: 64 per 4x4x4 bucket, buckets cover the grid."""
    nb = (n_grid + 3) // 4
    return 64 * nb * nb * nb


def bin_workspace(n: int, n_grid: int, dtype=torch.float32, device="cuda:0"):
    """An initialised workspace for bin_particles of up to n particles."""
    L = _lib.lib()
    code = _lib.dtype_code(dtype)
    nbytes = L.pmpm_workspace_bytes(n, n_grid, 1, code)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
    _lib.check(L.pmpm_workspace_init(ws.data_ptr(), nbytes, n, n_grid, 1, code, _lib.stream_ptr(torch.device(device))),
               "pmpm_workspace_init")
    return ws


def bin_particles(x: torch.Tensor, n_grid: int, grid_lim: float = 1.0, stable: bool = True, ws=None):
    """Counting sort of particles by cell, cells numbered bucket-major (4x4x4
    blocks).  Returns (key, perm, cell_start) int32 device tensors; with
    stable=True perm is the stable argsort of key.  ``ws`` (from
    bin_workspace, for this n) is reused across calls as the step reuses
    its own."""
    x = x.contiguous()
    n, dev = x.shape[0], x.device
    L = _lib.lib()
    code = _lib.dtype_code(x.dtype)
    if ws is None:
        ws = bin_workspace(n, n_grid, x.dtype, dev)
    nbytes = ws.numel()
    st = _lib.stream_ptr(dev)
    key = torch.empty(n, device=dev, dtype=torch.int32)
    perm = torch.empty(n, device=dev, dtype=torch.int32)
    start = torch.empty(n_cells(n_grid) + 1, device=dev, dtype=torch.int32)
    inv_dx = 1.0 / (grid_lim / n_grid)
    _lib.check(L.pmpm_bin(x.data_ptr(), code, n, n_grid, inv_dx, int(stable), key.data_ptr(), perm.data_ptr(),
                          start.data_ptr(), ws.data_ptr(), nbytes, st), "pmpm_bin")
    return key, perm, start
