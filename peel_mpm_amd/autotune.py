"""Launch-shape autotuner for the native step.

The step's kernels take their launch shape from a process-wide setting of
the library (``pmpm_set_launch_config``, include/pmpm.h):

* ``bucket_wgs``: work-groups of the persistent per-bucket kernels (P2G, G2P,
  the PCG drive and operator) -- how many buckets are in flight per CU;
* ``block_wgs``: work-groups of the per-node-block kernels (grid pass, PCG
  fold/update) -- also the number of partial dot products each PCG
  reduction sums (its reduction shape);
* ``p2g_pair``: particles a lane accumulates per LDS arena update in P2G
  (particles per block pass);
* ``p2g_waves``: wavefronts per bucket in P2G, each with its own LDS arena
  (2: smaller work-groups, more of them per CU; 4: more particles of a
  bucket in flight);
* ``const_threads``: work-group size of the per-particle constitutive
  kernel;
* ``g2p_threads``: work-group size of G2P (one bucket per work-group);
* ``cell_block``: the cell block of P2G and of the grid pass's fold -- 4
  (4x4x4 cells, a 7x7x7-node arena per bucket) or 8 (4x4x8 cells: the two
  buckets adjacent along z in one wavefront, their cells packed onto the
  lanes by particle count, into one 7x7x11-node arena; 21 % fewer arena
  nodes written and folded);
* ``p2g_split``: how a P2G work-group's wavefronts share a 4x4x4-cell
  bucket -- 1 (by particles: each takes every other pair of each cell), 2
  (by channels: one the quadratic, one the cubic, each with two arenas,
  the cells of two or more passes split over them and packed onto the
  lanes by load) or 3 (as 2 with the channels balanced: mv_x and mv_y on
  one wavefront, m, mv_z and the cubic pair on the other);
* ``const_stream``: where the explicit-damage step runs its constitutive
  kernel -- 1 (on the launch stream after the binning) or 2 (on a second
  stream beside the binning's scan and sort, its P2G records in particle
  order, read by P2G through the sort permutation).

The bucket (4x4x4 cells, one cell per lane of a 64-wide wavefront) stays
the unit of the binning and of the lane-per-cell scatter.

``tune`` runs a coordinate descent over these knobs on the caller's own
workload: for each knob in turn it times a few steps at every value (the
simulation state is restored before every trial, so tuning leaves the
simulation where it was) and keeps the fastest.  Results are cached per
workload shape (solver kind, particle count, grid, scene count, dtype,
damage solver) for the life of the process.  The reference has no tuner;
its step is a fixed sequence of torch operations.
"""
from __future__ import annotations

import ctypes as C
import time

import torch

from . import _lib

SPACE = {
    "bucket_wgs": (2048, 4096, 8192),
    "p2g_waves": (1, 2, 4),
    "g2p_threads": (128, 256, 512),
    "p2g_pair": (1, 2),
    "const_threads": (64, 128, 256),
    "block_wgs": (512, 1024, 2048),
    "cell_block": (4, 8),
    "p2g_split": (1, 2, 3),
    "const_stream": (1, 2),
}
DEFAULT = {"bucket_wgs": 2048, "block_wgs": 2048, "p2g_pair": 2, "const_threads": 256, "p2g_waves": 2,
           "g2p_threads": 256, "cell_block": 4, "p2g_split": 1, "const_stream": 1}
_CACHE: dict = {}


def get_config() -> dict:
    c = _lib.LaunchConfig()
    _lib.check(_lib.lib().pmpm_get_launch_config(C.byref(c)), "pmpm_get_launch_config")
    return {k: getattr(c, k) for k, _ in _lib.LaunchConfig._fields_}


def set_config(cfg: dict | None):
    """Set the launch shape (None: the defaults).  Raises on values the
    library rejects."""
    L = _lib.lib()
    if cfg is None:
        _lib.check(L.pmpm_set_launch_config(None), "pmpm_set_launch_config")
        return
    c = _lib.LaunchConfig(**{k: int(cfg.get(k, 0)) for k, _ in _lib.LaunchConfig._fields_})
    _lib.check(L.pmpm_set_launch_config(C.byref(c)), "pmpm_set_launch_config")


def _time_steps(step, restore, steps, device):
    restore()
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize(device)
    return (time.perf_counter() - t0) / steps


def tune(step, restore, device, space=None, steps=10, repeats=3, sweeps=1, log=None):
    """Coordinate descent over ``space`` (default SPACE).  ``step()`` advances
    the workload one step, ``restore()`` puts its state back.  Returns
    (best config, {config tuple: seconds per step}) and leaves the best
    config set."""
    space = space or SPACE
    best = dict(DEFAULT)
    seen = {}

    def measure(cfg):
        key = tuple(sorted(cfg.items()))
        if key not in seen:
            set_config(cfg)
            step_once = step
            _time_steps(step_once, restore, 1, device)          # first launch of a new shape
            seen[key] = min(_time_steps(step_once, restore, steps, device) for _ in range(repeats))
            if log:
                log(cfg, seen[key])
        return seen[key]

    measure(best)
    for _ in range(sweeps):
        for knob, values in space.items():
            for v in values:
                cand = dict(best, **{knob: v})
                if measure(cand) < measure(best):
                    best = cand
    set_config(best)
    restore()
    return best, seen


def _solver_key(sim, kind):
    return (kind, sim.n, sim.n_grid, getattr(sim, "n_scenes", 1), str(sim.dtype),
            getattr(sim, "damage_solver", "explicit"))


def tune_solver(sim, steps=10, space=None, log=None, use_cache=True):
    """Tune the launch shape on an AnisoMPM's current scene (its state is
    restored afterwards).  Returns the chosen config."""
    key = _solver_key(sim, "plain")
    if use_cache and key in _CACHE:
        set_config(_CACHE[key])
        return _CACHE[key]
    names = ("x", "v", "F", "C", "d", "lap", "D_hist", "phi", "phi_shear", "tau_last", "J_last")
    saved = {k: getattr(sim, k) for k in names if hasattr(sim, k)}
    step0 = sim._step
    t0 = step0 * sim.dt

    def restore():
        for k, v in saved.items():
            setattr(sim, k, v)          # the step never writes its inputs: the saved tensors are intact
        sim._step = step0

    def step():
        sim.substep(t0)

    best, _ = tune(step, restore, sim.device, space=space, steps=steps, log=log)
    _CACHE[key] = best
    return best


def tune_batched(b, steps=10, space=None, log=None, use_cache=True):
    """Tune the launch shape on a BatchedAnisoMPM before its graph is
    captured (tuning runs the step eagerly; the graph is captured with the
    chosen shape at the next substep).  Its state is restored afterwards."""
    if not b._built:
        b._build()
    key = _solver_key(b, "batched")
    if use_cache and key in _CACHE:
        set_config(_CACHE[key])
        b._graphs = [None, None]
        return _CACHE[key]
    use_graph, b.use_graph = b.use_graph, False
    cur0, step0 = b._cur, b._step
    saved = {k: v.clone() for k, v in b._bufs[cur0].items()}
    t0 = step0 * b.dt

    def restore():
        b._cur, b._step = cur0, step0
        for k, v in saved.items():
            b._bufs[cur0][k].copy_(v)
        b._bind()

    def step():
        b.substep(t0)

    try:
        best, _ = tune(step, restore, b.device, space=space, steps=steps, log=log)
    finally:
        b.use_graph = use_graph
        b._graphs = [None, None]        # recapture with the chosen shape
        restore()
    _CACHE[key] = best
    return best
