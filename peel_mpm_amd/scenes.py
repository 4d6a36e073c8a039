"""Synthetic scenes of the benchmark configurations (seeded, NumPy PCG64).

All randomness is ``numpy.random.default_rng(seed)`` drawn in float64 and
cast to the solver dtype, so the GPU step, the oracle and the CPU baseline
get identical inputs.  Particles are laid out cell by cell (x-major, then y,
then z), ``ppc`` uniform points per cell, as the reference's scene builders
lay out grids of points (src/test_block.py:41-47).

Configurations (sizes as in the benchmark definitions):
  config0  64^3 grid, 32x8x32-cell sheet, 8 ppc -> 65,536 particles, top
           layer pulled at (0, 0.5, 0), gravity, 200 steps at dt=1e-4
  config1  128^3 grid, 124x8x124-cell slab (the 2*dx particle clamp of the
           reference leaves 124 of 128 cells usable), 8 ppc -> 984,064
           particles, vMF(kappa=20) fibers about (1,0,0), log-normal E,
           500 steps at dt=5e-5
  config2  256^3 grid, two stacked 252x8x252 slabs, fibers (1,0,0) below and
           (0,0,1) above, sigma_c 50 at the interface layers and 200 in the
           bulk, top pulled at 1 m/s, bottom fixed, 300 steps at dt=2e-5
  config3  64 independent 64^3 scenes, log-uniform particle counts in
           [1e4, 2.5e5], random block shapes, fiber angles and E
  config4  config0 at 4 ppc (32,768 particles), 100 steps, for gradients
"""
from __future__ import annotations

import math

import numpy as np
import torch


def _cells_points(origin, cells, ppc, dx, rng):
    gx, gy, gz = (np.arange(o, o + c) for o, c in zip(origin, cells))
    cell = np.stack(np.meshgrid(gx, gy, gz, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    return (np.repeat(cell, ppc, 0) + rng.random((cell.shape[0] * ppc, 3))) * dx


def vmf_fibers(n, mean, kappa, rng):
    """Unit vectors from the von Mises-Fisher distribution on S^2 (Wood 1994,
    rejection-free in 3D: w = 1 + log(u + (1-u) e^{-2k}) / k)."""
    u = rng.random(n)
    w = 1.0 + np.log(u + (1.0 - u) * np.exp(-2.0 * kappa)) / kappa
    phi = 2.0 * math.pi * rng.random(n)
    s = np.sqrt(np.clip(1.0 - w * w, 0.0, None))
    local = np.stack([w, s * np.cos(phi), s * np.sin(phi)], 1)      # about +x
    mean = np.asarray(mean, np.float64)
    mean = mean / np.linalg.norm(mean)
    # rotation taking +x to mean
    ex = np.array([1.0, 0.0, 0.0])
    v = np.cross(ex, mean)
    c = float(ex @ mean)
    if np.linalg.norm(v) < 1e-12:
        R = np.eye(3) if c > 0 else np.diag([-1.0, -1.0, 1.0])
    else:
        K = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
        R = np.eye(3) + K + K @ K / (1.0 + c)
    return local @ R.T


def _lame(E, nu):
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


def _materials(x, dx, ppc, rho, E, nu, fibers, alpha, sigma_c, eta=0.01, zeta=1.0, residual=1e-3, l0_scale=0.5):
    """Per-particle parameter arrays as AnisoMPM.add_object builds them
    (anisompm.py:366-424), with sigma_c written directly."""
    n = x.shape[0]
    E = np.broadcast_to(np.asarray(E, np.float64), (n,))
    mu, lam = _lame(E, nu)
    vol = np.full(n, dx ** 3 / ppc)
    f = np.asarray(fibers, np.float64)
    a0 = f / (np.linalg.norm(f, axis=1, keepdims=True) + 1e-12)
    A0 = np.eye(3)[None] + alpha * a0[:, :, None] * a0[:, None, :]
    return {
        "x": x, "v": np.zeros_like(x), "F": np.tile(np.eye(3), (n, 1, 1)), "C": np.zeros((n, 3, 3)),
        "d": np.zeros(n), "lap": np.zeros(n), "D_hist": np.zeros(n),
        "vol0": vol, "mass": vol * rho, "mu": mu, "lam": lam,
        "sigma_c": np.broadcast_to(np.asarray(sigma_c, np.float64), (n,)).copy(),
        "eta": np.full(n, eta), "zeta": np.full(n, zeta), "l0": np.full(n, l0_scale * dx),
        "residual": np.full(n, residual), "A0": A0, "a0": a0, "alpha": np.full(n, alpha),
        "allow": np.ones(n, bool), "interface_mask": np.zeros(n, bool), "n_interface": np.zeros((n, 3)),
    }


def slab_block(n_grid=64, cells=(32, 8, 32), origin=(16, 8, 16), ppc=8, seed=0, E=1e5, nu=0.3, rho=1000.0,
               fiber=(1.0, 0.0, 0.0), alpha=-1.0, sigma_c=100.0, dt=1e-4, steps=200, pull=(0.0, 0.5, 0.0),
               gravity=(0.0, -9.8, 0.0), dtype=torch.float32):
    """One sheet, top particle layer (within dx/2 of the top face) pulled."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / n_grid
    x = _cells_points(origin, cells, ppc, dx, rng)
    arrays = _materials(x, dx, ppc, rho, E, nu, np.tile(fiber, (x.shape[0], 1)), alpha, sigma_c)
    y_top = (origin[1] + cells[1]) * dx
    top = x[:, 1] > y_top - 0.5 * dx
    return {"n_grid": n_grid, "dt": dt, "steps": steps, "gravity": gravity, "dtype": dtype, "arrays": arrays,
            "bc": [(top, pull)]}


def config0(dtype=torch.float32):
    return slab_block(dtype=dtype)


def config4(dtype=torch.float64):
    return slab_block(ppc=4, steps=100, dtype=dtype)


def config1(dtype=torch.float32, seed=0):
    rng = np.random.default_rng(seed)
    ng, dx, ppc = 128, 1.0 / 128, 8
    origin, cells = (2, 60, 2), (124, 8, 124)
    x = _cells_points(origin, cells, ppc, dx, rng)
    n = x.shape[0]
    fib = vmf_fibers(n, (1.0, 0.0, 0.0), 20.0, rng)
    E = 1e5 * np.exp(0.2 * rng.standard_normal(n))
    arrays = _materials(x, dx, ppc, 1000.0, E, 0.3, fib, -1.0, 100.0)
    top = x[:, 1] > (origin[1] + cells[1]) * dx - 0.5 * dx
    return {"n_grid": ng, "dt": 5e-5, "steps": 500, "gravity": (0.0, -9.8, 0.0), "dtype": dtype, "arrays": arrays,
            "bc": [(top, (0.0, 0.5, 0.0))]}


def config2(dtype=torch.float32, seed=0):
    rng = np.random.default_rng(seed)
    ng, dx, ppc = 256, 1.0 / 256, 8
    lo = _cells_points((2, 120, 2), (252, 8, 252), ppc, dx, rng)
    hi = _cells_points((2, 128, 2), (252, 8, 252), ppc, dx, rng)
    x = np.concatenate([lo, hi])
    fib = np.concatenate([np.tile([1.0, 0, 0], (lo.shape[0], 1)), np.tile([0, 0, 1.0], (hi.shape[0], 1))])
    y = x[:, 1]
    iface = np.abs(y - 128 * dx) < dx          # particle layers adjacent to the interface
    sig = np.where(iface, 50.0, 200.0)
    arrays = _materials(x, dx, ppc, 1000.0, 1e5, 0.3, fib, -1.0, sig)
    top = y > 136 * dx - 0.5 * dx
    bot = y < 120 * dx + 0.5 * dx
    arrays["allow"] = ~(top | bot)
    return {"n_grid": ng, "dt": 2e-5, "steps": 300, "gravity": (0.0, -9.8, 0.0), "dtype": dtype, "arrays": arrays,
            "bc": [(bot, (0.0, 0.0, 0.0)), (top, (0.0, 1.0, 0.0))]}


def config3_scene(s: int, dtype=torch.float32):
    """Scene s of the 64-scene batch (seed s)."""
    rng = np.random.default_rng(s)
    ng, dx, ppc = 64, 1.0 / 64, 8
    target = int(round(math.exp(rng.uniform(math.log(1e4), math.log(2.5e5)))))
    cells_total = max(1, round(target / ppc))
    # thickness 4..12 cells, remaining two edges as equal as the interior allows
    ty = int(rng.integers(4, 13))
    area = max(1, round(cells_total / ty))
    cx = int(np.clip(round(math.sqrt(area) * math.exp(rng.uniform(-0.3, 0.3))), 1, 60))
    cz = int(np.clip(round(area / cx), 1, 60))
    cells = (cx, ty, cz)
    origin = (2 + (60 - cx) // 2, 20, 2 + (60 - cz) // 2)
    x = _cells_points(origin, cells, ppc, dx, rng)
    th = rng.uniform(0.0, math.pi)
    fib = np.tile([math.cos(th), 0.0, math.sin(th)], (x.shape[0], 1))
    E = math.exp(rng.uniform(math.log(2e4), math.log(5e5)))
    arrays = _materials(x, dx, ppc, 1000.0, E, 0.3, fib, -1.0, 1e-3 * E)
    top = x[:, 1] > (origin[1] + ty) * dx - 0.5 * dx
    return {"n_grid": ng, "dt": 1e-4, "steps": 200, "gravity": (0.0, -9.8, 0.0), "dtype": dtype, "arrays": arrays,
            "bc": [(top, (0.0, 0.5, 0.0))]}


_STATE_PARAMS = ("x", "v", "F", "C", "d", "lap", "D_hist", "vol0", "mass", "mu", "lam", "sigma_c", "eta", "zeta",
                 "l0", "residual", "A0", "allow", "interface_mask", "n_interface")


def tensors(spec, device="cpu"):
    dt = spec["dtype"]
    out = {}
    for k in _STATE_PARAMS + ("a0", "alpha"):
        a = spec["arrays"][k]
        t = torch.from_numpy(np.ascontiguousarray(a))
        out[k] = (t if t.dtype == torch.bool else t.to(dt)).to(device)
    return out


def make_solver(cls, spec, device="cpu", **kw):
    """Instantiate ``cls`` (peel_mpm_amd.AnisoMPM or the oracle) on the scene,
    with the velocity boundary conditions attached."""
    from .boundaries import VelocityBC
    args = dict(n_grid=spec["n_grid"], grid_lim=1.0, dt=spec["dt"], gravity=spec["gravity"], dtype=spec["dtype"])
    if device != "cpu" or cls.__name__ == "AnisoMPM":
        args["device"] = device
    sim = cls(**args, **kw)
    t = tensors(spec, device)
    for k, v in t.items():
        setattr(sim, k, v)
    sim.n = t["x"].shape[0]
    for mask, vel in spec["bc"]:
        sim.particle_bc.append(VelocityBC(torch.from_numpy(mask).to(device), tuple(float(a) for a in vel)))
    return sim
