"""Fit material parameters to a target trajectory by gradient descent
through the hand-written adjoint.

``fit_materials`` recovers Young's modulus, Poisson ratio, the fiber angle
(rotation of the fibers about z) and the damage threshold sigma_c, each
shared by all particles of a scene, from the final positions of a target
rollout.  Parameters are optimised in unconstrained coordinates
(log E, logit-style nu in (0, 0.5), angle, log sigma_c) with Adam.
"""
from __future__ import annotations

import math

import torch

from .adjoint import material_tensors, rollout


def fiber_field(a0, angle):
    """Rotate unit fibers a0 (N,3) by ``angle`` (radians, a 0-dim tensor) about z."""
    c, s = torch.cos(angle), torch.sin(angle)
    x = c * a0[:, 0] - s * a0[:, 1]
    y = s * a0[:, 0] + c * a0[:, 1]
    return torch.stack([x, y, a0[:, 2]], 1)


def forward_positions(sim, E, nu, angle, sigma_c, steps, every=10, base_fibers=None):
    n = sim.n
    a0 = fiber_field(sim.a0 if base_fibers is None else base_fibers, angle)
    mu, lam, A0 = material_tensors(E.expand(n), nu.expand(n), a0, sim.alpha)
    return rollout(sim, sim.v.clone(), mu, lam, A0, steps, every=every, sigma_c=sigma_c.expand(n).contiguous())


def fit_materials(sim, target_x, steps, iters=20, lr=0.05, init=(1e5, 0.3, 0.0, 100.0), every=10, log=None):
    """Adam on L = sum |x_T - target|^2, scaled by its value at the initial
    guess so the optimiser sees O(1) losses.  Returns (params, loss history)."""
    dev, dt = sim.device, sim.dtype
    logE = torch.tensor(math.log(init[0]), device=dev, dtype=dt, requires_grad=True)
    nu_u = torch.tensor(math.log(init[1] / (0.5 - init[1])), device=dev, dtype=dt, requires_grad=True)
    angle = torch.tensor(init[2], device=dev, dtype=dt, requires_grad=True)
    logsc = torch.tensor(math.log(init[3]), device=dev, dtype=dt, requires_grad=True)
    opt = torch.optim.Adam([logE, nu_u, angle, logsc], lr=lr)
    base = sim.a0.clone()
    hist = []
    scale = None
    for it in range(iters):
        opt.zero_grad()
        E = torch.exp(logE)
        nu = 0.5 * torch.sigmoid(nu_u)
        xT = forward_positions(sim, E, nu, angle, torch.exp(logsc), steps, every, base)
        raw = ((xT - target_x) ** 2).sum()
        if scale is None:
            scale = float(raw.detach()) or 1.0
        loss = raw / scale
        loss.backward()
        opt.step()
        hist.append(float(loss))
        if log:
            log(it, float(loss), float(E), float(nu), float(angle), float(torch.exp(logsc)))
    params = {"E": float(torch.exp(logE)), "nu": float(0.5 * torch.sigmoid(nu_u)), "angle": float(angle),
              "sigma_c": float(torch.exp(logsc))}
    return params, hist
