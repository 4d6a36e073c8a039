"""Differentiable rollouts: the hand-written HIP adjoint behind autograd.

``StepRunner`` advances explicit state dicts with the native step and runs
the native backward step (``pmpm_substep_adjoint``).  ``rollout`` is a
``torch.autograd.Function`` over ``steps`` substeps: the forward keeps a
checkpoint every ``every`` steps; the backward recomputes each segment from
its checkpoint, then walks it in reverse with the adjoint step.  Inputs that
carry gradients: the initial velocity and the per-particle mu, lam, A0 and
damage threshold sigma_c, so
gradients with respect to E, nu and the fiber direction follow by ordinary
autograd through mu = E/(2(1+nu)), lam = E nu/((1+nu)(1-2nu)) and
A0 = I + alpha a0 a0^T (see ``material_tensors``).

The reference has no gradient code; its step differentiated by torch
autograd is NaN in every state (its Jacobi squares an overflowing zeta in
converged rotations).  The adjoint here differentiates the invariant
quantities instead: the polar rotation through (s_i + s_j) denominators and
the tensile stress through Daleckii-Krein divided differences, with the
f'(lambda) limit at equal eigenvalues, so it stays finite at F = I.

Supported configurations: damage_every = 1, f_clamp, device colliders
(sphere_collider / halfspace_collider, sticky or slip with friction, moving
or not), VelocityBC particle boundary conditions; not interface particles
(the directional stress split).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .boundaries import HalfspaceCollider, SphereCollider, VelocityBC
from .solver import AnisoMPM

STATE = ("x", "v", "F", "C", "d", "lap", "D_hist")
_FIXED = ("vol0", "mass", "eta", "zeta", "l0", "residual", "n_interface")


class _AdjIn(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in STATE]


PARAMS = ("mu", "lam", "A0", "sigma_c")


class _AdjOut(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in STATE + PARAMS]


def _lib_adj():
    L = _lib.lib()
    if not getattr(L, "_adj_ready", False):
        L.pmpm_adjoint_workspace_bytes.restype = C.c_size_t
        L.pmpm_adjoint_workspace_bytes.argtypes = [C.c_int64, C.c_int32, C.c_int32]
        L.pmpm_substep_adjoint.argtypes = [C.POINTER(_lib.StepParams), C.POINTER(_lib.ParticlesIn),
                                           C.POINTER(_lib.ParticlesOut), C.POINTER(_AdjIn), C.POINTER(_AdjOut),
                                           C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p]
        L._adj_ready = True
    return L


class StepRunner:
    """Functional native step over explicit state dicts, using the materials,
    configuration and VelocityBC list of an AnisoMPM ``template``.  Substep k
    runs at time time0 + k dt, with time0 (default) the template's own
    time, template._step * dt, as its next substep would.  With
    exact_sigma_eig the forward steps take the Cauchy eigenpairs from a
    second Jacobi as the template's do; the adjoint differentiates the closed
    form from the polar stretch, which is the same function (the two agree
    to rounding)."""

    def __init__(self, template: AnisoMPM, time0: float | None = None):
        sim = template
        if sim.damage_every != 1:
            raise NotImplementedError("the adjoint supports damage_every=1")
        if len(sim.colliders) > _lib.MAX_COLLIDERS or \
                not all(isinstance(c, (SphereCollider, HalfspaceCollider)) for c in sim.colliders):
            raise NotImplementedError("the adjoint takes device colliders (sphere_collider / halfspace_collider) only")
        if getattr(sim, "damage_solver", "explicit") != "explicit":
            raise NotImplementedError("the adjoint differentiates the explicit damage update only")
        if bool(sim.interface_mask.any()):
            raise NotImplementedError("the adjoint does not differentiate the directional (interface) stress")
        if not all(isinstance(b, VelocityBC) for b in sim.particle_bc):
            raise NotImplementedError("the adjoint takes VelocityBC particle boundary conditions only")
        self.sim = sim
        self.time0 = float(sim._step * sim.dt if time0 is None else time0)
        self.dev, self.dtype, self.n = sim.device, sim.dtype, sim.n
        self.do_damage = bool(sim.allow.any())
        self.allow = sim.allow.contiguous()
        self.imask = sim.interface_mask.contiguous()
        self.fixed = {k: getattr(sim, k).contiguous() for k in _FIXED}
        self.params = {k: getattr(sim, k).contiguous() for k in PARAMS}
        self._ws = sim._workspace()
        L = _lib_adj()
        nbytes = L.pmpm_adjoint_workspace_bytes(self.n, sim.n_grid, sim._code)
        self._aws = torch.empty(nbytes, dtype=torch.uint8, device=self.dev)
        self._ring = []             # one step workspace per step of a recomputed segment

    def _ring_ws(self, slot):
        L = _lib.lib()
        while len(self._ring) <= slot:
            nb = L.pmpm_workspace_bytes(self.n, self.sim.n_grid, 1, self.sim._code)
            w = torch.empty(nb, dtype=torch.uint8, device=self.dev)
            _lib.check(L.pmpm_workspace_init(w.data_ptr(), nb, self.n, self.sim.n_grid, 1, self.sim._code,
                                             _lib.stream_ptr(self.dev)), "pmpm_workspace_init")
            self._ring.append(w)
        return self._ring[slot]

    def set_params(self, mu, lam, A0, sigma_c=None):
        self.params = {"mu": mu.detach().contiguous(), "lam": lam.detach().contiguous(),
                       "A0": A0.detach().contiguous(),
                       "sigma_c": (self.sim.sigma_c if sigma_c is None else sigma_c).detach().contiguous()}

    def _step_params(self, k):
        sim = self.sim
        time = self.time0 + k * sim.dt
        bc_id, bc_vals = sim._velocity_bcs(time)
        p = _lib.StepParams()
        C.memmove(C.addressof(p), C.addressof(sim._step_params()), C.sizeof(p))
        p.do_damage = int(self.do_damage)
        p.dt_damage = sim.dt
        p.n_colliders = len(sim.colliders)
        for j, c in enumerate(sim.colliders):
            c.describe(time, p.colliders[j])
        p.export_active = 0
        p.n_vel_bc = len(bc_vals)
        for j, v in enumerate(bc_vals):
            p.vel_bc[j][:] = v
        p.stable_sort = 1
        p.keep_for_adjoint = 0
        return p, bc_id

    def _pin(self, s, bc_id):
        pin = _lib.ParticlesIn()
        pin.n = self.n
        for k in STATE:
            setattr(pin, k, s[k].data_ptr())
        for k, t in self.fixed.items():
            setattr(pin, k, t.data_ptr())
        for k, t in self.params.items():
            setattr(pin, k, t.data_ptr())
        pin.allow = self.allow.data_ptr()
        pin.interface_mask = self.imask.data_ptr()
        pin.bc_id = bc_id.data_ptr() if bc_id is not None else None
        return pin

    def _scratch_out(self):
        n, dev, dt = self.n, self.dev, self.dtype
        return {"x": torch.empty(n, 3, device=dev, dtype=dt), "v": torch.empty(n, 3, device=dev, dtype=dt),
                "F": torch.empty(n, 3, 3, device=dev, dtype=dt), "C": torch.empty(n, 3, 3, device=dev, dtype=dt),
                "lap": torch.empty(n, device=dev, dtype=dt), "d": torch.empty(n, device=dev, dtype=dt),
                "D_hist": torch.empty(n, device=dev, dtype=dt), "phi": torch.empty(n, device=dev, dtype=dt),
                "phi_shear": torch.empty(n, device=dev, dtype=dt),
                "tau_last": torch.empty(n, 3, 3, device=dev, dtype=dt), "J_last": torch.empty(n, device=dev, dtype=dt)}

    def step(self, s, k, slot=None):
        """State after substep k (time k*dt) from state s.  With ``slot`` the
        step runs on workspace ``slot`` of the segment ring and keeps what
        its adjoint needs there (step_adjoint with the same slot)."""
        p, bc_id = self._step_params(k)
        ws = self._ws if slot is None else self._ring_ws(slot)
        p.keep_for_adjoint = int(slot is not None)
        out = self._scratch_out()
        pout = _lib.ParticlesOut()
        for key, t in out.items():
            setattr(pout, key, t.data_ptr())
        _lib.check(_lib.lib().pmpm_substep(C.byref(p), C.byref(self._pin(s, bc_id)), C.byref(pout),
                                           ws.data_ptr(), ws.numel(), _lib.stream_ptr(self.dev)), "pmpm_substep")
        new = {k2: out[k2] for k2 in STATE}
        if not self.do_damage:
            new["d"], new["D_hist"] = s["d"], s["D_hist"]
        return new

    def step_adjoint(self, s_in, s_out, adj_out, k, grads, slot=None):
        """Adjoint of substep k: from the adjoints of s_out (adj_out) to those
        of s_in; parameter adjoints accumulate into grads (mu, lam, A0,
        sigma_c).  With ``slot``: the forward step(s_in, k, slot) was the last
        thing run on that workspace, and its binning and grid are reused
        instead of recomputed."""
        p, bc_id = self._step_params(k)
        ws = self._ws if slot is None else self._ring_ws(slot)
        p.keep_for_adjoint = int(slot is not None)
        scratch = self._scratch_out()
        scratch["v"], scratch["C"] = s_out["v"], s_out["C"]
        pfwd = _lib.ParticlesOut()
        for key, t in scratch.items():
            setattr(pfwd, key, t.data_ptr())
        gin = _AdjIn()
        for key in STATE:
            setattr(gin, key, adj_out[key].data_ptr())
        adj_in = {key: torch.empty_like(s_in[key]) for key in STATE}
        gout = _AdjOut()
        for key in STATE:
            setattr(gout, key, adj_in[key].data_ptr())
        for key in PARAMS:
            setattr(gout, key, grads[key].data_ptr())
        L = _lib_adj()
        _lib.check(L.pmpm_substep_adjoint(C.byref(p), C.byref(self._pin(s_in, bc_id)), C.byref(pfwd), C.byref(gin),
                                          C.byref(gout), ws.data_ptr(), ws.numel(), self._aws.data_ptr(),
                                          self._aws.numel(), _lib.stream_ptr(self.dev)), "pmpm_substep_adjoint")
        return adj_in

    def zeros_like_state(self, s):
        return {k: torch.zeros_like(s[k]) for k in STATE}


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, mu, lam, A0, sigma_c, runner, state0, steps, every):
        runner.set_params(mu, lam, A0, sigma_c)
        s = {k: state0[k].detach().contiguous() for k in STATE}
        s["v"] = v0.detach().contiguous()
        ckpt = {0: s}
        for k in range(steps):
            s = runner.step(s, k)
            if (k + 1) % every == 0 and k + 1 < steps:
                ckpt[k + 1] = s
        ctx.runner, ctx.steps, ctx.every, ctx.ckpt = runner, steps, every, ckpt
        ctx.final = s
        return s["x"]

    @staticmethod
    def backward(ctx, gx):
        runner, steps, every = ctx.runner, ctx.steps, ctx.every
        final = ctx.final
        adj = runner.zeros_like_state(final)
        adj["x"] = gx.contiguous()
        grads = {k: torch.zeros_like(v) for k, v in runner.params.items()}
        starts = list(range(0, steps, every))
        for s0 in reversed(starts):
            s1 = min(s0 + every, steps)
            states = [ctx.ckpt[s0]]
            for k in range(s0, s1):
                states.append(runner.step(states[-1], k, slot=k - s0))
            for k in reversed(range(s0, s1)):
                adj = runner.step_adjoint(states[k - s0], states[k - s0 + 1], adj, k, grads, slot=k - s0)
        ctx.adj0 = adj
        return adj["v"], grads["mu"], grads["lam"], grads["A0"], grads["sigma_c"], None, None, None, None


def rollout(sim: AnisoMPM, v0, mu, lam, A0, steps: int, every: int = 10, sigma_c=None, time0=None):
    """Final positions after ``steps`` substeps from ``sim``'s current state,
    differentiable in v0, mu, lam, A0 and sigma_c (the damage threshold;
    default: sim.sigma_c), with a checkpoint every ``every`` steps.  The
    substeps run at time0 + k dt (default time0: sim._step * dt), as
    sim.run_frame(steps, time0) would."""
    runner = StepRunner(sim, time0)
    state0 = {k: getattr(sim, k) for k in STATE}
    if sigma_c is None:
        sigma_c = sim.sigma_c
    return _Rollout.apply(v0, mu, lam, A0, sigma_c, runner, state0, steps, every)


def material_tensors(E, nu, a0, alpha):
    """mu, lam and A0 from Young's modulus, Poisson ratio and fibers as
    AnisoMPM.add_object builds them (anisompm.py:371-412), as torch ops."""
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    eye = torch.eye(3, device=a0.device, dtype=a0.dtype)
    A0 = eye + alpha.view(-1, 1, 1) * (a0.unsqueeze(2) * a0.unsqueeze(1))
    return mu, lam, A0


def state_vjp(sim: AnisoMPM, steps: int, x_bar, every: int = 10, reuse: bool = True, time0=None):
    """Adjoint of the whole initial state and of mu, lam, A0, sigma_c for the
    vector x_bar on the positions after ``steps`` substeps (no autograd
    graph), the substeps at time0 + k dt (default time0: sim._step * dt).
    reuse=False makes every adjoint step recompute its forward binning and
    grid instead of reusing the segment's workspaces."""
    runner = StepRunner(sim, time0)
    s = {k: getattr(sim, k).contiguous() for k in STATE}
    ckpt = {0: s}
    for k in range(steps):
        s = runner.step(s, k)
        if (k + 1) % every == 0 and k + 1 < steps:
            ckpt[k + 1] = s
    adj = runner.zeros_like_state(s)
    adj["x"] = x_bar.contiguous()
    grads = {k: torch.zeros_like(v) for k, v in runner.params.items()}
    for s0 in reversed(range(0, steps, every)):
        s1 = min(s0 + every, steps)
        states = [ckpt[s0]]
        for k in range(s0, s1):
            states.append(runner.step(states[-1], k, slot=(k - s0) if reuse else None))
        for k in reversed(range(s0, s1)):
            adj = runner.step_adjoint(states[k - s0], states[k - s0 + 1], adj, k, grads,
                                      slot=(k - s0) if reuse else None)
    return s, adj, grads
