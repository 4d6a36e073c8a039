"""Hand-written adjoint of the step (needs a GPU), in fp64:

* against autograd through the reference step (golden vectors: 3 substeps of
  a damaging scene; the reference's sym_eig swapped for a gradient-safe twin,
  since the reference's own autograd is NaN in every state) -- all state and
  parameter gradients within 1e-6 norm-wise relative error;
* against fp64 central finite differences of the loss along random
  directions: a 20-step elastic pulled-slab rollout (h = 1e-4) and 3
  damaging substeps (h = 1e-6..1e-4, below the damage update's
  relu/max/clamp kinks): 1e-4;
* the autograd.Function with checkpointing gives the same gradients for any
  checkpoint spacing.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from tests.golden_cases import load_npz, rel_err

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
STATE = ("x", "v", "F", "C", "d", "lap", "D_hist")
# fp32 rollout and adjoint against fp64 autograd gradients: measured 1.6e-5
# to 3.2e-5 per field (inputs rounded to fp32 included); held to 10x that
FP32_ADJ_TOL = 3e-4


def sim_from_golden(dtype=torch.float64, name="adjoint_rollout64.npz", **kw):
    from peel_mpm_amd import AnisoMPM
    z = load_npz(name)
    ng, lim, dt, steps = z["cfg"].tolist()
    sim = AnisoMPM(n_grid=int(ng), grid_lim=lim, dt=dt, device=DEV, dtype=dtype, **kw)
    for k in STATE + ("mu", "lam", "A0", "vol0", "mass", "sigma_c", "eta", "zeta", "l0", "residual", "allow",
                      "interface_mask", "n_interface"):
        t = torch.from_numpy(z["in_" + k].copy())
        setattr(sim, k, (t.to(dtype) if t.is_floating_point() else t).to(DEV))
    sim.n = sim.x.shape[0]
    return sim, z, int(steps)


def test_adjoint_matches_reference_autograd():
    from peel_mpm_amd.adjoint import state_vjp
    sim, z, steps = sim_from_golden()
    target = torch.from_numpy(z["target"]).to(DEV)
    # forward once to get x_T, then the VJP of L = |x_T - target|^2
    from peel_mpm_amd.adjoint import StepRunner
    runner = StepRunner(sim)
    s = {k: getattr(sim, k) for k in STATE}
    for k in range(steps):
        s = runner.step(s, k)
    loss = float(((s["x"] - target) ** 2).sum())
    assert abs(loss - float(z["loss"])) <= 1e-9 * abs(float(z["loss"]))
    _, adj, grads = state_vjp(sim, steps, 2.0 * (s["x"] - target))
    torch.cuda.synchronize()
    report = {}
    for k in STATE:
        report[k] = rel_err(adj[k], torch.from_numpy(z["grad_" + k]))
    for k in ("mu", "lam", "A0", "sigma_c"):
        report[k] = rel_err(grads[k], torch.from_numpy(z["grad_" + k]))
    print({k: f"{v:.2e}" for k, v in report.items()})
    for k, e in report.items():
        assert e <= 1e-6, (k, e, report)


def _golden_colliders(z):
    from peel_mpm_amd import halfspace_collider, sphere_collider
    sc, sv, sr = z["sph_c"], z["sph_v"], float(z["sph_r"])
    ho, hn = z["hs_o"], z["hs_n"].tolist()
    return [sphere_collider(lambda t: (sc + sv * t).tolist(), sr, mode="slip", friction=0.3),
            halfspace_collider(lambda t: ho.tolist(), hn, mode="sticky")]


def test_adjoint_with_fclamp_and_colliders_matches_reference_autograd():
    """The strain limit (f_clamp 0.97..1.03, active on most particles) and
    moving device colliders (a slip sphere with friction, a sticky
    half-space) on a damaging 3-step fp64 rollout: every state and parameter
    gradient within 1e-6 of autograd through the reference step (golden
    vectors, the reference's sym_eig swapped for a gradient-safe twin)."""
    from peel_mpm_amd.adjoint import StepRunner, state_vjp
    name = "adjoint_fclamp_colliders64.npz"
    z = load_npz(name)
    sim, z, steps = sim_from_golden(name=name, f_clamp=tuple(z["fclamp"].tolist()))
    sim.colliders.extend(_golden_colliders(z))
    target = torch.from_numpy(z["target"]).to(DEV)

    def forward(s_):
        r = StepRunner(s_)
        st = {k: getattr(s_, k) for k in STATE}
        for k in range(steps):
            st = r.step(st, k)
        return st["x"]
    xT = forward(sim)
    loss = float(((xT - target) ** 2).sum())
    assert abs(loss - float(z["loss"])) <= 1e-9 * abs(float(z["loss"])), (loss, float(z["loss"]))
    # the case exercises both: without either the loss moves far beyond rounding
    for drop in ("f_clamp", "colliders"):
        other, _, _ = sim_from_golden(name=name, **({} if drop == "f_clamp" else {"f_clamp": tuple(z["fclamp"].tolist())}))
        if drop == "f_clamp":
            other.colliders.extend(_golden_colliders(z))
        l2 = float(((forward(other) - target) ** 2).sum())
        print(f"without {drop}: loss {l2:.9e} vs {loss:.9e}", flush=True)
        assert abs(l2 - loss) > 1e-6 * loss
    _, adj, grads = state_vjp(sim, steps, 2.0 * (xT - target))
    torch.cuda.synchronize()
    report = {k: rel_err(adj[k], torch.from_numpy(z["grad_" + k])) for k in STATE}
    report.update({k: rel_err(grads[k], torch.from_numpy(z["grad_" + k])) for k in ("mu", "lam", "A0", "sigma_c")})
    print("f_clamp + colliders adjoint vs reference autograd", {k: f"{v:.2e}" for k, v in report.items()}, flush=True)
    for k, e in report.items():
        assert e <= 1e-6, (k, e, report)


def test_fp32_adjoint_tracks_the_reference_autograd():
    """The same rollout and loss in fp32 (the bench's precision): the
    hand-written adjoint against the reference's fp64 autograd gradients,
    to fp32's accuracy over the rollout."""
    from peel_mpm_amd.adjoint import StepRunner, state_vjp
    sim, z, steps = sim_from_golden(torch.float32)
    target = torch.from_numpy(z["target"]).float().to(DEV)
    runner = StepRunner(sim)
    s = {k: getattr(sim, k) for k in STATE}
    for k in range(steps):
        s = runner.step(s, k)
    _, adj, grads = state_vjp(sim, steps, 2.0 * (s["x"] - target))
    torch.cuda.synchronize()
    report = {}
    for k in STATE:
        report[k] = rel_err(adj[k].double(), torch.from_numpy(z["grad_" + k]))
    for k in ("mu", "lam", "A0", "sigma_c"):
        report[k] = rel_err(grads[k].double(), torch.from_numpy(z["grad_" + k]))
    print("fp32 adjoint vs reference fp64 autograd", {k: f"{v:.2e}" for k, v in report.items()}, flush=True)
    for k, e in report.items():
        assert e <= FP32_ADJ_TOL, (k, e, report)


def fd_scene(steps=20):
    from peel_mpm_amd import AnisoMPM, scenes
    spec = scenes.slab_block(n_grid=32, cells=(8, 4, 8), origin=(12, 10, 12), ppc=4, seed=3, steps=steps,
                             dtype=torch.float64)
    return scenes.make_solver(AnisoMPM, spec, device=DEV), spec


def loss_of(sim, v0, mu, lam, A0, steps, target):
    from peel_mpm_amd.adjoint import StepRunner
    runner = StepRunner(sim)
    runner.set_params(mu, lam, A0)
    s = {k: getattr(sim, k) for k in STATE}
    s["v"] = v0
    for k in range(steps):
        s = runner.step(s, k)
    return float(((s["x"] - target) ** 2).sum())


def test_runner_follows_the_solver_clock():
    """The adjoint's forward steps run at the solver's own time: from a
    solver already 3 substeps in, with a time-dependent pull and
    exact_sigma_eig, StepRunner's steps equal sim.run_frame's bit for bit."""
    from peel_mpm_amd import VelocityBC
    from peel_mpm_amd.adjoint import StepRunner
    sim, spec = fd_scene()
    sim.exact_sigma_eig = True
    sim.particle_bc[0] = VelocityBC(sim.particle_bc[0].mask, lambda t: (0.0, 0.5 + 3000.0 * t, 0.0))
    sim.run_frame(3)
    runner = StepRunner(sim)
    assert runner.time0 == 3 * sim.dt
    st = {k: getattr(sim, k) for k in STATE}
    for k in range(4):
        st = runner.step(st, k)
    sim.run_frame(4, time0=3 * sim.dt)
    torch.cuda.synchronize()
    for k in STATE:
        assert torch.equal(st[k], getattr(sim, k)), k


@pytest.mark.parametrize("which", ["v0", "E", "nu"])
def test_adjoint_vs_central_finite_differences_elastic(which):
    """20-step pulled slab with damage off (a smooth loss): <grad, delta> vs
    central differences at h = 1e-4 along a random direction."""
    from peel_mpm_amd.adjoint import material_tensors, rollout
    steps = 20
    sim, spec = fd_scene(steps)
    sim.allow[:] = False
    n = sim.n
    g = torch.Generator(device="cpu").manual_seed(7)
    E = torch.full((n,), 1e5, dtype=torch.float64, device=DEV)
    nu = torch.full((n,), 0.3, dtype=torch.float64, device=DEV)
    v0 = (0.05 * torch.randn(n, 3, generator=g, dtype=torch.float64)).to(DEV)
    target = (sim.x + 0.01 * torch.randn(n, 3, generator=g, dtype=torch.float64).to(DEV)).detach()
    leaves = {"v0": v0.clone().requires_grad_(True), "E": E.clone().requires_grad_(True),
              "nu": nu.clone().requires_grad_(True)}
    mu, lam, A0 = material_tensors(leaves["E"], leaves["nu"], sim.a0, sim.alpha)
    xT = rollout(sim, leaves["v0"], mu, lam, A0, steps, every=7)
    ((xT - target) ** 2).sum().backward()
    grad = leaves[which].grad
    assert bool(torch.isfinite(grad).all())
    scale = {"v0": 0.05, "E": 1e5, "nu": 0.3}[which]
    delta = torch.randn(grad.shape, generator=g, dtype=torch.float64).to(DEV) * scale
    h = 1e-4

    def L_at(eps):
        vals = {k: t.detach().clone() for k, t in leaves.items()}
        vals[which] = vals[which] + eps * delta
        m, l, A = material_tensors(vals["E"], vals["nu"], sim.a0, sim.alpha)
        return loss_of(sim, vals["v0"], m, l, A, steps, target)

    fd = (L_at(h) - L_at(-h)) / (2 * h)
    an = float((grad * delta).sum())
    print(which, "adjoint", an, "fd", fd)
    assert abs(an - fd) <= 1e-4 * abs(fd), (which, an, fd)


@pytest.mark.parametrize("which", ["mu", "lam", "A0", "sigma_c", "v", "F", "d"])
def test_adjoint_vs_central_finite_differences_damage(which):
    """3 damaging substeps from the golden scene's random state: directional
    derivative of the loss vs central differences (h = 1e-6 to 1e-4
    relative: small enough not to cross the damage update's relu/max/clamp
    kinks, large enough to stay above round-off)."""
    from peel_mpm_amd.adjoint import StepRunner, state_vjp
    sim, z, steps = sim_from_golden()
    target = torch.from_numpy(z["target"]).to(DEV)
    runner = StepRunner(sim)
    s = {k: getattr(sim, k) for k in STATE}
    for k in range(steps):
        s = runner.step(s, k)
    _, adj, grads = state_vjp(sim, steps, 2.0 * (s["x"] - target))
    grad = grads[which] if which in grads else adj[which]
    g = torch.Generator(device="cpu").manual_seed(11)
    base = getattr(sim, which)
    delta = torch.randn(base.shape, generator=g, dtype=torch.float64).to(DEV) * float(base.abs().mean())
    if which == "d":
        # half the particles sit on the clamp boundary d = 0, where torch's
        # (inclusive) clamp subgradient and a central difference differ
        delta = delta * (base > 0.01).to(delta.dtype)
    # mu, lam move the loss little over 3 steps: a larger step keeps the
    # difference above fp64 round-off of the loss
    h = {"mu": 1e-4, "lam": 1e-4, "d": 1e-5}.get(which, 1e-6)

    def L_at(eps):
        old = getattr(sim, which)
        setattr(sim, which, old + eps * delta)
        r = StepRunner(sim)
        st = {k: getattr(sim, k) for k in STATE}
        for k in range(steps):
            st = r.step(st, k)
        setattr(sim, which, old)
        return float(((st["x"] - target) ** 2).sum())

    fd = (L_at(h) - L_at(-h)) / (2 * h)
    an = float((grad * delta).sum())
    print(which, "adjoint", an, "fd", fd)
    assert abs(an - fd) <= 1e-4 * abs(fd), (which, an, fd)


def test_reused_forward_workspaces_give_the_recomputed_gradients():
    """The backward reuses each segment step's binning and grid (one
    workspace per step); recomputing them inside every adjoint step instead
    must give bit-identical gradients (the stable sort makes the recompute
    exact)."""
    from peel_mpm_amd.adjoint import state_vjp
    sim, z, steps = sim_from_golden()
    target = torch.from_numpy(z["target"]).to(DEV)
    xb = 2.0 * (sim.x - target)
    _, a1, g1 = state_vjp(sim, steps, xb, every=2, reuse=True)
    _, a2, g2 = state_vjp(sim, steps, xb, every=2, reuse=False)
    for k in STATE:
        assert torch.equal(a1[k], a2[k]), k
    for k in g1:
        assert torch.equal(g1[k], g2[k]), k


def test_checkpoint_spacing_does_not_change_gradients():
    from peel_mpm_amd.adjoint import material_tensors, rollout
    steps = 12
    sim, spec = fd_scene(steps)
    n = sim.n
    target = (sim.x + 0.01).detach()
    res = []
    for every in (1, 5, 12):
        E = torch.full((n,), 1e5, dtype=torch.float64, device=DEV, requires_grad=True)
        nu = torch.full((n,), 0.3, dtype=torch.float64, device=DEV)
        mu, lam, A0 = material_tensors(E, nu, sim.a0, sim.alpha)
        xT = rollout(sim, sim.v.clone(), mu, lam, A0, steps, every=every)
        ((xT - target) ** 2).sum().backward()
        res.append(E.grad.clone())
    for r in res[1:]:
        assert torch.equal(r, res[0])


FIT_STEPS = 20


def fit_target(sim, steps):
    import math
    from peel_mpm_amd import fit
    dev, dt = sim.device, sim.dtype
    T = lambda v: torch.tensor(v, device=dev, dtype=dt)
    with torch.no_grad():
        return fit.forward_positions(sim, T(1.3e5), T(0.28), T(math.radians(15.0)), T(100.0), steps).detach()


def test_fit_loss_gradients_vs_central_differences():
    """Gradient of the fitting loss in the optimiser's coordinates (log E,
    logit nu, fiber angle, log sigma_c) over a 20-step damaging rollout vs
    central differences at h = 1e-5: 1e-2.  The damage update's relu / max /
    clamp kinks, crossed by some particles inside +-h, bound the agreement
    here; the smooth elastic rollout above is held to 1e-4.

    Longer horizons are not tested: once the drive D~ reaches ~1e3 the
    explicit damage update's own Jacobian dd'/dd = 1 - dt (D~ + 1) / eta is
    about -17 at dt = 1e-4, eta = 0.01, so step-to-step sensitivities grow
    geometrically while the relu / max keep the trajectory bounded.  The loss
    is then piecewise smooth on pieces far narrower than any usable h, and a
    pointwise derivative (this adjoint's, or autograd's) no longer predicts a
    finite difference."""
    import math
    from peel_mpm_amd import fit
    sim, spec = fd_scene(FIT_STEPS)
    target = fit_target(sim, FIT_STEPS)
    dev, dt = sim.device, sim.dtype
    base = sim.a0.clone()

    def L(p):
        logE, nu_u, ang, logsc = p
        xT = fit.forward_positions(sim, torch.exp(logE), 0.5 * torch.sigmoid(nu_u), ang, torch.exp(logsc),
                                   FIT_STEPS, 10, base)
        return ((xT - target) ** 2).sum()

    p0 = [torch.tensor(v, device=dev, dtype=dt) for v in (math.log(1e5), math.log(0.3 / 0.2), 0.0, math.log(100.0))]
    p = [t.clone().requires_grad_(True) for t in p0]
    L(p).backward()
    h = 1e-5
    for i, name in enumerate(("logE", "nu_u", "angle", "log_sigma_c")):
        pp = [t.clone() for t in p0]
        pm = [t.clone() for t in p0]
        pp[i] = pp[i] + h
        pm[i] = pm[i] - h
        with torch.no_grad():
            fd = (float(L(pp)) - float(L(pm))) / (2 * h)
        an = float(p[i].grad)
        print(name, "adjoint", an, "fd", fd)
        assert abs(an - fd) <= 1e-2 * abs(fd), (name, an, fd)


def test_fit_materials_reduces_the_loss():
    """Recover (E, nu, fiber angle, sigma_c) from a target made with E=1.3e5,
    nu=0.28 and fibers rotated 15 degrees about z over a 20-step damaging
    rollout: the loss falls and E, nu move toward the target's values."""
    import math
    from peel_mpm_amd import fit
    sim, spec = fd_scene(FIT_STEPS)
    target = fit_target(sim, FIT_STEPS)
    params, hist = fit.fit_materials(sim, target, FIT_STEPS, iters=15, lr=0.05)
    print(params, hist)
    assert all(math.isfinite(h) for h in hist)
    assert hist[-1] < 0.5 * hist[0]
    assert min(hist) < 0.1 * hist[0]
    assert params["E"] > 1.15e5 and params["nu"] < 0.3
