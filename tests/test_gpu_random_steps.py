"""Randomised one-step parity of the HIP step against the oracle, in fp64
(needs a GPU).

The golden cases toggle each branch of the step on fixed small scenes; these
draw whole scenes at random per seed -- grid size, particle count (down to a
single particle), positions across the whole padded domain, velocities,
deformation gradients near identity, affine fields, damage, history and
Laplacian, per-particle materials (mu, lam, sigma_c, eta, zeta, l0, k_r,
structural tensor), the allow mask, interface particles with their normals,
damage_every, f_clamp, grid damping and gravity -- and hold one step of
every state field and diagnostic to 1e-10 of the oracle (1e-9 for the
diagnostics), as the fp64 golden cases are held.

Two conventions of the reference's API are kept: interface normals are zero
outside the interface mask (add_object zeroes them and
set_interface_normals writes masked rows only, anisompm.py:520-529; the
step reads a particle's normal only inside the mask), and the Laplacian,
which is zero up to roundoff for one particle's own constant field, is
measured against its natural scale inv_dx^2 |d| as well as its own norm.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import cpu_step as O
from tests.golden_cases import DIAG_KEYS, PARAM_KEYS, STATE_KEYS

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def random_scene(seed):
    g = np.random.default_rng(seed)
    ng = int(g.choice([16, 24, 32, 40]))
    n = int(g.choice([1, 2, 37, int(g.integers(200, 3000))]))
    dx = 1.0 / ng
    lo, hi = 2 * dx, 1.0 - 2 * dx
    if g.random() < 0.5:                       # one blob somewhere in the box
        c = g.uniform(lo + 0.1, hi - 0.1, 3)
        x = np.clip(c + 0.08 * g.normal(size=(n, 3)), lo, hi)
    else:                                      # spread over the whole padded domain
        x = g.uniform(lo, hi, (n, 3))
    a0 = g.normal(size=(n, 3))
    a0 /= np.linalg.norm(a0, axis=1, keepdims=True)
    alpha = g.choice([0.0, -1.0, 0.5])
    E = g.uniform(2e3, 2e5, n)
    nu = g.uniform(0.2, 0.4, n)
    vol = np.full(n, dx ** 3 / 8)
    iface = g.random(n) < 0.3
    nrm = g.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm[~iface] = 0.0
    t = {
        "x": x, "v": 0.3 * g.normal(size=(n, 3)),
        "F": np.eye(3) + 0.04 * g.normal(size=(n, 3, 3)), "C": 2.0 * g.normal(size=(n, 3, 3)),
        "d": 0.4 * g.random(n), "lap": 50.0 * g.normal(size=n), "D_hist": 2.0 * g.random(n),
        "vol0": vol, "mass": vol * g.uniform(300, 1500, n),
        "mu": E / (2 * (1 + nu)), "lam": E * nu / ((1 + nu) * (1 - 2 * nu)),
        "sigma_c": g.uniform(0.02, 0.3, n) * E, "eta": g.uniform(0.005, 0.5, n), "zeta": g.uniform(0.5, 2.0, n),
        "l0": g.uniform(0.3, 0.8, n) * dx, "residual": g.uniform(1e-3, 1e-2, n),
        "A0": np.eye(3) + alpha * a0[:, :, None] * a0[:, None, :],
        "allow": g.random(n) < 0.8, "interface_mask": iface, "n_interface": nrm,
    }
    cfg = dict(n_grid=ng, grid_lim=1.0, dt=float(g.choice([5e-5, 1e-4])), grid_damp=float(g.choice([1.0, 0.999])),
               f_clamp=None if g.random() < 0.5 else (0.9, 1.1), damage_every=int(g.choice([1, 2])),
               gravity=tuple(float(v) for v in g.normal(0.0, 5.0, 3)))
    step0 = int(g.integers(0, 4))
    tens = {k: torch.from_numpy(np.ascontiguousarray(v)) for k, v in t.items()}
    for k in ("allow", "interface_mask"):
        tens[k] = tens[k].bool()
    return cfg, tens, step0


@pytest.mark.parametrize("seed", range(12))
def test_random_scene_one_step_fp64(seed):
    from peel_mpm_amd import AnisoMPM
    cfg, tens, step0 = random_scene(seed)
    n = tens["x"].shape[0]
    print(f"random scene seed={seed} n={n} cfg={cfg} step={step0}", flush=True)
    gpu = AnisoMPM(device=DEV, dtype=torch.float64, **cfg)
    for k, v in tens.items():
        setattr(gpu, k, v.to(DEV))
    gpu.n = n
    ora = O.OracleMPM(dtype=torch.float64, **cfg).load({k: tens[k] for k in STATE_KEYS + PARAM_KEYS})
    gpu._step = ora._step = step0
    t = 0.37
    gpu.substep(t)
    ora.substep(t)
    errs = compare(gpu, ora, cfg, 1e-10, 1e-9, seed)
    print({k: f"{e:.1e}" for k, e in errs.items()}, flush=True)


def compare(gpu, ora, cfg, tol_state, tol_diag, case, keys=STATE_KEYS + DIAG_KEYS):
    """Norm-wise relative error of each field in ``keys``, held to its
    tolerance."""
    errs = {}
    for k in keys:
        if not hasattr(ora, k):
            continue
        a, b = getattr(gpu, k).double().cpu(), getattr(ora, k).double().cpu()
        assert a.shape == b.shape, (case, k, a.shape, b.shape)
        scale = float(b.norm())
        if k == "lap":
            scale = max(scale, cfg["n_grid"] ** 2 * float(ora.d.double().norm()))
        errs[k] = float((a - b).norm()) / scale if scale > 0 else float(a.norm())
        assert errs[k] <= (tol_diag if k in DIAG_KEYS else tol_state), (case, k, errs)
    return errs


@pytest.mark.parametrize("seed", range(12))
def test_random_scene_one_pcg_step_fp64(seed):
    """The implicit damage mode on the same random scenes (their first 1,000
    particles: the oracle's direct solve grows with the active grid): the
    GPU's PCG solve (to 1e-13 on the preconditioned residual) against the
    oracle's direct solve of the same backward-Euler system."""
    from peel_mpm_amd import AnisoMPM
    cfg, tens, step0 = random_scene(seed)
    tens = {k: v[:1000].contiguous() for k, v in tens.items()}
    n = tens["x"].shape[0]
    print(f"random pcg scene seed={seed} n={n} step={step0}", flush=True)
    gpu = AnisoMPM(device=DEV, dtype=torch.float64, damage_solver="pcg", pcg_tol=1e-13, pcg_max_iter=1000, **cfg)
    for k, v in tens.items():
        setattr(gpu, k, v.to(DEV))
    gpu.n = n
    ora = O.OracleMPM(dtype=torch.float64, damage_solver="implicit", **cfg).load(
        {k: tens[k] for k in STATE_KEYS + PARAM_KEYS})
    gpu._step = ora._step = step0
    gpu.substep(0.37)
    ora.substep(0.37)
    errs = compare(gpu, ora, cfg, 1e-9, 1e-8, seed)
    print({k: f"{e:.1e}" for k, e in errs.items()}, gpu.last_damage_solve(), flush=True)


@pytest.mark.parametrize("dtype,solver", [(torch.float32, "explicit"), (torch.float64, "explicit"),
                                          (torch.float64, "pcg")])
def test_empty_scene_then_a_scene(dtype, solver):
    """A solver whose only object has no particles steps as the oracle does
    on empty arrays (every output empty, nothing active); the first object
    added after those steps then steps like a fresh scene (the workspace
    grows from zero particles)."""
    from peel_mpm_amd import AnisoMPM
    cfg, tens, step0 = random_scene(9)
    print(f"empty scene dtype={dtype} solver={solver}", flush=True)
    kw = dict(damage_solver="pcg", pcg_tol=1e-13, pcg_max_iter=1000) if solver == "pcg" else {}
    gpu = AnisoMPM(device=DEV, dtype=dtype, **kw, **cfg)
    gpu.add_object(torch.zeros(0, 3), torch.zeros(0))
    assert gpu.n == 0
    ora = O.OracleMPM(dtype=dtype, damage_solver="implicit" if solver == "pcg" else "explicit", **cfg).load(
        {k: tens[k][:0].to(dtype) if tens[k].is_floating_point() else tens[k][:0] for k in STATE_KEYS + PARAM_KEYS})
    for s in range(3):
        gpu.substep(s * cfg["dt"])
        ora.substep(s * cfg["dt"])
    torch.cuda.synchronize()
    compare(gpu, ora, cfg, 0.0, 0.0, "empty")
    assert gpu._step == 3 and gpu.last_step_stats() == (0, 0)
    tens = {k: v[:1000].contiguous() for k, v in tens.items()}
    if dtype == torch.float32:
        tens = {k: v.float() if v.is_floating_point() else v for k, v in tens.items()}
    gpu.add_object(torch.zeros(0, 3), torch.zeros(0))
    for k, v in tens.items():
        setattr(gpu, k, v.to(DEV))
    gpu.n = tens["x"].shape[0]
    ora = O.OracleMPM(dtype=dtype, damage_solver="implicit" if solver == "pcg" else "explicit", **cfg).load(
        {k: tens[k] for k in STATE_KEYS + PARAM_KEYS})
    gpu._step = ora._step = step0
    gpu.substep(0.37)
    ora.substep(0.37)
    if dtype == torch.float32:       # the kinematic state, held as the fp32 golden cases hold it
        errs = compare(gpu, ora, cfg, 1e-5, 1e-4, (dtype, solver), keys=("x", "v", "F", "C"))
    else:
        errs = compare(gpu, ora, cfg, 1e-9, 1e-8, (dtype, solver))
    print(gpu.n, {k: f"{e:.1e}" for k, e in errs.items()}, flush=True)
