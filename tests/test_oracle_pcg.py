"""Pins of the implicit-damage oracle (oracle/pcg_damage.py), which has no
counterpart in the reference: closed forms, brute force and a second
solver, each chosen to catch a dropped term, a wrong index or a transposed
operand."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

from oracle import cpu_step as O
from oracle import pcg_damage as P


def scene(n=400, ng=16, seed=0):
    rng = np.random.default_rng(seed)
    dx = 1.0 / ng
    x = (4 + 6 * rng.random((n, 3))) * dx
    d = rng.random(n) * 0.5
    D = rng.random(n) * 20.0
    eta = np.full(n, 0.01) * (1 + rng.random(n))
    l0 = np.full(n, 0.5 * dx)
    return x, d, D, eta, l0, ng, 1.0 / dx


def test_cubic_values_match_the_step_and_slopes_match_differences():
    rng = np.random.default_rng(1)
    x = rng.random((200, 3)) * 0.8 + 0.1
    inv_dx = 20.0
    base, N, dN = P.cubic_values_and_slopes(x, inv_dx)
    b2, _, wc, _ = O.cubic_weights(torch.from_numpy(x), inv_dx)
    assert np.array_equal(base, b2.numpy())
    np.testing.assert_allclose(N, wc.numpy(), rtol=0, atol=1e-15)
    np.testing.assert_allclose(N.sum(1), 1.0, atol=1e-14)          # partition of unity
    np.testing.assert_allclose(dN.sum(1), 0.0, atol=1e-14)         # its derivative
    h = 1e-6
    _, Np, _ = P.cubic_values_and_slopes(x + h, inv_dx)
    _, Nm, _ = P.cubic_values_and_slopes(x - h, inv_dx)
    fd = (Np - Nm) / (2 * h * inv_dx)                              # d/dX in grid units
    ok = np.abs(((x * inv_dx) % 1.0) - 0.5) < 0.49                 # away from knots
    np.testing.assert_allclose(fd[ok[:, None, :].repeat(4, 1)], dN[ok[:, None, :].repeat(4, 1)], atol=1e-7)


def test_operator_is_spd_and_annihilates_constants():
    x, d, D, eta, l0, ng, inv_dx = scene()
    # a long step (c = dt/eta ~ 1) so the Laplacian term is of the mass term's order
    A, b, m, act, Nmat = P.assemble(x, d, D, eta, l0, 1e-2, ng, inv_dx)
    K = A - P.sp.diags(m)
    assert abs(A - A.T).max() < 1e-12 * abs(A).max()
    np.testing.assert_allclose(K @ np.ones(len(act)), 0.0, atol=1e-9 * abs(K).max())
    ev = np.linalg.eigvalsh(A.toarray())
    assert ev.min() > 0
    # the Laplacian term is not lost: K is far from zero
    assert abs(K).max() > 1e-2 * m.max()


def test_single_particle_stiffness_brute_force():
    ng, inv_dx = 12, 12.0
    x = np.array([[0.43, 0.51, 0.47]])
    eta, l0, dt = np.array([0.02]), np.array([0.04]), 1e-3
    A, b, m, act, Nmat = P.assemble(x, np.array([0.3]), np.array([5.0]), eta, l0, dt, ng, inv_dx)
    K = (A - P.sp.diags(m)).toarray()
    # brute force: nodes and gradients by direct loops over the 64 nodes
    X = x[0] * inv_dx
    base = np.floor(X).astype(int) - 1

    def Nw(t):
        t = abs(t)
        return (2 / 3 - t * t + 0.5 * t ** 3) if t < 1 else ((2 - t) ** 3 / 6 if t < 2 else 0.0)

    def dNw(t):
        s, t = np.sign(t), abs(t)
        return s * ((-2 * t + 1.5 * t * t) if t < 1 else (-0.5 * (2 - t) ** 2 if t < 2 else 0.0))

    nodes, grads = [], []
    for a in range(4):
        for bb in range(4):
            for e in range(4):
                i = base + np.array([a, bb, e])
                r = X - i
                g = inv_dx * np.array([dNw(r[0]) * Nw(r[1]) * Nw(r[2]), Nw(r[0]) * dNw(r[1]) * Nw(r[2]),
                                       Nw(r[0]) * Nw(r[1]) * dNw(r[2])])
                nodes.append((i[0] * ng + i[1]) * ng + i[2])
                grads.append(-g)      # d/dx of N(X - i) w.r.t. the particle = grad of the nodal shape at x
    c = dt / eta[0]
    kap = c * l0[0] ** 2
    pos = {n: k for k, n in enumerate(act)}
    Kb = np.zeros_like(K)
    for i, gi in zip(nodes, grads):
        for j, gj in zip(nodes, grads):
            Kb[pos[i], pos[j]] += kap * gi @ gj
    np.testing.assert_allclose(K, Kb, atol=1e-12 * abs(Kb).max())


def test_uniform_field_closed_form():
    x, d, D, eta, l0, ng, inv_dx = scene(seed=2)
    d0, D0, e0 = 0.2, 7.0, 0.015
    n = x.shape[0]
    dt = 1e-4
    dn_p, dn, act = P.implicit_update(x, np.full(n, d0), np.full(n, D0), np.full(n, e0), l0, np.ones(n, bool),
                                      dt, ng, inv_dx)
    c = dt / e0
    want = (d0 + c * D0) / (1 + c * (1 + D0))
    np.testing.assert_allclose(dn, want, rtol=1e-12)
    np.testing.assert_allclose(dn_p, want, rtol=1e-12)


def test_direct_solve_matches_conjugate_gradients():
    x, d, D, eta, l0, ng, inv_dx = scene(seed=3)
    A, b, m, act, Nmat = P.assemble(x, d, D, eta, l0, 5e-4, ng, inv_dx)
    xd = P.solve(A, b)
    xc, info = spla.cg(A, b, rtol=1e-13, maxiter=500, M=P.sp.diags(1.0 / A.diagonal()))
    assert info == 0
    np.testing.assert_allclose(xd, xc, rtol=1e-9, atol=1e-12)
    assert np.linalg.norm(A @ xd - b) <= 1e-12 * np.linalg.norm(b)


def test_irreversibility_and_allow_gate():
    x, d, D, eta, l0, ng, inv_dx = scene(seed=4)
    D = np.zeros_like(D)                   # no drive: the field only diffuses and decays
    allow = np.arange(len(d)) % 3 != 0
    dn_p, _, _ = P.implicit_update(x, d, D, eta, l0, allow, 1e-4, ng, inv_dx)
    assert np.all(dn_p[allow] >= d[allow])
    assert np.array_equal(dn_p[~allow], d[~allow])
    assert np.all((dn_p >= 0) & (dn_p <= 1))


@pytest.mark.parametrize("dtype", [torch.float64])
def test_oracle_step_implicit_mode_runs_and_bounds_damage(dtype):
    from peel_mpm_amd import scenes
    spec = scenes.slab_block(n_grid=16, cells=(4, 2, 4), origin=(6, 6, 6), ppc=2, seed=1, sigma_c=1.0, dtype=dtype)
    sim = scenes.make_solver(O.OracleMPM, spec, device="cpu")
    sim.damage_solver = "implicit"
    d0 = sim.d.clone()
    for s in range(3):
        sim.substep(s * sim.dt)
    assert bool(torch.isfinite(sim.d).all())
    assert bool((sim.d >= d0).all()) and float(sim.d.max()) <= 1.0
    assert float(sim.d.max()) > 0.0        # sigma_c = 1 drives damage at once
