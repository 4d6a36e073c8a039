"""Pin the oracle (CPU): against the reference's golden vectors bit-for-bit,
and independently against closed forms, LAPACK and brute force, so that a
dropped term, a wrong sign/index or a transposed operand in the oracle fails."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from oracle import cpu_step as O
from tests.golden_cases import (DIAG_KEYS, STATE_KEYS, STEP_CASES, attach_boundaries, case_config,
                                case_dtype, case_tensors, load_npz)


# ------------------------------------------------------------ golden vectors
def test_sym_eig_matches_reference_bits():
    z = load_npz("sym_eig.npz")
    for tag in ("32", "64"):
        ev, V = O.sym_eig(torch.from_numpy(z["A" + tag]))
        assert np.array_equal(ev.numpy(), z["evals" + tag], equal_nan=True)
        assert np.array_equal(V.numpy(), z["V" + tag], equal_nan=True)


@pytest.mark.parametrize("ng", [64, 24, 48])
def test_weights_and_node_index_match_reference_bits(ng):
    z = load_npz("weights.npz")
    x = torch.from_numpy(z[f"x_{ng}"])
    inv_dx = 1.0 / (1.0 / ng)
    base, fx, w = O.quad_weights(x, inv_dx)
    bc, f, wc, wc2 = O.cubic_weights(x, inv_dx)
    assert np.array_equal(base.numpy(), z[f"base_{ng}"])
    assert np.array_equal(fx.numpy(), z[f"fx_{ng}"])
    assert np.array_equal(w.numpy(), z[f"w_{ng}"])
    assert np.array_equal(bc.numpy(), z[f"basec_{ng}"])
    assert np.array_equal(f.numpy(), z[f"fc_{ng}"])
    assert np.array_equal(wc.numpy(), z[f"wc_{ng}"])
    assert np.array_equal(wc2.numpy(), z[f"wc2_{ng}"])
    o27, o64 = O.stencil_offsets(3), O.stencil_offsets(4)
    i27 = torch.stack([O.node_index(base + o27[k], ng) for k in range(27)], 1)
    i64 = torch.stack([O.node_index(bc + o64[k], ng) for k in range(64)], 1)
    assert np.array_equal(i27.numpy(), z[f"idx27_{ng}"])
    assert np.array_equal(i64.numpy(), z[f"idx64_{ng}"])


def run_oracle_case(name):
    rec = load_npz(f"step_{name}.npz")
    cfg, time0, nsteps = case_config(rec)
    sim = O.OracleMPM(dtype=case_dtype(rec), **cfg).load(case_tensors(rec))
    sim._step = int(rec["in__step"])
    attach_boundaries(sim, rec, O.SphereCollider, O.HalfspaceCollider)
    for s in range(nsteps):
        sim.substep(time0 + s * sim.dt)
    return sim, rec


@pytest.mark.parametrize("name", STEP_CASES)
def test_substep_matches_reference_bits(name):
    sim, rec = run_oracle_case(name)
    for k in STATE_KEYS + DIAG_KEYS:
        if "out_" + k not in rec:
            assert not hasattr(sim, k), k          # phi/phi_shear only exist after a damage step
            continue
        got = getattr(sim, k).numpy()
        assert np.array_equal(got, rec["out_" + k]), (name, k, np.abs(got - rec["out_" + k]).max())


@pytest.mark.parametrize("tag", ["32", "64"])
def test_trajectory_matches_reference_bits(tag):
    rec = load_npz(f"traj_slab{tag}.npz")
    dtype = torch.float32 if tag == "32" else torch.float64
    sim = O.OracleMPM(n_grid=32, grid_lim=1.0, dt=1e-4, dtype=dtype)
    sim.load({k: torch.from_numpy(rec["in_" + k].copy()) for k in O.STATE_KEYS + O.PARAM_KEYS})
    top = torch.from_numpy(rec["top"])
    sim.particle_bc.append(lambda s, t: s.v.__setitem__(top, torch.tensor([0.0, 0.5, 0.0], dtype=dtype)))
    marks = sorted({int(k[1:].split("_")[0]) for k in rec if k.startswith("s") and k[1].isdigit()})
    for s in range(max(marks)):
        sim.substep(s * sim.dt)
        if s + 1 in marks:
            for k in O.STATE_KEYS:
                assert np.array_equal(getattr(sim, k).numpy(), rec[f"s{s + 1}_{k}"]), (s + 1, k)


# ------------------------------------------------------------ independent pins
def test_sym_eig_against_lapack_and_la_selftest_recipe():
    """The reference's own LA self-test (anisompm.py:871-916): fp32 Jacobi vs
    fp64 LAPACK, incl. pure shear and 3I, tolerance 1e-3; fp64 Jacobi much tighter."""
    g = torch.Generator().manual_seed(0)
    B = torch.randn(4000, 3, 3, dtype=torch.float64, generator=g)
    A = 0.5 * (B + B.transpose(1, 2))
    A[0] = torch.tensor([[0., 2, 0], [2, 0, 0], [0, 0, 0]], dtype=torch.float64)
    A[1] = 3.0 * torch.eye(3, dtype=torch.float64)
    ref_ev = torch.from_numpy(np.linalg.eigh(A.numpy())[0])
    for dt, tol in ((torch.float32, 1e-3), (torch.float64, 1e-10)):
        ev, Q = O.sym_eig(A.to(dt))
        ev, Q = ev.double(), Q.double()
        assert (ev - ref_ev).abs().max() < tol
        assert (Q @ torch.diag_embed(ev) @ Q.transpose(1, 2) - A).abs().max() < tol
        assert (Q.transpose(1, 2) @ Q - torch.eye(3, dtype=torch.float64)).abs().max() < tol
        assert bool((ev[:, 1:] >= ev[:, :-1]).all())
    # the polar rotation built the solver's way is orthogonal and equals U V^T
    Q1 = torch.linalg.qr(torch.randn(500, 3, 3, dtype=torch.float64, generator=g)).Q
    Q2 = torch.linalg.qr(torch.randn(500, 3, 3, dtype=torch.float64, generator=g)).Q
    sv = 0.2 + 7.8 * torch.rand(500, 3, dtype=torch.float64, generator=g)
    F = Q1 @ torch.diag_embed(sv) @ Q2.transpose(1, 2)
    F = F[torch.linalg.det(F) > 0]
    lam, V = O.sym_eig(F.transpose(1, 2) @ F)
    W = F @ V
    R = (W / (W.norm(dim=1, keepdim=True) + 1e-12)) @ V.transpose(1, 2)
    U, _, Vh = torch.linalg.svd(F)
    assert (R - U @ Vh).abs().max() < 1e-8


def test_sym_eig_sanitises_nan_and_inf():
    A = torch.tensor([[[float("nan"), 1.0, 0.0], [1.0, 2.0, 0.0], [0.0, 0.0, float("inf")]]])
    ev, V = O.sym_eig(A)
    # NaN -> 0 leaves the block [[0,1],[1,2]] (eigenvalues 1 -+ sqrt 2); inf -> FLT_MAX
    assert np.allclose(ev[0, :2].numpy(), [1 - math.sqrt(2), 1 + math.sqrt(2)], atol=1e-6)
    assert float(ev[0, 2]) == float(torch.finfo(torch.float32).max)
    assert np.isfinite(V.numpy()).all()


@pytest.mark.parametrize("ng", [64, 24])
def test_weights_closed_forms(ng):
    g = torch.Generator().manual_seed(ng)
    x = torch.rand(5000, 3, dtype=torch.float64, generator=g) * 0.8 + 0.1
    inv_dx = float(ng)
    base, fx, w = O.quad_weights(x, inv_dx)
    assert bool(((fx >= 0.5) & (fx < 1.5)).all())
    assert torch.allclose(w.sum(1), torch.ones_like(fx), atol=1e-14)             # partition of unity
    nodes = torch.arange(3, dtype=torch.float64).view(1, 3, 1)
    assert torch.allclose((w * (nodes - fx.unsqueeze(1))).sum(1), torch.zeros_like(fx), atol=1e-14)
    bc, f, wc, wc2 = O.cubic_weights(x, inv_dx)
    assert torch.equal(bc + 1, torch.floor(x * inv_dx).long())
    assert torch.allclose(wc.sum(1), torch.ones_like(f), atol=1e-14)
    assert torch.allclose(wc2.sum(1), torch.zeros_like(f), atol=1e-14)
    nodes4 = torch.arange(4, dtype=torch.float64).view(1, 4, 1)
    assert torch.allclose((wc * (nodes4 - 1.0 - f.unsqueeze(1))).sum(1), torch.zeros_like(f), atol=1e-13)
    # second derivative of the cubic weights by central differences in f
    h = 1e-4
    _, _, wp, _ = O.cubic_weights(x + h / inv_dx, inv_dx)
    _, _, wm, _ = O.cubic_weights(x - h / inv_dx, inv_dx)
    inner = ((f > 2 * h) & (f < 1 - 2 * h)).unsqueeze(1).expand_as(wc)
    fd = (wp - 2 * wc + wm) / (h * h)
    assert torch.allclose(fd[inner], wc2[inner], atol=1e-5)
    # exact values at f = 0: (1/6, 2/3, 1/6, 0)
    _, _, w0, _ = O.cubic_weights(torch.full((1, 3), 10.0 / ng, dtype=torch.float64), inv_dx)
    assert torch.allclose(w0[0, :, 0], torch.tensor([1 / 6, 2 / 3, 1 / 6, 0.0], dtype=torch.float64))


def test_node_index_brute_force():
    ng = 7
    nodes = torch.tensor([[i, j, k] for i in range(-2, 9, 3) for j in range(-1, 9, 4) for k in (-5, 0, 3, 6, 7, 40)])
    got = O.node_index(nodes, ng)
    for r, (i, j, k) in enumerate(nodes.tolist()):
        ci, cj, ck = (min(max(a, 0), ng - 1) for a in (i, j, k))
        assert int(got[r]) == ci * ng * ng + cj * ng + ck


def brute_p2g(x, mass, v, affine, ng, dx):
    """Python-loop scatter on a tiny grid, written from the MLS-MPM formula."""
    m = np.zeros((ng, ng, ng)); mv = np.zeros((ng, ng, ng, 3))
    for p in range(x.shape[0]):
        X = x[p] / dx
        b = np.floor(X - 0.5).astype(int)
        f = X - b
        w = [0.5 * (1.5 - f) ** 2, 0.75 - (f - 1.0) ** 2, 0.5 * (f - 0.5) ** 2]
        for a in range(3):
            for bb in range(3):
                for c in range(3):
                    wk = w[a][0] * w[bb][1] * w[c][2]
                    dpos = (np.array([a, bb, c]) - f) * dx
                    i, j, k = np.clip(b + [a, bb, c], 0, ng - 1)
                    m[i, j, k] += wk * mass[p]
                    mv[i, j, k] += wk * (mass[p] * v[p] + affine[p] @ dpos)
    return m.reshape(-1), mv.reshape(-1, 3)


def make_sim(n=60, ng=8, seed=0, dtype=torch.float64, **kw):
    g = torch.Generator().manual_seed(seed)
    sim = O.OracleMPM(n_grid=ng, grid_lim=1.0, dt=1e-4, dtype=dtype, **kw)
    t = {}
    t["x"] = 0.3 + 0.4 * torch.rand(n, 3, dtype=dtype, generator=g)
    t["v"] = torch.randn(n, 3, dtype=dtype, generator=g)
    t["F"] = torch.eye(3, dtype=dtype) + 0.05 * torch.randn(n, 3, 3, dtype=dtype, generator=g)
    t["C"] = torch.randn(n, 3, 3, dtype=dtype, generator=g)
    t["d"] = 0.5 * torch.rand(n, dtype=dtype, generator=g)
    t["lap"] = torch.randn(n, dtype=dtype, generator=g)
    t["D_hist"] = torch.zeros(n, dtype=dtype)
    t["vol0"] = torch.full((n,), 1e-4, dtype=dtype)
    t["mass"] = 0.05 + 0.1 * torch.rand(n, dtype=dtype, generator=g)
    t["mu"] = torch.full((n,), 3e3, dtype=dtype)
    t["lam"] = torch.full((n,), 5e3, dtype=dtype)
    t["sigma_c"] = torch.full((n,), 50.0, dtype=dtype)
    t["eta"] = torch.full((n,), 0.01, dtype=dtype)
    t["zeta"] = torch.ones(n, dtype=dtype)
    t["l0"] = torch.full((n,), 0.5 / ng, dtype=dtype)
    t["residual"] = torch.full((n,), 1e-3, dtype=dtype)
    t["A0"] = torch.eye(3, dtype=dtype).expand(n, 3, 3).clone()
    t["allow"] = torch.ones(n, dtype=torch.bool)
    t["interface_mask"] = torch.zeros(n, dtype=torch.bool)
    t["n_interface"] = torch.zeros(n, 3, dtype=dtype)
    return sim.load(t)


def test_p2g_brute_force_and_conservation():
    sim = make_sim()
    tau = torch.randn(sim.n, 3, 3, dtype=torch.float64)
    _, _, _, m_grid, mv_grid = sim._p2g(tau)
    coeff = (-sim.dt * sim.vol0 * 4.0 * sim.inv_dx ** 2).view(-1, 1, 1)
    affine = coeff * tau + sim.mass.view(-1, 1, 1) * sim.C
    m_bf, mv_bf = brute_p2g(sim.x.numpy(), sim.mass.numpy(), sim.v.numpy(), affine.numpy(), sim.n_grid, sim.dx)
    assert np.allclose(m_grid.numpy(), m_bf, rtol=1e-12, atol=1e-15)
    assert np.allclose(mv_grid.numpy(), mv_bf, rtol=1e-12, atol=1e-13)
    # mass and momentum are conserved by the transfer (the affine part sums to zero)
    assert math.isclose(float(m_grid.sum()), float(sim.mass.sum()), rel_tol=1e-12)
    assert torch.allclose(mv_grid.sum(0), (sim.mass.view(-1, 1) * sim.v).sum(0), rtol=1e-10, atol=1e-12)


def test_rest_and_free_fall_invariants():
    sim = make_sim(n=200, ng=16, seed=3, gravity=(0.0, -9.8, 0.0))
    sim.F = torch.eye(3, dtype=torch.float64).expand(sim.n, 3, 3).clone()
    sim.C.zero_(); sim.v.zero_(); sim.d.zero_(); sim.lap.zero_()
    for s in range(5):
        sim.substep(s * sim.dt)
    assert torch.allclose(sim.v, torch.tensor([0.0, -9.8 * 5 * sim.dt, 0.0], dtype=torch.float64).expand_as(sim.v),
                          atol=1e-12)
    assert (sim.F - torch.eye(3, dtype=torch.float64)).abs().max() < 1e-12
    assert float(sim.d.abs().max()) == 0.0


def test_constitutive_closed_forms():
    # tau = 0 at F = I; fixed-corotated stress of a pure stretch diag(s)
    sim = make_sim(n=4, ng=8)
    sim.F = torch.diag_embed(torch.tensor([[1.0, 1.0, 1.0], [1.1, 1.0, 1.0], [0.9, 1.2, 1.05],
                                           [1.3, 1.3, 1.3]], dtype=torch.float64))
    sim.d.zero_(); sim.lap.zero_(); sim.allow[:] = False
    tau, _ = sim._constitutive()
    # R carries the reference's 1e-12 column-norm epsilon: |F - R| = 1e-12 at F = I in fp64
    assert float(tau[0].abs().max()) < 2 * 3e3 * 2e-12
    for p in range(1, 4):
        s = torch.diagonal(sim.F[p])
        J = float(s.prod())
        expect = torch.diag(2 * 3e3 * (s - 1) * s + 5e3 * (J - 1) * J)
        g = (1 - float(sim.d[p])) ** 2 * (1 - 1e-3) + 1e-3
        assert torch.allclose(tau[p], g * expect, rtol=1e-10, atol=1e-8)


def test_damage_single_particle_closed_form():
    """Uniaxial stretch, A0 = I: phi = (sigma_1^2)/sigma_c^2, then the explicit
    update d' = d + dt/eta * relu((1-d) D - (d - l0^2 lap))."""
    sim = make_sim(n=1, ng=8)
    s1 = 1.02
    sim.F = torch.diag(torch.tensor([s1, 1.0, 1.0], dtype=torch.float64)).unsqueeze(0)
    d0, lap0 = 0.1, 3.0
    sim.d[:] = d0; sim.lap[:] = lap0
    sim._constitutive()
    J = s1
    t11 = 2 * 3e3 * (s1 - 1) * s1 + 5e3 * (J - 1) * J
    t22 = 5e3 * (J - 1) * J
    sig = np.array([t11, t22, t22]) / J
    phi = float((np.clip(sig, 0, None) ** 2).sum() / 50.0 ** 2)
    assert math.isclose(float(sim.phi[0]), phi, rel_tol=1e-10)
    D = max(0.0, phi - 1.0)
    dd = d0 + sim.dt / 0.01 * max(0.0, (1 - d0) * D - (d0 - (0.5 / 8) ** 2 * lap0))
    assert math.isclose(float(sim.d[0]), min(max(dd, 0.0), 1.0), rel_tol=1e-9)


def test_directional_split_reconstruction_and_release():
    """d=0 rebuilds the corotated stress; d=1 keeps only the in-plane block
    (the reference's Gates 2-3, 1_fd_check.py:79-120)."""
    g = torch.Generator().manual_seed(5)
    n = 64
    F = torch.eye(3, dtype=torch.float64) + 0.2 * torch.randn(n, 3, 3, dtype=torch.float64, generator=g)
    T = torch.randn(n, 3, 3, dtype=torch.float64, generator=g)
    tau_c = T + T.transpose(1, 2)
    n0 = torch.nn.functional.normalize(torch.randn(n, 3, dtype=torch.float64, generator=g), dim=1)
    res = torch.full((n,), 1e-3, dtype=torch.float64)
    t0 = O.directional_tau(tau_c, F, torch.zeros(n, dtype=torch.float64), n0, res)
    assert (t0 - tau_c).abs().max() < 1e-12
    t1 = O.directional_tau(tau_c, F, torch.ones(n, dtype=torch.float64), n0, res, tension_gate=False)
    m = torch.nn.functional.normalize(torch.einsum('pij,pj->pi', F, n0), dim=1)
    P = torch.einsum('pi,pj->pij', m, m)
    Q = torch.eye(3, dtype=torch.float64) - P
    assert torch.allclose(Q @ t1 @ Q, Q @ tau_c @ Q, atol=1e-12)
    assert torch.allclose(P @ t1 @ P, 1e-3 * (P @ tau_c @ P), atol=1e-9)
    assert torch.allclose(P @ t1 @ Q, 1e-3 * (P @ tau_c @ Q), atol=1e-9)


def test_collider_resolve_closed_forms():
    v = torch.tensor([[1.0, -2.0, 0.5]], dtype=torch.float64)
    n = torch.tensor([[0.0, 1.0, 0.0]], dtype=torch.float64)
    vc = torch.zeros(3, dtype=torch.float64)
    assert torch.allclose(O.collider_resolve(v, vc, n, "slip", 0.0), torch.tensor([[1.0, 0.0, 0.5]], dtype=torch.float64))
    # Coulomb friction: tangential speed reduced by mu*|v_n| (here to zero: 0.4*2 > |vt|)
    out = O.collider_resolve(v, vc, n, "slip", 0.4)
    vt = math.hypot(1.0, 0.5)
    assert torch.allclose(out, torch.tensor([[1.0, 0.0, 0.5]], dtype=torch.float64) * max(0.0, 1 - 0.8 / vt))
    # separating velocity is untouched
    up = torch.tensor([[1.0, 2.0, 0.5]], dtype=torch.float64)
    assert torch.allclose(O.collider_resolve(up, vc, n, "slip", 0.4), up)
    assert torch.allclose(O.collider_resolve(v, vc + 1, n, "sticky", 0.0), torch.ones(1, 3, dtype=torch.float64))
