"""Option A on the device (peel_mpm_amd.variational) against the oracle's
restatement of the reference's energy / directional_pk1 (needs a GPU).

* fp64: energy and stress within 1e-13 (norm-wise) of the oracle, which takes
  R from an SVD where the kernel takes it from the Jacobi eigenvectors of
  F^T F; fp32 within 2e-6;
* the reference's finite-difference gate (1_fd_check.py:53-76) on the
  device energy itself: its central differences match the device stress to
  1e-6 relative at d = 0, 0.5, 0.9;
* per-particle parameters (N,) against scalars; stress-only and energy-only
  launches against the joint one; tau = P F^T;
* Option B through the step's own device code (pmpm_directional_kirchhoff):
  the reference's Gate 2 and Gate 3 (1_fd_check.py:79-120) against its
  recorded run, and the oracle's split at random d, with and without the
  tension gate, per-particle residuals.
"""
from __future__ import annotations

import pytest
import torch

from oracle import variational as O
from tests.golden_cases import rel_err
from tests.test_oracle_variational import (GATE_RES, PARAMS, RECORDED_KEPT, RECORDED_RECON, corotated, gate_inputs,
                                           kept_fractions)

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def inputs(n, seed, dtype):
    g = torch.Generator().manual_seed(seed)
    F = torch.eye(3, dtype=torch.float64).repeat(n, 1, 1) + 0.2 * torch.randn(n, 3, 3, generator=g, dtype=torch.float64)
    F = F[torch.linalg.det(F) > 0.2]
    n0 = torch.randn(F.shape[0], 3, generator=g, dtype=torch.float64)
    n0 = n0 / n0.norm(dim=-1, keepdim=True)
    d = torch.rand(F.shape[0], generator=g, dtype=torch.float64)
    return F.to(dtype), n0.to(dtype), d.to(dtype)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 2e-6)])
def test_matches_the_oracle(dtype, tol):
    from peel_mpm_amd import variational as V
    F, n0, d = inputs(100_000, 1, dtype)
    print(f"option A {dtype} n={F.shape[0]}", flush=True)
    psi, P = V.energy_and_pk1(F.to(DEV), n0.to(DEV), d.to(DEV), **PARAMS)
    F64, n064, d64 = F.double(), n0.double(), d.double()
    want_psi = O.energy(F64, n064, d64, **PARAMS)
    want_P = O.directional_pk1(F64, n064, d64, **PARAMS)
    e_psi, e_P = rel_err(psi, want_psi), rel_err(P, want_P)
    print(f"rel err psi {e_psi:.2e} P {e_P:.2e}", flush=True)
    assert e_psi <= tol and e_P <= tol


def test_reference_fd_gate_on_the_device_energy():
    from peel_mpm_amd import variational as V
    F, n0 = gate_inputs()
    F, n0 = F.to(DEV), n0.to(DEV)
    eps = 1e-6
    worst = 0.0
    for dv in (0.0, 0.5, 0.9):
        d = torch.full((F.shape[0],), dv, dtype=torch.float64, device=DEV)
        P = V.directional_pk1(F, n0, d, **PARAMS)
        P_fd = torch.zeros_like(F)
        for i in range(3):
            for j in range(3):
                Fp, Fm = F.clone(), F.clone()
                Fp[:, i, j] += eps
                Fm[:, i, j] -= eps
                P_fd[:, i, j] = (V.energy(Fp, n0, d, **PARAMS) - V.energy(Fm, n0, d, **PARAMS)) / (2 * eps)
        worst = max(worst, ((P - P_fd).abs().max() / P.abs().max()).item())
    print(f"device FD gate worst rel err {worst:.3e}", flush=True)
    assert worst < 1e-6


def test_per_particle_parameters_and_partial_outputs():
    from peel_mpm_amd import variational as V
    F, n0, d = (t.to(DEV) for t in inputs(5000, 2, torch.float64))
    n = F.shape[0]
    g = torch.Generator(device=DEV).manual_seed(3)
    mu = 5.0 + 10.0 * torch.rand(n, generator=g, device=DEV, dtype=torch.float64)
    lam = 10.0 + 20.0 * torch.rand(n, generator=g, device=DEV, dtype=torch.float64)
    psi, P = V.energy_and_pk1(F, n0, d, mu, lam, 5.0, 3.0)
    want_psi = O.energy(F.cpu(), n0.cpu(), d.cpu(), mu.cpu(), lam.cpu(), 5.0, 3.0)
    want_P = O.directional_pk1(F.cpu(), n0.cpu(), d.cpu(), mu.cpu()[:, None, None], lam.cpu(), 5.0, 3.0)
    assert rel_err(psi, want_psi) <= 1e-11 and rel_err(P, want_P) <= 1e-11
    assert torch.equal(V.energy(F, n0, d, mu, lam, 5.0, 3.0), psi)
    assert torch.equal(V.directional_pk1(F, n0, d, mu, lam, 5.0, 3.0), P)
    tau = V.kirchhoff(F, n0, d, mu, lam, 5.0, 3.0)
    assert torch.allclose(tau, P @ F.transpose(1, 2), rtol=1e-13, atol=1e-13)


def test_option_b_gates_on_the_step_code():
    from peel_mpm_amd import variational as V
    F, n0 = gate_inputs(64, seed=1)
    tau_c = corotated(F)
    d0 = torch.zeros(F.shape[0], dtype=torch.float64)
    split = V.directional_kirchhoff(tau_c.to(DEV), F.to(DEV), d0.to(DEV), n0.to(DEV), residual=GATE_RES).cpu()
    err = (split - tau_c).abs().max().item()
    print(f"gate 2 reconstruction {err:.2e} (recorded {RECORDED_RECON:.2e})", flush=True)
    assert err < 1e-9 and err < 100 * RECORDED_RECON
    F, n0 = gate_inputs(64, seed=2)
    tau_c = corotated(F)
    d1 = torch.ones(F.shape[0], dtype=torch.float64)
    tau1 = V.directional_kirchhoff(tau_c.to(DEV), F.to(DEV), d1.to(DEV), n0.to(DEV), residual=GATE_RES,
                                   tension_gate=False).cpu()
    kept = kept_fractions(tau_c, tau1, F, n0)
    print(f"gate 3 kept {kept} (recorded {RECORDED_KEPT})", flush=True)
    for got, want in zip(kept, RECORDED_KEPT):
        assert abs(got - want) < 1e-9


@pytest.mark.parametrize("gate", [True, False])
def test_option_b_matches_the_oracle(gate):
    from oracle.cpu_step import directional_tau
    from peel_mpm_amd import variational as V
    F, n0, d = inputs(20_000, 5, torch.float64)
    tau_c = corotated(F)
    res = 1e-3 + 0.01 * torch.rand(F.shape[0], dtype=torch.float64, generator=torch.Generator().manual_seed(6))
    got = V.directional_kirchhoff(tau_c.to(DEV), F.to(DEV), d.to(DEV), n0.to(DEV), residual=res.to(DEV),
                                  tension_gate=gate)
    want = directional_tau(tau_c, F, d, n0, res, tension_gate=gate)
    assert rel_err(got, want) <= 1e-13


def test_reference_gate_functions_over_the_device_split():
    """reconstruction_check / release_check with the reference's signatures,
    on CPU tensors as 1_fd_check.py passes them."""
    from peel_mpm_amd import variational as V
    F, n0 = gate_inputs(64, seed=1)
    U, S, Vh = torch.linalg.svd(F)
    err = V.reconstruction_check(F, U @ Vh, torch.linalg.det(F), n0, mu=10.0, lam=20.0, verbose=False)
    assert err < 1e-9
    F, n0 = gate_inputs(64, seed=2)
    U, S, Vh = torch.linalg.svd(F)
    ok, nk, ik, sk = V.release_check(F, U @ Vh, torch.linalg.det(F), n0, mu=10.0, lam=20.0, res=GATE_RES,
                                     verbose=False)
    assert ok
    for got, want in zip((nk, ik, sk), RECORDED_KEPT):
        assert abs(got - want) < 1e-9


def test_cpu_inputs_as_the_reference_gate_passes_them():
    """1_fd_check.py's Gate 1 calls energy / directional_pk1 on CPU fp64
    tensors: they compute on the GPU and return CPU tensors, equal to the
    CUDA-path results."""
    from peel_mpm_amd import variational as V
    F, n0 = gate_inputs(64, seed=0)
    d = torch.full((F.shape[0],), 0.5, dtype=torch.float64)
    psi, P = V.energy(F, n0, d, **PARAMS), V.directional_pk1(F, n0, d, **PARAMS)
    assert psi.device.type == "cpu" and P.device.type == "cpu"
    psi_g, P_g = V.energy_and_pk1(F.to(DEV), n0.to(DEV), d.to(DEV), **PARAMS)
    assert torch.equal(psi, psi_g.cpu()) and torch.equal(P, P_g.cpu())
