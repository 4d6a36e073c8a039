"""CPU pins of the Option A oracle (oracle/variational.py), which restates
the reference's energy / directional_pk1 (anisompm.py:129-187), and the
argument checks of its device entry point (no GPU work).

* the reference's own gate (1_fd_check.py:53-76): P equals the central
  difference of the energy to 1e-6 relative at d = 0, 0.5, 0.9 on its random
  near-identity F; its recorded run (out/validate.jsonl) reached 2.96e-10;
* closed forms: uniaxial stretch along n0 (matrix + mode I, mode I off in
  compression) and simple shear across n0 (mode II = gamma_II / 2 g^2);
* zero energy and stress at F = I and at a pure rotation; frame invariance
  Psi(QF) = Psi(F), P(QF) = Q P(F);
* Option B (the step's projection split, oracle/cpu_step.directional_tau)
  on the reference's Gate 2 and Gate 3 samples (1_fd_check.py:79-120):
  the d=0 reconstruction and the d=1 kept fractions against the values its
  recorded run wrote to out/validate.jsonl.
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from oracle import variational as O

PARAMS = dict(mu=10.0, lam=20.0, gamma_I=5.0, gamma_II=3.0)
RECORDED_FD_REL = 2.964987481764463e-10      # optionA_max_rel_err in the reference's out/validate.jsonl
# the reference's recorded Option B gates (out/validate.jsonl): d=0 reconstruction error, and at d=1 the
# kept fractions of the normal traction, in-plane stress and interfacial shear, residual k_r = 0.005
RECORDED_RECON = 1.4210854715202004e-14
RECORDED_KEPT = (0.00499999999993469, 0.9999999994358679, 0.005000003081451347)
GATE_RES = 0.005


def gate_inputs(n=64, seed=0):
    """The reference gate's samples (1_fd_check.py:55-61)."""
    torch.manual_seed(seed)
    F = torch.eye(3, dtype=torch.float64).repeat(n, 1, 1) + 0.15 * torch.randn(n, 3, 3, dtype=torch.float64)
    F = F[torch.linalg.det(F) > 0.2]
    n0 = torch.randn(F.shape[0], 3, dtype=torch.float64)
    return F, n0 / n0.norm(dim=-1, keepdim=True)


def fd_gradient(energy, F, n0, d, eps=1e-6):
    P = torch.zeros_like(F)
    for i in range(3):
        for j in range(3):
            Fp, Fm = F.clone(), F.clone()
            Fp[:, i, j] += eps
            Fm[:, i, j] -= eps
            P[:, i, j] = (energy(Fp, n0, d, **PARAMS) - energy(Fm, n0, d, **PARAMS)) / (2 * eps)
    return P


def test_reference_fd_gate():
    F, n0 = gate_inputs()
    worst = 0.0
    for dv in (0.0, 0.5, 0.9):
        d = torch.full((F.shape[0],), dv, dtype=torch.float64)
        P = O.directional_pk1(F, n0, d, **PARAMS)
        err = (P - fd_gradient(O.energy, F, n0, d)).abs().max().item() / P.abs().max().item()
        worst = max(worst, err)
    assert worst < 1e-6 and worst < 100 * RECORDED_FD_REL, worst


def test_uniaxial_stretch_closed_form():
    n0 = torch.tensor([[1.0, 0.0, 0.0]], dtype=torch.float64)
    mu, lam, g1 = PARAMS["mu"], PARAMS["lam"], PARAMS["gamma_I"]
    for s, dv in ((1.2, 0.3), (0.8, 0.3)):
        F = torch.diag(torch.tensor([s, 1.0, 1.0], dtype=torch.float64))[None]
        d = torch.tensor([dv], dtype=torch.float64)
        gI = (1 - dv) ** 2 + 1e-4
        opening = max(s - 1.0, 0.0)
        want = (mu + 0.5 * lam) * (s - 1) ** 2 + gI * 0.5 * g1 * opening ** 2
        assert math.isclose(float(O.energy(F, n0, d, **PARAMS)), want, rel_tol=1e-12)
        P = O.directional_pk1(F, n0, d, **PARAMS)[0]
        # dPsi/ds along the stretch: 2 mu (s-1) + lam (s-1) + gI g1 <s-1>
        assert math.isclose(float(P[0, 0]), 2 * mu * (s - 1) + lam * (s - 1) + gI * g1 * opening, rel_tol=1e-12)


def test_simple_shear_mode_two():
    g = 0.3
    F = torch.eye(3, dtype=torch.float64)[None].clone()
    F[0, 0, 1] = g                                   # x += g y: shear across the normal e_y
    n0 = torch.tensor([[0.0, 1.0, 0.0]], dtype=torch.float64)
    d = torch.tensor([0.0], dtype=torch.float64)
    zero = dict(PARAMS, mu=0.0, lam=0.0, gamma_I=0.0)
    # I4 = 1 + g^2, I5 = g^2 + (1 + g^2)^2: I5 - I4^2 = g^2
    assert math.isclose(float(O.energy(F, n0, d, **zero)), (1 + 1e-4) * 0.5 * PARAMS["gamma_II"] * g * g,
                        rel_tol=1e-12)


def random_rotation(n, seed):
    q, r = torch.linalg.qr(torch.randn(n, 3, 3, generator=torch.Generator().manual_seed(seed), dtype=torch.float64))
    q = q * torch.sign(torch.diagonal(r, dim1=1, dim2=2))[:, None, :]
    q[torch.linalg.det(q) < 0, :, 0] *= -1
    return q


def test_rest_rotation_and_frame_invariance():
    F, n0 = gate_inputs(32, seed=3)
    d = torch.rand(F.shape[0], dtype=torch.float64)
    Q = random_rotation(F.shape[0], 4)
    I = torch.eye(3, dtype=torch.float64).expand_as(F).contiguous()
    assert O.energy(I, n0, d, **PARAMS).abs().max() < 1e-12
    assert O.directional_pk1(I, n0, d, **PARAMS).abs().max() < 1e-12
    assert O.energy(Q, n0, d, **PARAMS).abs().max() < 1e-12
    QF = Q @ F
    assert torch.allclose(O.energy(QF, n0, d, **PARAMS), O.energy(F, n0, d, **PARAMS), rtol=1e-11, atol=1e-11)
    assert torch.allclose(O.directional_pk1(QF, n0, d, **PARAMS), Q @ O.directional_pk1(F, n0, d, **PARAMS),
                          rtol=1e-10, atol=1e-10)


def test_entry_point_rejects_bad_arguments_before_any_gpu_work():
    from peel_mpm_amd import _lib
    L = _lib.lib()
    p = C.c_void_p(16)                   # never dereferenced: every call below fails validation
    assert L.pmpm_directional_pk1(7, 4, p, p, p, None, 1.0, None, 1.0, None, 1.0, None, 1.0, p, p, None) != 0
    assert L.pmpm_directional_pk1(_lib.PMPM_F64, -1, p, p, p, None, 1.0, None, 1.0, None, 1.0, None, 1.0, p, p,
                                  None) != 0
    assert L.pmpm_directional_pk1(_lib.PMPM_F64, 4, None, p, p, None, 1.0, None, 1.0, None, 1.0, None, 1.0, p, p,
                                  None) != 0
    assert L.pmpm_directional_pk1(_lib.PMPM_F64, 4, p, p, p, None, 1.0, None, 1.0, None, 1.0, None, 1.0, None, None,
                                  None) != 0
    assert b"null" in L.pmpm_last_error()
    assert L.pmpm_directional_pk1(_lib.PMPM_F64, 0, None, None, None, None, 1.0, None, 1.0, None, 1.0, None, 1.0,
                                  None, None, None) == 0


def corotated(F, mu=10.0, lam=20.0):
    """The plain corotated Kirchhoff stress of the gates (anisompm.py:230-241)."""
    U, S, Vh = torch.linalg.svd(F)
    R = U @ Vh
    J = torch.linalg.det(F)
    I3 = torch.eye(3, dtype=F.dtype).expand_as(F)
    return 2.0 * mu * (F - R) @ F.transpose(1, 2) + (lam * (J - 1.0) * J).view(-1, 1, 1) * I3


def kept_fractions(tau_c, tau1, F, n0):
    """release_check's measures (anisompm.py:279-291): normal, in-plane, shear."""
    I3 = torch.eye(3, dtype=F.dtype).expand_as(F)
    m = torch.einsum('pij,pj->pi', F, n0)
    m = m / (m.norm(dim=1, keepdim=True) + 1e-9)
    Pm = torch.einsum('pi,pj->pij', m, m)
    Qm = I3 - Pm
    nrm = lambda t: t.reshape(F.shape[0], -1).norm(dim=1)
    tnc = torch.einsum('pi,pij,pj->p', m, tau_c, m).abs()
    tn1 = torch.einsum('pi,pij,pj->p', m, tau1, m).abs()
    ip = (nrm(Qm @ tau1 @ Qm) / (nrm(Qm @ tau_c @ Qm) + 1e-9)).min().item()
    sh = (nrm(Pm @ tau1 @ Qm + Qm @ tau1 @ Pm) / (nrm(Pm @ tau_c @ Qm + Qm @ tau_c @ Pm) + 1e-9)).max().item()
    return (tn1 / (tnc + 1e-9)).max().item(), ip, sh


def test_option_b_gates_against_the_recorded_run():
    from oracle.cpu_step import directional_tau
    F, n0 = gate_inputs(64, seed=1)                    # Gate 2 samples (1_fd_check.py:81-88)
    tau_c = corotated(F)
    d0 = torch.zeros(F.shape[0], dtype=torch.float64)
    err = (directional_tau(tau_c, F, d0, n0, 0.0) - tau_c).abs().max().item()
    assert err < 1e-9 and err < 100 * RECORDED_RECON, err
    F, n0 = gate_inputs(64, seed=2)                    # Gate 3 samples (1_fd_check.py:103-111)
    tau_c = corotated(F)
    d1 = torch.ones(F.shape[0], dtype=torch.float64)
    kept = kept_fractions(tau_c, directional_tau(tau_c, F, d1, n0, GATE_RES, tension_gate=False), F, n0)
    for got, want in zip(kept, RECORDED_KEPT):
        assert abs(got - want) < 1e-9, (kept, RECORDED_KEPT)
