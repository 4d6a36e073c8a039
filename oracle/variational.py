"""Oracle: a CPU restatement of the reference's Option A directional damage
pair, for checking only.

TEST INFRASTRUCTURE.  Only ``tests/`` may import this module, and only as
the checker; ``peel_mpm_amd`` never imports it.

``energy`` and ``directional_pk1`` follow ``anisompm_py/src/anisompm.py:
129-187`` with torch on the CPU (SVD polar factor, determinant, inverse, in
the reference's order).  Pinned by ``tests/test_oracle_variational.py``:
the reference's own finite-difference gate (``1_fd_check.py:53-76``: P is
the central-difference gradient of the energy to 1e-6 relative at d = 0,
0.5, 0.9, against the 2.96e-10 its recorded run reached in
``out/validate.jsonl``), zero energy and stress at F = I, frame invariance
(Psi(QF) = Psi(F), P(QF) = Q P(F)), the pure-matrix limit gamma = 0 against
the fixed-corotated closed form 2 mu (F - R) + lam (J - 1) J F^-T, and the
mode-I term's switch-off in compression.
"""
from __future__ import annotations

import torch


def gI(d):
    return (1.0 - d) ** 2 + 1e-4


def gII(d):
    return (1.0 - d) ** 2 + 1e-4


def energy(F, n0, d, mu, lam, gamma_I, gamma_II):
    """anisompm.py:137-159."""
    U, S, Vh = torch.linalg.svd(F)
    R = U @ Vh
    J = torch.linalg.det(F)
    psi_mat = mu * ((F - R) ** 2).sum((-1, -2)) + 0.5 * lam * (J - 1) ** 2
    Fn = torch.einsum('pij,pj->pi', F, n0)
    I4 = (Fn * Fn).sum(-1)
    lam_n = torch.sqrt(I4)
    C = torch.einsum('pki,pkj->pij', F, F)
    Cn = torch.einsum('pij,pj->pi', C, n0)
    I5 = (Cn * Cn).sum(-1)
    psi_I = 0.5 * gamma_I * torch.clamp(lam_n - 1.0, min=0.0) ** 2
    psi_II = 0.5 * gamma_II * (I5 - I4 ** 2)
    return psi_mat + gI(d) * psi_I + gII(d) * psi_II


def directional_pk1(F, n0, d, mu, lam, gamma_I, gamma_II):
    """anisompm.py:162-187."""
    U, S, Vh = torch.linalg.svd(F)
    R = U @ Vh
    J = torch.linalg.det(F)
    FinvT = torch.linalg.inv(F).transpose(-1, -2)
    P_mat = 2 * mu * (F - R) + (lam * (J - 1) * J)[..., None, None] * FinvT
    Fn = torch.einsum('pij,pj->pi', F, n0)
    C = torch.einsum('pki,pkj->pij', F, F)
    Cn = torch.einsum('pij,pj->pi', C, n0)
    I4 = (Fn * Fn).sum(-1)
    lam_n = torch.sqrt(I4)
    opening = torch.clamp(lam_n - 1.0, min=0.0)
    P_I = gamma_I * (opening / lam_n)[..., None, None] * torch.einsum('pi,pj->pij', Fn, n0)
    FCn = torch.einsum('pij,pj->pi', F, Cn)
    P_II = gamma_II * (torch.einsum('pi,pj->pij', Fn, Cn) + torch.einsum('pi,pj->pij', FCn, n0)
                       - 2 * I4[..., None, None] * torch.einsum('pi,pj->pij', Fn, n0))
    return P_mat + gI(d)[..., None, None] * P_I + gII(d)[..., None, None] * P_II
