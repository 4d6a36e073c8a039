"""Option A directional damage on the GPU: the reference's variationally
consistent energy / stress pair (``anisompm_py/src/anisompm.py:129-187``).

    psi = energy(F, n0, d, mu, lam, gamma_I, gamma_II)           # (N,)
    P   = directional_pk1(F, n0, d, mu, lam, gamma_I, gamma_II)  # (N,3,3), dPsi/dF
    tau = kirchhoff(F, n0, d, mu, lam, gamma_I, gamma_II)        # P F^T

The same names, arguments and results as the reference's functions: F
(N,3,3), n0 (N,3), d (N,) CUDA tensors of one dtype (fp32 or fp64); mu, lam,
gamma_I, gamma_II scalars or (N,) tensors.  One kernel (``pmpm_directional_pk1``)
computes both per particle.  The step keeps the reference's Option B
(the projection split in ``k_const``), as the reference's solver does;
this pair is the differentiable alternative the reference validates with
its finite-difference gate (``1_fd_check.py``), which the tests repeat on
the device.
"""
from __future__ import annotations

import torch

from . import _lib


def gI(d):
    """Mode-I degradation (anisompm.py:129-130)."""
    return (1.0 - d) ** 2 + 1e-4


def gII(d):
    """Mode-II degradation (anisompm.py:133-134)."""
    return (1.0 - d) ** 2 + 1e-4


def _param(v, n, like):
    if isinstance(v, torch.Tensor) and v.dim() > 0:
        t = v.to(device=like.device, dtype=like.dtype).reshape(-1).contiguous()
        if t.numel() != n:
            raise ValueError(f"per-particle parameter has {t.numel()} values for {n} particles")
        return t, t.data_ptr(), 0.0
    return None, None, float(v)


def _run(F, n0, d, mu, lam, gamma_I, gamma_II, want_P, want_psi):
    if not (F.is_cuda and n0.is_cuda and d.is_cuda):
        raise ValueError("the Option A kernels run on the GPU only")
    if F.dim() != 3 or F.shape[1:] != (3, 3) or n0.shape != (F.shape[0], 3) or d.shape != (F.shape[0],):
        raise ValueError("expected F (N,3,3), n0 (N,3), d (N,)")
    if not (F.dtype == n0.dtype == d.dtype):
        raise TypeError("F, n0 and d must share one dtype")
    n = F.shape[0]
    F, n0, d = F.contiguous(), n0.contiguous(), d.contiguous()
    keep = []
    args = []
    for v in (mu, lam, gamma_I, gamma_II):
        t, ptr, s = _param(v, n, F)
        keep.append(t)
        args += [ptr, s]
    P = torch.empty_like(F) if want_P else None
    psi = torch.empty_like(d) if want_psi else None
    ptr = lambda t: None if t is None else t.data_ptr()
    _lib.check(_lib.lib().pmpm_directional_pk1(_lib.dtype_code(F.dtype), n, F.data_ptr(), n0.data_ptr(), d.data_ptr(),
                                               *args, ptr(P), ptr(psi), _lib.stream_ptr(F.device)),
               "pmpm_directional_pk1")
    return P, psi


def energy(F, n0, d, mu, lam, gamma_I, gamma_II):
    """Option A energy density per particle (anisompm.py:137-159)."""
    return _run(F, n0, d, mu, lam, gamma_I, gamma_II, False, True)[1]


def directional_pk1(F, n0, d, mu, lam, gamma_I, gamma_II):
    """Option A first Piola-Kirchhoff stress dPsi/dF (anisompm.py:162-187)."""
    return _run(F, n0, d, mu, lam, gamma_I, gamma_II, True, False)[0]


def energy_and_pk1(F, n0, d, mu, lam, gamma_I, gamma_II):
    """Both from one launch: (psi (N,), P (N,3,3))."""
    P, psi = _run(F, n0, d, mu, lam, gamma_I, gamma_II, True, True)
    return psi, P


def kirchhoff(F, n0, d, mu, lam, gamma_I, gamma_II):
    """tau = P F^T, the Kirchhoff stress of the variational route
    (anisompm.py:166-168)."""
    P = directional_pk1(F, n0, d, mu, lam, gamma_I, gamma_II)
    return torch.matmul(P, F.transpose(1, 2).contiguous())
