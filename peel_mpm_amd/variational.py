"""Directional damage on the GPU: the reference's Option A, a variationally
consistent energy / stress pair (``anisompm_py/src/anisompm.py:129-187``),
and its Option B projection split (``:190-227``) as a standalone call.

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
the device.  ``directional_kirchhoff`` runs the step's own Option B code on
given stresses, for the reference's reconstruction and release gates.
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


def directional_kirchhoff(tau_c, F, d, n0, residual=1e-3, tension_gate=True):
    """Option B, the projection split the step applies at interface
    particles (anisompm.py:190-227 as called at :693-703): tau_c split by
    the deformed normal, normal and shear blocks degraded by
    g(d) = (1 - d)^2 (1 - k_r) + k_r, the normal one only in tension when
    ``tension_gate``.  ``residual`` k_r: a scalar or (N,) tensor.  Runs the
    step's own device code (``pmpm_directional_kirchhoff``)."""
    if not (tau_c.is_cuda and F.is_cuda and d.is_cuda and n0.is_cuda):
        raise ValueError("the directional split runs on the GPU only")
    n = F.shape[0]
    if tau_c.shape != (n, 3, 3) or F.shape != (n, 3, 3) or d.shape != (n,) or n0.shape != (n, 3):
        raise ValueError("expected tau_c, F (N,3,3), d (N,), n0 (N,3)")
    if not (tau_c.dtype == F.dtype == d.dtype == n0.dtype):
        raise TypeError("tau_c, F, d and n0 must share one dtype")
    tau_c, F, d, n0 = tau_c.contiguous(), F.contiguous(), d.contiguous(), n0.contiguous()
    r, rptr, rs = _param(residual, n, F)
    out = torch.empty_like(tau_c)
    _lib.check(_lib.lib().pmpm_directional_kirchhoff(_lib.dtype_code(F.dtype), n, tau_c.data_ptr(), F.data_ptr(),
                                                     d.data_ptr(), n0.data_ptr(), rptr, rs, int(bool(tension_gate)),
                                                     out.data_ptr(), _lib.stream_ptr(F.device)),
               "pmpm_directional_kirchhoff")
    return out


def kirchhoff(F, n0, d, mu, lam, gamma_I, gamma_II):
    """tau = P F^T, the Kirchhoff stress of the variational route
    (anisompm.py:166-168)."""
    P = directional_pk1(F, n0, d, mu, lam, gamma_I, gamma_II)
    return torch.matmul(P, F.transpose(1, 2).contiguous())
