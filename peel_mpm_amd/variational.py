"""Directional damage on the GPU: the reference's Option A, a variationally
consistent energy / stress pair (``anisompm_py/src/anisompm.py:129-187``),
and its Option B projection split (``:190-227``) as a standalone call.

    psi = energy(F, n0, d, mu, lam, gamma_I, gamma_II)           # (N,)
    P   = directional_pk1(F, n0, d, mu, lam, gamma_I, gamma_II)  # (N,3,3), dPsi/dF
    tau = kirchhoff(F, n0, d, mu, lam, gamma_I, gamma_II)        # P F^T

The same names, arguments and results as the reference's functions: F
(N,3,3), n0 (N,3), d (N,) tensors of one dtype (fp32 or fp64); mu, lam,
gamma_I, gamma_II scalars or (N,) tensors.  The kernels run on the GPU;
inputs on the CPU are moved there and the results returned to the CPU.  One kernel (``pmpm_directional_pk1``)
computes both per particle.  The step keeps the reference's Option B
(the projection split in ``k_const``), as the reference's solver does;
this pair is the differentiable alternative the reference validates with
its finite-difference gate (``1_fd_check.py``), which the tests repeat on
the device.  ``directional_kirchhoff`` runs the step's own Option B code on
given stresses, and ``reconstruction_check`` / ``release_check`` are the
reference's gates over it, so ``1_fd_check.py`` runs against this module
with its import line switched.
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


def _corotated(F, R, J, mu, lam):
    I3 = torch.eye(3, device=F.device, dtype=F.dtype).expand_as(F)
    mu_ = mu.view(-1, 1, 1) if torch.is_tensor(mu) and mu.dim() == 1 else mu
    lam_ = lam.view(-1) if torch.is_tensor(lam) and lam.dim() == 1 else lam
    return 2.0 * mu_ * (F - R) @ F.transpose(1, 2) + (lam_ * (J - 1.0) * J).view(-1, 1, 1) * I3


def _on_gpu(*ts):
    dev = torch.device("cuda", torch.cuda.current_device())
    return [t.to(dev) for t in ts]


def _param(v, n, like):
    if isinstance(v, torch.Tensor) and v.dim() > 0:
        t = v.to(device=like.device, dtype=like.dtype).reshape(-1).contiguous()
        if t.numel() != n:
            raise ValueError(f"per-particle parameter has {t.numel()} values for {n} particles")
        return t, t.data_ptr(), 0.0
    return None, None, float(v)


def _run(F, n0, d, mu, lam, gamma_I, gamma_II, want_P, want_psi):
    if F.dim() != 3 or F.shape[1:] != (3, 3) or n0.shape != (F.shape[0], 3) or d.shape != (F.shape[0],):
        raise ValueError("expected F (N,3,3), n0 (N,3), d (N,)")
    if not (F.dtype == n0.dtype == d.dtype):
        raise TypeError("F, n0 and d must share one dtype")
    home = F.device
    if not F.is_cuda:                 # computed on the GPU, returned where the inputs live
        F, n0, d = _on_gpu(F, n0, d)
    n = F.shape[0]
    F, n0, d = F.to(n0.device).contiguous(), n0.contiguous(), d.to(n0.device).contiguous()
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
    back = lambda t: None if t is None else t.to(home)
    return back(P), back(psi)


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


def reconstruction_check(F, R, J, n0, mu, lam, verbose=True):
    """Option B's d = 0 gate (anisompm.py:230-245) on the device split: the
    maximum |split(d=0) - corotated| of the plain corotated stress built
    from F, R, J.  Tensors may live on any device; the split runs on the
    GPU."""
    tau_c = _corotated(F, R, J, mu, lam)
    tc, Fd, nd = _on_gpu(tau_c, F, n0)
    d0 = torch.zeros(F.shape[0], device=Fd.device, dtype=F.dtype)
    split = directional_kirchhoff(tc, Fd, d0, nd, residual=0.0).to(F.device)
    err = (split - tau_c).abs().max().item()
    if verbose:
        print(f"reconstruction_check: max|split(d=0) - corotated| = {err:.2e}")
    return err


def release_check(F, R, J, n0, mu, lam, res=1e-3, tol=0.05, verbose=True):
    """Option B's d = 1 gate (anisompm.py:248-292) on the device split, tension
    gate off: the kept fractions of the normal traction, the in-plane stress
    and the interfacial shear.  Returns (ok, normal_kept, inplane_kept,
    shear_kept)."""
    tau_c = _corotated(F, R, J, mu, lam)
    tc, Fd, nd = _on_gpu(tau_c, F, n0)
    d1 = torch.ones(F.shape[0], device=Fd.device, dtype=F.dtype)
    tau1 = directional_kirchhoff(tc, Fd, d1, nd, residual=res, tension_gate=False).to(F.device)
    I3 = torch.eye(3, device=F.device, dtype=F.dtype).expand_as(F)
    m = torch.einsum('pij,pj->pi', F, n0)
    m = m / (m.norm(dim=1, keepdim=True) + 1e-9)
    Pm = torch.einsum('pi,pj->pij', m, m)
    Qm = I3 - Pm
    nrm = lambda t: t.reshape(F.shape[0], -1).norm(dim=1)
    tnc = torch.einsum('pi,pij,pj->p', m, tau_c, m).abs()
    tn1 = torch.einsum('pi,pij,pj->p', m, tau1, m).abs()
    normal_kept = (tn1 / (tnc + 1e-9)).max().item()
    inplane_kept = (nrm(Qm @ tau1 @ Qm) / (nrm(Qm @ tau_c @ Qm) + 1e-9)).min().item()
    shear_kept = (nrm(Pm @ tau1 @ Qm + Qm @ tau1 @ Pm) / (nrm(Pm @ tau_c @ Qm + Qm @ tau_c @ Pm) + 1e-9)).max().item()
    ok = (normal_kept < res + tol) and (shear_kept < res + tol) and (inplane_kept > 1.0 - tol)
    if verbose:
        print(f"release_check: d=1 normal_kept={normal_kept:.3f} shear_kept={shear_kept:.3f} "
              f"inplane_kept={inplane_kept:.3f} -> {'PASS' if ok else 'FAIL'}")
    return ok, normal_kept, inplane_kept, shear_kept


def kirchhoff(F, n0, d, mu, lam, gamma_I, gamma_II):
    """tau = P F^T, the Kirchhoff stress of the variational route
    (anisompm.py:166-168)."""
    P = directional_pk1(F, n0, d, mu, lam, gamma_I, gamma_II)
    return torch.matmul(P, F.transpose(1, 2).contiguous())
