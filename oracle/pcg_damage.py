"""Oracle of the implicit (PCG) damage mode -- an extension with no
counterpart in the reference.

TEST INFRASTRUCTURE, like ``cpu_step``: only ``tests/``, ``smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker.

The reference advances the phase field explicitly
(anisompm.py:679-685):  d <- d + dt/eta * relu((1-d) D - (d - l0^2 lap d)),
with lap d from the previous step's cubic pass (:762-791).  The implicit
mode takes the same equation,  eta d' = (1-d) D - d + l0^2 lap d,  with the
drive D = D_hist after this step's history max (:679), one backward-Euler
step on the cubic B-spline grid of the reference's Laplacian pass, with
particles as quadrature points and the nodal mass lumped:

    A d_n = b,   A = diag(m) + K,
    m_i = sum_p N_ip (1 + c_p (1 + D_p)),           c_p = dt_damage / eta_p
    K_ij = sum_p c_p l0_p^2 grad N_ip . grad N_jp,
    b_i  = sum_p N_ip (d_p + c_p D_p),

over the cubic nodes with sum_p N_ip > 1e-12 (the reference's active-node
test for the cubic grid, :778), node indices clamped to the grid as in
:586-590.  The new particle damage is the interpolant of d_n, clamped to
[0, 1], made irreversible and gated by ``allow`` exactly as the explicit
update's result is (:684-685).

A is symmetric positive definite: diag(m) > 0 on active nodes and K is a
Gram matrix.  K 1 = 0 (the gradients of a partition of unity sum to zero),
so a uniform d with a uniform drive D solves to the closed form
(d + c D) / (1 + c (1 + D)) at every node.

Everything here is assembled explicitly (scipy.sparse) in float64 and
solved directly; the GPU solves the same system matrix-free by Jacobi-
preconditioned CG to a relative residual tolerance.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def cubic_values_and_slopes(x: np.ndarray, inv_dx: float):
    """Cubic B-spline weights N (P, 4, 3) and their derivatives dN/dX
    (P, 4, 3, in grid units) at the nodes base+0..3 of each axis, with
    base = floor(X) - 1 (the reference's cubic stencil, anisompm.py:559-584)."""
    X = x * inv_dx
    fl = np.floor(X)
    f = X - fl
    o = 1.0 - f
    N = np.stack([o ** 3 / 6.0, 2.0 / 3.0 - f ** 2 + 0.5 * f ** 3, 2.0 / 3.0 - o ** 2 + 0.5 * o ** 3, f ** 3 / 6.0], 1)
    dN = np.stack([-0.5 * o ** 2, -2.0 * f + 1.5 * f ** 2, 2.0 * o - 1.5 * o ** 2, 0.5 * f ** 2], 1)
    return fl.astype(np.int64) - 1, N, dN


def assemble(x, d, D, eta, l0, dt_damage, n_grid, inv_dx):
    """The system of the implicit step.  Returns (A, b, m, active node ids,
    Nmat): A and b restricted to the active nodes, Nmat the (P, active)
    interpolation matrix."""
    x = np.asarray(x, np.float64)
    P = x.shape[0]
    base, N, dN = cubic_values_and_slopes(x, inv_dx)
    c = dt_damage / np.asarray(eta, np.float64)
    kap = c * np.asarray(l0, np.float64) ** 2
    rows, cols, wv, gx, gy, gz = [], [], [], [], [], []
    ng = n_grid
    for a in range(4):
        for b in range(4):
            for e in range(4):
                node = base + np.array([a, b, e])
                nc = np.clip(node, 0, ng - 1)
                lin = (nc[:, 0] * ng + nc[:, 1]) * ng + nc[:, 2]
                rows.append(np.arange(P))
                cols.append(lin)
                wv.append(N[:, a, 0] * N[:, b, 1] * N[:, e, 2])
                gx.append(inv_dx * dN[:, a, 0] * N[:, b, 1] * N[:, e, 2])
                gy.append(inv_dx * N[:, a, 0] * dN[:, b, 1] * N[:, e, 2])
                gz.append(inv_dx * N[:, a, 0] * N[:, b, 1] * dN[:, e, 2])
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    G3 = ng ** 3
    Nfull = sp.csr_matrix((np.concatenate(wv), (rows, cols)), shape=(P, G3))   # duplicates (clamped) sum
    wsum = np.asarray(Nfull.sum(0)).ravel()
    act = np.nonzero(wsum > 1e-12)[0]
    Nmat = Nfull[:, act]
    Gm = sp.vstack([sp.csr_matrix((np.concatenate(g), (rows, cols)), shape=(P, G3))[:, act] for g in (gx, gy, gz)])
    kap3 = np.concatenate([kap, kap, kap])
    K = (Gm.T @ sp.diags(kap3) @ Gm).tocsr()
    D = np.asarray(D, np.float64)
    m = Nmat.T @ (1.0 + c * (1.0 + D))
    b = Nmat.T @ (np.asarray(d, np.float64) + c * D)
    A = (sp.diags(m) + K).tocsr()
    return A, b, m, act, Nmat


def solve(A, b):
    """Direct sparse solve (SuperLU)."""
    if A.shape[0] == 0:
        return np.zeros(0)
    return spla.spsolve(A.tocsc(), b)


def implicit_update(x, d, D, eta, l0, allow, dt_damage, n_grid, inv_dx):
    """New particle damage of the implicit step: interpolate the nodal
    solution, clamp to [0, 1], keep the larger of old and new where
    ``allow`` (anisompm.py:684-685 applied to the implicit value).  Returns
    (d_new, nodal solution, active node ids)."""
    A, b, m, act, Nmat = assemble(x, d, D, eta, l0, dt_damage, n_grid, inv_dx)
    dn = solve(A, b)
    di = np.clip(Nmat @ dn, 0.0, 1.0)
    d = np.asarray(d, np.float64)
    allow = np.asarray(allow, bool)
    return np.where(allow, np.maximum(di, d), d), dn, act
