"""Oracle: a CPU restatement of the AnisoMPM substep, for checking only.

TEST INFRASTRUCTURE.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or the timed CPU baseline).  The product path (``peel_mpm_amd``) never
imports it and has no CPU fallback.

What it restates (all citations are ``anisompm_py/src/anisompm.py``):

* ``sym_eig``            -- cyclic 3x3 Jacobi, 5 sweeps, ascending sort  (:48-95)
* ``quad_weights``       -- quadratic B-spline weights, int64 base       (:544-557)
* ``cubic_weights``      -- cubic B-spline weights + 2nd derivatives     (:559-584)
* ``node_index``         -- clamp-to-edge linear node index              (:586-590)
* ``directional_tau``    -- projection split of the Kirchhoff stress     (:190-227)
* ``collider_resolve``   -- sticky / slip(+friction) projection          (:811-825)
* ``OracleMPM.substep``  -- the whole step, in the reference's op order  (:593-803)

The arithmetic is written with the same torch operations in the same order
as the reference so that, on the CPU and with the same torch build, results
are bit-identical to the reference's; ``tests/test_oracle.py`` checks exactly
that against the committed golden vectors, and additionally pins every
function against closed forms, LAPACK and brute force.  The program structure
(state dict, phase functions, the six-entry symmetric Jacobi) is this file's
own.
"""
from __future__ import annotations

import math

import torch

STATE_KEYS = ("x", "v", "F", "C", "d", "lap", "D_hist")
PARAM_KEYS = ("vol0", "mass", "mu", "lam", "sigma_c", "eta", "zeta", "l0", "residual",
              "A0", "allow", "interface_mask", "n_interface")


# ------------------------------------------------------------------ linear algebra
def sym_eig(A: torch.Tensor, sweeps: int = 5):
    """Batched cyclic Jacobi on symmetric 3x3 matrices (anisompm.py:48-95).

    The input is symmetrised and NaN/inf-sanitised first.  The matrix is
    tracked by its six distinct entries (the reference keeps it exactly
    symmetric, so this loses nothing); rotation pairs are visited in the order
    (0,1), (0,2), (1,2).  Returns eigenvalues ascending and eigenvectors as
    the matching columns of V."""
    S = torch.nan_to_num(0.5 * (A + A.transpose(1, 2)))
    n = S.shape[0]
    a = {(i, j): S[:, i, j].clone() for i in range(3) for j in range(i, 3)}

    def get(i, j):
        return a[(i, j)] if i <= j else a[(j, i)]

    def put(i, j, val):
        a[(i, j) if i <= j else (j, i)] = val

    V = torch.eye(3, dtype=A.dtype).expand(n, 3, 3).contiguous()
    cols = [V[:, :, k].clone() for k in range(3)]
    for _ in range(sweeps):
        for p, q in ((0, 1), (0, 2), (1, 2)):
            r = 3 - p - q
            app, aqq, apq = get(p, p), get(q, q), get(p, q)
            apr, aqr = get(p, r), get(q, r)
            zeta = (aqq - app) / (2.0 * apq)
            sgn = torch.where(zeta >= 0, torch.ones_like(zeta), -torch.ones_like(zeta))
            t = sgn / (zeta.abs() + torch.sqrt(1.0 + zeta * zeta))
            t = torch.where(apq.abs() > 1e-30, t, torch.zeros_like(t))
            c = 1.0 / torch.sqrt(1.0 + t * t)
            s = c * t
            put(p, p, c * c * app - 2.0 * s * c * apq + s * s * aqq)
            put(q, q, s * s * app + 2.0 * s * c * apq + c * c * aqq)
            put(p, q, torch.zeros_like(apq))
            put(p, r, c * apr - s * aqr)
            put(q, r, s * apr + c * aqr)
            vp, vq = cols[p], cols[q]
            cols[p] = c.unsqueeze(1) * vp - s.unsqueeze(1) * vq
            cols[q] = s.unsqueeze(1) * vp + c.unsqueeze(1) * vq
    evals = torch.stack([a[(0, 0)], a[(1, 1)], a[(2, 2)]], 1)
    V = torch.stack(cols, 2)
    evals, order = torch.sort(evals, dim=1)
    V = torch.gather(V, 2, order.unsqueeze(1).expand(-1, 3, -1))
    return evals, V


# ------------------------------------------------------------------ kernels
def quad_weights(x: torch.Tensor, inv_dx: float):
    """Quadratic B-spline weights (anisompm.py:544-557): base (N,3) int64,
    fx (N,3) in [0.5, 1.5), w (N, node 0..2, axis)."""
    Xp = x * inv_dx
    base = torch.floor(Xp - 0.5).to(torch.long)
    fx = Xp - base.to(x.dtype)
    w = torch.stack([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2], dim=1)
    return base, fx, w


def cubic_weights(x: torch.Tensor, inv_dx: float):
    """Cubic B-spline weights and 2nd derivatives in grid units
    (anisompm.py:559-584): base = floor(Xp)-1, f = frac(Xp)."""
    Xp = x * inv_dx
    base = torch.floor(Xp).to(torch.long) - 1
    f = Xp - (base + 1).to(x.dtype)
    one = 1.0 - f
    w = torch.stack([(1.0 / 6.0) * one ** 3,
                     2.0 / 3.0 - f ** 2 + 0.5 * f ** 3,
                     2.0 / 3.0 - one ** 2 + 0.5 * one ** 3,
                     (1.0 / 6.0) * f ** 3], dim=1)
    w2 = torch.stack([one, 3.0 * f - 2.0, 1.0 - 3.0 * f, f], dim=1)
    return base, f, w, w2


def node_index(nodes: torch.Tensor, ng: int):
    """Clamp each integer node coordinate to [0, ng-1], then (i*ng+j)*ng+k
    (anisompm.py:586-590)."""
    c = nodes.clamp(0, ng - 1)
    return (c[:, 0] * ng + c[:, 1]) * ng + c[:, 2]


def stencil_offsets(width: int):
    """(width^3, 3) offsets in meshgrid 'ij' order: k = ox*w*w + oy*w + oz
    (anisompm.py:329-343)."""
    r = torch.arange(width)
    gx, gy, gz = torch.meshgrid(r, r, r, indexing="ij")
    return torch.stack([gx.reshape(-1), gy.reshape(-1), gz.reshape(-1)], 1)


# ------------------------------------------------------------------ stress pieces
def directional_tau(tau_c, F, d, n0, res, tension_gate=True):
    """Projection split of the corotated Kirchhoff stress by the deformed
    interface normal m = F n0/|F n0| (anisompm.py:190-227):
    tau = Q tau Q + gI(d) P tau P + gII(d) (P tau Q + Q tau P), with
    g(d) = (1-d)^2 (1-res) + res for both modes and the normal degradation
    dropped where sigma_nn <= 0 (tension gate)."""
    m = torch.einsum('pij,pj->pi', F, n0)
    m = m / (m.norm(dim=1, keepdim=True) + 1e-9)
    P = torch.einsum('pi,pj->pij', m, m)
    I3 = torch.eye(3, dtype=tau_c.dtype).expand_as(tau_c)
    Q = I3 - P
    t_perp = P @ tau_c @ P
    t_shear = P @ tau_c @ Q + Q @ tau_c @ P
    t_par = Q @ tau_c @ Q
    g = ((1.0 - d) ** 2 * (1.0 - res) + res).view(-1, 1, 1)
    gI = g
    if tension_gate:
        snn = torch.einsum('pi,pij,pj->p', m, tau_c, m).view(-1, 1, 1)
        gI = torch.where(snn > 0.0, gI, torch.ones_like(gI))
    return t_par + gI * t_perp + g * t_shear


def collider_resolve(v, vcol, n, mode, friction):
    """Grid-velocity projection against a collider with outward normal n
    (anisompm.py:811-825)."""
    if mode == "sticky":
        return vcol.expand_as(v)
    vrel = v - vcol
    vn = (vrel * n).sum(1, keepdim=True)
    inward = vn.clamp(max=0.0)
    vtan = vrel - vn * n
    if friction > 0.0:
        vtmag = vtan.norm(dim=1, keepdim=True)
        vtan = vtan * (1.0 + friction * inward / (vtmag + 1e-9)).clamp(min=0.0)
    return vcol + vtan + (vn - inward) * n


class SphereCollider:
    """Sphere obstacle; node inside when |pos-c| < r, outward normal radial
    (anisompm.py:828-844).  ``center_fn(time)`` gives the centre."""

    def __init__(self, center_fn, radius, mode="sticky", friction=0.0):
        self.center_fn, self.radius, self.mode, self.friction = center_fn, radius, mode, friction

    def __call__(self, pos, v, time):
        c = torch.as_tensor(self.center_fn(time), dtype=pos.dtype)
        rel = pos - c
        dist = rel.norm(dim=1)
        inside = dist < self.radius
        if not inside.any():
            return v
        vcol = (torch.as_tensor(self.center_fn(time + 1e-4), dtype=pos.dtype) - c) / 1e-4
        n = rel[inside] / (dist[inside].unsqueeze(1) + 1e-9)
        out = v.clone()
        out[inside] = collider_resolve(v[inside], vcol, n, self.mode, self.friction)
        return out


class HalfspaceCollider:
    """Half-space obstacle on the side opposite the outward normal; node
    inside when (pos-o).n < 0 (anisompm.py:847-867).  The normal is
    normalised in float64 first, as the reference does."""

    def __init__(self, origin_fn, normal, mode="slip", friction=0.0):
        nrm = torch.tensor(normal, dtype=torch.float64)
        self.normal = (nrm / nrm.norm()).tolist()
        self.origin_fn, self.mode, self.friction = origin_fn, mode, friction

    def __call__(self, pos, v, time):
        n = torch.as_tensor(self.normal, dtype=pos.dtype)
        o = torch.as_tensor(self.origin_fn(time), dtype=pos.dtype)
        inside = ((pos - o) * n).sum(1) < 0.0
        if not inside.any():
            return v
        vcol = (torch.as_tensor(self.origin_fn(time + 1e-4), dtype=pos.dtype) - o) / 1e-4
        out = v.clone()
        out[inside] = collider_resolve(v[inside], vcol, n.unsqueeze(0), self.mode, self.friction)
        return out


# ------------------------------------------------------------------ the step
class OracleMPM:
    """State holder plus ``substep``; attribute names follow the reference
    class (anisompm.py:303-463) so fixtures and scenes load directly."""

    def __init__(self, n_grid=96, grid_lim=1.0, dt=1e-4, gravity=(0.0, -9.8, 0.0),
                 grid_damp=1.0, f_clamp=None, damage_every=1, dtype=torch.float32, damage_solver="explicit"):
        self.dtype = dtype
        # "explicit": the reference's update; "implicit": the PCG mode's
        # backward-Euler grid solve (oracle/pcg_damage.py), solved directly
        self.damage_solver = damage_solver
        self.n_grid = int(n_grid)
        self.grid_lim = float(grid_lim)
        self.dx = grid_lim / n_grid
        self.inv_dx = 1.0 / self.dx
        self.dt = float(dt)
        self.gravity = torch.tensor(gravity, dtype=dtype)
        self.grid_damp = float(grid_damp)
        self.f_clamp = f_clamp
        self.damage_every = max(1, int(damage_every))
        self._step = 0
        self.off27 = stencil_offsets(3)
        self.off64 = stencil_offsets(4)
        self.colliders = []
        self.particle_bc = []
        self.n = 0

    def load(self, tensors: dict):
        """Take particle state and parameters from a dict of CPU tensors."""
        for k in STATE_KEYS + PARAM_KEYS:
            t = torch.as_tensor(tensors[k])
            if t.dtype not in (torch.bool, torch.int64, torch.int32):
                t = t.to(self.dtype)
            setattr(self, k, t.clone())
        self.allow = self.allow.bool()
        self.interface_mask = self.interface_mask.bool()
        self.n = self.x.shape[0]
        return self

    # -- phase 1: stress, damage and degradation (anisompm.py:614-706)
    def _constitutive(self):
        F = self.F
        lamC, V = sym_eig(F.transpose(1, 2) @ F)
        S = torch.sqrt(torch.clamp(lamC, min=1e-12))
        Vt = V.transpose(1, 2)
        W = F @ V
        R = (W / (W.norm(dim=1, keepdim=True) + 1e-12)) @ Vt
        if self.f_clamp is not None:
            Sc = S.clamp(self.f_clamp[0], self.f_clamp[1])
            F = F @ (V @ torch.diag_embed(Sc / S.clamp(min=1e-12)) @ Vt)
            self.F = F
            S = Sc
        J = (S[:, 0] * S[:, 1] * S[:, 2]).clamp(min=1e-6)
        I3 = torch.eye(3, dtype=self.dtype).expand_as(F)
        tau_dev = 2.0 * self.mu.view(-1, 1, 1) * (F - R) @ F.transpose(1, 2)
        tau_vol = (self.lam * (J - 1.0) * J).view(-1, 1, 1) * I3

        if bool(self.allow.any()) and self._step % self.damage_every == 0:
            self._damage(tau_dev + tau_vol, J, R)

        g = (1.0 - self.d) ** 2 * (1.0 - self.residual) + self.residual
        gv = g.view(-1, 1, 1)
        tau = gv * tau_dev + torch.where((J >= 1.0).view(-1, 1, 1), gv, torch.ones_like(gv)) * tau_vol
        if bool(self.interface_mask.any()):
            tau_dir = directional_tau(tau_dev + tau_vol, F, self.d, self.n_interface,
                                      self.residual.view(-1))
            tau = torch.where(self.interface_mask.view(-1, 1, 1), tau_dir, tau)
        self.tau_last = tau
        self.J_last = J
        return tau, I3

    def _damage(self, tau_full, J, R):
        """Structural-tensor drive, history max and explicit phase-field
        update with the one-step-lagged Laplacian (anisompm.py:646-685)."""
        sigma = tau_full / J.view(-1, 1, 1)
        sigma = 0.5 * (sigma + sigma.transpose(1, 2))
        ev, Q = sym_eig(sigma)
        sig_p = Q @ torch.diag_embed(ev.clamp(min=0.0)) @ Q.transpose(1, 2)
        A = R @ self.A0 @ R.transpose(1, 2)
        M = A @ sig_p
        phi = (M @ M).diagonal(dim1=1, dim2=2).sum(1) / (self.sigma_c ** 2)
        self.phi = phi
        Sn = torch.einsum('pij,pj->pi', sig_p, self.n_interface)
        tn = (Sn * self.n_interface).sum(1, keepdim=True)
        self.phi_shear = ((Sn - tn * self.n_interface).norm(dim=1) / self.sigma_c) ** 2
        self.D_hist = torch.maximum(self.D_hist, self.zeta * torch.relu(phi - 1.0))
        if self.damage_solver == "implicit":
            from . import pcg_damage
            dn, _, _ = pcg_damage.implicit_update(self.x.double().numpy(), self.d.double().numpy(),
                                                  self.D_hist.double().numpy(), self.eta.double().numpy(),
                                                  self.l0.double().numpy(), self.allow.numpy(),
                                                  self.dt * self.damage_every, self.n_grid, self.inv_dx)
            self.d = torch.from_numpy(dn).to(self.dtype)
            return
        ddot = torch.relu((1.0 - self.d) * self.D_hist - (self.d - (self.l0 ** 2) * self.lap))
        dnew = torch.clamp(self.d + (self.dt * self.damage_every / self.eta) * ddot, min=0.0, max=1.0)
        self.d = torch.where(self.allow, torch.maximum(dnew, self.d), self.d)

    # -- phase 2: quadratic P2G (anisompm.py:707-724)
    def _p2g(self, tau):
        ng3 = self.n_grid ** 3
        base, fx, w = quad_weights(self.x, self.inv_dx)
        m_grid = torch.zeros(ng3, dtype=self.dtype)
        mv_grid = torch.zeros(ng3, 3, dtype=self.dtype)
        coeff = (-self.dt * self.vol0 * 4.0 * self.inv_dx * self.inv_dx).view(-1, 1, 1)
        affine = coeff * tau + self.mass.view(-1, 1, 1) * self.C
        mv = self.mass.view(-1, 1) * self.v
        for k in range(27):
            o = self.off27[k]
            idx = node_index(base + o, self.n_grid)
            wk = w[:, o[0], 0] * w[:, o[1], 1] * w[:, o[2], 2]
            dpos = (o.to(self.dtype) - fx) * self.dx
            mv_grid.index_add_(0, idx, wk.unsqueeze(1) * (mv + (affine @ dpos.unsqueeze(-1)).squeeze(-1)))
            m_grid.index_add_(0, idx, wk * self.mass)
        return base, fx, w, m_grid, mv_grid

    # -- phase 3: grid velocity and colliders (anisompm.py:726-746)
    def _grid(self, m_grid, mv_grid, time):
        ng = self.n_grid
        active = m_grid > 1e-12
        inv_m = torch.zeros_like(m_grid)
        inv_m[active] = 1.0 / m_grid[active]
        gv = mv_grid * inv_m.unsqueeze(1) + self.dt * self.gravity.unsqueeze(0)
        if self.grid_damp != 1.0:
            gv = gv * self.grid_damp
        gv[~active] = 0.0
        if self.colliders:
            lin = torch.nonzero(active, as_tuple=False).squeeze(1)
            pos = torch.stack([lin // (ng * ng), (lin // ng) % ng, lin % ng], 1).to(self.dtype) * self.dx
            vs = gv[lin]
            for col in self.colliders:
                vs = col(pos, vs, time)
            gv[lin] = vs
        return gv

    # -- phase 4: quadratic G2P (anisompm.py:748-760)
    def _g2p(self, base, fx, w, gv):
        v_new = torch.zeros_like(self.v)
        C_new = torch.zeros_like(self.C)
        for k in range(27):
            o = self.off27[k]
            g = gv[node_index(base + o, self.n_grid)]
            wk = w[:, o[0], 0] * w[:, o[1], 1] * w[:, o[2], 2]
            dpos = (o.to(self.dtype) - fx) * self.dx
            v_new = v_new + wk.unsqueeze(1) * g
            C_new = C_new + wk.unsqueeze(1).unsqueeze(2) * (g.unsqueeze(2) * dpos.unsqueeze(1))
        return v_new, C_new * (4.0 * self.inv_dx * self.inv_dx)

    # -- phase 5: cubic damage scatter, normalise, Laplacian gather (anisompm.py:762-791)
    def _laplacian(self):
        ng3 = self.n_grid ** 3
        bc, _, wc, wc2 = cubic_weights(self.x, self.inv_dx)
        dsum = torch.zeros(ng3, dtype=self.dtype)
        wsum = torch.zeros(ng3, dtype=self.dtype)
        idx = [node_index(bc + self.off64[k], self.n_grid) for k in range(64)]
        for k in range(64):
            o = self.off64[k]
            wk = wc[:, o[0], 0] * wc[:, o[1], 1] * wc[:, o[2], 2]
            dsum.index_add_(0, idx[k], wk * self.d)
            wsum.index_add_(0, idx[k], wk)
        gd = torch.zeros_like(dsum)
        nz = wsum > 1e-12
        gd[nz] = dsum[nz] / wsum[nz]
        lap = torch.zeros(self.n, dtype=self.dtype)
        for k in range(64):
            o = self.off64[k]
            lw = self.inv_dx * self.inv_dx * (wc2[:, o[0], 0] * wc[:, o[1], 1] * wc[:, o[2], 2]
                                              + wc[:, o[0], 0] * wc2[:, o[1], 1] * wc[:, o[2], 2]
                                              + wc[:, o[0], 0] * wc[:, o[1], 1] * wc2[:, o[2], 2])
            lap = lap + lw * gd[idx[k]]
        return lap

    def substep(self, time: float = 0.0):
        for bc in self.particle_bc:
            bc(self, time)
        tau, I3 = self._constitutive()
        base, fx, w, m_grid, mv_grid = self._p2g(tau)
        gv = self._grid(m_grid, mv_grid, time)
        v_new, C_new = self._g2p(base, fx, w, gv)
        lap = self._laplacian()
        self.v, self.C, self.lap = v_new, C_new, lap
        self.x = self.x + self.dt * self.v
        self.F = (I3 + self.dt * self.C) @ self.F
        pad = 2.0 * self.dx
        self.x = self.x.clamp(pad, self.grid_lim - pad)
        self._step += 1

    def run_frame(self, n_sub: int, time0: float = 0.0):
        for s in range(n_sub):
            self.substep(time0 + s * self.dt)
