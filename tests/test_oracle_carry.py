"""CPU pins of the Gaussian-carry oracle (anisompm_py/src/run_orange.py:116-128)
and of the C entry point's argument checks (no GPU work).

The oracle restates the reference's torch operations; these checks hold it
to what the mathematics fixes, so that a transposed F, F^T cov F, a wrong
frame map or a misplaced tint fails one of them:
* F = I leaves offsets and covariances unchanged;
* a rotation keeps |F off0| and the covariance's eigenvalues;
* a diagonal stretch scales offset i by s_i and covariance ij by s_i s_j;
* the offset enters linearly;
* an independent float64 numpy.einsum evaluation of F off0 and F cov F^T;
* sim_to_world inverts the world -> sim placement
  (anisompm_py/src/orange_setup.py:126-135);
* the crack tint's closed form.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from oracle import gaussian_carry as O

GL, SC, CEN = 1.0, 0.37, (0.1, -0.2, 0.3)


def scene(G=500, N=80, seed=0, dtype=torch.float64):
    g = torch.Generator().manual_seed(seed)
    x = torch.rand(N, 3, generator=g, dtype=dtype)
    F = torch.eye(3, dtype=dtype) + 0.2 * torch.randn(N, 3, 3, generator=g, dtype=dtype)
    d = torch.rand(N, generator=g, dtype=dtype)
    bind = torch.randint(0, N, (G,), generator=g)
    off0 = 0.01 * torch.randn(G, 3, generator=g, dtype=dtype)
    A = torch.randn(G, 3, 3, generator=g, dtype=dtype)
    cov0 = 1e-4 * (A @ A.transpose(1, 2) + 0.1 * torch.eye(3, dtype=dtype))
    color = torch.rand(G, 3, generator=g, dtype=dtype)
    return x, F, d, bind, off0, cov0, color


def world(x):
    return O.sim_to_world(x, GL, SC, CEN)


def test_sim_to_world_inverts_the_placement():
    w = torch.randn(100, 3, dtype=torch.float64)
    sim = ((w - torch.tensor(CEN, dtype=torch.float64)) * SC + 0.5) * GL      # orange_setup.py:126-127
    assert torch.allclose(O.sim_to_world(sim, GL, SC, CEN), w, atol=1e-12)


def test_identity_gradient_leaves_splats_unchanged():
    x, F, d, bind, off0, cov0, color = scene()
    F = torch.eye(3, dtype=F.dtype).expand_as(F).clone()
    pos, cov, col = O.carry(x, F, d, bind, off0, cov0, color, GL, SC, CEN)
    assert torch.allclose(pos, world(x)[bind] + off0, atol=1e-15)
    assert torch.equal(cov, cov0)
    assert torch.equal(col, color)


def test_rotation_keeps_offset_length_and_covariance_spectrum():
    x, F, d, bind, off0, cov0, color = scene()
    Q, _ = torch.linalg.qr(torch.randn(F.shape[0], 3, 3, dtype=F.dtype))
    pos, cov, _ = O.carry(x, Q, d, bind, off0, cov0, color, GL, SC, CEN)
    rel = pos - world(x)[bind]
    assert torch.allclose(rel.norm(dim=1), off0.norm(dim=1), rtol=1e-12)
    assert torch.allclose(torch.linalg.eigvalsh(cov), torch.linalg.eigvalsh(cov0), rtol=1e-10, atol=1e-18)


def test_diagonal_stretch_scales_each_axis():
    x, F, d, bind, off0, cov0, color = scene()
    s = 0.5 + torch.rand(F.shape[0], 3, dtype=F.dtype)
    pos, cov, _ = O.carry(x, torch.diag_embed(s), d, bind, off0, cov0, color, GL, SC, CEN)
    sg = s[bind]
    assert torch.allclose(pos - world(x)[bind], sg * off0, rtol=1e-12, atol=1e-14)
    assert torch.allclose(cov, sg[:, :, None] * sg[:, None, :] * cov0, rtol=1e-12, atol=1e-20)


def test_offset_enters_linearly():
    x, F, d, bind, off0, cov0, color = scene()
    a, b = off0, 0.01 * torch.randn_like(off0)
    pa = O.carry(x, F, d, bind, a, cov0, color, GL, SC, CEN)[0] - world(x)[bind]
    pb = O.carry(x, F, d, bind, b, cov0, color, GL, SC, CEN)[0] - world(x)[bind]
    pab = O.carry(x, F, d, bind, a + b, cov0, color, GL, SC, CEN)[0] - world(x)[bind]
    assert torch.allclose(pab, pa + pb, atol=1e-15)


def test_matches_independent_einsum_evaluation():
    x, F, d, bind, off0, cov0, color = scene(seed=3)
    pos, cov, _ = O.carry(x, F, d, bind, off0, cov0, color, GL, SC, CEN)
    Fn, bn = F.numpy(), bind.numpy()
    Fg = Fn[bn]
    xw = (x.numpy() / GL - 0.5) / SC + np.array(CEN)
    ref_pos = xw[bn] + np.einsum("gij,gj->gi", Fg, off0.numpy())
    ref_cov = np.einsum("gij,gjk,glk->gil", Fg, cov0.numpy(), Fg)
    assert np.allclose(pos.numpy(), ref_pos, rtol=1e-13, atol=1e-16)
    assert np.allclose(cov.numpy(), ref_cov, rtol=1e-12, atol=1e-20)


def test_crack_tint_darkens_by_damage():
    x, F, d, bind, off0, cov0, color = scene()
    _, _, col = O.carry(x, F, d, bind, off0, cov0, color, GL, SC, CEN, crack_tint=0.4)
    assert torch.allclose(col, color * (1.0 - 0.4 * d[bind])[:, None], rtol=1e-15)
    d1 = torch.ones_like(d)
    _, _, col1 = O.carry(x, F, d1, bind, off0, cov0, color, GL, SC, CEN, crack_tint=1.0)
    assert torch.equal(col1, torch.zeros_like(color))


def test_entry_point_rejects_bad_arguments_before_any_gpu_work():
    from peel_mpm_amd import _lib
    L = _lib.lib()
    cen = (C.c_double * 3)(*CEN)
    dummy = C.c_void_p(16)              # never dereferenced: every call below fails validation or is empty
    # nothing to do
    assert L.pmpm_gaussian_carry(_lib.PMPM_F32, 0, 10, None, None, None, None, None, None, None, GL, SC, cen, 0.0,
                                 None, None, None, None) == 0
    # null binding / outputs
    assert L.pmpm_gaussian_carry(_lib.PMPM_F32, 5, 10, None, dummy, dummy, None, dummy, dummy, None, GL, SC, cen, 0.0,
                                 dummy, dummy, None, None) != 0
    assert b"null" in L.pmpm_last_error()
    # splats bound to no particles
    assert L.pmpm_gaussian_carry(_lib.PMPM_F32, 5, 0, dummy, dummy, dummy, None, dummy, dummy, None, GL, SC, cen, 0.0,
                                 dummy, dummy, None, None) != 0
    # a tint needs d; a colour output needs colours
    assert L.pmpm_gaussian_carry(_lib.PMPM_F32, 5, 10, dummy, dummy, dummy, None, dummy, dummy, dummy, GL, SC, cen, 0.5,
                                 dummy, dummy, dummy, None) != 0
    assert L.pmpm_gaussian_carry(_lib.PMPM_F32, 5, 10, dummy, dummy, dummy, None, dummy, dummy, None, GL, SC, cen, 0.0,
                                 dummy, dummy, dummy, None) != 0
    # frame map, dtype
    assert L.pmpm_gaussian_carry(_lib.PMPM_F32, 5, 10, dummy, dummy, dummy, None, dummy, dummy, None, 0.0, SC, cen, 0.0,
                                 dummy, dummy, None, None) != 0
    assert L.pmpm_gaussian_carry(7, 5, 10, dummy, dummy, dummy, None, dummy, dummy, None, GL, SC, cen, 0.0,
                                 dummy, dummy, None, None) != 0
