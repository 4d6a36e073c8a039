"""Oracle: a CPU restatement of the Gaussian carry, for checking only.

TEST INFRASTRUCTURE.  Only ``tests/`` and ``bench.py``'s CPU baseline may
import this module, and only as the checker; ``peel_mpm_amd`` never imports
it.

It restates the splat motion of ``deform_and_render``
(``anisompm_py/src/run_orange.py:116-128``) with the reference's torch
operations in the reference's order:

* ``xp_world = sim_to_world(x)[idx]``, where ``sim_to_world(y) =
  (y / grid_lim - 0.5) / scale_to_sim + center``
  (``anisompm_py/src/orange_setup.py:134-135``);
* ``g_pos = xp_world + bmm(F[idx], off0[..., None])[..., 0]``;
* ``g_cov = F[idx] @ cov0 @ F[idx]^T``;
* with ``crack_tint > 0``: ``color * (1 - crack_tint * d[idx])``.

Pinned by ``tests/test_oracle_carry.py`` against closed forms (identity and
rotation deformation gradients, diagonal stretch, linearity in the offset,
the tint) and an independent ``numpy.einsum`` evaluation.
"""
from __future__ import annotations

import torch


def sim_to_world(y: torch.Tensor, grid_lim: float, scale_to_sim: float, center) -> torch.Tensor:
    """anisompm_py/src/orange_setup.py:134-135."""
    return (y / grid_lim - 0.5) / scale_to_sim + torch.as_tensor(center, device=y.device, dtype=y.dtype)


def carry(x, F, d, bind, off0, cov0, color, grid_lim, scale_to_sim, center, crack_tint=0.0):
    """(g_pos, g_cov, color) of anisompm_py/src/run_orange.py:119-127."""
    idx = bind
    Fg = F[idx]                                                        # :120
    xp_world = sim_to_world(x, grid_lim, scale_to_sim, center)[idx]    # :121
    g_pos = xp_world + torch.bmm(Fg, off0.unsqueeze(-1)).squeeze(-1)  # :122
    g_cov = Fg @ cov0 @ Fg.transpose(1, 2)                            # :123
    if crack_tint > 0.0:                                               # :125-127
        dmg = d[idx].unsqueeze(1)
        color = color * (1.0 - crack_tint * dmg)
    return g_pos, g_cov, color
