"""Gaussian carry on the GPU: the splat motion that feeds the renderer.

The reference moves every Gaussian of a splat model with the affine motion
of the particle it is bound to, in ``deform_and_render``
(``anisompm_py/src/run_orange.py:116-128``), as five torch operations
(gathers, a ``bmm``, two matrix products).  ``carry_gaussians`` does the
same in one HIP kernel (``pmpm_gaussian_carry``).  The renderer itself
(``render_gaussians``) stays the reference's: call it on what this returns.

    g_pos, g_cov, color = carry_gaussians(sim, data, crack_tint=0.3)
    img = render_gaussians(cam, g_pos, g_cov, color, opacity, ...)

``data`` holds the binding as ``prepare_orange`` builds it
(``anisompm_py/src/orange_setup.py:146-170``): ``g_bind`` (G,) int64,
``g_off0`` (G,3), ``g_cov0`` (G,3,3), ``g_color`` (G,3), and the frame map
``grid_lim``, ``scale_to_sim``, ``center``.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib


def carry(x, F, d, bind, off0, cov0, color, grid_lim: float, scale_to_sim: float, center, crack_tint: float = 0.0,
          with_color: bool = True):
    """Splat positions (G,3), covariances (G,3,3) and colours (G,3) (or None
    without ``with_color``) from the particle state x (N,3), F (N,3,3), d (N,)."""
    dev, dt = x.device, x.dtype
    if dev.type != "cuda":
        raise ValueError("the Gaussian carry runs on the GPU only")
    G, N = int(bind.shape[0]), int(x.shape[0])
    code = _lib.dtype_code(dt)
    bind = bind.to(dev, torch.int64).contiguous()
    off0 = off0.to(dev, dt).contiguous()
    cov0 = cov0.to(dev, dt).contiguous()
    col = color.to(dev, dt).contiguous() if with_color else None
    x, F = x.contiguous(), F.contiguous()
    dd = d.contiguous() if (with_color and crack_tint > 0.0) else None
    pos = torch.empty(G, 3, device=dev, dtype=dt)
    cov = torch.empty(G, 3, 3, device=dev, dtype=dt)
    cout = torch.empty(G, 3, device=dev, dtype=dt) if with_color else None
    # (a tuple of Python floats keeps its double precision; a tensor its own)
    cen = (C.c_double * 3)(*[float(c) for c in torch.as_tensor(center, dtype=torch.float64).flatten().tolist()])
    ptr = lambda t: None if t is None else t.data_ptr()
    _lib.check(_lib.lib().pmpm_gaussian_carry(code, G, N, ptr(bind), ptr(x), ptr(F), ptr(dd), ptr(off0), ptr(cov0),
                                              ptr(col), float(grid_lim), float(scale_to_sim), cen, float(crack_tint),
                                              ptr(pos), ptr(cov), ptr(cout), _lib.stream_ptr(dev)),
               "pmpm_gaussian_carry")
    return pos, cov, cout


def carry_gaussians(sim, data: dict, crack_tint: float = 0.0):
    """``deform_and_render``'s splat motion (anisompm_py/src/run_orange.py:119-127)
    for an AnisoMPM ``sim`` and a ``prepare_orange``-style ``data`` dict."""
    return carry(sim.x, sim.F, sim.d, data["g_bind"], data["g_off0"], data["g_cov0"], data["g_color"],
                 float(data["grid_lim"]), float(data["scale_to_sim"]), data["center"], crack_tint)
