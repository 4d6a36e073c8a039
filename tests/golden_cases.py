"""Load the committed golden step fixtures into solver objects.

Shared by the oracle tests (CPU) and the HIP parity tests (GPU): both build
their solver from the same fixture, with the same colliders and particle BC.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from tests.conftest import GOLDEN

STATE_KEYS = ("x", "v", "F", "C", "d", "lap", "D_hist")
PARAM_KEYS = ("vol0", "mass", "mu", "lam", "sigma_c", "eta", "zeta", "l0", "residual",
              "A0", "allow", "interface_mask", "n_interface")
DIAG_KEYS = ("phi", "phi_shear", "tau_last", "J_last")

STEP_CASES = ("plain32", "plain64", "fclamp_damp_every2_on", "fclamp_damp_every2_off",
              "noallow32", "directional32", "colliders_bc32", "edges32", "rest32")


def load_npz(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def case_config(rec):
    ng, lim, dt, damp, lo, hi, every, time0, nsteps = rec["cfg"].tolist()
    return dict(n_grid=int(ng), grid_lim=lim, dt=dt, grid_damp=damp,
                f_clamp=None if lo < 0 else (lo, hi), damage_every=int(every),
                gravity=tuple(rec["gravity"].tolist())), time0, int(nsteps)


def case_dtype(rec):
    return torch.float64 if rec["in_x"].dtype == np.float64 else torch.float32


def case_tensors(rec):
    return {k: torch.from_numpy(rec["in_" + k].copy()) for k in STATE_KEYS + PARAM_KEYS}


def attach_boundaries(sim, rec, sphere_collider, halfspace_collider, device="cpu"):
    """Recreate the fixture's colliders and particle BC on ``sim`` using the
    given collider factories (the oracle's or the product's)."""
    if "sph_c" in rec:
        sc, sv, sr = rec["sph_c"], rec["sph_v"], float(rec["sph_r"])
        sim.colliders.append(sphere_collider(lambda t: (sc + sv * t).tolist(), sr, mode="sticky"))
        ho, hv, hn = rec["hs_o"], rec["hs_v"], rec["hs_n"].tolist()
        sim.colliders.append(halfspace_collider(lambda t: (ho + hv * t).tolist(), hn, mode="slip",
                                                friction=0.4))
        ho2, hn2 = rec["hs2_o"], rec["hs2_n"].tolist()
        sim.colliders.append(halfspace_collider(lambda t: ho2.tolist(), hn2, mode="slip",
                                                friction=0.0))
    if "bc_top" in rec:
        top = torch.from_numpy(rec["bc_top"]).to(device)
        bot = torch.from_numpy(rec["bc_bot"]).to(device)
        pull = torch.from_numpy(rec["bc_pull"]).to(device)
        ramp = float(rec["bc_ramp"])

        def grip(s, t):
            s.v[bot] = 0.0
            s.v[top] = min(t / ramp, 1.0) * pull
        sim.particle_bc.append(grip)


def rel_err(a, b):
    """Norm-wise relative error ||a-b|| / ||b|| (absolute when ||b|| = 0)."""
    a = torch.as_tensor(a).double().cpu()
    b = torch.as_tensor(b).double().cpu()
    nb = float(b.norm())
    diff = float((a - b).norm())
    return diff / nb if nb > 0 else diff
