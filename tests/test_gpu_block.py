"""The reference's physics-behaviour tests (anisompm_py/src/test_block.py) on
the device step, built with the device grid_block (needs a GPU).

* drop & settle (test_block.py:62-81): a 28^3-particle block dropped on a
  slip floor for 16 frames stays finite with |v|max < 50;
* notched bar pulled apart (test_block.py:85-136), isotropic and with fibres
  across the bar, with the reference's grips as a host closure: a crack
  through the notch (max d > 0.99, none far from it, d_notch > 2 (d_far +
  0.05)) along the trajectory of the oracle (the reference step restated on
  the CPU) on the same scene: mean damage in the notch band at frames 26,
  39 and 51, and the maximum.

Same scenes, materials, time steps and frame counts as the reference's
script.  Its notch gate also asks for d_notch > 0.5; the reference step
itself reaches 0.144 on this scene in 52 frames (the crack is a few
particles wide in a band of 0.12, and its script prints "check" rather than
failing), so the test holds the device to the reference's trajectory
instead.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def test_drop_and_settle():
    from peel_mpm_amd import AnisoMPM, halfspace_collider
    from peel_mpm_amd.builders import grid_block
    sim = AnisoMPM(n_grid=48, grid_lim=1.0, dt=4e-4, gravity=(0, -9.8, 0), device=DEV)
    pts = grid_block(lo=[0.38, 0.45, 0.38], hi=[0.62, 0.69, 0.62], n_per_dim=[28, 28, 28])
    vol = np.full(len(pts), (0.24 ** 3) / len(pts))
    sim.add_object(torch.tensor(pts), torch.tensor(vol), rho=500, E=1e4, nu=0.3, fibers=None, alpha=0.0,
                   allow_damage=False)
    sim.colliders.append(halfspace_collider(lambda t: [0, 0.30, 0], [0, 1, 0], mode="slip", friction=0.3))
    n_sub = int((1 / 24) / sim.dt)
    for f in range(16):
        sim.run_frame(n_sub, f / 24)
        if f % 4 == 0 or f == 15:
            print(f"drop frame {f:2d} y min {float(sim.x[:, 1].min()):.3f} |v|max {float(sim.v.norm(dim=1).max()):.3f}",
                  flush=True)
    assert not bool(torch.isnan(sim.x).any())
    assert float(sim.v.norm(dim=1).max()) < 50
    assert float(sim.x[:, 1].min()) > 0.30 - 2 * sim.dx          # resting on the floor, not through it


# the oracle on the same scenes, fp32, 138 substeps per frame: frame -> (mean d in |x - 0.5| < 0.06, max d)
ORACLE_BAR = {
    False: {26: (0.0007, 0.107), 39: (0.0420, 0.995), 51: (0.1443, 0.997)},
    True: {26: (0.0006, 0.098), 39: (0.0326, 0.994), 51: (0.1227, 0.997)},
}


@pytest.mark.parametrize("anisotropic", [False, True])
def test_notched_bar_localises_at_the_notch(anisotropic):
    from peel_mpm_amd import AnisoMPM
    from peel_mpm_amd.builders import grid_block
    sim = AnisoMPM(n_grid=96, grid_lim=1.0, dt=1.5e-4, gravity=(0, 0, 0), device=DEV)
    pts = grid_block(lo=[0.25, 0.45, 0.45], hi=[0.75, 0.55, 0.55], n_per_dim=[96, 20, 20])
    notch = (np.abs(pts[:, 0] - 0.5) < 0.008) & (pts[:, 1] > 0.515)
    pts = pts[~notch]
    vol = np.full(len(pts), (0.5 * 0.1 * 0.1) / len(pts))
    fib, alpha = None, 0.0
    if anisotropic:
        fib = torch.zeros(len(pts), 3)
        fib[:, 1] = 1.0
        alpha = -1.0
    sim.add_object(torch.tensor(pts), torch.tensor(vol), rho=1000, E=3e3, nu=0.3, fibers=fib, alpha=alpha,
                   percentage=0.35, eta=0.5, l0_scale=0.7, residual=0.01, allow_damage=True)
    x0 = sim.x.clone()
    left, right = x0[:, 0] < 0.30, x0[:, 0] > 0.70
    sim.allow[left] = False
    sim.allow[right] = False
    pull_max = 0.08

    def gripper(s, t):
        vmag = pull_max * min(t / 0.4, 1.0)
        s.v[left, 0] = -vmag
        s.v[left, 1] = 0
        s.v[left, 2] = 0
        s.v[right, 0] = vmag
        s.v[right, 1] = 0
        s.v[right, 2] = 0
    sim.particle_bc.append(gripper)
    xc = x0[:, 0].cpu().numpy()
    notch_reg, far_reg = np.abs(xc - 0.5) < 0.06, np.abs(xc - 0.5) > 0.16
    n_sub = int((1 / 48) / sim.dt)
    print(f"notched bar anisotropic={anisotropic}: {sim.n} particles, 52 frames x {n_sub} substeps", flush=True)
    for f in range(52):
        sim.run_frame(n_sub, f / 48)
        if f in ORACLE_BAR[anisotropic]:
            d = sim.d.cpu().numpy()
            dn, dm = float(d[notch_reg].mean()), float(d.max())
            want_n, want_m = ORACLE_BAR[anisotropic][f]
            print(f"frame {f:2d} d_notch {dn:.4f} (oracle {want_n}) max {dm:.3f} (oracle {want_m})", flush=True)
            assert abs(dn - want_n) <= 0.05 * want_n + 2e-4 and abs(dm - want_m) <= 0.01, (f, dn, dm)
    d = sim.d.cpu().numpy()
    assert not bool(torch.isnan(sim.x).any())
    dn, df = d[notch_reg].mean(), d[far_reg].mean()
    assert df < 1e-3 and d.max() > 0.99 and dn > 2 * (df + 0.05), (dn, df)
