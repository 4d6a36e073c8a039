"""Gaussian carry on the GPU (pmpm_gaussian_carry) against the oracle's
restatement of anisompm_py/src/run_orange.py:116-128 (needs a GPU).

Norm-wise relative error per output: 1e-6 in fp32 (a handful of roundings
on O(1) positions and covariance products), 1e-13 in fp64.
"""
from __future__ import annotations

import pytest
import torch

from oracle import gaussian_carry as O
from tests.golden_cases import rel_err
from tests.test_oracle_carry import CEN, GL, SC, scene

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TOL = {torch.float32: 1e-6, torch.float64: 1e-13}


@pytest.mark.parametrize("tint", [0.0, 0.35])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_carry_matches_oracle(dtype, tint):
    from peel_mpm_amd.carry import carry
    x, F, d, bind, off0, cov0, color = scene(G=20000, N=3000, seed=5, dtype=dtype)
    print(f"carry dtype={dtype} tint={tint} G={bind.shape[0]}", flush=True)
    pos, cov, col = carry(x.to(DEV), F.to(DEV), d.to(DEV), bind.to(DEV), off0, cov0, color, GL, SC, CEN, tint)
    torch.cuda.synchronize()
    rp, rc, rcol = O.carry(x, F, d, bind, off0, cov0, color, GL, SC, CEN, tint)
    for name, a, b in (("pos", pos, rp), ("cov", cov, rc), ("color", col, rcol)):
        assert rel_err(a, b) <= TOL[dtype], (name, rel_err(a, b))


def test_carry_of_a_stepped_scene_and_without_colour():
    """Splats bound to the particles of a small peel slab after three steps."""
    from peel_mpm_amd import AnisoMPM, scenes
    from peel_mpm_amd.carry import carry, carry_gaussians
    spec = scenes.slab_block(n_grid=32, cells=(12, 4, 12), origin=(10, 8, 10), ppc=8, seed=1)
    sim = scenes.make_solver(AnisoMPM, spec, device=DEV)
    for s in range(3):
        sim.substep(s * spec["dt"])
    g = torch.Generator().manual_seed(7)
    G = 4 * sim.n
    data = {"g_bind": torch.randint(0, sim.n, (G,), generator=g).to(DEV),
            "g_off0": (0.3 * sim.dx * torch.randn(G, 3, generator=g)).to(DEV),
            "g_cov0": ((0.2 * sim.dx) ** 2 * torch.eye(3).expand(G, 3, 3)).contiguous().to(DEV),
            "g_color": torch.rand(G, 3, generator=g).to(DEV), "grid_lim": 1.0, "scale_to_sim": 0.8,
            "center": torch.tensor([0.0, 0.1, -0.1])}
    pos, cov, col = carry_gaussians(sim, data, crack_tint=0.5)
    torch.cuda.synchronize()
    cpu = lambda t: t.detach().cpu()
    rp, rc, rcol = O.carry(cpu(sim.x), cpu(sim.F), cpu(sim.d), cpu(data["g_bind"]), cpu(data["g_off0"]),
                           cpu(data["g_cov0"]), cpu(data["g_color"]), 1.0, 0.8, data["center"], 0.5)
    assert rel_err(pos, rp) <= 1e-6 and rel_err(cov, rc) <= 1e-6 and rel_err(col, rcol) <= 1e-6
    p2, c2, none = carry(sim.x, sim.F, sim.d, data["g_bind"], data["g_off0"], data["g_cov0"], data["g_color"],
                         1.0, 0.8, data["center"], with_color=False)
    assert none is None and torch.equal(p2, pos) and torch.equal(c2, cov)
