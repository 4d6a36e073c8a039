"""Launch-shape autotuning (needs a GPU): the knobs change how work is laid
out on the chip, not the arithmetic -- bucket and block work-group counts
and the constitutive block size leave the step bit-identical; pairing
particles in P2G only reorders floating-point sums; tuning leaves the
simulation's state where it was."""
from __future__ import annotations

import pytest
import torch

from tests.golden_cases import rel_err

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
STATE = ("x", "v", "F", "C", "d", "lap", "D_hist")


def run(cfg, steps=3, spec=None, **kw):
    from peel_mpm_amd import AnisoMPM, autotune, scenes
    autotune.set_config(cfg)
    if spec is None:
        spec = scenes.slab_block(n_grid=32, cells=(16, 4, 16), origin=(8, 10, 8), ppc=8, seed=2, sigma_c=20.0)
    sim = scenes.make_solver(AnisoMPM, spec, device=DEV, **kw)
    for s in range(steps):
        sim.substep(s * spec["dt"])
    torch.cuda.synchronize()
    autotune.set_config(None)
    return sim


def test_work_group_counts_leave_the_step_bit_identical():
    ref = run(None)
    for cfg in ({"bucket_wgs": 1024, "block_wgs": 512}, {"bucket_wgs": 4096, "const_threads": 64},
                {"block_wgs": 1, "bucket_wgs": 8, "const_threads": 128}, {"g2p_threads": 128},
                {"g2p_threads": 512, "bucket_wgs": 8}):
        # (the number of P2G waves changes how a bucket's particles are split
        # among per-wave arenas, hence the summation order: compared below)
        sim = run(cfg)
        for k in STATE:
            assert torch.equal(getattr(sim, k), getattr(ref, k)), (cfg, k, rel_err(getattr(sim, k), getattr(ref, k)))


def test_single_particle_p2g_matches_paired_to_rounding():
    a = run({"p2g_pair": 2})
    for cfg in ({"p2g_pair": 1}, {"p2g_waves": 4}, {"p2g_waves": 4, "p2g_pair": 1}, {"p2g_waves": 1},
                {"p2g_waves": 1, "p2g_pair": 1}):
        b = run(cfg)
        for k in ("x", "v", "F", "C"):
            assert rel_err(getattr(b, k), getattr(a, k)) <= 1e-6, (cfg, k)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_cell_block_8_matches_4_to_rounding(dtype):
    """The 4x4x8-cell P2G block (two buckets' arenas summed into one, folded
    by the grid pass) against the 4x4x4 one: the same sums in another order.
    Grids of 8 and 9 buckets per edge (an odd count leaves the last block
    half empty), explicit and PCG damage."""
    from peel_mpm_amd import scenes
    tol = 1e-6 if dtype == torch.float32 else 1e-12
    for n_grid, cells, origin in ((32, (16, 4, 16), (8, 10, 8)), (36, (20, 5, 18), (6, 9, 7))):
        spec = scenes.slab_block(n_grid=n_grid, cells=cells, origin=origin, ppc=8, seed=3, sigma_c=20.0, dtype=dtype)
        for solver in ("explicit", "pcg"):
            print(f"cell_block 8 vs 4: n_grid {n_grid} {solver} {dtype}", flush=True)
            a = run({"cell_block": 4}, spec=spec, damage_solver=solver)
            b = run({"cell_block": 8}, spec=spec, damage_solver=solver)
            for k in ("x", "v", "F", "C", "d", "lap"):
                e = rel_err(getattr(b, k), getattr(a, k))
                assert e <= (tol if k not in ("d", "lap") else 100 * tol), (n_grid, solver, k, e)


def test_cell_block_8_at_the_domain_edges_matches_the_golden():
    """The reference's edge case (particles near and past the domain faces:
    clamp-to-edge node indices, arena positions beyond the grid folded onto
    the edge nodes) with cell_block 8, against the reference's fp32 output."""
    from peel_mpm_amd import autotune
    from tests.golden_cases import load_npz
    from tests.test_gpu_parity import gpu_sim_from_case
    rec = load_npz("step_edges32.npz")
    autotune.set_config({"cell_block": 8})
    try:
        sim, time0, nsteps = gpu_sim_from_case(rec)
        for s in range(nsteps):
            sim.substep(time0 + s * sim.dt)
        torch.cuda.synchronize()
    finally:
        autotune.set_config(None)
    for k in ("x", "v", "F", "C"):
        e = rel_err(getattr(sim, k), torch.from_numpy(rec["out_" + k]))
        assert e <= 1e-5, (k, e)


def test_batched_cell_block_8_matches_plain_bit_for_bit():
    from peel_mpm_amd import AnisoMPM, VelocityBC, autotune, scenes
    from peel_mpm_amd.batched import BatchedAnisoMPM
    autotune.set_config({"cell_block": 8})
    try:
        specs = [scenes.slab_block(n_grid=36, cells=(6 + 3 * s, 4, 9 - s), origin=(8, 8 + s, 9), ppc=8, seed=s)
                 for s in range(3)]
        b = BatchedAnisoMPM(n_grid=36, dt=1e-4, device=DEV, use_graph=True)
        for sp in specs:
            b.add_scene(scenes.tensors(sp, DEV), [VelocityBC(torch.from_numpy(m).to(DEV), tuple(v)) for m, v in sp["bc"]])
        plain = [scenes.make_solver(AnisoMPM, sp, device=DEV) for sp in specs]
        for s in range(4):
            b.substep(s * 1e-4)
            for p in plain:
                p.substep(s * 1e-4)
        torch.cuda.synchronize()
        for i, p in enumerate(plain):
            for k in STATE:
                assert torch.equal(b.scene(i, k), getattr(p, k)), (i, k)
    finally:
        autotune.set_config(None)


def test_pcg_reduction_shape_changes_only_rounding():
    a = run({"block_wgs": 2048}, damage_solver="pcg")
    b = run({"block_wgs": 64}, damage_solver="pcg")
    for k in ("x", "v", "d"):
        assert rel_err(getattr(b, k), getattr(a, k)) <= 1e-6, k


def test_tune_solver_restores_state_and_picks_a_valid_config():
    from peel_mpm_amd import AnisoMPM, autotune, scenes
    spec = scenes.slab_block(n_grid=32, cells=(16, 4, 16), origin=(8, 10, 8), ppc=8, seed=2)
    sim = scenes.make_solver(AnisoMPM, spec, device=DEV)
    sim.substep(0.0)
    before = {k: getattr(sim, k).clone() for k in STATE}
    step0 = sim._step
    small = {"bucket_wgs": (1024, 2048), "p2g_pair": (1, 2), "const_threads": (128, 256), "block_wgs": (1024, 2048),
             "p2g_waves": (2, 4)}
    best = autotune.tune_solver(sim, steps=2, space=small, use_cache=False)
    assert autotune.get_config() == best
    assert sim._step == step0
    for k in STATE:
        assert torch.equal(getattr(sim, k), before[k]), k
    sim.substep(step0 * spec["dt"])
    assert bool(torch.isfinite(sim.x).all())
    autotune.set_config(None)


def test_tune_batched_then_graph_replays_match_eager():
    from peel_mpm_amd import VelocityBC, autotune, scenes
    from peel_mpm_amd.batched import BatchedAnisoMPM

    def build(use_graph):
        b = BatchedAnisoMPM(n_grid=32, dt=1e-4, device=DEV, use_graph=use_graph)
        for s in range(3):
            sp = scenes.slab_block(n_grid=32, cells=(6 + 2 * s, 4, 8), origin=(8, 8 + s, 9), ppc=8, seed=s)
            b.add_scene(scenes.tensors(sp, DEV), [VelocityBC(torch.from_numpy(m).to(DEV), tuple(v)) for m, v in sp["bc"]])
        return b

    g = build(True)
    small = {"bucket_wgs": (1024, 2048), "p2g_pair": (2,), "const_threads": (256,), "block_wgs": (1024, 2048)}
    best = autotune.tune_batched(g, steps=2, space=small, use_cache=False)
    e = build(False)
    for s in range(4):
        g.substep(s * 1e-4)
        e.substep(s * 1e-4)
        torch.cuda.synchronize()
        for k in STATE:
            assert torch.equal(getattr(g, k), getattr(e, k)), (best, s, k)
    autotune.set_config(None)
