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


LAYOUTS = [{"cell_block": 8}, {"p2g_split": 2}, {"p2g_split": 2, "p2g_pair": 1}, {"p2g_split": 3},
           {"p2g_split": 3, "p2g_pair": 1}]


def uneven_spec(dtype, seed=5):
    """A slab whose particles are redrawn uniformly over its box (Poisson
    cell counts, empty cells among full ones) with 40 of them piled into one
    cell and 12 into a cell of the next bucket along z: lanes of a packed
    wavefront then hold several cells of unequal load.  The pile drives
    many particles to the damage threshold, where the last bit of a stress
    decides the relu of the update; so damage is frozen (allow off) at a
    random field, which the cubic scatter and the Laplacian still carry."""
    import numpy as np
    from peel_mpm_amd import scenes
    spec = scenes.slab_block(n_grid=32, cells=(12, 5, 14), origin=(9, 9, 7), ppc=6, seed=seed, sigma_c=20.0,
                             dtype=dtype)
    x = spec["arrays"]["x"]
    rng = np.random.default_rng(seed)
    lo, hi = x.min(0), x.max(0)
    x[:] = lo + (hi - lo) * rng.random(x.shape)
    dx = 1.0 / 32
    x[:40] = (np.array([13.0, 11.0, 10.0]) + 0.1 + 0.8 * rng.random((40, 3))) * dx
    x[40:52] = (np.array([13.0, 11.0, 14.0]) + 0.1 + 0.8 * rng.random((12, 3))) * dx
    spec["arrays"]["d"] = 0.5 * rng.random(x.shape[0])
    spec["arrays"]["allow"] = np.zeros(x.shape[0], dtype=bool)
    return spec


@pytest.mark.parametrize("cfg", LAYOUTS, ids=lambda c: "-".join(f"{k}{v}" for k, v in c.items()))
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_block_layouts_match_the_4x4x4_bucket_to_rounding(cfg, dtype):
    """The other P2G layouts against one 4x4x4-cell bucket per wavefront:
    the 4x4x8-cell block (two buckets' arenas summed into one, folded by the
    grid pass) and two or four buckets packed onto one wavefront's lanes by
    load -- the same sums in another order.  Grids of 8 and 9 buckets per
    edge (an odd count leaves the last block or group half empty), a slab
    with uneven cells, explicit and PCG damage."""
    from peel_mpm_amd import scenes
    tol = 1e-6 if dtype == torch.float32 else 1e-12
    cases = [(32, (16, 4, 16), (8, 10, 8)), (36, (20, 5, 18), (6, 9, 7)), "uneven"]
    for case in cases:
        if case == "uneven":
            spec = uneven_spec(dtype)
        else:
            n_grid, cells, origin = case
            spec = scenes.slab_block(n_grid=n_grid, cells=cells, origin=origin, ppc=8, seed=3, sigma_c=20.0, dtype=dtype)
        for solver in ("explicit", "pcg"):
            print(f"{cfg} vs 4x4x4: {case} {solver} {dtype}", flush=True)
            a = run({"cell_block": 4}, spec=spec, damage_solver=solver)
            b = run(cfg, spec=spec, damage_solver=solver)
            for k in ("x", "v", "F", "C", "d", "lap"):
                e = rel_err(getattr(b, k), getattr(a, k))
                assert e <= (tol if k not in ("d", "lap") else 100 * tol), (case, solver, k, e)


@pytest.mark.parametrize("cfg", LAYOUTS, ids=lambda c: "-".join(f"{k}{v}" for k, v in c.items()))
def test_block_layouts_at_the_domain_edges_match_the_golden(cfg):
    """The reference's edge case (particles near and past the domain faces:
    clamp-to-edge node indices, arena positions beyond the grid folded onto
    the edge nodes) with each layout, against the reference's fp32 output."""
    from peel_mpm_amd import autotune
    from tests.golden_cases import load_npz
    from tests.test_gpu_parity import gpu_sim_from_case
    rec = load_npz("step_edges32.npz")
    autotune.set_config(cfg)
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


@pytest.mark.parametrize("cfg", LAYOUTS, ids=lambda c: "-".join(f"{k}{v}" for k, v in c.items()))
def test_batched_block_layouts_match_plain_bit_for_bit(cfg):
    """Blocks and packed groups are fixed by bucket position within a scene,
    so a scene's sums run in the same order batched or alone."""
    from peel_mpm_amd import AnisoMPM, VelocityBC, autotune, scenes
    from peel_mpm_amd.batched import BatchedAnisoMPM
    autotune.set_config(cfg)
    try:
        specs = [scenes.slab_block(n_grid=36, cells=(6 + 3 * s, 4 + s, 9 - s), origin=(8, 8 + s, 9), ppc=8, seed=s)
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
