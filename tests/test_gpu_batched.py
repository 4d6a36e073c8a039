"""Batched multi-scene step vs the plain step scene by scene, and HIP-graph
replay vs eager (needs a GPU).  With the stable particle sort the step is
deterministic: the batched launch does the same per-bucket work in the same
order as the single-scene step, and a graph replay the same as an eager
launch, so results must agree bit for bit."""
from __future__ import annotations

import pytest
import torch

from tests.golden_cases import rel_err

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
KEYS = ("x", "v", "F", "C", "d", "lap", "D_hist")


def small_scenes(k=4, dtype=torch.float32):
    from peel_mpm_amd import scenes
    specs = []
    for s in range(k):
        sp = scenes.slab_block(n_grid=32, cells=(6 + 2 * s, 4, 8 - s), origin=(8, 8 + s, 9), ppc=8, seed=s,
                               E=1e5 * (1 + s), sigma_c=100.0 * (1 + s), dtype=dtype)
        specs.append(sp)
    return specs


def build_batched(specs, **kw):
    from peel_mpm_amd import VelocityBC, scenes
    from peel_mpm_amd.batched import BatchedAnisoMPM
    b = BatchedAnisoMPM(n_grid=specs[0]["n_grid"], dt=specs[0]["dt"], gravity=specs[0]["gravity"], device=DEV,
                        dtype=specs[0]["dtype"], **kw)
    for sp in specs:
        t = scenes.tensors(sp, DEV)
        bcs = [VelocityBC(torch.from_numpy(m).to(DEV), tuple(v)) for m, v in sp["bc"]]
        b.add_scene(t, bcs)
    return b


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_batched_matches_plain_scene_by_scene(dtype):
    from peel_mpm_amd import AnisoMPM, scenes
    specs = small_scenes(dtype=dtype)
    print(f"batched vs plain {dtype}", flush=True)
    batch = build_batched(specs)
    plain = [scenes.make_solver(AnisoMPM, sp, device=DEV) for sp in specs]
    # one scene without damage-allowed particles: its damage must stay untouched
    plain[2].allow[:] = False
    batch._parts[2][0]["allow"] = torch.zeros_like(batch._parts[2][0]["allow"])
    for s in range(6):
        t = s * specs[0]["dt"]
        batch.substep(t)
        for p in plain:
            p.substep(t)
    torch.cuda.synchronize()
    for i, p in enumerate(plain):
        for k in KEYS:
            assert torch.equal(batch.scene(i, k), getattr(p, k)), (i, k, rel_err(batch.scene(i, k), getattr(p, k)))
    assert float(batch.scene(2, "d").abs().max()) == 0.0


@pytest.mark.parametrize("use_graph", [False, True])
def test_ragged_batch_with_empty_scenes(use_graph):
    """Empty scenes first, between and last in the batch: the others match
    the plain step scene by scene bit for bit, and the empty ones stay
    empty."""
    from peel_mpm_amd import AnisoMPM, VelocityBC, scenes
    from peel_mpm_amd.batched import BatchedAnisoMPM
    specs = small_scenes(2)
    print(f"ragged batch use_graph={use_graph}", flush=True)
    b = BatchedAnisoMPM(n_grid=specs[0]["n_grid"], dt=specs[0]["dt"], gravity=specs[0]["gravity"], device=DEV,
                        use_graph=use_graph)
    empty = {k: v[:0] for k, v in scenes.tensors(specs[0], DEV).items()}
    order = [None, 0, None, 1, None]
    for i in order:
        if i is None:
            b.add_scene(empty)
        else:
            bcs = [VelocityBC(torch.from_numpy(m).to(DEV), tuple(v)) for m, v in specs[i]["bc"]]
            b.add_scene(scenes.tensors(specs[i], DEV), bcs)
    plain = [scenes.make_solver(AnisoMPM, sp, device=DEV) for sp in specs]
    for s in range(5):
        t = s * specs[0]["dt"]
        b.substep(t)
        for p in plain:
            p.substep(t)
    torch.cuda.synchronize()
    assert b.n == sum(p.n for p in plain)
    for slot, i in enumerate(order):
        for k in KEYS:
            if i is None:
                assert b.scene(slot, k).shape[0] == 0
            else:
                assert torch.equal(b.scene(slot, k), getattr(plain[i], k)), (slot, k)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_graph_replays_match_eager(dtype):
    specs = small_scenes(3, dtype)
    print(f"graph vs eager {dtype}", flush=True)
    eager = build_batched(specs)
    graph = build_batched(specs, use_graph=True)
    for s in range(8):
        t = s * specs[0]["dt"]
        eager.substep(t)
        graph.substep(t)
        torch.cuda.synchronize()
        # eager work between replays, every replay checked
        for k in KEYS:
            assert torch.equal(getattr(graph, k), getattr(eager, k)), (s, k, rel_err(getattr(graph, k), getattr(eager, k)))


def _plain_with(specs, **kw):
    from peel_mpm_amd import AnisoMPM, scenes
    return [scenes.make_solver(AnisoMPM, sp, device=DEV, **kw) for sp in specs]


@pytest.mark.parametrize("use_graph", [False, True])
def test_moving_collider_and_timed_velocity_follow_the_replay(use_graph):
    """A sphere moving with time and a time-dependent pull: the batched
    step (eager or replayed from one captured graph) equals the plain step
    scene by scene, which evaluates both on the host every step."""
    from peel_mpm_amd import VelocityBC, sphere_collider
    specs = small_scenes(3)
    print(f"moving collider use_graph={use_graph}", flush=True)
    centre = lambda t: (0.45 + 40.0 * t, 0.42 - 25.0 * t, 0.5)
    pull = lambda t: (0.0, 0.5 + 2000.0 * t, 0.1 * (t > 3e-4))
    b = build_batched(specs, use_graph=use_graph)
    b.colliders.append(sphere_collider(centre, 0.08, mode="slip", friction=0.3))
    for i, sp in enumerate(specs):
        b._parts[i][1][0].velocity = pull
    plain = _plain_with(specs)
    for p in plain:
        p.colliders.append(sphere_collider(centre, 0.08, mode="slip", friction=0.3))
        p.particle_bc[0] = VelocityBC(p.particle_bc[0].mask, pull)
    for s in range(8):
        t = s * specs[0]["dt"]
        b.substep(t)
        for p in plain:
            p.substep(t)
        torch.cuda.synchronize()
        for i, p in enumerate(plain):
            for k in KEYS:
                assert torch.equal(b.scene(i, k), getattr(p, k)), (s, i, k, rel_err(b.scene(i, k), getattr(p, k)))
    if use_graph:
        assert b._graphs[0] is not None and b._graphs[1] is not None


@pytest.mark.parametrize("use_graph", [False, True])
def test_allow_cleared_between_steps_stops_that_scene_damage(use_graph):
    """Clearing one scene's allow in place after a few steps: from then on
    that scene skips the damage update (D_hist and d frozen, as the
    reference's allow.any() gate does, anisompm.py:647) and matches the plain
    step given the same edit."""
    specs = small_scenes(3)
    print(f"allow cleared use_graph={use_graph}", flush=True)
    b = build_batched(specs, use_graph=use_graph)
    plain = _plain_with(specs)
    for s in range(8):
        t = s * specs[0]["dt"]
        if s == 3:
            b.allow[b.offsets[1]:b.offsets[2]] = False
            plain[1].allow[:] = False
            frozen = b.scene(1, "D_hist").clone(), b.scene(1, "d").clone()
        b.substep(t)
        for p in plain:
            p.substep(t)
        torch.cuda.synchronize()
        for i, p in enumerate(plain):
            for k in KEYS:
                assert torch.equal(b.scene(i, k), getattr(p, k)), (s, i, k)
    assert torch.equal(b.scene(1, "D_hist"), frozen[0]) and torch.equal(b.scene(1, "d"), frozen[1])
    assert float(b.scene(0, "D_hist").max()) > 0.0


@pytest.mark.parametrize("use_graph", [False, True])
def test_damage_every_two(use_graph):
    specs = small_scenes(2)
    print(f"damage_every=2 use_graph={use_graph}", flush=True)
    b = build_batched(specs, use_graph=use_graph, damage_every=2)
    plain = _plain_with(specs, damage_every=2)
    for s in range(6):
        t = s * specs[0]["dt"]
        b.substep(t)
        for p in plain:
            p.substep(t)
    torch.cuda.synchronize()
    for i, p in enumerate(plain):
        for k in KEYS:
            assert torch.equal(b.scene(i, k), getattr(p, k)), (i, k)


def test_config3_all_64_scenes_graph_vs_plain_and_momentum():
    """The whole config3 mix (64 scenes of 64^3, 4.8 M particles), batched
    and replayed from one graph, against the plain step scene by scene bit
    for bit over 5 steps.  Each step also conserves momentum exactly up to
    rounding: the total particle momentum after the step equals the
    momentum the transfer starts from (velocity BC applied) plus the gravity
    impulse M dt g (quadratic weights reproduce linear functions, so P2G,
    the grid update and G2P change momentum by gravity alone here)."""
    from peel_mpm_amd import scenes
    specs = [scenes.config3_scene(s) for s in range(64)]
    print(f"config3: {sum(sp['arrays']['x'].shape[0] for sp in specs)} particles", flush=True)
    b = build_batched(specs, use_graph=True)
    plain = _plain_with(specs)
    b.substep(0.0)                       # build + capture; compared below from step 1
    for p in plain:
        p.substep(0.0)
    mass = b.mass.double()
    M = float(mass.sum())
    g = torch.tensor(specs[0]["gravity"], dtype=torch.float64, device=DEV)
    bcv = torch.zeros(b.n, 3, dtype=torch.float64, device=DEV)
    has_bc = b._bc_id > 0
    bcv[has_bc] = torch.tensor(b._bc_fns[0].value(0.0), dtype=torch.float64, device=DEV)
    for s in range(1, 6):
        t = s * 1e-4
        v0 = torch.where(has_bc.unsqueeze(1), bcv, b.v.double())
        p_start = (mass.unsqueeze(1) * v0).sum(0)
        b.substep(t)
        for p in plain:
            p.substep(t)
        torch.cuda.synchronize()
        for i, p in enumerate(plain):
            for k in KEYS:
                assert torch.equal(b.scene(i, k), getattr(p, k)), (s, i, k)
        p_end = (mass.unsqueeze(1) * b.v.double()).sum(0)
        expect = p_start + M * 1e-4 * g
        err = float((p_end - expect).norm() / expect.norm())
        print(f"step {s}: momentum rel err {err:.2e}", flush=True)
        assert err < 1e-5, err
    assert bool(torch.isfinite(b.x).all()) and bool(torch.isfinite(b.F).all())
    assert b.n == sum(sp["arrays"]["x"].shape[0] for sp in specs)
