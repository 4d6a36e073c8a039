"""AnisoMPM.run_frame replayed from a captured HIP graph (use_graph=True)
against the eager frame loop, bit for bit: with a moving sphere collider, a
time-dependent pull, damage_every 2, odd and even frame lengths, and the
caller editing velocities and allow in place between frames."""
from __future__ import annotations

import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
KEYS = ("x", "v", "F", "C", "d", "lap", "D_hist", "tau_last", "J_last", "phi", "phi_shear")


def _sim(dtype, **kw):
    from peel_mpm_amd import AnisoMPM, VelocityBC, scenes, sphere_collider
    spec = scenes.slab_block(n_grid=32, cells=(12, 4, 10), origin=(10, 8, 11), ppc=8, seed=4, dtype=dtype)
    centre = lambda t: (0.50 + 30.0 * t, 0.43 - 20.0 * t, 0.5)
    pull = lambda t: (0.0, 0.5 + 1500.0 * t, 0.05)
    sim = scenes.make_solver(AnisoMPM, spec, device=DEV, **kw)
    sim.particle_bc[0] = VelocityBC(sim.particle_bc[0].mask, pull)
    sim.colliders.append(sphere_collider(centre, 0.07, mode="slip", friction=0.2))
    return sim


def _pair(dtype, **kw):
    return [_sim(dtype, use_graph=graph, **kw) for graph in (False, True)]


def _same(a, b, label):
    torch.cuda.synchronize()
    for k in KEYS:
        assert torch.equal(getattr(a, k), getattr(b, k)), f"{label}: {k}"


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_graph_frames_match_eager_frames(dtype):
    print(f"graph frames {dtype}", flush=True)
    eager, graph = _pair(dtype, damage_every=2)
    t = 0.0
    for f, n_sub in enumerate((5, 4, 5, 3, 6)):
        if f == 2:
            for s in (eager, graph):
                s.v[:100] += 0.01                  # in-place edits between frames
                s.allow[::3] = False
        eager.run_frame(n_sub, t)
        graph.run_frame(n_sub, t)
        t += n_sub * eager.dt
        _same(eager, graph, f"frame {f}")
        assert eager._step == graph._step
    assert graph._graph is not None and len(graph._graph["graphs"]) >= 1
    assert float(graph.D_hist.max()) > 0.0


def test_graph_frame_falls_back_for_host_hooks_and_recaptures_on_new_tensors():
    eager, graph = _pair(torch.float32)
    hook = lambda sim, time: sim.v.__setitem__(slice(0, 10), 0.0)
    for s in (eager, graph):
        s.particle_bc.append(hook)          # a plain callable: the frame runs eagerly
    eager.run_frame(3, 0.0)
    graph.run_frame(3, 0.0)
    _same(eager, graph, "host hook")
    assert graph._graph is None
    for s in (eager, graph):
        s.particle_bc.pop()
        s.mu = s.mu * 1.1                   # a replaced parameter tensor: captured again
    eager.run_frame(4, 3 * eager.dt)
    graph.run_frame(4, 3 * graph.dt)
    _same(eager, graph, "after new mu")
    for s in (eager, graph):
        s.mu = s.mu * 0.9
    eager.run_frame(4, 7 * eager.dt)
    graph.run_frame(4, 7 * graph.dt)
    _same(eager, graph, "after second mu")


@pytest.mark.parametrize("mode", ["eager", "graph", "pcg"])
def test_frame_end_diagnostics_match_every_substep(mode):
    """Diagnostics written only where a frame's final values come from
    (frame_diagnostics="end", the default) leave the same state and the same
    tau_last, J_last, phi and phi_shear as writing them on every substep.
    With damage_every 2 the first and third frames end on a substep that
    skips damage, so phi comes from an earlier substep; the last frame runs
    with allow cleared, updates no damage, and phi keeps the values of the
    frame before (anisompm.py:647-673)."""
    print(f"frame diagnostics {mode}", flush=True)
    kw = {"damage_every": 2, "use_graph": mode == "graph"}
    if mode == "pcg":
        kw["damage_solver"] = "pcg"
    end, every = _sim(torch.float32, **kw), _sim(torch.float32, frame_diagnostics="every", **kw)
    assert end._diag_substeps(4) == (3, 2)
    t = 0.0
    for f, n_sub in enumerate((4, 5, 3, 4)):
        if f == 3:
            for s in (end, every):
                s.allow[:] = False
            phi_before = end.phi.clone()
        end.run_frame(n_sub, t)
        every.run_frame(n_sub, t)
        t += n_sub * end.dt
        _same(end, every, f"{mode} frame {f}")
    assert torch.equal(end.phi, phi_before)
    assert float(end.D_hist.max()) > 0.0 and float(end.tau_last.abs().max()) > 0.0
    if mode == "graph":
        assert end._graph is not None and every._graph is not None
