"""Benchmark of the MI355X-native AnisoMPM step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload config1]
                    [--cpu-steps 12] [--no-cpu-baseline] [--no-batched] [--no-adjoint] [--no-pcg]
                    [--no-config2] [--no-carry] [--no-builders] [--no-substep-loop]
                    [--tune | --launch-config JSON]

Prints ONE JSON line.  A step is one substep of the whole scene (binning,
stress/damage, P2G, grid update, G2P, advection) on particle state already
resident in HBM; ``value`` = particles x steps / wall seconds of the K timed
steps (bracketed by torch.cuda.synchronize()).  The K steps run as
AnisoMPM.run_frame calls of 10 substeps (the reference's frame loop,
peel_test.py:271-272) from the scene's initial state: by default K is the
config's whole run (500 steps for config1), over which the cost per step
grows as the slab tears and spreads.  A frame writes the diagnostics
tau_last, J_last, phi and phi_shear only on the substep whose values it
leaves (what a caller can read).  ``substep_loop`` times K single substep
calls that write them every step.  The headline runs the
reference's explicit damage update (the parity mode); the config's implicit
damage solve (PCG, tol 1e-6, <= 50 iterations) is the ``pcg`` leg, with its
own rooflines.

Algorithmic bytes follow one fixed convention (fp32):
  whole step   289 B per particle (27 state values read and written, 18
               material values and the allow byte read) + 32 B per active
               quadratic node (m, mv written then read) + 16 B per active
               cubic node (sum w d, sum w written then read)
  stress+P2G   189 B per particle (state and materials read, d and D_hist
               written) + 16 B per quadratic and 8 B per cubic node written
  G2P          100 B per particle (x, v, F, C, lap written) + 16 B per
               quadratic and 8 B per cubic node read
  grid update  32 B per quadratic and 16 B per cubic node (its own read and
               write; outside the step sum)
  PCG          48 B per active cubic node + 16 B per particle per operator
               application (the operator goes through the particles)
  diagnostics  phi, phi_shear, tau_last, J_last (48 B per particle written
               each step) are outside the convention and reported apart
The convention counts every material field as a per-particle read.  A field
that holds one value on every particle goes to the kernels as that value
(config1: vol0, mass, sigma_c, eta, zeta, l0, residual), so 28 of its 73
material bytes per particle are not moved; the measured ``traffic`` shows
what is.
``roofline`` is for the dominant kernel group, timed live with HIP events
recorded by libpmpm on the launch stream (pmpm_timing): its algorithmic
bytes per launch / its mean launch time, against MI355X's 8 TB/s HBM peak.
``traffic`` is the measured HBM bytes of that group (rocprofv3 --pmc,
committed under profiles/), given only when the committed measurement was
taken at the launch shape this run uses.  The plain step runs at a pinned
launch shape (the autotuner's pick for this scene) unless --tune asks for a
live search.  Forward+adjoint adds 834 B per particle-step for the backward
pass (segment recompute + adjoint kernels).
``cpu_baseline`` times the oracle (an op-for-op torch-CPU restatement of the
reference step, bit-identical to it on the CPU) on this host, on a bounded
sample of the same scene.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 1M fp32, batched vs plain; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
KERNELS = ("binning", "stress_p2g", "grid_update", "g2p")
PCG_GROUP, PCG_ITER_GROUP = 4, 5    # pmpm_timing groups: the implicit damage solve, one CG iteration

# The plain step's pinned launch shape, and the shape profiles/ were taken
# at: the fastest over config1's whole 500-step run in a coordinate sweep of
# the autotuner's knobs.  autotune.tune_solver on the scene's first steps
# picks p2g_waves 1 (the intact slab's uniform buckets); over the run, two
# wavefronts per bucket measured 0.272 against 0.286 ms per step, the two
# wavefronts split by channel (p2g_split 2, p2g_waves then unused) 0.2607
# against 0.2625 ms, and the balanced channel split (p2g_split 3: mv_x, mv_y
# on one wavefront, m, mv_z and the cubic pair on the other) 0.2528 against
# 0.2559 ms; the constitutive kernel on a second stream beside the binning
# (const_stream 2, with 256-thread work-groups) 0.2462-0.2467 against
# 0.2527 ms on the launch stream after it.
PINNED_LAUNCH = {"bucket_wgs": 8192, "block_wgs": 1024, "p2g_pair": 2, "const_threads": 256, "p2g_waves": 2,
                 "g2p_threads": 256, "cell_block": 4, "p2g_split": 3, "const_stream": 2}

FRAME = 10                          # substeps per run_frame call in the timed legs

STEP_B_PER_PARTICLE = 289
DIAG_B_PER_PARTICLE = 48
ADJ_B_PER_PARTICLE = 834


def step_bytes(n, n_q, n_c):
    return STEP_B_PER_PARTICLE * n + 32 * n_q + 16 * n_c


def pcg_apply_bytes(n, n_c):
    return 48 * n_c + 16 * n


def step_roofline(nbytes, seconds_per_step):
    gbs = nbytes / seconds_per_step / 1e9
    return {"algorithmic_bytes": nbytes, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}


def algorithmic_bytes(kernel, n, n_q, n_c):
    return {"binning": 20 * n, "stress_p2g": 189 * n + 16 * n_q + 8 * n_c, "grid_update": 32 * n_q + 16 * n_c,
            "g2p": 100 * n + 16 * n_q + 8 * n_c}[kernel]


def make_scene(name):
    from peel_mpm_amd import scenes
    return {"config0": scenes.config0, "config1": scenes.config1, "config2": scenes.config2}[name]()


def frame_length(steps, frame):
    """Substeps per run_frame call: the largest divisor of steps that is at
    most frame, so every timed frame has one length (one captured graph)."""
    f = max(1, min(frame, steps))
    while steps % f:
        f -= 1
    return f


def time_gpu(spec, steps, warmup, tune=True, prep=None, frame=FRAME, use_graph=False, count_cg=False, **solver_kw):
    """Time ``steps`` substeps of the scene from its initial state (the
    config's whole run at the default ``steps``), run as
    run_frame calls of frame_length(steps, frame) substeps (eager launches;
    use_graph replays a captured HIP graph per frame, which measured slower
    at this size), then the same run again with live per-group timing.  The
    warm-up (``warmup`` steps, at least one, after any tuning) runs first
    and the initial state is restored before each timed run, so both time
    steps 0..steps-1 of the scene: its cost grows as the slab tears and
    spreads.  Returns the
    solver, wall seconds, ms per launch and launches of each timing group,
    active nodes (mean over the run), and the launch shape used (with the
    frame length).  count_cg: a third, untimed pass of the same run as
    single substeps reads the CG iterations each damage solve took (the
    device's count) into groups["pcg_executed"] = (iterations, solves)."""
    import torch
    from peel_mpm_amd import AnisoMPM, _lib, autotune, scenes
    sim = scenes.make_solver(AnisoMPM, spec, device="cuda:0", use_graph=use_graph, **solver_kw)
    if prep:
        prep(sim)
    names = ("x", "v", "F", "C", "d", "lap", "D_hist", "tau_last", "J_last", "phi", "phi_shear")
    initial = {k: getattr(sim, k).clone() for k in names if hasattr(sim, k)}

    def restart():
        for k, t in initial.items():
            setattr(sim, k, t.clone())
        sim._step = 0

    L = _lib.lib()
    t_tune = time.perf_counter()
    launch = autotune.tune_solver(sim, steps=10) if tune else autotune.get_config()
    t_tune = time.perf_counter() - t_tune
    f = frame_length(steps, frame)
    dt = spec["dt"]

    def run(k):
        for s in range(0, k, f):
            sim.run_frame(min(f, k - s), s * dt)

    run(max(warmup, 1))
    restart()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # per-group live timing over the same run, eager (HIP events are not
    # recorded inside a replayed graph), and the active nodes of each
    # frame's last substep, averaged over the run
    restart()
    sim.use_graph = False
    L.pmpm_timing(1)
    qs, cs = [], []
    for s in range(0, steps, f):
        sim.run_frame(f, s * dt)
        q, c = sim.last_step_stats()
        qs.append(q)
        cs.append(c)
    torch.cuda.synchronize()
    n_q, n_c = round(sum(qs) / len(qs)), round(sum(cs) / len(cs))
    ms = (C.c_double * 7)()
    cnt = (C.c_int64 * 7)()
    _lib.check(L.pmpm_timing_read(ms, cnt, 7), "pmpm_timing_read")
    L.pmpm_timing(0)
    sim.use_graph = use_graph
    groups = {k: (ms[i], cnt[i]) for i, k in enumerate(KERNELS + ("pcg_solve", "pcg_iteration"))}
    # with const_stream 2 the constitutive kernel is timed apart, on the
    # second stream (group 6), beside the binning: stress+P2G per step is
    # still its duration plus P2G's (the PCG mode keeps one stream)
    groups["stress_p2g"] = (ms[1] + ms[6], cnt[1])
    if count_cg:
        restart()
        ran = 0
        for s in range(steps):
            sim.substep(s * dt)
            ran += sim.last_damage_solve()[0]
        groups["pcg_executed"] = (ran, steps)
    graphed = bool(use_graph and sim._graph_eligible())
    return sim, wall, groups, (n_q, n_c), {"config": launch, "source": "autotuned" if tune else "pinned",
                                           "tuning_s": round(t_tune, 2), "frame": f, "graph": graphed}


def time_substeps(spec, steps, warmup):
    """The same run as single substep calls (frames of one substep, so every
    step writes the diagnostics): ms per step and per-group times."""
    sim, wall, groups, _, _ = time_gpu(spec, steps, warmup, tune=False, frame=1)
    return {"ms_per_step": wall / steps * 1e3, "steps": steps, "value": sim.n * steps / wall,
            "unit": "particle-steps/s", "kernel_ms": {k: round(groups[k][0] / max(groups[k][1], 1), 5) for k in KERNELS}}


def time_batched(n_scenes, steps, warmup):
    """The 64-scene mix: plain AnisoMPM scene by scene vs one batched,
    graph-captured launch set per step."""
    import torch
    from peel_mpm_amd import AnisoMPM, VelocityBC, autotune, scenes
    from peel_mpm_amd.batched import BatchedAnisoMPM
    autotune.set_config(None)
    specs = [scenes.config3_scene(s) for s in range(n_scenes)]
    total = sum(sp["arrays"]["x"].shape[0] for sp in specs)
    plain = [scenes.make_solver(AnisoMPM, sp, device="cuda:0") for sp in specs]
    for s in range(warmup):
        for p in plain:
            p.substep(s * 1e-4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        for p in plain:
            p.substep((warmup + s) * 1e-4)
    torch.cuda.synchronize()
    plain_wall = time.perf_counter() - t0
    del plain
    torch.cuda.empty_cache()
    b = BatchedAnisoMPM(n_grid=64, dt=1e-4, gravity=(0.0, -9.8, 0.0), device="cuda:0", use_graph=True)
    for sp in specs:
        b.add_scene(scenes.tensors(sp, "cuda:0"),
                    [VelocityBC(torch.from_numpy(m).to("cuda:0"), tuple(v)) for m, v in sp["bc"]])
    t_tune = time.perf_counter()
    launch = autotune.tune_batched(b, steps=6)         # before the graph is captured
    t_tune = time.perf_counter() - t_tune
    for s in range(warmup):
        b.substep(s * 1e-4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        b.substep((warmup + s) * 1e-4)
    torch.cuda.synchronize()
    batched_wall = time.perf_counter() - t0
    finite = bool(torch.isfinite(b.x).all())
    n_q, n_c = b.last_step_stats()
    step_bytes = STEP_B_PER_PARTICLE * total + 32 * n_q + 16 * n_c
    return {"scenes": n_scenes, "particles": total, "steps": steps,
            "value": total * steps / batched_wall, "plain_scene_by_scene": total * steps / plain_wall,
            "speedup_vs_plain": plain_wall / batched_wall, "ms_per_step": batched_wall / steps * 1e3,
            "plain_ms_per_step": plain_wall / steps * 1e3, "graph": True, "finite": finite,
            "active_nodes": {"quadratic": n_q, "cubic": n_c},
            "step_roofline": dict(step_roofline(step_bytes, batched_wall / steps), peak=HBM_PEAK_GBS),
            "plain_step_roofline": dict(step_roofline(step_bytes, plain_wall / steps), peak=HBM_PEAK_GBS),
            "autotune": {"config": launch, "tuning_s": round(t_tune, 2)}}


def time_adjoint(steps, every):
    """The differentiable form, config4 as specified: the config-0 scene at
    32,768 particles, ``steps`` substeps; the loss is the squared distance of
    the final positions to a target made by the same forward pass with
    E = 1.3e5, nu = 0.28 and the fibres turned 15 degrees about z;
    gradients with respect to per-particle E, nu, fibre direction and the
    initial velocity.  Times the forward rollout (checkpoint every
    ``every``) against its backward pass (segment recompute + adjoint
    steps)."""
    import math

    import torch
    from peel_mpm_amd import AnisoMPM, scenes
    from peel_mpm_amd.adjoint import material_tensors, rollout
    spec = scenes.config4(dtype=torch.float32)
    sim = scenes.make_solver(AnisoMPM, spec, device="cuda:0")
    n, dev = sim.n, sim.device
    c, s_ = math.cos(math.radians(15.0)), math.sin(math.radians(15.0))
    Rz = torch.tensor([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]], device=dev)
    with torch.no_grad():
        mu_t, lam_t, A0_t = material_tensors(torch.full((n,), 1.3e5, device=dev), torch.full((n,), 0.28, device=dev),
                                             sim.a0 @ Rz.T, sim.alpha)
        target = rollout(sim, sim.v, mu_t, lam_t, A0_t, steps, every=every).clone()
    E = torch.full((n,), 1e5, device=dev, requires_grad=True)
    nu = torch.full((n,), 0.3, device=dev, requires_grad=True)
    a0 = sim.a0.clone().requires_grad_(True)
    v0 = sim.v.clone().requires_grad_(True)
    res, loss = {}, None
    for rep in range(2):                   # first pass warms up allocations and code objects
        for t in (E, nu, a0, v0):
            t.grad = None
        mu, lam, A0 = material_tensors(E, nu, a0, sim.alpha)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xT = rollout(sim, v0, mu, lam, A0, steps, every=every)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss = ((xT - target) ** 2).sum()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        res = {"fwd": t1 - t0, "bwd": t2 - t1}
    finite = all(bool(torch.isfinite(t.grad).all()) for t in (E, nu, a0, v0))
    n_q, n_c = sim.last_step_stats()
    fwd_bytes = step_bytes(n, n_q, n_c) * steps
    bwd_bytes = ADJ_B_PER_PARTICLE * n * steps
    tot_s = res["fwd"] + res["bwd"]
    return {"workload": f"config4: {n} particles, 64^3 grid, {steps} steps, checkpoint every {every}; target from "
                        "E 1.3e5, nu 0.28, fibres turned 15 deg about z; gradients w.r.t. E, nu, fibre, v0",
            "loss": float(loss),
            "roofline": {"bound": "hbm", "algorithmic_bytes": fwd_bytes + bwd_bytes,
                         "achieved_GBs": (fwd_bytes + bwd_bytes) / tot_s / 1e9, "peak": HBM_PEAK_GBS,
                         "frac": (fwd_bytes + bwd_bytes) / tot_s / 1e9 / HBM_PEAK_GBS,
                         "backward_frac": bwd_bytes / res["bwd"] / 1e9 / HBM_PEAK_GBS},
            "forward_ms_per_step": res["fwd"] / steps * 1e3, "backward_ms_per_step": res["bwd"] / steps * 1e3,
            "adjoint_forward_ratio": res["bwd"] / res["fwd"],
            "forward_plus_adjoint_particle_steps_per_s": n * steps / (res["fwd"] + res["bwd"]),
            "grads_finite": finite}


def time_pcg(spec, steps, warmup, label, prep=None):
    """The implicit (PCG) damage mode on ``spec``: whole-step rate and
    rooflines of the step, of the solve and of one CG iteration (HIP events
    around the solve's kernels and around each iteration's)."""
    import torch
    sim, wall, groups, (n_q, n_c), launch = time_gpu(spec, steps, warmup, tune=False, prep=prep, count_cg=True,
                                                     damage_solver="pcg", pcg_tol=1e-6, pcg_max_iter=50)
    it, rel = sim.last_damage_solve()
    n = sim.n
    solve_ms, solves = groups["pcg_solve"]
    iter_ms, issued = groups["pcg_iteration"]       # issued iterations (those past convergence exit early)
    ran, ran_solves = groups["pcg_executed"]
    it_per_step = ran / max(ran_solves, 1)
    # the iterations' launch time over the iterations that ran: no figure when
    # every solve converged at its Jacobi start
    per_iter_ms = iter_ms / ran if ran > 0 else None
    applies = it_per_step + 1                       # the initial residual's operator application
    ab = pcg_apply_bytes(n, n_c)
    sb = step_bytes(n, n_q, n_c) + applies * ab
    ms_step = wall / steps * 1e3
    return {"workload": f"{label}: {n} particles, damage_solver='pcg' (tol 1e-6 on the preconditioned residual, "
                        "<= 50 iterations)",
            "value": n * steps / wall, "unit": "particle-steps/s", "ms_per_step": ms_step,
            "solve_ms_per_step": solve_ms / max(solves, 1), "iterations_per_step": it_per_step,
            "issued_iterations_per_step": issued / max(solves, 1),
            "ms_per_iteration": per_iter_ms, "bytes_per_iteration": ab,
            "last_iterations": it, "last_residual": rel,
            "step_roofline": dict(step_roofline(sb, ms_step * 1e-3), peak=HBM_PEAK_GBS),
            "iteration_roofline": (dict(step_roofline(ab, per_iter_ms * 1e-3), peak=HBM_PEAK_GBS)
                                   if per_iter_ms else None),
            "active_nodes": {"quadratic": n_q, "cubic": n_c},
            "kernel_ms": {k: round(v[0] / max(v[1], 1), 5) for k, v in groups.items() if k != "pcg_executed"},
            "finite": bool(torch.isfinite(sim.d).all())}


def stiff_damage(sim):
    """A stiffer damage system on the same scene: eta 1e-4 (c = dt/eta =
    0.5) and l0 = 2 dx, so the solve takes many CG iterations."""
    sim.eta.fill_(1e-4)
    sim.l0.fill_(2.0 * sim.dx)


def time_small_frame(steps):
    """The host floor of a small scene: config4's scene (32,768 particles,
    fp32) stepped by run_frame eagerly (one substep call per step: argument
    set-up plus eight kernel launches from the host) and replayed from a
    captured graph (use_graph=True: one replay per frame), at the launch
    shape the autotuner picks for it.  Also the host time one eager
    substep call takes to enqueue its work."""
    import torch
    from peel_mpm_amd import AnisoMPM, autotune, scenes
    spec = scenes.config4(dtype=torch.float32)
    out = {"workload": f"config4 scene forward, {steps} substeps per frame, fp32"}
    for graph in (False, True):
        sim = scenes.make_solver(AnisoMPM, spec, device="cuda:0", use_graph=graph)
        launch = autotune.tune_solver(sim, steps=10)
        sim.run_frame(steps, 0.0)                     # warm-up (and capture)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.run_frame(steps, steps * spec["dt"])
        torch.cuda.synchronize()
        per = (time.perf_counter() - t0) / steps
        key = "graph" if graph else "eager"
        out[f"{key}_us_per_step"] = per * 1e6
        out[f"{key}_particle_steps_per_s"] = sim.n * steps / (per * steps)
        out["launch"] = launch
        if not graph:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for s in range(20):
                sim.substep(s * spec["dt"])
            out["eager_host_us_per_call"] = (time.perf_counter() - t0) / 20 * 1e6
            torch.cuda.synchronize()
    return out


CARRY_B_PER_SPLAT = 180    # bind 8, x 12, F 36, d 4, off0 12, cov0 36, color 12 in; pos 12, cov 36, color 12 out


def time_carry(spec, splats_per_particle, launches):
    """The Gaussian carry (pmpm_gaussian_carry) after a few steps of the
    scene: splats_per_particle splats per particle, each bound to a random
    particle (seeded; a worst-case gather), with the crack tint on."""
    import torch
    from peel_mpm_amd import AnisoMPM, scenes
    from peel_mpm_amd.carry import carry_gaussians
    sim = scenes.make_solver(AnisoMPM, spec, device="cuda:0")
    for s in range(3):
        sim.substep(s * spec["dt"])
    g = torch.Generator(device="cuda:0").manual_seed(0)
    G = splats_per_particle * sim.n
    data = {"g_bind": torch.randint(0, sim.n, (G,), generator=g, device="cuda:0"),
            "g_off0": 0.3 * sim.dx * torch.randn(G, 3, generator=g, device="cuda:0"),
            "g_cov0": ((0.2 * sim.dx) ** 2 * torch.eye(3, device="cuda:0")).expand(G, 3, 3).contiguous(),
            "g_color": torch.rand(G, 3, generator=g, device="cuda:0"), "grid_lim": 1.0, "scale_to_sim": 1.0,
            "center": (0.0, 0.0, 0.0)}
    for _ in range(2):
        carry_gaussians(sim, data, crack_tint=0.3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(launches):
        carry_gaussians(sim, data, crack_tint=0.3)
    torch.cuda.synchronize()
    per = (time.perf_counter() - t0) / launches
    gbs = CARRY_B_PER_SPLAT * G / per / 1e9
    return {"workload": f"{G} splats bound at random to the {sim.n} particles, crack tint on", "splats": G,
            "value": G / per, "unit": "splats/s", "ms_per_launch": per * 1e3,
            "roofline": {"bound": "hbm", "algorithmic_bytes": CARRY_B_PER_SPLAT * G, "achieved": gbs,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}}


def orange_cloud(n, seed=0):
    """A synthetic stand-in for the orange's Gaussian centres (fp32): 70 %
    in a thick peel shell, the rest through the core, 5 cm across."""
    import numpy as np
    g = np.random.default_rng(seed)
    v = g.normal(size=(n, 3))
    r = np.where(g.random(n) < 0.7, 0.9 + 0.1 * g.random(n), g.random(n) ** (1 / 3) * 0.9)
    return (0.05 * r[:, None] * v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def time_builders(repeats, cpu):
    """The scene builders on the GPU (peel_mpm_amd.builders) at the
    reference's scales: fill_interior on 3.93 M fp32 points placed as the
    orange pipeline places them (fp32 centre and scale, 62 % of the box,
    voxels of half a 64-grid cell), and the peel ball (R 0.28, spacing
    1/128).  Each call includes its host synchronisations.  With ``cpu``,
    the oracle (the reference's NumPy / SciPy code) times the same calls on
    this host as the baseline."""
    import numpy as np
    import torch
    from peel_mpm_amd.builders import ball_particles, fill_interior
    pts = orange_cloud(3_930_000)
    lo, hi = pts.min(0), pts.max(0)
    center = np.float32(0.5) * (lo + hi)
    scale = np.float32(0.62) / (hi - lo).max()
    vox = (1.0 / 64) / 2.0
    wd = torch.from_numpy(pts).cuda()
    res = {}

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / repeats, r

    t, (x, _, _) = timed(lambda: fill_interior(wd, center, scale, vox, as_tensor=True))
    res["fill_interior"] = {"workload": f"{len(pts)} fp32 points (synthetic orange-sized cloud), voxel 1/128",
                            "points_in": len(pts), "particles_out": int(x.shape[0]), "ms_per_call": t * 1e3,
                            "value": len(pts) / t, "unit": "input points/s"}
    t, xb = timed(lambda: ball_particles((0.5, 0.5, 0.5), 0.28, 1.0 / 128, as_tensor=True))
    res["ball_particles"] = {"workload": "peel ball: R 0.28, spacing 1/128, jitter 0.3", "particles_out": int(xb.shape[0]),
                             "ms_per_call": t * 1e3, "value": xb.shape[0] / t, "unit": "particles/s"}
    if cpu:
        from oracle import scene_builders as O
        t0 = time.perf_counter()
        O.fill_interior(pts, center, scale, vox)
        tc = time.perf_counter() - t0
        res["fill_interior"]["cpu_baseline"] = {"ms_per_call": tc * 1e3, "value": len(pts) / tc, "unit": "input points/s",
                                                "cores": 1, "kind": "port",
                                                "sample": "oracle (the reference's NumPy + SciPy fill_interior), one call"}
        t0 = time.perf_counter()
        O.ball_particles((0.5, 0.5, 0.5), 0.28, 1.0 / 128)
        tc = time.perf_counter() - t0
        res["ball_particles"]["cpu_baseline"] = {"ms_per_call": tc * 1e3, "value": xb.shape[0] / tc,
                                                 "unit": "particles/s", "cores": 1, "kind": "port",
                                                 "sample": "oracle (the reference's NumPy ball_particles), one call"}
    return res


OPTA_B_PER_PARTICLE = 92   # F 36, n0 12, d 4 in; P 36, psi 4 out (fp32)


def time_option_a(n, launches, cpu):
    """Option A (peel_mpm_amd.variational: energy and dPsi/dF in one
    launch) on n particles of random near-identity F (I + 0.15 N(0,1), det
    > 0.2), random unit normals and damage: per-particle rate and the HBM
    roofline at 92 B per particle.  With ``cpu``, the oracle (the
    reference's torch functions) on a 100 K sample on this host."""
    import torch
    from peel_mpm_amd import variational as V
    g = torch.Generator(device="cuda:0").manual_seed(0)
    F = torch.eye(3, device="cuda:0").repeat(n, 1, 1) + 0.15 * torch.randn(n, 3, 3, generator=g, device="cuda:0")
    F[torch.linalg.det(F) <= 0.2] = torch.eye(3, device="cuda:0")
    n0 = torch.randn(n, 3, generator=g, device="cuda:0")
    n0 = n0 / n0.norm(dim=1, keepdim=True)
    d = torch.rand(n, generator=g, device="cuda:0")
    prm = dict(mu=10.0, lam=20.0, gamma_I=5.0, gamma_II=3.0)
    for _ in range(2):
        V.energy_and_pk1(F, n0, d, **prm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(launches):
        V.energy_and_pk1(F, n0, d, **prm)
    torch.cuda.synchronize()
    per = (time.perf_counter() - t0) / launches
    gbs = OPTA_B_PER_PARTICLE * n / per / 1e9
    out = {"workload": f"{n} particles, random near-identity F, fp32; per call through the Python wrapper "
                       "(output allocation included)", "value": n / per, "unit": "particles/s",
           "ms_per_launch": per * 1e3,
           "roofline": {"bound": "hbm", "algorithmic_bytes": OPTA_B_PER_PARTICLE * n, "achieved": gbs,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}}
    if cpu:
        from oracle import variational as O
        m = 100_000
        Fc, nc, dc = F[:m].double().cpu(), n0[:m].double().cpu(), d[:m].double().cpu()
        t0 = time.perf_counter()
        O.energy(Fc, nc, dc, **prm)
        O.directional_pk1(Fc, nc, dc, **prm)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": m / tc, "unit": "particles/s", "cores": torch.get_num_threads(), "kind": "port",
                               "sample": f"oracle (the reference's torch energy + directional_pk1, fp64) on {m} of "
                                         f"the particles, {tc:.2f} s"}
    return out


OBS_B_PER_PARTICLE = 42     # phi, phi_shear, d 12, x 12, region + grip 2 in; phi_pp, psh_pp 16 read and written


def time_observe(spec, frames):
    """The peel observables (PeelObserver, one pmpm_observe_frame per frame)
    on the scene's state after a few steps: interface = the middle fifth of
    the slab's height, peel above it, the grip = its top tenth."""
    import torch
    from peel_mpm_amd import AnisoMPM, scenes
    from peel_mpm_amd.observe import PeelObserver
    sim = scenes.make_solver(AnisoMPM, spec, device="cuda:0")
    for s in range(3):
        sim.substep(s * spec["dt"])
    y = sim.x[:, 1]
    t = (y - y.min()) / (y.max() - y.min())
    interface, peel = (t > 0.4) & (t <= 0.6), t > 0.6
    obs = PeelObserver(sim, interface, ~interface & ~peel, peel, t > 0.9, frames=frames + 2)
    obs.observe(0)
    obs.observe(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(frames):
        obs.observe(2 + f)
    torch.cuda.synchronize()
    per = (time.perf_counter() - t0) / frames
    gbs = OBS_B_PER_PARTICLE * sim.n / per / 1e9
    return {"workload": f"{sim.n} particles, three regions and a grip", "ms_per_frame": per * 1e3,
            "value": sim.n / per, "unit": "particles/s",
            "roofline": {"bound": "hbm", "algorithmic_bytes": OBS_B_PER_PARTICLE * sim.n, "achieved": gbs,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}}


def copy_calibration(mib):
    """A plain device-to-device copy of ``mib`` MiB, three times: the
    reference kernel whose FETCH_SIZE / WRITE_SIZE counters scale the
    counters of the step's kernels to bytes (run under rocprofv3 --pmc)."""
    import torch
    x = torch.ones(mib * 2 ** 18, device="cuda:0", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    return {"copy_bytes": mib * 2 ** 20, "repeats": 3}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_config1.md")


def measured_traffic(workload, kernel, launch):
    """HBM bytes per launch of kernel group ``kernel`` on ``workload`` from
    the committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE,
    copy-kernel scaled), and the launch shape it was measured at; the bytes
    are None when the summary does not cover the group or was taken at
    another launch shape than ``launch``."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None, None
    m = re.search(r"^launch_config\s+(\{.*\})\s*$", text, re.M)
    shape = json.loads(m.group(1)) if m else None
    m = re.search(r"^traffic_bytes_per_launch\s+" + re.escape(workload) + r"\s+" + re.escape(kernel) + r"\s+(\d+)",
                  text, re.M)
    if not m or shape != launch:
        return None, shape
    return int(m.group(1)), shape


def time_cpu(spec, steps):
    import torch
    from oracle import cpu_step as O
    from peel_mpm_amd import scenes
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    torch.set_num_threads(threads)
    sim = scenes.make_solver(O.OracleMPM, spec, device="cpu")
    sim.substep(0.0)                       # warm-up
    t0 = time.perf_counter()
    for s in range(steps):
        sim.substep((1 + s) * spec["dt"])
    wall = time.perf_counter() - t0
    return sim.n * steps / wall, torch.get_num_threads(), wall


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="default: the workload's configured run (500 for config1)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="config1", choices=("config0", "config1", "config2"))
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched-scenes", type=int, default=64)
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=100)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-pcg", action="store_true")
    ap.add_argument("--no-substep-loop", action="store_true", help="skip the single-substep-call leg")
    ap.add_argument("--no-carry", action="store_true", help="skip the Gaussian-carry and peel-observable legs")
    ap.add_argument("--no-builders", action="store_true", help="skip the scene-builder and Option A legs")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the launch shape live instead of the pinned one (the batched form always tunes)")
    ap.add_argument("--no-config2", action="store_true", help="skip the 8M-particle config2 scene")
    ap.add_argument("--launch-config", default=None, metavar="JSON",
                    help="fixed launch shape, e.g. '{\"bucket_wgs\": 4096}' (default: the pinned shape)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the plain run, copy a MIB-MiB buffer (counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the step has no multi-GPU path in this build (--gpus {args.gpus}); "
              "scenes shard across GPUs as independent processes", file=sys.stderr)
        sys.exit(2)

    from peel_mpm_amd import autotune
    tune = args.tune
    if not tune:
        autotune.set_config(json.loads(args.launch_config) if args.launch_config else PINNED_LAUNCH)
    spec = make_scene(args.workload)
    if args.steps is None:
        args.steps = spec["steps"]
    sim, wall, groups, (n_q, n_c), launch = time_gpu(spec, args.steps, args.warmup, tune=tune)
    n = sim.n
    finite = bool(__import__("torch").isfinite(sim.x).all())
    del sim
    per_kernel = {k: groups[k][0] / max(groups[k][1], 1) for k in KERNELS}
    value = n * args.steps / wall
    dominant = max(per_kernel, key=per_kernel.get)
    alg = algorithmic_bytes(dominant, n, n_q, n_c)
    achieved = alg / (per_kernel[dominant] * 1e-3) / 1e9
    traffic, traffic_shape = measured_traffic(args.workload, dominant, launch["config"])
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n} particles, {spec['n_grid']}^3 grid, dt={spec['dt']}, explicit "
                               "damage update (the reference's; the config's PCG damage solve is the pcg leg), "
                               f"steps 0..{args.steps - 1} from the initial state as run_frame({launch['frame']}) frames"
                               f"{' replayed from a HIP graph' if launch['graph'] else ''}, diagnostics written on "
                               "each frame's last substep",
                   "particles": n, "n_grid": spec["n_grid"], "dt": spec["dt"], "form": "plain single-scene",
                   "damage": "explicit", "frame": launch["frame"]},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "algorithmic_bytes": alg,
                     "launch_ms": per_kernel[dominant],
                     "diagnostics_bytes_excluded": DIAG_B_PER_PARTICLE * n,
                     "traffic_launch_config": traffic_shape},
        "kernel_ms": {k: round(v, 5) for k, v in per_kernel.items()},
        "step_roofline": step_roofline(step_bytes(n, n_q, n_c), wall / args.steps),
        "active_nodes": {"quadratic": n_q, "cubic": n_c},
        "launch": launch,
        "finite": finite,
    }
    if not args.no_substep_loop:
        out["substep_loop"] = time_substeps(spec, args.steps, args.warmup)
    if args.copy_calibration:
        out["copy_calibration"] = copy_calibration(args.copy_calibration)
    if not args.no_pcg:
        out["pcg"] = time_pcg(spec, args.steps, 2, args.workload)
        out["pcg_stiff"] = time_pcg(spec, 10, 1, f"{args.workload}, eta 1e-4 and l0 = 2 dx", prep=stiff_damage)
    if not args.no_carry:
        out["gaussian_carry"] = time_carry(spec, 4, 20)
        out["peel_observables"] = time_observe(spec, 50)
    if not args.no_builders:
        out["scene_builders"] = time_builders(3, cpu=not args.no_cpu_baseline)
        out["option_a"] = time_option_a(984_064, 20, cpu=not args.no_cpu_baseline)
    if not args.no_config2 and args.workload != "config2":
        import torch
        s2 = make_scene("config2")
        k2 = s2["steps"]
        sim2, wall2, g2, (q2, c2), tuned2 = time_gpu(s2, k2, 2, tune=tune)
        n2 = sim2.n
        out["config2"] = {"workload": f"config2: {n2} particles, 256^3 grid, two slabs, dt=2e-05, explicit damage, "
                                      f"its {k2}-step run as run_frame({tuned2['frame']}) frames",
                          "value": n2 * k2 / wall2, "unit": "particle-steps/s", "ms_per_step": wall2 / k2 * 1e3,
                          "kernel_ms": {k: round(g2[k][0] / max(g2[k][1], 1), 5) for k in KERNELS},
                          "step_roofline": step_roofline(step_bytes(n2, q2, c2), wall2 / k2),
                          "active_nodes": {"quadratic": q2, "cubic": c2},
                          "finite": bool(torch.isfinite(sim2.x).all()), "launch": tuned2}
        del s2, sim2
        torch.cuda.empty_cache()
    if not args.no_batched:
        out["batched"] = time_batched(args.batched_scenes, args.batched_steps, 2)
    if not args.no_adjoint:
        out["adjoint"] = time_adjoint(args.adjoint_steps, 10)
        out["small_scene_frame"] = time_small_frame(100)
    if not args.no_cpu_baseline and args.cpu_steps > 0:
        cv, cores, cwall = time_cpu(spec, args.cpu_steps)
        out["cpu_baseline"] = {"value": cv, "unit": "particle-steps/s", "cores": cores, "kind": "port",
                               "sample": f"oracle (torch-CPU restatement of the reference step) on the same "
                                         f"{args.workload} scene: 1 warm-up + {args.cpu_steps} timed steps, "
                                         f"{cwall:.1f} s"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
