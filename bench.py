"""Benchmark of the MI355X-native AnisoMPM step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload config1]
                    [--cpu-steps 12] [--no-cpu-baseline] [--no-batched] [--no-adjoint] [--no-pcg]
                    [--no-config2] [--no-carry] [--no-builders] [--no-tune]

Prints ONE JSON line.  A step is one AnisoMPM.substep of the whole scene
(binning, stress/damage, P2G, grid update, G2P, advection) on particle state
already resident in HBM; ``value`` = particles x steps / wall seconds of the
K timed steps (bracketed by torch.cuda.synchronize()).

``roofline`` is for the dominant kernel, timed live with HIP events recorded
by libpmpm on the launch stream (pmpm_timing): achieved = its algorithmic
bytes per launch / its mean launch time, against MI355X's 8 TB/s HBM peak.
Algorithmic bytes per launch (fp32, damage on):
  stress+P2G  238 B per particle + 24 B per active grid node
  G2P         152 B per particle + 16 B per active grid node
  grid update  40 B per active grid node
  binning      32 B per particle
Whole-step rooflines (``step_roofline``, and the batched form's) use
289 B per particle + 32 B per active quadratic node + 16 B per active cubic
node; forward+adjoint adds 834 B per particle-step for the backward pass
(segment recompute + adjoint kernels).
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
PCG_GROUP = 4          # pmpm_timing group of the implicit damage solve


STEP_B_PER_PARTICLE = 289
ADJ_B_PER_PARTICLE = 834


def step_roofline(step_bytes, seconds_per_step):
    gbs = step_bytes / seconds_per_step / 1e9
    return {"algorithmic_bytes": step_bytes, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}


def algorithmic_bytes(kernel, n, n_act):
    return {"binning": 32 * n, "stress_p2g": 238 * n + 24 * n_act, "grid_update": 40 * n_act,
            "g2p": 152 * n + 16 * n_act}[kernel]


def make_scene(name):
    from peel_mpm_amd import scenes
    return {"config0": scenes.config0, "config1": scenes.config1, "config2": scenes.config2}[name]()


def time_gpu(spec, steps, warmup, tune=True):
    import torch
    from peel_mpm_amd import AnisoMPM, _lib, autotune, scenes
    sim = scenes.make_solver(AnisoMPM, spec, device="cuda:0")
    L = _lib.lib()
    t_tune = time.perf_counter()
    launch = autotune.tune_solver(sim, steps=10) if tune else autotune.get_config()
    t_tune = time.perf_counter() - t_tune
    for s in range(warmup):
        sim.substep(s * spec["dt"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        sim.substep((warmup + s) * spec["dt"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # per-kernel live timing over a second pass of the same number of steps
    L.pmpm_timing(1)
    for s in range(steps):
        sim.substep((warmup + steps + s) * spec["dt"])
    torch.cuda.synchronize()
    ms = (C.c_double * 4)()
    cnt = (C.c_int64 * 4)()
    _lib.check(L.pmpm_timing_read(ms, cnt, 4), "pmpm_timing_read")
    L.pmpm_timing(0)
    n_q, n_c = sim.last_step_stats()
    finite = bool(torch.isfinite(sim.x).all())
    per_kernel = {k: (ms[i] / max(cnt[i], 1)) for i, k in enumerate(KERNELS)}
    return sim.n, wall, per_kernel, (n_q, n_c), finite, {"config": launch, "tuning_s": round(t_tune, 2)}


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
    """The differentiable form: config4 (32,768 particles), a forward rollout
    of ``steps`` substeps vs its backward pass (segment recompute from the
    checkpoints every ``every`` steps + the adjoint steps)."""
    import torch
    from peel_mpm_amd import AnisoMPM, scenes
    from peel_mpm_amd.adjoint import material_tensors, rollout
    spec = scenes.config4(dtype=torch.float32)
    sim = scenes.make_solver(AnisoMPM, spec, device="cuda:0")
    n = sim.n
    E = torch.full((n,), 1e5, device="cuda:0", requires_grad=True)
    nu = torch.full((n,), 0.3, device="cuda:0", requires_grad=True)
    a0 = sim.a0.clone().requires_grad_(True)
    v0 = sim.v.clone().requires_grad_(True)
    target = sim.x.detach() + 0.01
    res = {}
    for rep in range(2):                   # first pass warms up allocations and code objects
        mu, lam, A0 = material_tensors(E, nu, a0, sim.alpha)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xT = rollout(sim, v0, mu, lam, A0, steps, every=every)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ((xT - target) ** 2).sum().backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        res = {"fwd": t1 - t0, "bwd": t2 - t1}
    finite = all(bool(torch.isfinite(t.grad).all()) for t in (E, nu, a0, v0))
    n_q, n_c = sim.last_step_stats()
    fwd_bytes = (STEP_B_PER_PARTICLE * n + 32 * n_q + 16 * n_c) * steps
    bwd_bytes = ADJ_B_PER_PARTICLE * n * steps
    tot_s = res["fwd"] + res["bwd"]
    return {"workload": f"config4: {n} particles, 64^3 grid, {steps} steps, checkpoint every {every}",
            "roofline": {"bound": "hbm", "algorithmic_bytes": fwd_bytes + bwd_bytes,
                         "achieved_GBs": (fwd_bytes + bwd_bytes) / tot_s / 1e9, "peak": HBM_PEAK_GBS,
                         "frac": (fwd_bytes + bwd_bytes) / tot_s / 1e9 / HBM_PEAK_GBS,
                         "backward_frac": bwd_bytes / res["bwd"] / 1e9 / HBM_PEAK_GBS},
            "forward_ms_per_step": res["fwd"] / steps * 1e3, "backward_ms_per_step": res["bwd"] / steps * 1e3,
            "adjoint_forward_ratio": res["bwd"] / res["fwd"],
            "forward_plus_adjoint_particle_steps_per_s": n * steps / (res["fwd"] + res["bwd"]),
            "grads_finite": finite}


def time_pcg(spec, steps, warmup):
    """The implicit (PCG) damage mode on the same scene: whole-step rate and
    the solve's own time per step (HIP events around its kernels)."""
    import torch
    from peel_mpm_amd import AnisoMPM, _lib, scenes
    sim = scenes.make_solver(AnisoMPM, spec, device="cuda:0", damage_solver="pcg", pcg_tol=1e-6, pcg_max_iter=50)
    L = _lib.lib()
    for s in range(warmup):
        sim.substep(s * spec["dt"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        sim.substep((warmup + s) * spec["dt"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    it, rel = sim.last_damage_solve()
    L.pmpm_timing(1)
    for s in range(steps):
        sim.substep((warmup + steps + s) * spec["dt"])
    torch.cuda.synchronize()
    ms = (C.c_double * 5)()
    cnt = (C.c_int64 * 5)()
    _lib.check(L.pmpm_timing_read(ms, cnt, 5), "pmpm_timing_read")
    L.pmpm_timing(0)
    return {"workload": "same scene, damage_solver='pcg' (tol 1e-6 on the preconditioned residual, <= 50 iterations)",
            "value": sim.n * steps / wall, "unit": "particle-steps/s", "ms_per_step": wall / steps * 1e3,
            "solve_ms_per_step": ms[PCG_GROUP] / max(cnt[PCG_GROUP], 1), "last_iterations": it,
            "last_residual": rel, "finite": bool(torch.isfinite(sim.d).all())}


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


def measured_traffic(workload, kernel):
    """HBM bytes per launch of ``kernel`` on ``workload`` from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-kernel scaled);
    None when no summary covers them."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch\s+" + re.escape(workload) + r"\s+" + re.escape(kernel) + r"\s+(\d+)",
                  text, re.M)
    return int(m.group(1)) if m else None


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
    ap.add_argument("--steps", type=int, default=50)
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
    ap.add_argument("--no-carry", action="store_true", help="skip the Gaussian-carry and peel-observable legs")
    ap.add_argument("--no-builders", action="store_true", help="skip the scene-builder and Option A legs")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape (no autotuning)")
    ap.add_argument("--no-config2", action="store_true", help="skip the 8M-particle config2 scene")
    ap.add_argument("--launch-config", default=None, metavar="JSON",
                    help="fixed launch shape, e.g. '{\"bucket_wgs\": 4096}' (implies --no-tune)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the plain run, copy a MIB-MiB buffer (counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the step has no multi-GPU path in this build (--gpus {args.gpus}); "
              "scenes shard across GPUs as independent processes", file=sys.stderr)
        sys.exit(2)

    if args.launch_config:
        from peel_mpm_amd import autotune
        autotune.set_config(json.loads(args.launch_config))
        args.no_tune = True
    spec = make_scene(args.workload)
    n, wall, per_kernel, (n_q, n_c), finite, tuned = time_gpu(spec, args.steps, args.warmup, tune=not args.no_tune)
    value = n * args.steps / wall
    dominant = max(per_kernel, key=per_kernel.get)
    n_act = max(n_q, n_c)
    alg = algorithmic_bytes(dominant, n, n_act)
    achieved = alg / (per_kernel[dominant] * 1e-3) / 1e9
    step_bytes = STEP_B_PER_PARTICLE * n + 32 * n_q + 16 * n_c
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n} particles, {spec['n_grid']}^3 grid, dt={spec['dt']}",
                   "particles": n, "n_grid": spec["n_grid"], "dt": spec["dt"], "form": "plain single-scene"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.workload, dominant),
                     "algorithmic_bytes": alg},
        "kernel_ms": {k: round(v, 5) for k, v in per_kernel.items()},
        "step_roofline": step_roofline(step_bytes, wall / args.steps),
        "active_nodes": {"quadratic": n_q, "cubic": n_c},
        "autotune": tuned,
        "finite": finite,
    }
    if args.copy_calibration:
        out["copy_calibration"] = copy_calibration(args.copy_calibration)
    if not args.no_pcg:
        out["pcg"] = time_pcg(spec, min(args.steps, 20), 2)
    if not args.no_carry:
        out["gaussian_carry"] = time_carry(spec, 4, 20)
        out["peel_observables"] = time_observe(spec, 50)
    if not args.no_builders:
        out["scene_builders"] = time_builders(3, cpu=not args.no_cpu_baseline)
        out["option_a"] = time_option_a(984_064, 20, cpu=not args.no_cpu_baseline)
    if not args.no_config2 and args.workload != "config2":
        import torch
        s2 = make_scene("config2")
        n2, wall2, pk2, (q2, c2), fin2, tuned2 = time_gpu(s2, 10, 2, tune=not args.no_tune)
        out["config2"] = {"workload": f"config2: {n2} particles, 256^3 grid, two slabs, dt=2e-05", "value": n2 * 10 / wall2,
                          "unit": "particle-steps/s", "ms_per_step": wall2 / 10 * 1e3,
                          "kernel_ms": {k: round(v, 5) for k, v in pk2.items()},
                          "active_nodes": {"quadratic": q2, "cubic": c2}, "finite": fin2, "autotune": tuned2}
        del s2
        torch.cuda.empty_cache()
    if not args.no_batched:
        out["batched"] = time_batched(args.batched_scenes, args.batched_steps, 2)
    if not args.no_adjoint:
        out["adjoint"] = time_adjoint(args.adjoint_steps, 10)
    if not args.no_cpu_baseline and args.cpu_steps > 0:
        cv, cores, cwall = time_cpu(spec, args.cpu_steps)
        out["cpu_baseline"] = {"value": cv, "unit": "particle-steps/s", "cores": cores, "kind": "port",
                               "sample": f"oracle (torch-CPU restatement of the reference step) on the same "
                                         f"{args.workload} scene: 1 warm-up + {args.cpu_steps} timed steps, "
                                         f"{cwall:.1f} s"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
