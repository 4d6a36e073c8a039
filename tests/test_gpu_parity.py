"""HIP step vs the oracle and the reference's golden vectors (needs a GPU).

Tolerances.  Integer work (cells, node indices, bucket keys, the stable sort
permutation) is compared bit-exactly.  Floating fields use the norm-wise
relative error ||a-b||/||b|| per field per step:

* one step from a golden input against the reference's fp32 output: every
  field 1e-5, the diagnostics (phi, phi_shear, tau_last, J_last) included
  (measured worst over the cases: 3.5e-6, D_hist);
* kinematic state x, v, F, C against the fp32 reference: 1e-5;
* over several steps, stress-derived fields (d, D_hist, lap, phi, phi_shear, tau_last) are
  ill-conditioned in fp32 at small strain: the corotated stress is a
  difference F - R of O(1) numbers, so two fp32 implementations that round
  differently disagree by about ulp/strain.  They are held to the fp32 noise
  floor instead: against the fp64 oracle on the same input, the HIP result
  must be within NOISE_X times the fp32 reference's own error (plus 1e-6),
  and within 1e-5 of it where the floor is below that;
* fp64 runs against the fp64 reference: 1e-10 per step (1e-7 for the
  20-step trajectory's damage field, which thresholds at relu/max).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import cpu_step as O
from tests.golden_cases import (DIAG_KEYS, PARAM_KEYS, STATE_KEYS, STEP_CASES, attach_boundaries, case_config,
                                case_dtype, case_tensors, load_npz, rel_err)

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TOL_STATE = {torch.float32: 1e-5, torch.float64: 1e-10}
TOL_DIAG = TOL_STATE
KINEMATIC = ("x", "v", "F", "C", "J_last")
SENSITIVE = ("d", "D_hist", "lap", "phi", "phi_shear", "tau_last")
NOISE_X = 3.0


def gpu_sim_from_case(rec, host_colliders=False, **kw):
    from peel_mpm_amd import AnisoMPM, halfspace_collider, sphere_collider
    cfg, time0, nsteps = case_config(rec)
    sim = AnisoMPM(device=DEV, dtype=case_dtype(rec), **cfg, **kw)
    for k, v in case_tensors(rec).items():
        setattr(sim, k, v.to(DEV))
    sim.n = sim.x.shape[0]
    sim._step = int(rec["in__step"])
    attach_boundaries(sim, rec, sphere_collider, halfspace_collider, device=DEV)
    if host_colliders:
        sim.colliders = [(lambda pos, v, t, c=c: c(pos, v, t)) for c in sim.colliders]
    return sim, time0, nsteps


def compare(sim, expect: dict, dtype, label, keys=STATE_KEYS + DIAG_KEYS):
    errs = {}
    for k in keys:
        if k not in expect:
            continue
        got = getattr(sim, k)
        errs[k] = rel_err(got, expect[k])
        tol = TOL_DIAG[dtype] if k in DIAG_KEYS else TOL_STATE[dtype]
        assert errs[k] <= tol, f"{label}: {k} rel err {errs[k]:.3e} > {tol:.0e}  (all: {errs})"
    return errs


# ---------------------------------------------------------------- function level
def test_stencil_cells_and_indices_bit_exact():
    from peel_mpm_amd import kernels
    z = load_npz("weights.npz")
    for ng in (64, 24, 48):
        x = torch.from_numpy(z[f"x_{ng}"]).to(DEV)
        print(f"stencil ng={ng} n={x.shape[0]}", flush=True)
        s = kernels.stencil(x, ng)
        torch.cuda.synchronize()
        assert np.array_equal(s["base_q"].cpu().numpy(), z[f"base_{ng}"])
        assert np.array_equal(s["base_c"].cpu().numpy(), z[f"basec_{ng}"])
        assert np.array_equal(s["idx27"].cpu().numpy(), z[f"idx27_{ng}"])
        assert np.array_equal(s["idx64"].cpu().numpy(), z[f"idx64_{ng}"])
        for k, g in (("w", "w"), ("wc", "wc"), ("wc2", "wc2")):
            np.testing.assert_allclose(s[k].cpu().numpy(), z[f"{g}_{ng}"], rtol=0, atol=2e-7)


def expected_cell_keys(x: torch.Tensor, ng: int):
    """Cell key from the oracle's cubic base cell (c = base_c + 1 = floor(x/dx)):
    4x4x4-cell bucket, then the cell inside it, both clamped to the grid."""
    base_c, _, _, _ = O.cubic_weights(x.cpu(), float(ng))
    c = (base_c + 1).clamp(-4, ng + 4)
    nb = (ng + 3) // 4
    b = torch.div(c, 4, rounding_mode="floor").clamp(0, nb - 1)
    l = (c - 4 * b).clamp(0, 3)
    bucket = (b[:, 0] * nb + b[:, 1]) * nb + b[:, 2]
    return (bucket * 64 + l[:, 0] * 16 + l[:, 1] * 4 + l[:, 2]).numpy()


@pytest.mark.parametrize("ng,n,seed", [(64, 65536, 0), (24, 1000, 1), (48, 300000, 2), (128, 1, 3), (32, 0, 4)])
def test_binning_stable_sort_bit_exact(ng, n, seed):
    from peel_mpm_amd import kernels
    print(f"bin ng={ng} n={n}", flush=True)
    g = np.random.default_rng(seed)
    x = g.uniform(-0.1, 1.1, (n, 3)).astype(np.float32)
    if n > 10:
        x[: n // 2] = np.sort(g.uniform(0.2, 0.8, (n // 2, 3)), axis=0)   # spatially coherent runs
    xt = torch.from_numpy(x).to(DEV)
    key, perm, start = kernels.bin_particles(xt, ng, stable=True)
    torch.cuda.synchronize()
    want = expected_cell_keys(torch.from_numpy(x), ng)
    assert np.array_equal(key.cpu().numpy(), want)
    assert np.array_equal(perm.cpu().numpy(), np.argsort(want, kind="stable"))
    counts = np.bincount(want, minlength=kernels.n_cells(ng))
    assert np.array_equal(start.cpu().numpy(), np.concatenate([[0], np.cumsum(counts)]))


def test_binning_stable_over_a_reused_workspace():
    """One workspace reused over inputs in cell order, shuffled, and
    reversed by cells or by particles (each cell's particles then one run
    of one wavefront, or spread over many), as the step reuses its own:
    every call the stable argsort."""
    from peel_mpm_amd import kernels
    ng, ncell, ppc = 64, 5000, 8           # 8 particles per cell: no cell straddles a wavefront in cell order
    g = np.random.default_rng(21)
    flat = g.choice((ng - 4) ** 3, size=ncell, replace=False)
    c = np.stack([flat // (ng - 4) ** 2, (flat // (ng - 4)) % (ng - 4), flat % (ng - 4)], 1) + 2
    x = ((np.repeat(c, ppc, 0) + g.uniform(0.05, 0.95, (ncell * ppc, 3))) / ng).astype(np.float32)
    x = x[np.argsort(expected_cell_keys(torch.from_numpy(x), ng), kind="stable")]
    grp = np.arange(ncell * ppc).reshape(ncell, ppc)
    inputs = {"cell order": x, "shuffled": x[g.permutation(len(x))], "cell order again": x,
              "cells reversed": x[grp[::-1].reshape(-1)], "particles reversed": x[::-1].copy()}
    ws = kernels.bin_workspace(len(x), ng)
    for name, xi in inputs.items():
        print(f"bin reuse {name}", flush=True)
        key, perm, _ = kernels.bin_particles(torch.from_numpy(np.ascontiguousarray(xi)).to(DEV), ng, stable=True, ws=ws)
        want = expected_cell_keys(torch.from_numpy(xi), ng)
        assert np.array_equal(key.cpu().numpy(), want), name
        assert np.array_equal(perm.cpu().numpy(), np.argsort(want, kind="stable")), name


def test_binning_piles_take_every_sort_path():
    """Cells and buckets far above the usual fill, shuffled: a cell of 40
    particles (sorted in LDS by insertion), a bucket of 2,600 particles (more
    than the LDS stage holds: sorted in place in global memory), cells of
    2-16 (the register network) and singletons -- each still the stable
    argsort."""
    from peel_mpm_amd import kernels
    ng = 32
    g = np.random.default_rng(33)
    parts = [
        (np.array([10, 11, 12]) + g.uniform(0.1, 0.9, (40, 3))) / ng,          # one cell of 40
        (np.array([20, 20, 20]) + g.uniform(0.0, 4.0, (2600, 3))) / ng,        # one bucket of 2,600
        (np.array([21, 21, 21]) + g.uniform(0.2, 0.8, (300, 3))) / ng,         # and a pile in it
        g.uniform(0.0, 1.0, (5000, 3)),
    ]
    x = np.concatenate(parts).astype(np.float32)
    x = x[g.permutation(len(x))]
    key, perm, start = kernels.bin_particles(torch.from_numpy(x).to(DEV), ng, stable=True)
    torch.cuda.synchronize()
    want = expected_cell_keys(torch.from_numpy(x), ng)
    counts = np.bincount(want, minlength=kernels.n_cells(ng))
    assert counts.max() > 16 and np.bincount(want // 64).max() > 2048
    assert np.array_equal(key.cpu().numpy(), want)
    assert np.array_equal(perm.cpu().numpy(), np.argsort(want, kind="stable"))
    assert np.array_equal(start.cpu().numpy(), np.concatenate([[0], np.cumsum(counts)]))


def test_binning_unstable_is_a_permutation_grouped_by_bucket():
    from peel_mpm_amd import kernels
    g = np.random.default_rng(9)
    x = torch.from_numpy(g.uniform(0, 1, (200000, 3)).astype(np.float32)).to(DEV)
    key, perm, start = kernels.bin_particles(x, 64, stable=False)
    k, p = key.cpu().numpy(), perm.cpu().numpy()
    assert np.array_equal(np.sort(p), np.arange(p.size))
    assert np.all(np.diff(k[p]) >= 0)


def test_sym_eig_against_reference_golden_and_invariants():
    from peel_mpm_amd import kernels
    z = load_npz("sym_eig.npz")
    for tag, dt, tol in (("32", torch.float32, 2e-6), ("64", torch.float64, 1e-13)):
        A = torch.from_numpy(z["A" + tag]).to(DEV)
        ev, V = kernels.sym_eig(A)
        ev, V = ev.cpu().double(), V.cpu().double()
        ref = torch.from_numpy(z["evals" + tag]).double()
        scale = ref.abs().amax(1, keepdim=True).clamp(min=1.0)
        ok = torch.isfinite(ref).all(1)
        assert ((ev - ref).abs() / scale)[ok].max() < tol
        # eigenvector columns at ties are arbitrary: compare the reconstruction instead
        As = torch.nan_to_num(0.5 * (torch.from_numpy(z["A" + tag]).double() + torch.from_numpy(z["A" + tag]).double().transpose(1, 2)))
        fin = ok & (As.abs().amax((1, 2)) < 1e30)
        recon = V @ torch.diag_embed(ev) @ V.transpose(1, 2)
        assert ((recon - As).abs().amax((1, 2)) / scale[:, 0])[fin].max() < 10 * tol
        assert (V.transpose(1, 2) @ V - torch.eye(3, dtype=torch.float64)).abs().amax((1, 2))[fin].max() < 10 * tol


# ---------------------------------------------------------------- one step
@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("name", STEP_CASES)
def test_substep_matches_reference_golden(name, exact):
    rec = load_npz(f"step_{name}.npz")
    print(f"step case={name} exact_sigma={exact}", flush=True)
    sim, time0, nsteps = gpu_sim_from_case(rec, exact_sigma_eig=exact)
    for s in range(nsteps):
        sim.substep(time0 + s * sim.dt)
    torch.cuda.synchronize()
    expect = {k[4:]: torch.from_numpy(v) for k, v in rec.items() if k.startswith("out_")}
    errs = compare(sim, expect, case_dtype(rec), name)
    print(name, {k: f"{v:.2e}" for k, v in errs.items()})
    if "out_phi" not in rec:
        assert not hasattr(sim, "phi")


def test_host_colliders_match_device_colliders():
    rec = load_npz("step_colliders_bc32.npz")
    a, t0, _ = gpu_sim_from_case(rec)
    b, _, _ = gpu_sim_from_case(rec, host_colliders=True)
    a.substep(t0)
    b.substep(t0)
    for k in ("x", "v", "C", "F"):
        assert rel_err(getattr(b, k), getattr(a, k)) < 1e-6, k


def test_velocity_bc_descriptor_matches_closure():
    from peel_mpm_amd import VelocityBC
    rec = load_npz("step_colliders_bc32.npz")
    a, t0, _ = gpu_sim_from_case(rec)
    b, _, _ = gpu_sim_from_case(rec)
    top = torch.from_numpy(rec["bc_top"]).to(DEV)
    bot = torch.from_numpy(rec["bc_bot"]).to(DEV)
    pull = rec["bc_pull"].astype(np.float64)
    ramp = float(rec["bc_ramp"])
    b.particle_bc = [VelocityBC(bot, (0.0, 0.0, 0.0)), VelocityBC(top, lambda t: (min(t / ramp, 1.0) * pull))]
    a.substep(t0)
    b.substep(t0)
    for k in ("x", "v", "C", "F", "d"):
        assert rel_err(getattr(b, k), getattr(a, k)) < 1e-6, k


@pytest.mark.parametrize("tag", ["32", "64"])
def test_trajectory_matches_reference_golden(tag):
    """20 steps of a pulled slab.  fp32: kinematic fields within 1e-5 per step
    taken of the fp32 reference; every field's drift from the fp64 reference
    within NOISE_X times the fp32 reference's own drift.  fp64: vs fp64."""
    rec = load_npz(f"traj_slab{tag}.npz")
    r64 = load_npz("traj_slab64.npz")
    from peel_mpm_amd import AnisoMPM
    dtype = torch.float32 if tag == "32" else torch.float64
    sim = AnisoMPM(n_grid=32, grid_lim=1.0, dt=1e-4, device=DEV, dtype=dtype)
    for k in STATE_KEYS + PARAM_KEYS:
        setattr(sim, k, torch.from_numpy(rec["in_" + k].copy()).to(DEV))
    sim.n = sim.x.shape[0]
    top = torch.from_numpy(rec["top"]).to(DEV)
    sim.particle_bc.append(lambda s, t: s.v.__setitem__(top, torch.tensor([0.0, 0.5, 0.0], dtype=dtype, device=DEV)))
    marks = sorted({int(k[1:].split("_")[0]) for k in rec if k.startswith("s") and k[1].isdigit()})
    report = {}
    for s in range(max(marks)):
        sim.substep(s * sim.dt)
        if s + 1 not in marks:
            continue
        torch.cuda.synchronize()
        for k in STATE_KEYS:
            mine = getattr(sim, k)
            ref = torch.from_numpy(rec[f"s{s + 1}_{k}"])
            e = rel_err(mine, ref)
            report[(s + 1, k)] = e
            if dtype == torch.float64:
                assert e <= (1e-7 if k in SENSITIVE else TOL_STATE[dtype] * (s + 1)), (s + 1, k, e)
            elif k in KINEMATIC:
                assert e <= TOL_STATE[dtype] * (s + 1), (s + 1, k, e)
    if dtype == torch.float32:
        for k in STATE_KEYS:
            ref64 = torch.from_numpy(r64[f"s20_{k}"])
            floor = rel_err(torch.from_numpy(rec[f"s20_{k}"]), ref64)
            mine = rel_err(getattr(sim, k), ref64)
            print(f"traj step 20 {k}: HIP-vs-fp64 {mine:.2e}  ref32-vs-fp64 {floor:.2e}")
            assert mine <= NOISE_X * floor + 1e-6, (k, mine, floor)
    print({f"{a}:{b}": f"{v:.1e}" for (a, b), v in report.items()})


# ---------------------------------------------------------------- larger scenes vs the oracle
def slab(n_grid=64, cells=(32, 8, 32), origin=(16, 8, 16), ppc=8, seed=0, dtype=torch.float32):
    from peel_mpm_amd import scenes
    return scenes.slab_block(n_grid=n_grid, cells=cells, origin=origin, ppc=ppc, seed=seed, dtype=dtype)


def test_cfg0_scene_five_steps_vs_oracle():
    """The 64^3 peel block (65,536 particles), 5 steps, each compared on its own
    input: the fp32 oracle and the fp64 oracle both start from the HIP state."""
    from peel_mpm_amd import AnisoMPM, scenes
    spec = scenes.config0()
    gpu = scenes.make_solver(AnisoMPM, spec, device=DEV)
    o32 = scenes.make_solver(O.OracleMPM, spec, device="cpu")
    spec64 = dict(spec, dtype=torch.float64)
    o64 = scenes.make_solver(O.OracleMPM, spec64, device="cpu")
    for s in range(5):
        for k in STATE_KEYS:
            src = getattr(gpu, k).detach().cpu()
            setattr(o32, k, src.clone())
            setattr(o64, k, src.double())
        o32._step = o64._step = gpu._step
        t = s * spec["dt"]
        gpu.substep(t)
        o32.substep(t)
        o64.substep(t)
        torch.cuda.synchronize()
        line = {}
        for k in STATE_KEYS + DIAG_KEYS:
            if not hasattr(o32, k):
                continue
            mine, ref32, ref64 = getattr(gpu, k), getattr(o32, k), getattr(o64, k)
            e32 = rel_err(mine, ref32)
            floor = rel_err(ref32, ref64)
            e64 = rel_err(mine, ref64)
            line[k] = f"{e32:.1e}/{e64:.1e}/{floor:.1e}"
            if k in SENSITIVE:
                assert e64 <= NOISE_X * floor + 1e-6, (s, k, e64, floor)
                assert e32 <= max(1e-5, (NOISE_X + 1) * floor), (s, k, e32, floor)
            else:
                assert e32 <= TOL_STATE[torch.float32], (s, k, e32)
        print(f"cfg0 step {s} (vs ref32 / vs ref64 / ref32 floor):", line, flush=True)


def test_rest_and_free_fall_at_scale():
    from peel_mpm_amd import AnisoMPM, scenes
    spec = scenes.config0()
    spec["bc"] = []
    sim = scenes.make_solver(AnisoMPM, spec, device=DEV)
    for s in range(10):
        sim.substep(s * spec["dt"])
    torch.cuda.synchronize()
    vy = -9.8 * 10 * spec["dt"]
    # free fall: v = -g t; sideways velocity and strain stay at fp32 rounding level
    # (sum_i w_i (x_i - x_p) = 0 holds only to rounding, so C and then F pick up ~1e-7)
    assert torch.allclose(sim.v[:, 1], torch.full_like(sim.v[:, 1], vy), rtol=1e-5, atol=1e-6)
    assert float(sim.v[:, 0].abs().max()) < 1e-5 and float(sim.v[:, 2].abs().max()) < 1e-5
    assert float((sim.F - torch.eye(3, device=DEV)).abs().max()) < 1e-5
    assert float(sim.d.abs().max()) == 0.0


def test_free_fall_at_config2_scale():
    """The 8M-particle, 256^3 two-slab scene with its boundary conditions and
    damage off: free fall -- a size-independent check of binning, P2G, grid
    and G2P at full size.  Total momentum is conserved to rounding (mean
    v_y = -g t to 1e-5); per particle, v_y = -g t to the fp32 noise of the
    affine transfer: C = 4/dx^2 sum w v (x_i - x_p) vanishes only to
    rounding, ~ulp * 4/dx (1024 at this grid), and feeds back ~C dx into v,
    so 1e-6 absolute (measured 1.6e-7); F stays I, d stays 0."""
    from peel_mpm_amd import AnisoMPM, scenes
    spec = scenes.config2()
    spec["bc"] = []
    spec["arrays"]["allow"][:] = False
    sim = scenes.make_solver(AnisoMPM, spec, device=DEV)
    steps = 3
    for s in range(steps):
        sim.substep(s * spec["dt"])
    torch.cuda.synchronize()
    vy = -9.8 * steps * spec["dt"]
    err = (sim.v[:, 1] - vy).abs()
    worst = int(err.argmax())
    print("config2:", sim.n, "particles; max |v_y - g t|", float(err.max()), "over 1e-7:", int((err > 1e-7).sum()),
          "worst x", sim.x[worst].tolist(), "v", sim.v[worst].tolist(), "C", sim.C[worst].flatten().tolist(), flush=True)
    m = sim.mass.double()
    mean_vy = float((m * sim.v[:, 1].double()).sum() / m.sum())
    assert abs(mean_vy - vy) <= 1e-5 * abs(vy), (mean_vy, vy)
    assert float(err.max()) <= 1e-6
    assert float(sim.v[:, 0].abs().max()) < 1e-6 and float(sim.v[:, 2].abs().max()) < 1e-6
    assert float((sim.F - torch.eye(3, device=DEV)).abs().max()) < 1e-5
    assert float(sim.d.abs().max()) == 0.0
    q, c = sim.last_step_stats()
    assert q > 0 and c > 0


def test_config2_step_with_bcs_and_damage_vs_oracle():
    """config2 as specified (8.1 M particles, 256^3, the bottom layer fixed,
    the top pulled at 1 m/s, sigma_c 50 at the interface layers and 200 in
    the bulk, damage on): 30 steps on the device load the interface, then one
    step of the device against the fp32 and fp64 oracles from the same
    state, judged as the config0 steps are (kinematic fields within 1e-5 of
    the fp32 oracle; the damage-side fields within 3x the fp32 oracle's own
    distance from the fp64 one)."""
    from peel_mpm_amd import AnisoMPM, scenes
    spec = scenes.config2()
    gpu = scenes.make_solver(AnisoMPM, spec, device=DEV)
    for s in range(30):
        gpu.substep(s * spec["dt"])
    torch.cuda.synchronize()
    print(f"config2: {gpu.n} particles; after 30 steps max D_hist {float(gpu.D_hist.max()):.3e}, "
          f"max d {float(gpu.d.max()):.3e}", flush=True)
    assert float(gpu.D_hist.max()) > 0.0
    o32 = scenes.make_solver(O.OracleMPM, spec, device="cpu")
    o64 = scenes.make_solver(O.OracleMPM, dict(spec, dtype=torch.float64), device="cpu")
    for k in STATE_KEYS:
        src = getattr(gpu, k).detach().cpu()
        setattr(o32, k, src.clone())
        setattr(o64, k, src.double())
    o32._step = o64._step = gpu._step
    t = 30 * spec["dt"]
    gpu.substep(t)
    o32.substep(t)
    o64.substep(t)
    torch.cuda.synchronize()
    line = {}
    for k in STATE_KEYS + DIAG_KEYS:
        mine, ref32, ref64 = getattr(gpu, k), getattr(o32, k), getattr(o64, k)
        e32, floor, e64 = rel_err(mine, ref32), rel_err(ref32, ref64), rel_err(mine, ref64)
        line[k] = f"{e32:.1e}/{e64:.1e}/{floor:.1e}"
        if k in SENSITIVE:
            assert e64 <= NOISE_X * floor + 1e-6, (k, e64, floor)
            assert e32 <= max(1e-5, (NOISE_X + 1) * floor), (k, e32, floor)
        else:
            assert e32 <= TOL_STATE[torch.float32], (k, e32)
    print("config2 step 31 (vs ref32 / vs ref64 / ref32 floor):", line, flush=True)
