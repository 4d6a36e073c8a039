"""Binning carried from one step to the next (needs a GPU): G2P counts the
next step's cells from the positions it writes, and the next step skips its
count pass.  With the stable sort the binning is the same permutation either
way, so a carried run equals a run that counts every step afresh bit for
bit: with a moving collider and a pulled layer (particles cross cells), with
particles past the grid's faces, in the PCG damage mode, through run_frame
with in-place edits of x between frames, and when a claim the library cannot
verify is refused."""
from __future__ import annotations

import ctypes as C

import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
KEYS = ("x", "v", "F", "C", "d", "lap", "D_hist", "tau_last", "J_last", "phi", "phi_shear")


def _sim(dtype=torch.float32, **kw):
    from peel_mpm_amd import AnisoMPM, VelocityBC, scenes, sphere_collider
    spec = scenes.slab_block(n_grid=32, cells=(12, 4, 10), origin=(10, 8, 11), ppc=8, seed=4, sigma_c=20.0, dtype=dtype)
    centre = lambda t: (0.50 + 30.0 * t, 0.43 - 20.0 * t, 0.5)
    pull = lambda t: (0.0, 0.5 + 1500.0 * t, 0.05)
    sim = scenes.make_solver(AnisoMPM, spec, device=DEV, **kw)
    sim.particle_bc[0] = VelocityBC(sim.particle_bc[0].mask, pull)
    sim.colliders.append(sphere_collider(centre, 0.07, mode="slip", friction=0.2))
    return sim


def _steps(sim, k, carry, t0=0.0):
    for s in range(k):
        if not carry:
            sim._carry = None               # no claim: the step counts afresh
        sim.substep(t0 + s * sim.dt)


def _same(a, b, label):
    torch.cuda.synchronize()
    for k in KEYS:
        assert torch.equal(getattr(a, k), getattr(b, k)), f"{label}: {k}"


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_carried_binning_matches_counting_afresh(dtype):
    print(f"carried vs fresh binning {dtype}", flush=True)
    a, b = _sim(dtype), _sim(dtype)
    for chunk in range(3):
        _steps(a, 10, True, chunk * 10 * a.dt)
        _steps(b, 10, False, chunk * 10 * b.dt)
        _same(a, b, f"after {10 * (chunk + 1)} steps")
    assert a.carried_steps() == 29 and b.carried_steps() == 0
    # the particles moved across cells: the binning did change along the run
    assert float((torch.floor(a.x * a.inv_dx) != torch.floor(_sim(dtype).x * a.inv_dx)).any(1).float().sum()) > 0


def test_carried_binning_through_frames_and_in_place_edits():
    print("carried binning through run_frame", flush=True)
    a, b = _sim(), _sim()
    t = 0.0
    for f, n_sub in enumerate((4, 5, 3, 6)):
        if f in (1, 3):
            for s in (a, b):
                s.x[:40] += 0.3 * s.dx      # in place: the next step must count afresh
        a.run_frame(n_sub, t)
        for s in range(n_sub):              # b: the same substeps, counting every step
            b._carry = None
            b.substep(t + s * b.dt, _diag=s == n_sub - 1, _phi=s == n_sub - 1)
        t += n_sub * a.dt
        _same(a, b, f"frame {f}")
    # 18 substeps; the first and the two after the edits count afresh
    assert a.carried_steps() == 15, a.carried_steps()


def test_carried_binning_with_particles_past_the_grid():
    """The reference's edge scene (stencils past the domain faces): the
    off-grid flag travels with the carried count."""
    from tests.golden_cases import load_npz
    from tests.test_gpu_parity import gpu_sim_from_case
    rec = load_npz("step_edges32.npz")
    print("carried binning, edge scene", flush=True)
    (a, time0, _), (b, _, _) = gpu_sim_from_case(rec), gpu_sim_from_case(rec)
    for s in range(8):
        for sim, carry in ((a, True), (b, False)):
            if not carry:
                sim._carry = None
            sim.substep(time0 + s * sim.dt)
    _same(a, b, "edge scene")
    assert a.carried_steps() == 7


def test_carried_binning_in_the_pcg_damage_mode():
    print("carried binning, PCG damage", flush=True)
    a, b = _sim(damage_solver="pcg"), _sim(damage_solver="pcg")
    _steps(a, 6, True)
    _steps(b, 6, False)
    _same(a, b, "pcg")
    assert a.carried_steps() == 5


def _abi_step(sim, time, prebinned, x=None):
    """One pmpm_substep through the C ABI with an explicit claim (and
    optionally another x tensor of the same values)."""
    from peel_mpm_amd import _lib
    p, pin, pout, new, _ = sim._prepare(time)
    p.prebinned = prebinned
    if x is not None:
        pin.x = x.data_ptr()
    ws = sim._workspace()
    _lib.check(_lib.lib().pmpm_substep(C.byref(p), C.byref(pin), C.byref(pout), ws.data_ptr(), ws.numel(),
                                       _lib.stream_ptr(sim.device)), "pmpm_substep")
    for k, t in new.items():
        setattr(sim, k, t)
    sim._step += 1


def test_claims_the_library_cannot_verify_are_refused():
    print("refused carry claims", flush=True)
    a, b = _sim(), _sim()
    # a claim on a workspace that carries nothing
    _abi_step(a, 0.0, 1)
    _abi_step(b, 0.0, 0)
    _same(a, b, "claim before any carried count")
    assert a.carried_steps() == 0
    # a claim with another tensor of the same positions: counted afresh
    xa, xb = a.x.clone(), b.x.clone()
    _abi_step(a, a.dt, 1, xa)
    _abi_step(b, b.dt, 0, xb)
    _same(a, b, "claim with another x tensor")
    assert a.carried_steps() == 0
    # the honest claim is taken up and gives the same step
    _abi_step(a, 2 * a.dt, 1)
    _abi_step(b, 2 * b.dt, 0)
    _same(a, b, "honest claim")
    assert a.carried_steps() == 1
