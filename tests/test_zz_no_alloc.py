"""pmpm_substep_noalloc (the step on the library's own scratch, no
allocation) against pmpm_substep on a caller workspace, through the C ABI:
bit for bit on consecutive calls, the second with a smaller scene on a
smaller grid (the scratch then holds the first call's contents at other
offsets), in fp32 and fp64 and in the PCG damage mode; a step too large for
the scratch is refused before anything is enqueued."""
from __future__ import annotations

import ctypes as C

import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
KEYS = ("x", "v", "F", "C", "d", "D_hist", "lap", "phi", "phi_shear", "tau_last", "J_last")


def _pair(spec, **kw):
    from peel_mpm_amd import AnisoMPM, scenes
    return scenes.make_solver(AnisoMPM, spec, device=DEV, **kw), scenes.make_solver(AnisoMPM, spec, device=DEV, **kw)


def _noalloc_step(sim, time):
    """One substep of sim through pmpm_substep_noalloc (the solver's own
    argument preparation, then the no-allocation entry point)."""
    from peel_mpm_amd import _lib
    L = _lib.lib()
    p, pin, pout, new, device_cols = sim._prepare(time)
    assert device_cols
    rc = L.pmpm_substep_noalloc(C.byref(p), C.byref(pin), C.byref(pout), _lib.stream_ptr(sim.device))
    _lib.check(rc, "pmpm_substep_noalloc")
    for k, t in new.items():
        setattr(sim, k, t)
    sim._step += 1


def _check(a, b, label):
    torch.cuda.synchronize()
    for k in KEYS:
        if hasattr(a, k):
            assert torch.equal(getattr(a, k), getattr(b, k)), f"{label}: {k} differs"


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_matches_the_workspace_step_on_consecutive_calls(dtype):
    from peel_mpm_amd import scenes
    ng = 64 if dtype == torch.float32 else 48        # fp64 on 64^3 needs more than the 64 MiB scratch
    big = scenes.slab_block(n_grid=ng, cells=(20, 6, 20), origin=(14, 10, 14), ppc=8, seed=0, dtype=dtype)
    small = scenes.slab_block(n_grid=32, cells=(8, 4, 8), origin=(10, 8, 10), ppc=8, seed=1, dtype=dtype)
    print(f"noalloc {dtype}: {big['arrays']['x'].shape[0]} then {small['arrays']['x'].shape[0]} particles", flush=True)
    ref_big, na_big = _pair(big)
    ref_small, na_small = _pair(small)
    for s in range(3):
        t = s * big["dt"]
        ref_big.substep(t)
        _noalloc_step(na_big, t)
        _check(ref_big, na_big, f"big step {s}")
        ref_small.substep(t)
        _noalloc_step(na_small, t)
        _check(ref_small, na_small, f"small step {s}")


def test_noalloc_pcg_damage_mode():
    from peel_mpm_amd import scenes
    spec = scenes.slab_block(n_grid=32, cells=(10, 4, 10), origin=(10, 8, 10), ppc=8, seed=2)
    print("noalloc pcg", flush=True)
    ref, na = _pair(spec, damage_solver="pcg", pcg_tol=1e-6, pcg_max_iter=20)
    for s in range(3):
        t = s * spec["dt"]
        ref.substep(t)
        _noalloc_step(na, t)
        _check(ref, na, f"pcg step {s}")


def test_noalloc_refuses_a_step_beyond_the_scratch():
    from peel_mpm_amd import _lib, scenes
    L = _lib.lib()
    spec = scenes.slab_block(n_grid=32, cells=(4, 4, 4), origin=(10, 8, 10), ppc=8, seed=0)
    a, b = _pair(spec)
    a.n_grid = 256          # a 256^3 grid alone needs far more than 64 MiB of workspace
    p, pin, pout, new, _ = a._prepare(0.0)
    assert L.pmpm_workspace_bytes(a.n, 256, 1, 0) > _lib.NOALLOC_SCRATCH_BYTES
    rc = L.pmpm_substep_noalloc(C.byref(p), C.byref(pin), C.byref(pout), _lib.stream_ptr(a.device))
    assert rc == -3 and b"scratch" in L.pmpm_last_error()
    p.keep_for_adjoint = 1
    p.n_grid = 32
    rc = L.pmpm_substep_noalloc(C.byref(p), C.byref(pin), C.byref(pout), _lib.stream_ptr(a.device))
    assert rc == -1 and b"keep_for_adjoint" in L.pmpm_last_error()
    # nothing was enqueued by the refused calls: the scratch still serves a step
    a.n_grid = 32
    a._p_key = None
    b.substep(0.0)
    _noalloc_step(a, 0.0)
    _check(b, a, "after refusals")
