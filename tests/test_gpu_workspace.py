"""The caller's workspace: one set up for max_n particles serves any smaller
particle count at the same results, and a call whose shape the workspace was
not set up for is refused on the host before anything is enqueued."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _spec():
    from peel_mpm_amd import scenes
    return scenes.slab_block(n_grid=32, cells=(12, 4, 12), origin=(10, 8, 10), ppc=8, seed=0, steps=3)


def _ws(n, n_grid, dtype=torch.float32, n_scenes=1, bytes_for=None):
    """A workspace set up for (n, n_grid, n_scenes, dtype), holding the bytes
    of bytes_for particles (default n)."""
    from peel_mpm_amd import _lib
    L = _lib.lib()
    code = _lib.dtype_code(dtype)
    nbytes = L.pmpm_workspace_bytes(max(n, bytes_for or 0), n_grid, n_scenes, code)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=DEV)
    _lib.check(L.pmpm_workspace_init(ws.data_ptr(), nbytes, n, n_grid, n_scenes, code,
                                     _lib.stream_ptr(torch.device(DEV))), "pmpm_workspace_init")
    return ws


def test_larger_workspace_gives_identical_steps():
    """Steps on a workspace set up for 3x the particles (earlier steps of a
    bigger scene leave it dirty) equal steps on an exact-size one, bit for
    bit."""
    from peel_mpm_amd import AnisoMPM, scenes
    spec = _spec()
    a = scenes.make_solver(AnisoMPM, spec, device=DEV)
    b = scenes.make_solver(AnisoMPM, spec, device=DEV)
    big = scenes.make_solver(AnisoMPM, scenes.slab_block(n_grid=32, cells=(16, 6, 16), origin=(8, 6, 8), ppc=8,
                                                         seed=1), device=DEV)
    assert big.n > 2 * a.n
    ws = _ws(big.n, 32)
    big._ws, big._ws_n = ws, big.n
    for s in range(2):
        big.substep(s * big.dt)
    b._ws, b._ws_n = ws, b.n            # b now steps on the big, dirty workspace
    for s in range(3):
        a.substep(s * a.dt)
        b.substep(s * b.dt)
        torch.cuda.synchronize()
        for k in ("x", "v", "F", "C", "d", "D_hist", "lap", "tau_last", "J_last"):
            assert torch.equal(getattr(a, k), getattr(b, k)), f"step {s}: {k} differs"


def test_bin_on_a_larger_workspace_is_the_stable_argsort():
    """pmpm_bin with fewer particles than its workspace was set up for (the
    advisor's case: the arena and later sections used to move with n)."""
    from peel_mpm_amd import kernels
    g = np.random.default_rng(3)
    ws = kernels.bin_workspace(50_000, 48, device=DEV)
    for n in (50_000, 7_001, 123, 20_000):
        x = torch.from_numpy(g.uniform(0.05, 0.95, (n, 3))).float().to(DEV)
        key, perm, start = kernels.bin_particles(x, 48, ws=ws)
        k = key.cpu().numpy()
        assert np.array_equal(perm.cpu().numpy(), np.argsort(k, kind="stable")), n
        st = start.cpu().numpy()
        assert st[-1] == n and np.all(np.diff(st) >= 0)
        assert np.array_equal(np.bincount(k, minlength=st.size - 1), np.diff(st))


def test_mismatched_shapes_are_refused():
    from peel_mpm_amd import AnisoMPM, _lib, kernels, scenes
    spec = _spec()
    sim = scenes.make_solver(AnisoMPM, spec, device=DEV)
    # set up for too few particles (the buffer itself is big enough)
    sim._ws, sim._ws_n = _ws(sim.n // 2, 32, bytes_for=sim.n), sim.n
    with pytest.raises(RuntimeError, match="at most"):
        sim.substep(0.0)
    # another grid
    sim._ws, sim._ws_n = _ws(sim.n, 48), sim.n
    with pytest.raises(RuntimeError, match="another grid"):
        sim.substep(0.0)
    # another dtype (fp64 layout for an fp32 step)
    sim._ws, sim._ws_n = _ws(sim.n, 32, torch.float64), sim.n
    with pytest.raises(RuntimeError, match="another grid"):
        sim.substep(0.0)
    # another scene count
    sim._ws, sim._ws_n = _ws(sim.n, 32, n_scenes=2), sim.n
    with pytest.raises(RuntimeError, match="another grid"):
        sim.substep(0.0)
    x = torch.rand(100, 3, device=DEV)
    with pytest.raises(RuntimeError, match="workspace"):
        kernels.bin_particles(x, 32, ws=kernels.bin_workspace(50, 32, device=DEV))  # ESIZE: sized by n
    # nothing ran: the solver still steps once given its own workspace
    sim._ws = None
    sim.substep(0.0)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(sim.x).all())
    q, c = C.c_int64(), C.c_int64()
    L = _lib.lib()
    assert L.pmpm_last_step_stats(sim._ws.data_ptr(), sim.n, 48, 1, 0, C.byref(q), C.byref(c), None) == -1
