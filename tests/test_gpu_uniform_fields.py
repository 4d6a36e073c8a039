"""Uniform material fields (needs a GPU): a field that holds one value, bit
for bit, on every particle goes to the step as that value instead of an
array (add_object fills each field from one scalar).  The step computes
the same bits either way: compared with the array path on a scene whose
every field is uniform, on one with per-particle moduli and fibres, in the
PCG damage mode, in fp64, through graph frames with an in-place edit that
breaks a field's uniformity, and with the arrays of the uniform fields
overwritten by NaN (so a kernel that read them would show it)."""
from __future__ import annotations

import ctypes as C

import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
KEYS = ("x", "v", "F", "C", "d", "lap", "D_hist", "tau_last", "J_last", "phi", "phi_shear")


def _spec(dtype=torch.float32, varied=False, seed=3):
    import numpy as np
    from peel_mpm_amd import scenes
    spec = scenes.slab_block(n_grid=32, cells=(12, 4, 10), origin=(10, 8, 11), ppc=8, seed=seed, sigma_c=20.0,
                             dtype=dtype)
    if varied:
        a = spec["arrays"]
        rng = np.random.default_rng(seed)
        n = a["x"].shape[0]
        E = 1e5 * np.exp(0.2 * rng.standard_normal(n))
        a["mu"], a["lam"] = E / 2.6, E * 0.3 / (1.3 * 0.4)
        f = scenes.vmf_fibers(n, (1.0, 0.0, 0.0), 20.0, rng)
        a["A0"] = np.eye(3)[None] - f[:, :, None] * f[:, None, :]
    return spec


def _pair(spec, **kw):
    from peel_mpm_amd import AnisoMPM, scenes
    a = scenes.make_solver(AnisoMPM, spec, device=DEV, **kw)
    b = scenes.make_solver(AnisoMPM, spec, device=DEV, **kw)
    b._uniform_fields = lambda: (0, {})          # b: every field read as an array
    return a, b


def _same(a, b, label):
    torch.cuda.synchronize()
    for k in KEYS:
        assert torch.equal(getattr(a, k), getattr(b, k)), f"{label}: {k}"


def _mask(names):
    from peel_mpm_amd import _lib
    return sum(1 << _lib.UNIFORM_FIELDS.index(k) for k in names)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_uniform_fields_match_the_arrays_bit_for_bit(dtype):
    from peel_mpm_amd import _lib
    print(f"uniform fields, every field uniform, {dtype}", flush=True)
    a, b = _pair(_spec(dtype))
    for s in range(8):
        a.substep(s * a.dt)
        b.substep(s * b.dt)
    _same(a, b, "uniform scene")
    assert a._uniform_fields()[0] == _mask(_lib.UNIFORM_FIELDS)
    assert float(a.D_hist.max()) > 0.0


def test_per_particle_moduli_and_fibres_stay_arrays():
    from peel_mpm_amd import _lib
    print("uniform fields, varied moduli and fibres", flush=True)
    a, b = _pair(_spec(varied=True))
    assert a._uniform_fields()[0] == _mask(set(_lib.UNIFORM_FIELDS) - {"mu", "lam", "A0"})
    for s in range(8):
        a.substep(s * a.dt)
        b.substep(s * b.dt)
    _same(a, b, "varied scene")


def test_uniform_fields_in_the_pcg_damage_mode():
    print("uniform fields, PCG damage", flush=True)
    a, b = _pair(_spec(varied=True), damage_solver="pcg")
    for s in range(5):
        a.substep(s * a.dt)
        b.substep(s * b.dt)
    _same(a, b, "pcg")


def test_graph_frames_recapture_when_an_edit_breaks_uniformity():
    from peel_mpm_amd import _lib
    print("uniform fields, graph frames with an in-place edit", flush=True)
    a, b = _pair(_spec(), use_graph=True)
    b.use_graph = False
    t = 0.0
    for f in range(3):
        if f == 1:
            for s in (a, b):
                s.sigma_c[:60] = 5.0               # in place: sigma_c is no longer uniform
        a.run_frame(4, t)
        b.run_frame(4, t)
        t += 4 * a.dt
        _same(a, b, f"frame {f}")
    assert not a._uniform_fields()[0] & _mask(["sigma_c"])
    assert a._uniform_fields()[0] & _mask(["eta", "zeta", "l0"])


def test_the_step_reads_the_values_not_the_arrays():
    """The arrays of the uniform fields replaced by NaN for one step: the
    result is the array path's, so no kernel read them."""
    from peel_mpm_amd import _lib
    print("uniform fields, NaN arrays", flush=True)
    a, b = _pair(_spec())
    p, pin, pout, new, _ = a._prepare(0.0)
    assert pin.uniform_mask == _mask(_lib.UNIFORM_FIELDS)
    junk = {k: torch.full_like(getattr(a, k), float("nan")) for k in _lib.UNIFORM_FIELDS}
    for k, t in junk.items():
        setattr(pin, k, t.data_ptr())
    ws = a._workspace()
    _lib.check(_lib.lib().pmpm_substep(C.byref(p), C.byref(pin), C.byref(pout), ws.data_ptr(), ws.numel(),
                                       _lib.stream_ptr(a.device)), "pmpm_substep")
    for k, t in new.items():
        setattr(a, k, t)
    b.substep(0.0)
    _same(a, b, "NaN arrays")


def test_batched_scenes_take_the_fields_uniform_across_the_batch():
    """A batch of scenes sharing every field but the moduli: the uniform
    fields go as values, graph-replayed, equal to the array path."""
    from peel_mpm_amd import VelocityBC, _lib, scenes
    from peel_mpm_amd.batched import BatchedAnisoMPM
    print("uniform fields, batched", flush=True)

    def build(arrays_only):
        b = BatchedAnisoMPM(n_grid=32, dt=1e-4, device=DEV, use_graph=True)
        for s in range(3):
            sp = scenes.slab_block(n_grid=32, cells=(6 + 2 * s, 4, 8), origin=(8, 8 + s, 9), ppc=8, seed=s,
                                   E=1e5 * (1 + 0.3 * s), sigma_c=20.0)
            b.add_scene(scenes.tensors(sp, DEV), [VelocityBC(torch.from_numpy(m).to(DEV), tuple(v)) for m, v in sp["bc"]])
        if arrays_only:
            b._uf.state = lambda owner: (0, {})
        return b

    a, b = build(False), build(True)
    for s in range(5):
        a.substep(s * 1e-4)
        b.substep(s * 1e-4)
    torch.cuda.synchronize()
    for k in ("x", "v", "F", "C", "d", "lap", "D_hist"):
        assert torch.equal(getattr(a, k), getattr(b, k)), k
    assert a._uf.state(a)[0] == _mask(set(_lib.UNIFORM_FIELDS) - {"mu", "lam"})
