"""CPU checks of the C boundary: the library loads, exports every entry point
include/pmpm.h declares, and the ctypes structs match the C layout."""
from __future__ import annotations

import ctypes as C
import os
import re

import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "pmpm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pmpm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_step_entry_points():
    names = declared_functions()
    for must in ("pmpm_substep", "pmpm_bin", "pmpm_sym_eig", "pmpm_workspace_bytes", "pmpm_last_error"):
        assert must in names


def test_library_loads_and_exports_every_declared_symbol():
    from peel_mpm_amd import _lib
    L = _lib.lib()                      # also checks the struct sizes against the library's
    for name in declared_functions():
        assert hasattr(L, name), name
    assert set(_lib.EXPORTS) <= set(declared_functions())
    assert L.pmpm_version().startswith(b"pmpm")


def test_workspace_size_and_argument_errors():
    from peel_mpm_amd import _lib
    L = _lib.lib()
    small = L.pmpm_workspace_bytes(1000, 32, 1, _lib.PMPM_F32)
    big = L.pmpm_workspace_bytes(1 << 20, 128, 1, _lib.PMPM_F32)
    assert 0 < small < big
    assert L.pmpm_workspace_bytes(1000, 32, 1, _lib.PMPM_F64) > small
    assert L.pmpm_workspace_bytes(1000, 32, 8, _lib.PMPM_F32) > small
    # argument validation fails before touching a device
    p, pin, pout = _lib.StepParams(), _lib.ParticlesIn(), _lib.ParticlesOut()
    p.dtype, p.n_grid = 7, 32
    rc = L.pmpm_substep(C.byref(p), C.byref(pin), C.byref(pout), None, 0, None)
    assert rc == -1 and b"dtype" in L.pmpm_last_error()
    p.dtype = _lib.PMPM_F32
    pin.n = 10
    rc = L.pmpm_substep(C.byref(p), C.byref(pin), C.byref(pout), None, 0, None)
    assert rc == -3 and b"workspace" in L.pmpm_last_error()
    p.n_colliders = 99
    rc = L.pmpm_substep(C.byref(p), C.byref(pin), C.byref(pout), None, 1 << 30, None)
    assert rc == -1 and b"collider" in L.pmpm_last_error()


def test_workspace_not_set_up_is_refused_before_any_launch():
    """A workspace pmpm_workspace_init never set up is refused by every call
    that takes one, on the host, before anything is enqueued (the address
    below is never dereferenced)."""
    from peel_mpm_amd import _lib
    L = _lib.lib()
    fake = 0x7F0000000000
    p, pin, pout = _lib.StepParams(), _lib.ParticlesIn(), _lib.ParticlesOut()
    p.dtype, p.n_grid = _lib.PMPM_F32, 32
    pin.n = 0
    for fn in (L.pmpm_substep, L.pmpm_substep_begin, L.pmpm_substep_end):
        rc = fn(C.byref(p), C.byref(pin), C.byref(pout), fake, 1 << 30, None)
        assert rc == -1 and b"pmpm_workspace_init" in L.pmpm_last_error()
    rc = L.pmpm_bin(None, _lib.PMPM_F32, 0, 32, 32.0, 1, None, None, None, fake, 1 << 30, None)
    assert rc == -1 and b"pmpm_workspace_init" in L.pmpm_last_error()
    q, c = C.c_int64(), C.c_int64()
    rc = L.pmpm_last_step_stats(fake, 0, 32, 1, _lib.PMPM_F32, C.byref(q), C.byref(c), None)
    assert rc == -1 and b"pmpm_workspace_init" in L.pmpm_last_error()
    # a shape the workspace layout cannot take is refused by init itself
    assert L.pmpm_workspace_init(fake, 1 << 30, 100, 2, 1, _lib.PMPM_F32, None) == -1
    assert L.pmpm_workspace_init(fake, 1 << 30, -1, 32, 1, _lib.PMPM_F32, None) == -1


def test_diagnostic_outputs_are_skipped_in_pairs():
    """tau_last/J_last, and phi/phi_shear when damage runs, may be NULL only
    together; a half-skipped pair is refused on the host (every case below
    fails there, so the fake addresses are never dereferenced)."""
    from peel_mpm_amd import _lib
    L = _lib.lib()
    fake = 0x7F0000000000
    p, pin, pout = _lib.StepParams(), _lib.ParticlesIn(), _lib.ParticlesOut()
    p.dtype, p.n_grid = _lib.PMPM_F32, 32
    pin.n = 10
    for name in _lib._IN_PTRS:
        if name != "bc_id":
            setattr(pin, name, fake)
    for name, _ in _lib.ParticlesOut._fields_:
        setattr(pout, name, fake + 4096)
    pout.J_last = None
    rc = L.pmpm_substep(C.byref(p), C.byref(pin), C.byref(pout), None, 1 << 40, None)
    assert rc == -1 and b"tau_last and J_last" in L.pmpm_last_error()
    pout.tau_last, pout.J_last = None, fake + 4096
    rc = L.pmpm_substep(C.byref(p), C.byref(pin), C.byref(pout), None, 1 << 40, None)
    assert rc == -1 and b"tau_last and J_last" in L.pmpm_last_error()
    pout.J_last = None
    p.do_damage = 1
    pout.phi_shear = None
    rc = L.pmpm_substep(C.byref(p), C.byref(pin), C.byref(pout), None, 1 << 40, None)
    assert rc == -1 and b"phi and phi_shear" in L.pmpm_last_error()


def test_product_refuses_cpu_device():
    import torch
    from peel_mpm_amd import AnisoMPM
    with pytest.raises(ValueError):
        AnisoMPM(n_grid=16, device="cpu", dtype=torch.float32)


def test_launch_config_round_trip_and_validation():
    """The autotuner's knobs (host-side library state, no kernel runs)."""
    from peel_mpm_amd import autotune
    autotune.set_config(None)
    assert autotune.get_config() == autotune.DEFAULT
    cfg = {"bucket_wgs": 1024, "block_wgs": 512, "p2g_pair": 1, "const_threads": 128, "p2g_waves": 4, "g2p_threads": 512,
           "cell_block": 8, "p2g_split": 2, "const_stream": 1}
    autotune.set_config(cfg)
    assert autotune.get_config() == cfg
    for bad in ({"bucket_wgs": 1001}, {"bucket_wgs": 32768}, {"block_wgs": 4096}, {"p2g_pair": 3},
                {"const_threads": 512}, {"p2g_waves": 3}, {"g2p_threads": 64}, {"cell_block": 6},
                {"p2g_split": 4}, {"const_stream": 3}):
        with pytest.raises(RuntimeError):
            autotune.set_config(bad)
    # a rejected setting leaves the previous one in place
    assert autotune.get_config()["bucket_wgs"] == 1024
    autotune.set_config(None)
    assert autotune.get_config() == autotune.DEFAULT
