"""ctypes binding of libpmpm.so (declared in include/pmpm.h).

This is the FFI layer a Python caller of the reference step would add: the
reference's step is the Python method ``AnisoMPM.substep``
(anisompm_py/src/anisompm.py:593-803) and has no native boundary of its own.
Loading fails loudly when the library is missing: there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpmpm.so")

PMPM_F32, PMPM_F64 = 0, 1
NOALLOC_SCRATCH_BYTES = 64 << 20
MAX_COLLIDERS, MAX_VEL_BC = 8, 8
SPHERE, HALFSPACE = 1, 2
STICKY, SLIP = 0, 1
DAMAGE_EXPLICIT, DAMAGE_PCG = 0, 1

EXPORTS = ("pmpm_workspace_bytes", "pmpm_workspace_init", "pmpm_substep", "pmpm_substep_begin",
           "pmpm_substep_end", "pmpm_grid_views", "pmpm_last_step_stats", "pmpm_bin", "pmpm_stencil",
           "pmpm_sym_eig", "pmpm_last_error", "pmpm_version", "pmpm_abi_sizes", "pmpm_timing", "pmpm_timing_read",
           "pmpm_adjoint_workspace_bytes", "pmpm_substep_adjoint", "pmpm_damage_solve_info",
           "pmpm_set_launch_config", "pmpm_get_launch_config", "pmpm_gaussian_carry", "pmpm_observe_workspace_bytes",
           "pmpm_observe_frame", "pmpm_compact_workspace_bytes", "pmpm_ball_particles", "pmpm_grid_block",
           "pmpm_fill_workspace_bytes", "pmpm_fill_bounds", "pmpm_fill_interior", "pmpm_directional_pk1",
           "pmpm_directional_kirchhoff", "pmpm_substep_noalloc")


class Pcg64(C.Structure):
    _fields_ = [("state_hi", C.c_uint64), ("state_lo", C.c_uint64), ("inc_hi", C.c_uint64), ("inc_lo", C.c_uint64)]


class Collider(C.Structure):
    _fields_ = [("kind", C.c_int32), ("mode", C.c_int32), ("friction", C.c_double), ("radius", C.c_double),
                ("p0", C.c_double * 3), ("p1", C.c_double * 3), ("normal", C.c_double * 3)]


class StepParams(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("n_grid", C.c_int32),
                ("grid_lim", C.c_double), ("dx", C.c_double), ("inv_dx", C.c_double), ("dt", C.c_double),
                ("gravity", C.c_double * 3), ("grid_damp", C.c_double),
                ("use_fclamp", C.c_int32), ("fclamp_lo", C.c_double), ("fclamp_hi", C.c_double),
                ("do_damage", C.c_int32), ("dt_damage", C.c_double),
                ("n_colliders", C.c_int32), ("colliders", Collider * MAX_COLLIDERS),
                ("n_vel_bc", C.c_int32), ("vel_bc", (C.c_double * 3) * MAX_VEL_BC),
                ("stable_sort", C.c_int32), ("export_active", C.c_int32), ("exact_sigma_eig", C.c_int32),
                ("n_scenes", C.c_int32), ("scene_offsets", C.c_void_p), ("scene_damage", C.c_void_p),
                ("damage_solver", C.c_int32), ("pcg_tol", C.c_double), ("pcg_max_iter", C.c_int32),
                ("keep_for_adjoint", C.c_int32), ("boundary_table", C.c_void_p),
                ("bin_next", C.c_int32), ("prebinned", C.c_int32)]


class BoundaryTable(C.Structure):
    """pmpm_boundary_table: colliders and BC velocities read by the kernels
    when they run (StepParams.boundary_table holds its device address)."""
    _fields_ = [("colliders", Collider * MAX_COLLIDERS), ("vel_bc", (C.c_double * 3) * MAX_VEL_BC)]


_IN_PTRS = ("x", "v", "F", "C", "d", "lap", "D_hist", "vol0", "mass", "mu", "lam", "sigma_c", "eta", "zeta", "l0",
            "residual", "A0", "allow", "interface_mask", "n_interface", "bc_id")
_OUT_PTRS = ("x", "v", "F", "C", "lap", "d", "D_hist", "phi", "phi_shear", "tau_last", "J_last")


class LaunchConfig(C.Structure):
    _fields_ = [("bucket_wgs", C.c_int32), ("block_wgs", C.c_int32), ("p2g_pair", C.c_int32),
                ("const_threads", C.c_int32), ("p2g_waves", C.c_int32), ("g2p_threads", C.c_int32),
                ("cell_block", C.c_int32), ("p2g_split", C.c_int32), ("const_stream", C.c_int32)]


class FrameIn(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("phi", "phi_shear", "d", "x", "tau_last", "J_last", "region", "cap")]


class FrameObs(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("phi_pp", "psh_pp", "d_e", "phipp_e", "pshpp_e", "x_e", "tau_e", "J_e",
                                          "peaks", "onset", "log")]


class ParticlesIn(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(k, C.c_void_p) for k in _IN_PTRS] + \
        [("uniform_mask", C.c_uint32), ("uniform_value", C.c_double * 9), ("uniform_A0", C.c_double * 9)]


# the material fields a step may take as one value (bit k of uniform_mask)
UNIFORM_FIELDS = ("vol0", "mass", "mu", "lam", "sigma_c", "eta", "zeta", "l0", "residual", "A0")


class ParticlesOut(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in _OUT_PTRS]


_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m peel_mpm_amd.build` "
                          "(the HIP step has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, sz, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_size_t, C.c_double
    L.pmpm_workspace_bytes.restype = sz
    L.pmpm_workspace_bytes.argtypes = [i64, i32, i32, i32]
    L.pmpm_workspace_init.argtypes = [vp, sz, i64, i32, i32, i32, vp]
    for f in (L.pmpm_substep, L.pmpm_substep_begin, L.pmpm_substep_end):
        f.argtypes = [C.POINTER(StepParams), C.POINTER(ParticlesIn), C.POINTER(ParticlesOut), vp, sz, vp]
    L.pmpm_substep_noalloc.argtypes = [C.POINTER(StepParams), C.POINTER(ParticlesIn), C.POINTER(ParticlesOut), vp]
    L.pmpm_grid_views.argtypes = [i64, i32, i32, i32, C.POINTER(sz), C.POINTER(sz)]
    L.pmpm_last_step_stats.argtypes = [vp, i64, i32, i32, i32, C.POINTER(i64), C.POINTER(i64), vp]
    L.pmpm_bin.argtypes = [vp, i32, i64, i32, dbl, i32, vp, vp, vp, vp, sz, vp]
    L.pmpm_stencil.argtypes = [vp, i32, i64, i32, dbl, vp, vp, vp, vp, vp, vp, vp, vp]
    L.pmpm_sym_eig.argtypes = [vp, i32, i64, vp, vp, vp]
    L.pmpm_damage_solve_info.argtypes = [vp, i64, i32, i32, i32, C.POINTER(i32), C.POINTER(dbl), vp]
    L.pmpm_carried_steps.argtypes = [vp, C.POINTER(i64)]
    L.pmpm_set_launch_config.argtypes = [C.POINTER(LaunchConfig)]
    L.pmpm_get_launch_config.argtypes = [C.POINTER(LaunchConfig)]
    L.pmpm_last_error.restype = C.c_char_p
    L.pmpm_version.restype = C.c_char_p
    L.pmpm_abi_sizes.argtypes = [C.POINTER(sz), i32]
    L.pmpm_timing.argtypes = [i32]
    L.pmpm_timing_read.argtypes = [C.POINTER(C.c_double), C.POINTER(i64), i32]
    L.pmpm_observe_workspace_bytes.restype = sz
    L.pmpm_observe_workspace_bytes.argtypes = []
    L.pmpm_observe_frame.argtypes = [i32, i64, i32, C.POINTER(FrameIn), C.POINTER(FrameObs), vp, sz, vp]
    L.pmpm_gaussian_carry.argtypes = [i32, i64, i64, vp, vp, vp, vp, vp, vp, vp, dbl, dbl, C.POINTER(dbl), dbl, vp, vp,
                                      vp, vp]
    L.pmpm_directional_pk1.argtypes = [i32, i64, vp, vp, vp, vp, dbl, vp, dbl, vp, dbl, vp, dbl, vp, vp, vp]
    L.pmpm_directional_kirchhoff.argtypes = [i32, i64, vp, vp, vp, vp, vp, dbl, i32, vp, vp]
    L.pmpm_compact_workspace_bytes.restype = sz
    L.pmpm_compact_workspace_bytes.argtypes = [i64]
    L.pmpm_ball_particles.argtypes = [vp, i64, dbl, dbl, C.POINTER(dbl), C.POINTER(Pcg64), vp, i64, C.POINTER(i64), vp, sz,
                                      vp]
    L.pmpm_grid_block.argtypes = [vp, vp, vp, i64, i64, i64, vp, vp]
    L.pmpm_fill_workspace_bytes.restype = sz
    L.pmpm_fill_workspace_bytes.argtypes = [C.POINTER(i64)]
    L.pmpm_fill_bounds.argtypes = [i32, i32, vp, i64, C.POINTER(dbl), dbl, dbl, C.POINTER(dbl), C.POINTER(i64), vp, sz, vp]
    L.pmpm_fill_interior.argtypes = [i32, i32, vp, i64, C.POINTER(dbl), dbl, dbl, C.POINTER(dbl), C.POINTER(i64), dbl,
                                     C.POINTER(Pcg64), vp, i64, C.POINTER(i64), vp, sz, vp]
    sizes = (sz * 8)()
    L.pmpm_abi_sizes(sizes, 8)
    mine = (C.sizeof(Collider), C.sizeof(StepParams), C.sizeof(ParticlesIn), C.sizeof(ParticlesOut), C.sizeof(FrameIn),
            C.sizeof(FrameObs), C.sizeof(Pcg64), C.sizeof(BoundaryTable))
    if tuple(sizes) != mine:
        raise ImportError(f"libpmpm ABI mismatch: library struct sizes {tuple(sizes)} != binding {mine}")
    _lib = L
    return L


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().pmpm_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def dtype_code(dtype) -> int:
    import torch
    if dtype == torch.float32:
        return PMPM_F32
    if dtype == torch.float64:
        return PMPM_F64
    raise TypeError(f"unsupported dtype {dtype}: the native step runs in float32 or float64")


def stream_ptr(device) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream
