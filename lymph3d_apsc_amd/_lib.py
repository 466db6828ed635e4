"""ctypes binding of liblymph_gpu.so (C ABI declared in include/lymph_gpu.h).

The shared library is built in-tree (lymph3d_apsc_amd/liblymph_gpu.so) by
`__graft_entry__.build()` / `make -C lymph3d_apsc_amd/csrc`.  There is no CPU
fallback: if the library is missing, or no HIP device is visible when a
context is created, the calls raise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblymph_gpu.so")

# exit / error codes (src/Poly_global.f90:130-173 where they exist)
EXIT_INSTAB = 2
EXIT_SETUP = 4
EXIT_SINGULARMTX = 5
EXIT_NEIGHBOUR = 18
ERR_ARG = 101
ERR_HIP = 102
ERR_STATE = 103

LAYOUT_POLY, LAYOUT_FORTRAN = 0, 1
PROB_NULL, PROB_REF_MMS, PROB_CUBE_WAVE, PROB_RICKER = 0, 1, 2, 3
FN_ZERO, FN_REF_VEL, FN_REF_UEX, FN_CUBE_U = 0, 1, 2, 3
(SLOT_RHS_STAT, SLOT_RHS_DYN, SLOT_U_PREV, SLOT_U_CUR, SLOT_U_NEXT, SLOT_V0,
 SLOT_UEX) = range(7)

# every symbol include/lymph_gpu.h declares
EXPORTS = [
    "lymph_last_error", "lymph_version", "lymph_device_count", "lymph_ctx_create",
    "lymph_ctx_destroy", "lymph_ctx_info", "lymph_ctx_set_layout", "lymph_ctx_set_stream",
    "lymph_ctx_get_stream", "lymph_get_pattern", "lymph_build_operator",
    "lymph_export_operator", "lymph_build_rhs", "lymph_project", "lymph_set_slot",
    "lymph_get_slot", "lymph_slot_ptr", "lymph_first_step", "lymph_step",
    "lymph_first_step_dev", "lymph_step_dev", "lymph_step_n", "lymph_time_steps",
    "lymph_error_l2", "lymph_error_dg", "lymph_error_slots", "lymph_run", "lymph_check_finite",
    "lymph_build_rhs_fn", "lymph_project_fn",
    "lymph_time_function", "lymph_autotune", "lymph_set_step_variant", "lymph_set_receivers",
    "lymph_sample_receivers", "lymph_step_n_record", "lymph_get_traces",
    "lymph_set_apply_precision", "lymph_set_source_groups", "lymph_ctx_set_keep_adg", "lymph_ctx_set_build", "lymph_release_fp64_operator", "lymph_set_symmetric_storage",
    "lymph_add_point_sources", "lymph_add_point_sources_noalloc", "lymph_add_double_couples", "lymph_set_ricker", "lymph_ricker",
    "lymph_capture_solve", "lymph_replay", "lymph_sync",
]


class LymphMesh(C.Structure):
    _fields_ = [
        ("n_node", C.c_int32), ("n_tet", C.c_int32), ("n_poly", C.c_int32),
        ("xyz", C.c_void_p), ("tet_vert", C.c_void_p), ("tet_poly", C.c_void_p),
        ("neigh", C.c_void_p), ("bc_space", C.c_void_p), ("normal", C.c_void_p),
        ("area", C.c_void_p), ("bbox", C.c_void_p), ("hk", C.c_void_p),
        ("rho", C.c_void_p), ("lam", C.c_void_p), ("mu", C.c_void_p),
    ]


class LymphError(RuntimeError):
    def __init__(self, code, where, msg):
        super().__init__(f"{where} failed with code {code}: {msg}")
        self.code = code


_LIB = None
_vp, _d, _i = C.c_void_p, C.c_double, C.c_int


def lib():
    """Load liblymph_gpu.so (raises if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " or `make -C lymph3d_apsc_amd/csrc` (the HIP path has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    sig = {
        "lymph_last_error": (C.c_char_p, []),
        "lymph_version": (C.c_char_p, []),
        "lymph_device_count": (_i, []),
        "lymph_ctx_create": (_i, [C.POINTER(LymphMesh), _i, _d, _d, C.POINTER(_vp)]),
        "lymph_ctx_destroy": (_i, [_vp]),
        "lymph_ctx_info": (_i, [_vp, _vp]),
        "lymph_ctx_set_layout": (_i, [_vp, _i]),
        "lymph_ctx_set_stream": (_i, [_vp, _vp]),
        "lymph_ctx_get_stream": (_vp, [_vp]),
        "lymph_get_pattern": (_i, [_vp, _vp, _vp]),
        "lymph_build_operator": (_i, [_vp]),
        "lymph_export_operator": (_i, [_vp, _vp, _vp, _vp, _vp]),
        "lymph_build_rhs": (_i, [_vp, _i, _i, _vp]),
        "lymph_project": (_i, [_vp, _i, _i, _vp]),
        "lymph_set_slot": (_i, [_vp, _i, _vp]),
        "lymph_get_slot": (_i, [_vp, _i, _vp]),
        "lymph_slot_ptr": (_vp, [_vp, _i]),
        "lymph_first_step": (_i, [_vp, _d, _d, _vp, _vp, _vp]),
        "lymph_step": (_i, [_vp, _d, _d, _vp, _vp, _vp]),
        "lymph_first_step_dev": (_i, [_vp, _d, _d, _vp, _vp, _vp]),
        "lymph_step_dev": (_i, [_vp, _d, _d, _vp, _vp, _vp]),
        "lymph_step_n": (_i, [_vp, _i, _i, _d, _d, _vp]),
        "lymph_time_steps": (_i, [_vp, _i, _i, _d, _d, _vp]),
        "lymph_error_l2": (_i, [_vp, _vp, _vp, _vp]),
        "lymph_error_dg": (_i, [_vp, _vp, _vp, _vp]),
        "lymph_error_slots": (_i, [_vp, _i, _i, _vp, _vp]),
        "lymph_check_finite": (_i, [_vp, _i, _vp]),
        "lymph_build_rhs_fn": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _vp]),
        "lymph_project_fn": (_i, [_vp, _vp, _vp, _i, _vp]),
        "lymph_run": (_i, [_vp, _i, _d, _d, _i, _vp, _vp]),
        "lymph_time_function": (_d, [_i, _d]),
        "lymph_autotune": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
        "lymph_set_step_variant": (_i, [_vp, _i, _vp]),
        "lymph_set_receivers": (_i, [_vp, _i, _vp, _vp, _i]),
        "lymph_sample_receivers": (_i, [_vp, _i, _vp]),
        "lymph_step_n_record": (_i, [_vp, _i, _i, _d, _d, _i, _vp]),
        "lymph_get_traces": (_i, [_vp, _vp, _vp]),
        "lymph_set_apply_precision": (_i, [_vp, _i]),
        "lymph_ctx_set_keep_adg": (_i, [_vp, _i]),
        "lymph_ctx_set_build": (_i, [_vp, _i]),
        "lymph_set_source_groups": (_i, [_vp, _i, _vp, _vp, _vp]),
        "lymph_release_fp64_operator": (_i, [_vp]),
        "lymph_set_symmetric_storage": (_i, [_vp, _i, _vp]),
        "lymph_add_point_sources": (_i, [_vp, _i, _vp, _vp, _vp, _i]),
        "lymph_add_point_sources_noalloc": (_i, [_vp, _i, _vp, _vp, _vp, _i]),
        "lymph_add_double_couples": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _i]),
        "lymph_set_ricker": (_i, [_vp, _d, _d]),
        "lymph_ricker": (_d, [_d, _d, _d]),
        "lymph_capture_solve": (_i, [_vp, _i, _d, _i, _i]),
        "lymph_replay": (_i, [_vp]),
        "lymph_sync": (_i, [_vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def check(rc, where):
    if rc != 0:
        raise LymphError(rc, where, lib().lymph_last_error().decode())


def device_count():
    return lib().lymph_device_count()


def ricker(f0, t0, t):
    return lib().lymph_ricker(f0, t0, t)


def time_function(problem, t):
    return lib().lymph_time_function(problem, t)


# Callback types of lymph_build_rhs_fn (include/lymph_gpu.h)
FORCING_FN = C.CFUNCTYPE(None, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                         C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                         C.c_void_p)
DIRICHLET_FN = C.CFUNCTYPE(None, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int32),
                           C.POINTER(C.c_double), C.c_void_p)
FIELD_FN = C.CFUNCTYPE(None, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)
NEUMANN_FN = C.CFUNCTYPE(None, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                         C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int32),
                         C.POINTER(C.c_double), C.c_void_p)
