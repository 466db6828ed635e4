"""lymph3d_apsc_amd -- MI355X-native matrix-free PolyDG leapfrog path of LYMPH3D.

The numerics live in liblymph_gpu.so (hand-written HIP kernels for gfx950
behind the C ABI of include/lymph_gpu.h); this package is the host side:
a ctypes binding (`_lib`), the mesh pipeline and the reference's input and
partition files (`mesh`), a mirror of the reference's `matrix_free` routines
(`matrix_free.Context`), the driver of src/Lymph3D.f90 (`driver`) and its
output files (`export`).
"""
from ._lib import (FN_CUBE_U, FN_REF_UEX, FN_REF_VEL, FN_ZERO, PROB_CUBE_WAVE, PROB_NULL,
                   PROB_REF_MMS, PROB_RICKER, LymphError, device_count, ricker, time_function)
from .matrix_free import Context
from .mesh import (Mesh, MeshError, cube_mesh, read_deck, read_elem_in_poly, write_deck,
                   write_elem_in_poly)

__all__ = [
    "Context", "Mesh", "MeshError", "LymphError", "cube_mesh", "read_deck", "write_deck",
    "write_elem_in_poly", "read_elem_in_poly", "device_count",
    "time_function", "ricker", "PROB_RICKER", "PROB_NULL", "PROB_REF_MMS", "PROB_CUBE_WAVE", "FN_ZERO", "FN_REF_VEL",
    "FN_REF_UEX", "FN_CUBE_U",
]
