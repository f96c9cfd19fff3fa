"""ctypes binding of libafar.so (include/afar.h)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libafar.so")
HEADER = os.path.join(HERE, "..", "include", "afar.h")

AFAR_OK = 0


class AfarError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("afar error %d: %s" % (code, msg))
        self.code = code


class Schemes(C.Structure):
    """afar_schemes: fvSchemes CUDA{} / gaussQuadrature{} keys plus extensions."""
    _fields_ = [
        ("R", C.c_double),
        ("Cv", C.c_double),
        ("maxCo", C.c_double),
        ("shapeFactor", C.c_double),
        ("stencilType", C.c_int32),
        ("recalcBasis", C.c_int32),
        ("cOrder", C.c_int32),
        ("fOrder", C.c_int32),
        ("precision", C.c_int32),
        ("exactRK", C.c_int32),
        ("tile", C.c_int32),
        ("boundaryFirstOrder", C.c_int32),
        ("reconLanes", C.c_int32),
        ("cellOrder", C.c_int32),
        ("faceQuadrature", C.c_int32),
    ]


P = C.POINTER


class FieldPointer(C.Structure):
    """afar_field_pointer: the host members of the reference's fieldPointer."""
    _fields_ = [
        ("CELL", P(C.c_double)),
        ("FACE", P(C.c_double)),
        ("cellVolume", P(C.c_double)),
        ("FD", P(C.c_double)),
        ("shockIndicator", P(C.c_int8)),
        ("cFD", P(C.c_double)),
        ("neighbour", P(C.c_int32)),
        ("boundaryFacesNum", P(C.c_int32)),
        ("boundaryFaceNeiLabel", P(C.c_int32)),
        ("boundaryFacesType", P(C.c_uint8)),
        ("stencilSize", P(C.c_uint8)),
        ("compactStencilSize", P(C.c_uint8)),
        ("RBFbasis", P(C.c_float)),
        ("extendStencil", P(C.c_int32)),
        ("extendStencilSize", P(C.c_uint16)),
        ("compactExtendStencilSize", P(C.c_uint16)),
        ("localStencil", P(C.c_uint16)),
        ("cellFaces", P(C.c_int32)),
        ("globalStencil", P(C.c_int32)),
        ("RBFbasis64", P(C.c_double)),
        ("faceShift", P(C.c_double)),
        ("boundaryValue", P(C.c_double)),
    ]


_lib = None

# name -> (argtypes, restype); every function returning int is checked.
_VP = C.c_void_p
_D = P(C.c_double)
_I32 = P(C.c_int32)
_I64 = P(C.c_int64)
_U8 = P(C.c_uint8)
_I8 = P(C.c_int8)
SIGNATURES = {
    "afar_last_error": ([], C.c_char_p),
    "afar_version": ([], C.c_char_p),
    "afar_schemes_default": ([P(Schemes)], None),
    "afar_mesh_from_polygons": ([C.c_int64, _D, C.c_int64, _I64, _I32, _I8, C.c_double, C.c_double, C.c_int64,
                                 _I32, _I32, C.c_int32, _I32, _D, C.c_double, C.c_int32, P(_VP)], C.c_int),
    "afar_mesh_generate_corner": ([C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                   C.c_double, C.c_double, _D, C.c_int32, P(_VP)], C.c_int),
    "afar_mesh_generate_riemann": ([C.c_int32, C.c_double, C.c_uint64, C.c_double, P(_VP)], C.c_int),
    "afar_mesh_generate_periodic": ([C.c_int32, C.c_double, C.c_double, C.c_uint64, C.c_double, P(_VP)], C.c_int),
    "afar_mesh_destroy": ([_VP], None),
    "afar_mesh_read_foam": ([C.c_char_p, C.c_int32, P(_VP)], C.c_int),
    "afar_mesh_write_foam": ([_VP, C.c_char_p], C.c_int),
    "afar_mesh_patch_info": ([_VP, C.c_int32, C.c_char_p, C.c_int32, _I32], C.c_int),
    "afar_foam_read_fields": ([_VP, C.c_char_p, C.c_char_p, _D, _D, _I32], C.c_int),
    "afar_mesh_set_patch": ([_VP, C.c_int32, C.c_int32, _D], C.c_int),
    "afar_mesh_sizes": ([_VP, _I64], C.c_int),
    "afar_mesh_cells": ([_VP, _D, _D], C.c_int),
    "afar_mesh_permutation": ([_VP, _I32], C.c_int),
    "afar_mesh_faces": ([_VP, _I32, _I32, _D, _D, _I32], C.c_int),
    "afar_mesh_polygons": ([_VP, _I32, _D], C.c_int),
    "afar_mesh_precompute": ([_VP, P(Schemes), C.c_char_p], C.c_int),
    "afar_mesh_precompute_gpu": ([_VP, P(Schemes)], C.c_int),
    "afar_mesh_stencils": ([_VP, _U8, _U8, _I32, _I32], C.c_int),
    "afar_mesh_rbf_basis": ([_VP, _D, _U8, _D], C.c_int),
    "afar_mesh_face_quadrature": ([_VP, C.POINTER(C.c_int32), _D, _D, _D, _D], C.c_int),
    "afar_mesh_gauss_points": ([_VP, C.c_int32, C.c_int32, _D, _D, C.c_int32, _I32], C.c_int),
    "afar_mesh_write_basis": ([_VP, C.c_char_p], C.c_int),
    "afar_pinv": ([C.c_int32, _D, _D], C.c_int),
    "afar_mesh_export_reference": ([_VP, P(Schemes), _D, P(FieldPointer)], C.c_int),
    "afar_solver_create": ([_VP, P(Schemes), _D, P(_VP)], C.c_int),
    "afar_solver_create_on_stream": ([_VP, P(Schemes), _D, _VP, P(_VP)], C.c_int),
    "afar_solver_set_delta_t": ([_VP, _D], C.c_int),
    "afar_solver_evolve": ([_VP, C.c_double], C.c_int),
    "afar_solver_evolve_auto": ([_VP, C.c_int32], C.c_int),
    "afar_solver_last_dt": ([_VP, _D], C.c_int),
    "afar_solver_synchronize": ([_VP], C.c_int),
    "afar_solver_query": ([_VP, _I32], C.c_int),
    "afar_solver_set_time": ([_VP, C.c_double, C.c_double], C.c_int),
    "afar_solver_get_time": ([_VP, _D, _D], C.c_int),
    "afar_solver_residual_norm": ([_VP, _D], C.c_int),
    "afar_batch_set_time": ([_VP, C.c_int32, C.c_double, C.c_double], C.c_int),
    "afar_batch_get_time": ([_VP, _D, _D], C.c_int),
    "afar_solver_write_field": ([_VP, _D, _D, _I8], C.c_int),
    "afar_solver_get_conserved": ([_VP, _D], C.c_int),
    "afar_solver_set_conserved": ([_VP, _D], C.c_int),
    "afar_solver_time_steps": ([_VP, C.c_int32, C.c_int32, _D], C.c_int),
    "afar_solver_profile": ([_VP, C.c_int32, _D, _I32], C.c_int),
    "afar_solver_time_steps_kernels": ([_VP, C.c_int32, _D, _D, _I32], C.c_int),
    "afar_solver_sizes": ([_VP, _I64], C.c_int),
    "afar_solver_run_stage": ([_VP, C.c_int32, C.c_double, C.c_int32], C.c_int),
    "afar_solver_fetch": ([_VP, C.c_char_p, _VP, C.c_size_t], C.c_int),
    "afar_solver_destroy": ([_VP], None),
    "afar_batch_create": ([C.c_int32, P(_VP), P(Schemes), P(_D), P(_VP)], C.c_int),
    "afar_batch_evolve": ([_VP, C.c_int32], C.c_int),
    "afar_batch_time_steps": ([_VP, C.c_int32, _D], C.c_int),
    "afar_batch_synchronize": ([_VP], C.c_int),
    "afar_batch_reset": ([_VP], C.c_int),
    "afar_batch_sizes": ([_VP, _I64], C.c_int),
    "afar_batch_cell_offsets": ([_VP, _I32], C.c_int),
    "afar_batch_last_dt": ([_VP, _D], C.c_int),
    "afar_batch_get_conserved": ([_VP, C.c_int32, _D], C.c_int),
    "afar_batch_write_field": ([_VP, C.c_int32, _D, _I8], C.c_int),
    "afar_batch_profile": ([_VP, C.c_int32, _D, _I32], C.c_int),
    "afar_batch_destroy": ([_VP], None),
    "afar_set_device_field_data": ([P(_VP), P(FieldPointer), C.c_int64, C.c_int64, C.c_int64], C.c_int),
    "afar_device_configure": ([_VP, C.c_int32, C.c_int32], C.c_int),
    "afar_device_set_stream": ([_VP, _VP], C.c_int),
    "afar_set_device_stencil_data": ([_VP, P(FieldPointer), C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                      C.c_int64, C.c_int64, C.c_int64], C.c_int),
    "afar_free_field_data": ([_VP], C.c_int),
    "afar_adjust_time_step": ([_VP, C.c_int64, C.c_double, C.c_double, C.c_double, _D], C.c_int),
    "afar_gpu_evolve": ([_VP, C.c_double, C.c_double, C.c_double, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                         C.c_int64, C.c_int64, C.c_int64], C.c_int),
    "afar_copy_field_data_device_to_host": ([P(FieldPointer), _VP, C.c_int64, C.c_int64], C.c_int),
    "afar_adjust_time_step_noalloc": ([_VP, _VP, _VP, _VP, _VP, C.c_int64, C.c_double, C.c_double, C.c_double, _VP,
                                       _VP], C.c_int),
    "afar_selftest_math": ([_D, _D, C.c_int32], C.c_int),
    "afar_host_math": ([_D, _D, C.c_int32], C.c_int),
    "afar_selftest_math_f32": ([C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int32], C.c_int),
    "afar_host_math_f32": ([C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int32], C.c_int),
    "afar_device_info": ([C.c_char_p, C.c_int32, _I32, _I64], C.c_int),
}


def lib():
    """Load libafar.so; raises if it has not been built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libafar.so is not built: run __graft_entry__.build() "
                              "or python afar_cudaeuler2dmixo3rk3solver_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(code):
    if code != AFAR_OK:
        raise AfarError(code, lib().afar_last_error().decode())


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P(ctype))


def dptr(a):
    return ptr(a, C.c_double)
