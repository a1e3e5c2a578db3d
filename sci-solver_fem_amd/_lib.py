"""ctypes binding of libfem_mi355x.so (the C ABI declared in include/fem_mi355x.h).

The library is loaded from this package's ``lib/`` folder.  There is no
fallback: a missing library raises ``ImportError`` from ``load()``.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libfem_mi355x.so"

FEM_OK = 0
FEM_ERR_ARG = 1
FEM_ERR_IO = 2
FEM_ERR_HIP = 3
FEM_ERR_EMPTY = 4
FEM_ERR_NOT_CONVERGED = 5
FEM_ERR_NUMERIC = 6
FEM_ERR_STATE = 7
FEM_ERR_CAPACITY = 8


class FemParams(C.Structure):
    _fields_ = [
        ("verbose", C.c_int),
        ("maxLevels", C.c_int),
        ("maxIters", C.c_int),
        ("preInnerIters", C.c_int),
        ("postInnerIters", C.c_int),
        ("postRelaxes", C.c_int),
        ("cycleIters", C.c_int),
        ("dsType", C.c_int),
        ("topSize", C.c_int),
        ("randMisParameters", C.c_int),
        ("partitionMaxSize", C.c_int),
        ("aggregatorType", C.c_int),
        ("convergeType", C.c_int),
        ("tolerance", C.c_double),
        ("cycleType", C.c_int),
        ("solverType", C.c_int),
        ("smootherWeight", C.c_double),
        ("proOmega", C.c_double),
        ("device", C.c_int),
        ("blockSize", C.c_int),
        ("precision", C.c_int),
        ("smootherType", C.c_int),
        ("aggregationRule", C.c_int),
        ("coarseSolve", C.c_int),
        ("seed", C.c_ulonglong),
        ("strengthThreshold", C.c_double),
        ("launchMode", C.c_int),
    ]


class FemSolveInfo(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("converged", C.c_int),
        ("rel_residual", C.c_double),
        ("setup_ms", C.c_double),
        ("solve_ms", C.c_double),
        ("num_levels", C.c_int),
        ("grid_complexity", C.c_double),
        ("operator_complexity", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = C.c_void_p
_I64 = C.c_int64
_PI64 = C.POINTER(C.c_int64)
_PI32 = C.POINTER(C.c_int32)
_PD = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol include/fem_mi355x.h declares
SIGNATURES = {
    "fem_params_default": (None, [C.POINTER(FemParams)]),
    "fem_create_from_files": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(_P)]),
    "fem_create_from_arrays": (C.c_int, [C.c_int, _I64, _P, _I64, _P, _P, C.c_double, C.c_int, C.c_int,
                                         C.POINTER(_P)]),
    "fem_destroy": (None, [_P]),
    "fem_release_cached_memory": (None, []),
    "fem_last_error": (C.c_char_p, [_P]),
    "fem_set_params": (C.c_int, [_P, C.POINTER(FemParams)]),
    "fem_get_params": (C.c_int, [_P, C.POINTER(FemParams)]),
    "fem_set_stream": (C.c_int, [_P, _P]),
    "fem_get_matrix_rows": (_I64, [_P]),
    "fem_get_matrix_nnz": (_I64, [_P]),
    "fem_get_num_elements": (_I64, [_P]),
    "fem_get_matrix_csr": (C.c_int, [_P, _P, _P, _P]),
    "fem_read_matlab_sparse": (C.c_int, [_P, C.c_char_p]),
    "fem_update_kappa": (C.c_int, [_P, _P]),
    "fem_setup": (C.c_int, [_P]),
    "fem_solve": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.POINTER(FemSolveInfo)]),
    "fem_adjoint_grad": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, C.POINTER(FemSolveInfo)]),
    "fem_element_grad": (C.c_int, [_P, _P, _P, _P, C.c_int]),
    "fem_apply_operator": (C.c_int, [_P, _P, _P, C.c_int, C.c_int]),
    "fem_apply_preconditioner": (C.c_int, [_P, _P, _P, C.c_int]),
    "fem_apply_preconditioner_noalloc": (C.c_int, [_P, _P, _P]),
    "fem_noalloc_max_rows": (C.c_int64, []),
    "fem_apply_level": (C.c_int, [_P, C.c_int, C.c_int, _P, _P]),
    "fem_get_level_smoother": (C.c_int, [_P, C.c_int, _PD, _PD, _PD, _PD]),
    "fem_apply_smoother_step": (C.c_int, [_P, C.c_int, C.c_int, C.c_double, C.c_double, _P, _P, _P, _P]),
    "fem_get_num_levels": (C.c_int, [_P]),
    "fem_get_level_info": (C.c_int, [_P, C.c_int, _PI64, _PI64]),
    "fem_get_level_matrix": (C.c_int, [_P, C.c_int, _P, _P, _P, _P]),
    "fem_get_level_prolongator": (C.c_int, [_P, C.c_int, _PI64, _P, _P, _P]),
    "fem_get_permutation": (C.c_int, [_P, _P]),
    "fem_device_aggregate": (C.c_int, [_I64, _P, _P, C.c_int, C.c_int, C.c_int, C.c_ulonglong, _P, _PI32]),
    "fem_condition_aggregates": (C.c_int, [_I64, _P, _P, _P, _I64, C.c_int, C.c_int, C.POINTER(C.c_int64)]),
    "fem_flood_fill_aggregate": (C.c_int, [_I64, _P, _P, C.c_int, C.c_ulonglong, _P, _P, _P, _P]),
    "fem_profile_enable": (C.c_int, [_P, C.c_int]),
    "fem_profile_get": (C.c_int, [_P, C.c_char_p, _PI64, _PD, _PD]),
    "fem_read_tetgen": (C.c_int, [C.c_char_p, _PI64, C.POINTER(_PD), _PI64, C.POINTER(_PI32),
                                  C.POINTER(_PI32)]),
    "fem_read_trimesh": (C.c_int, [C.c_char_p, _PI64, C.POINTER(_PD), _PI64, C.POINTER(_PI32)]),
    "fem_free": (None, [_P]),
    "fem_read_matlab_array": (C.c_int, [C.c_char_p, _P, _PI64]),
    "fem_write_matlab_array": (C.c_int, [C.c_char_p, _P, _I64]),
    "fem_read_matlab_sparse_coo": (C.c_int, [C.c_char_p, C.c_int, _PI64, _PI64, _PI64, _P, _P, _P]),
    "fem_write_vtk": (C.c_int, [C.c_char_p, C.c_int, _I64, _P, _I64, _P, _P, _I64]),
    "fem_version": (C.c_char_p, []),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library (raises ImportError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} not found: build it with `python sci-solver_fem_amd/build.py`")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
