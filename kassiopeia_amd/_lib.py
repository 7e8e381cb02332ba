"""ctypes binding of the C-ABI in ``include/kemhip.h`` (``kassiopeia_amd/libkemhip.so``).

Loading fails loudly when the library has not been built: there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libkemhip.so")

KEM_OK = 0
KEM_W_NOT_CONVERGED = 1
KEM_E_ARG, KEM_E_HIP, KEM_E_NAME, KEM_E_GEOM, KEM_E_STATE = -1, -2, -3, -4, -5
KEM_FP64, KEM_MIXED, KEM_MIXED_BF16 = 0, 1, 2
KEM_STORE_FP64, KEM_STORE_FP32, KEM_STORE_BF16 = 1, 2, 4
KEM_SOLVER_GMRES, KEM_SOLVER_ROBIN_HOOD, KEM_SOLVER_GAUSS = 0, 1, 2
KEM_TUNE_FAR_RANGE_FP64, KEM_TUNE_FAR_RANGE_MIXED, KEM_TUNE_GEMV_BLOCKS, KEM_TUNE_BMM_BLOCKS = 1, 2, 3, 4
KEM_TUNE_NEAR_BLOCKS = 5
KEM_TUNE_BMM_ROWS = 6
KEM_TUNE_ROBIN_HOOD_GROUPS = 7
KEM_TUNE_BMM_STREAM_BLOCKS = 8
KEM_TUNE_BMM_TILE_BLOCKS = 9


class KemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kemhip error {code}: {msg}")
        self.code = code


class SolveOpts(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("tolerance", ctypes.c_double),
        ("max_iterations", ctypes.c_int),
        ("restart", ctypes.c_int),
        ("check_interval", ctypes.c_int64),
        ("time_limit_s", ctypes.c_double),
        ("inner_tolerance", ctypes.c_double),
    ]


class SolveStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("sub_iterations", ctypes.c_int64),
        ("refinements", ctypes.c_int),
        ("residual", ctypes.c_double),
        ("true_residual", ctypes.c_double),
        ("seconds", ctypes.c_double),
    ]


class ZHParams(ctypes.Structure):
    """kem_zh_params (KZonalHarmonicParameters)"""
    _fields_ = [
        ("n_bifurcations", ctypes.c_int32),
        ("convergence_ratio", ctypes.c_double),
        ("proximity_to_source_point", ctypes.c_double),
        ("convergence_parameter", ctypes.c_double),
        ("n_central_coefficients", ctypes.c_int32),
        ("central_fractional_spacing", ctypes.c_int32),
        ("central_fractional_distance", ctypes.c_double),
        ("central_delta_z", ctypes.c_double),
        ("central_z1", ctypes.c_double),
        ("central_z2", ctypes.c_double),
        ("n_remote_coefficients", ctypes.c_int32),
        ("n_remote_source_points", ctypes.c_int32),
        ("remote_z1", ctypes.c_double),
        ("remote_z2", ctypes.c_double),
    ]


# name -> (restype, argtypes); the exported symbol set of include/kemhip.h
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
_INT = ctypes.c_int
SIGNATURES = {
    "kem_version": (_INT, []),
    "kem_ctx_create": (_INT, [_INT, ctypes.POINTER(_P)]),
    "kem_ctx_destroy": (None, [_P]),
    "kem_ctx_set_stream": (_INT, [_P, _P]),
    "kem_ctx_stream": (_P, [_P]),
    "kem_ctx_set_mixed_fp64_ratio": (_INT, [_P, _D]),
    "kem_dev_alloc": (_INT, [_P, _I64, ctypes.POINTER(_P)]),
    "kem_dev_free": (_INT, [_P, _P]),
    "kem_dev_upload": (_INT, [_P, _P, _P, _I64]),
    "kem_dev_download": (_INT, [_P, _P, _P, _I64]),
    "kem_ctx_set_tuning": (_INT, [_P, _INT, _I64]),
    "kem_ctx_get_tuning": (_INT, [_P, _INT, ctypes.POINTER(_I64)]),
    "kem_ctx_set_kernel_timing": (_INT, [_P, _INT]),
    "kem_ctx_kernel_time": (_INT, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]),
    "kem_ctx_sync": (_INT, [_P]),
    "kem_last_error": (ctypes.c_char_p, [_P]),
    "kem_set_integrator": (_INT, [_P, ctypes.c_char_p]),
    "kem_mesh_create": (_INT, [_P, _I64, _P, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "kem_mesh_create_groups": (_INT, [_P, _I64, _P, _P, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "kem_mesh_destroy": (None, [_P]),
    "kem_mesh_size": (_I64, [_P]),
    "kem_mesh_rhs": (_INT, [_P, _P]),
    "kem_mesh_set_values": (_INT, [_P, _P]),
    "kem_element_eval": (_INT, [_P, _P, _I64, _P, _P, _P]),
    "kem_field_eval": (_INT, [_P, _P, _P, _I64, _P, _INT, _P, _P]),
    "kem_field_eval_device": (_INT, [_P, _P, _P, _I64, _P, _P, _P, _INT, _P, _P, _P, _P]),
    "kem_field_eval_device_noalloc": (_INT, [_P, _P, _P, _I64, _P, _P, _P, _INT, _P, _P, _P, _P]),
    "kem_field_graph_create": (_INT, [_P, _P, _P, _I64, _P, _P, _P, _INT, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "kem_graph_launch": (_INT, [_P]),
    "kem_graph_destroy": (None, [_P]),
    "kem_field_eval_fields": (_INT, [_P, _I64, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P]),
    "kem_field_eval_sets": (_INT, [_P, _P, _I64, _P, _P, _I64, _P, _P, _P, _INT, _P, _P, _P, _P]),
    "kem_field_census": (_INT, [_P, _P, _I64, _P, _P, _P, _INT, _P]),
    "kem_field_census_paths": (_INT, [_P, _P, _I64, _P, _P, _P, _P]),
    "kem_zh_params_default": (None, [ctypes.POINTER(ZHParams)]),
    "kem_zh_create": (_INT, [_P, _P, _P, ctypes.POINTER(ZHParams), ctypes.POINTER(_P)]),
    "kem_zh_destroy": (None, [_P]),
    "kem_zh_info": (_INT, [_P, _P]),
    "kem_zh_field_eval_device": (_INT, [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _P]),
    "kem_zh_field_eval": (_INT, [_P, _I64, _P, _P, _P, _P]),
    "kem_field_plan": (_INT, [_INT, _INT, _INT, _INT, _I64, _I64, _I64, _I64, _I64, _P]),
    "kem_matrix_assemble": (_INT, [_P, _P, _INT, ctypes.POINTER(_P)]),
    "kem_matrix_destroy": (None, [_P]),
    "kem_matrix_dimension": (_I64, [_P]),
    "kem_matrix_get": (_INT, [_P, _I64, _I64, ctypes.POINTER(_D)]),
    "kem_matrix_copy": (_INT, [_P, _P]),
    "kem_matrix_multiply": (_INT, [_P, _INT, _P, _P]),
    "kem_solve_opts_default": (None, [ctypes.POINTER(SolveOpts)]),
    "kem_solve": (_INT, [_P, ctypes.POINTER(SolveOpts), _P, _P, ctypes.POINTER(SolveStats)]),
    "kem_solve_transposed": (_INT, [_P, ctypes.POINTER(SolveOpts), _P, _P, ctypes.POINTER(SolveStats)]),
    "kem_field_vjp": (_INT, [_P, _P, _P, _I64, _P, _P, _P, _P]),
    "kem_assemble_vjp": (_INT, [_P, _P, _P, _P, _P]),
    "kem_solve_batch": (_INT, [_P, ctypes.POINTER(SolveOpts), _I64, _P, _P, _P, ctypes.POINTER(SolveStats)]),
    "kem_matrix_multiply_batch": (_INT, [_P, _INT, _I64, _P, _P]),
    "kem_cache_error": (ctypes.c_char_p, []),
    "kem_md5_hex": (_INT, [_P, _I64, ctypes.c_char_p]),
    "kem_hash_doubles": (_INT, [_P, _I64, _INT, _D, ctypes.c_char_p]),
    "kem_hash_surfaces": (_INT, [_I64, _P, _P, _P, _P, _P, _P, _INT, _INT, _D, ctypes.c_char_p]),
    "kem_cache_save": (_INT, [ctypes.c_char_p, ctypes.c_char_p, _I64, _P, _P, _P, _P, _P, _P, _D, _INT, _D]),
    "kem_cache_find": (_INT, [ctypes.c_char_p, _I64, _P, _P, _P, _P, _P, _D, _INT, _D, _P, ctypes.POINTER(_INT),
                              ctypes.c_char_p, _I64]),
}

_lib = None


def load():
    """Load libkemhip.so (built by ``python -m kassiopeia_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError(f"{LIBPATH} is not built; run `python -m kassiopeia_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIBPATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib
