"""ctypes binding of libillum_hip.so (the C ABI declared in include/illum_hip.h).

The library is loaded from this package directory; there is no fallback: if it
is missing, importing the compute API raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libillum_hip.so")

ILLUM_OK = 0
ILLUM_ERR_ARG = 1
ILLUM_ERR_NAN = 2
ILLUM_ERR_ZERO_PIVOT = 3
ILLUM_ERR_OOM = 4
ILLUM_ERR_HIP = 5
ILLUM_ERR_IO = 6
ILLUM_ERR_CAPACITY = 7
ILLUM_ERR_NOT_CONVERGED = 8
ILLUM_ERR_SHAPE = 9

ILLUM_IN_ON_DEVICE = 1
ILLUM_OUT_ON_DEVICE = 2
ILLUM_ACCUMULATE = 4
ILLUM_ASYNC = 8
ILLUM_FF_KEEP_VISIBILITY = 16
ILLUM_SUN_PER_BODY = 32
ILLUM_TL_EAGER = 64
ILLUM_Q_HELD = 128
ILLUM_RAD_JACOBI_FP64 = 1
ILLUM_RAD_MIXED = 2
ILLUM_THERM_FP64 = 1
ILLUM_THERM_MIXED = 2

_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64
_D = C.c_double
_PP = C.POINTER(C.c_void_p)

# name -> argtypes (every function returns int status unless listed in _RET)
SIGNATURES = {
    "illum_last_error": [],
    "illum_version": [],
    "illum_ctx_create": [_I, _PP],
    "illum_ctx_destroy": [_P],
    "illum_ctx_synchronize": [_P],
    "illum_ctx_set_profiling": [_P, _I],
    "illum_ctx_reset_profile": [_P],
    "illum_ctx_get_profile": [_P, _I, _P, _P, _P, _I, _P],
    "illum_ctx_tuning": [_P, _P, _I, _P],
    "illum_dev_alloc": [_P, _I64, _PP],
    "illum_dev_free": [_P, _P],
    "illum_dev_copy": [_P, _P, _P, _I64, _I],
    "illum_synth_sizes": [_I, _P, _P],
    "illum_synth_ellipsoid": [_I, _D, _D, _D, _I, _D, _D, _D, _D, C.c_uint64, _P, _P],
    "illum_mesh_create": [_P, _P, _I64, _P, _I64, _P, _D, _PP],
    "illum_mesh_from_obj": [_P, C.c_char_p, _D, _PP],
    "illum_mesh_destroy": [_P],
    "illum_mesh_num_faces": [_P, _P],
    "illum_mesh_set_bodies": [_P, _P, _I],
    "illum_mesh_geometry": [_P, _P, _P, _P, _P, _P],
    "illum_mesh_triangles": [_P, _P],
    "illum_bvh_build": [_P, _I, _PP],
    "illum_bvh_destroy": [_P],
    "illum_bvh_tuning": [_P, _I, _P, _P, _P, _P],
    "illum_bvh_info": [_P, _P, _P, _P],
    "illum_bvh_export": [_P, _P, _P, _P, _P, _P],
    "illum_bvh_closest_hit": [_P, _P, _P, _I64, _P, _P],
    "illum_form_factors": [_P, _D, _P, _D, _PP],
    "illum_form_factors_ex": [_P, _D, _P, _D, _I, _PP],
    "illum_form_factors_work": [_P, _D, _P],
    "illum_form_factors_work_ex": [_P, _D, _P, _I],
    "illum_sparse_create_csr": [_P, _I64, _I64, _P, _P, _P, _PP],
    "illum_sparse_destroy": [_P],
    "illum_sparse_info": [_P, _P, _P],
    "illum_sparse_copy_csc": [_P, _P, _P, _P],
    "illum_sparse_copy_csr": [_P, _P, _P, _P],
    "illum_sparse_visibility_nnz": [_P, _P],
    "illum_sparse_copy_visibility_lists": [_P, _P, _P],
    "illum_sparse_visibility_words": [_P, _P],
    "illum_sparse_copy_visibility": [_P, _P],
    "illum_sparse_rays": [_P, _P],
    "illum_horizons": [_P, _P, _I, _D, _D, _P, _I],
    "illum_render": [_P, _I, _I, _I, _D, _D, _P, _D, _D, _P, _P, _I],
    "illum_fib_spiral": [_P, _I],
    "illum_direct_radiosity": [_P, _P, _I, _P, _I, _P, _I, _P, _I],
    "illum_radiosity_solve": [_P, _P, _I, _I, _D, _I, _P, _P, _P, _I],
    "illum_radiosity_solve_noalloc": [_P, _P, _I, _I, _D, _I, _P, _P, _P, _I],
    "illum_thermal_create": [_P, _I64, _I, _P, _P, _P, _P, _P, _P, _D, _D, _P, _I, _PP],
    "illum_thermal_destroy": [_P],
    "illum_thermal_info": [_P, _P, _P],
    "illum_thermal_step": [_P, _D, _P, _I],
    "illum_thermal_run": [_P, _D, _P, _I, _I],
    "illum_thermal_get": [_P, _P, _P, _P, _I],
    "illum_thermal_get_radiosity": [_P, _P, _I],
    "illum_timeloop_create": [_P, _P, _P, _P, _I, _P, _I, _I, _D, _I, _I, _PP],
    "illum_timeloop_run": [_P, _P, _I, _D, _I],
    "illum_timeloop_stats": [_P, _P, _P, _P, _P, _P],
    "illum_timeloop_get": [_P, _P, _P, _I],
    "illum_timeloop_destroy": [_P],
    "illum_knn": [_P, _P, _I64, _P, _I64, _I, _P, _P, _I],
    "illum_mesh_transfer": [_P, _P, _I64, _P, _I64, _P, _I, _I, _I, _P, _P, _P],
    "illum_conductionQ": [_I, _P, _D, _D, _D, _P, _P, _P, _D, _P, _D, _P],
}
_RET = {"illum_last_error": C.c_char_p}

_lib = None


class IllumError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[illum status {code}] {msg}")
        self.code = code


def load():
    """Load libillum_hip.so (raises ImportError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RET.get(name, C.c_int)
    _lib = lib
    return lib


def check(status):
    if status != ILLUM_OK:
        msg = load().illum_last_error()
        raise IllumError(status, msg.decode() if msg else "")
    return status


def call(name, *args):
    return check(getattr(load(), name)(*args))
