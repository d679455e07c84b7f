"""ctypes binding of libautopdex_amd.so (the C ABI in include/autopdex_amd.h).

The shared library is the only compute path of this package: if it is missing
or fails to load, every entry point raises -- there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libautopdex_amd.so")

APX_FP64, APX_FP32, APX_FP64_ROUNDED = 0, 1, 2
APX_PRECOND_JACOBI, APX_PRECOND_NONE, APX_PRECOND_LEARNED = 0, 1, 2
MODEL_POISSON_SOURCE, MODEL_POISSON_VARK, MODEL_NEO_HOOKE, MODEL_LINE_TRACTION = 1, 2, 3, 4


class PlanInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_dofs", C.c_int64), ("n_free", C.c_int64), ("nnz", C.c_int64),
                ("dim", C.c_int), ("dofs_per_node", C.c_int), ("n_sets", C.c_int), ("structured", C.c_int),
                ("n_colours", C.c_int * 8), ("build_seconds", C.c_double), ("device_bytes", C.c_int64)]


class KrylovStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("initial_residual", C.c_double),
                ("final_residual", C.c_double), ("seconds", C.c_double)]


class NewtonOptions(C.Structure):
    _fields_ = [("newton_tol", C.c_double), ("relative", C.c_int), ("maxiter", C.c_int),
                ("damping", C.c_double), ("precision", C.c_int), ("precond", C.c_int),
                ("lin_tol", C.c_double), ("lin_atol", C.c_double), ("lin_maxiter", C.c_int),
                ("reuse_tangent", C.c_int), ("forcing", C.c_int), ("reliable_delta", C.c_double),
                ("fp32_element_arith", C.c_int), ("krylov_fp64_vectors", C.c_int)]


class NewtonStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("diverged", C.c_int), ("residual_norm", C.c_double),
                ("initial_residual", C.c_double), ("krylov_iterations", C.c_int), ("fallbacks", C.c_int),
                ("seconds_total", C.c_double), ("seconds_residual", C.c_double),
                ("seconds_tangent", C.c_double), ("seconds_krylov", C.c_double),
                ("history", C.c_double * 64)]


_vp, _i64, _int, _dbl = C.c_void_p, C.c_int64, C.c_int, C.c_double
_SIGS = {
    "apx_last_error": (C.c_char_p, []),
    "apx_version": (_int, []),
    "apx_plan_create": (_int, [C.POINTER(_vp), _int, _i64, _vp, _int, _vp]),
    "apx_plan_add_set": (_int, [_vp, _vp, _i64, _int, _int, _vp, _int, C.POINTER(_int)]),
    "apx_plan_set_params": (_int, [_vp, _int, _vp, _int]),
    "apx_plan_bind_nodal_field": (_int, [_vp, _int, _int, _vp]),
    "apx_plan_finalize": (_int, [_vp]),
    "apx_plan_build_pattern": (_int, [_vp]),
    "apx_plan_destroy": (_int, [_vp]),
    "apx_plan_get_info": (_int, [_vp, C.POINTER(PlanInfo)]),
    "apx_plan_csr_pattern": (_int, [_vp, _vp, _vp]),
    "apx_plan_free_dofs": (_int, [_vp, _vp]),
    "apx_assemble_residual": (_int, [_vp, _vp, _vp, _vp]),
    "apx_assemble_tangent": (_int, [_vp, _vp, _vp, _int, _vp]),
    "apx_assemble_diagonal": (_int, [_vp, _vp, _vp, _vp]),
    "apx_tangent_coo_entries": (_int, [_vp, C.POINTER(_i64)]),
    "apx_assemble_tangent_coo": (_int, [_vp, _vp, _vp, _int, _vp]),
    "apx_spmv": (_int, [_vp, _vp, _vp, _vp, _int, _vp]),
    "apx_pcg": (_int, [_vp, _vp, _vp, _vp, _int, _int, _dbl, _dbl, _int, C.POINTER(KrylovStats), _vp]),
    "apx_pcg_noalloc": (_int, [_vp, _vp, _vp, _vp, _int, _int, _dbl, _dbl, _int, C.POINTER(KrylovStats), _vp]),
    "apx_newton": (_int, [_vp, _vp, _vp, C.POINTER(NewtonOptions), C.POINTER(NewtonStats), _vp]),
    "apx_unet_create": (_int, [C.POINTER(_vp), _int, _vp, _int, _int, _int, _int, _int, _int]),
    "apx_unet_destroy": (_int, [_vp]),
    "apx_unet_num_params": (_int, [_vp, C.POINTER(_i64), C.POINTER(_int)]),
    "apx_unet_layer": (_int, [_vp, _int] + [C.POINTER(_i64)] * 2 + [C.POINTER(_int)] * 6),
    "apx_unet_set_params": (_int, [_vp, _vp]),
    "apx_unet_get_params": (_int, [_vp, _vp]),
    "apx_unet_get_grads": (_int, [_vp, _vp]),
    "apx_unet_buffers": (_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "apx_unet_forward": (_int, [_vp, _vp]),
    "apx_unet_loss_backward": (_int, [_vp, _vp, C.POINTER(_dbl), _vp]),
    "apx_unet_adam": (_int, [_vp, C.c_float, C.c_float, C.c_float, C.c_float, _int, _vp]),
    "apx_unet_load_pair": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _int, _vp]),
    "apx_unet_apply": (_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "apx_unet_train_step": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _int, _dbl, C.POINTER(_dbl), _vp]),
    "apx_plan_attach_unet": (_int, [_vp, _vp, _vp]),
    "apx_unet_set_jacobi_skip": (_int, [_vp, _int]),
    "apx_unet_set_fused_tail": (_int, [_vp, _int, _vp]),
    "apx_fcg": (_int, [_vp, _vp, _vp, _vp, _int, _dbl, _dbl, _int, C.POINTER(KrylovStats), _vp]),
    "apx_set_timing": (_int, [_vp, _int]),
    "apx_get_timing": (_int, [_vp, C.c_char_p, C.POINTER(_dbl), C.POINTER(_i64), C.POINTER(_i64)]),
    "apx_set_timing_filter": (_int, [_vp, C.c_char_p]),
    "apx_get_timer": (_int, [_vp, _int, C.c_char_p, _int, C.POINTER(_dbl), C.POINTER(_i64), C.POINTER(_i64),
                             C.POINTER(_i64)]),
}

_lib = None


class NativeError(RuntimeError):
    pass


def lib():
    """Load (once) and return the native library; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(f"libautopdex_amd.so not built at {LIB_PATH}; run __graft_entry__.build() "
                              "(python -m autopdex_amd.build). There is no CPU fallback.")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise NativeError(f"autopdex_amd error {rc}: {lib().apx_last_error().decode()}")


def call(name, *args):
    check(getattr(lib(), name)(*args))
