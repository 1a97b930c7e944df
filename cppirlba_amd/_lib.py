"""ctypes binding of libirlba_hip.so (include/irlba_hip.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in this package.
"""

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libirlba_hip.so")

F64, F32, BF16 = 0, 1, 2

_lib = None


class Options(ctypes.Structure):
    _fields_ = [
        ("invariant_subspace_tolerance", ctypes.c_double),
        ("convergence_tolerance", ctypes.c_double),
        ("singular_value_ratio_tolerance", ctypes.c_double),
        ("extra_work", ctypes.c_int32),
        ("max_iterations", ctypes.c_int32),
        ("exact_for_small_matrix", ctypes.c_int32),
        ("exact_for_large_number", ctypes.c_int32),
        ("cap_number", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("initial", ctypes.POINTER(ctypes.c_double)),
        ("initial_length", ctypes.c_int64),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("forward_products", ctypes.c_int64),
        ("adjoint_products", ctypes.c_int64),
        ("forward_kernel_ms", ctypes.c_double),
        ("adjoint_kernel_ms", ctypes.c_double),
        ("forward_bytes", ctypes.c_double),
        ("adjoint_bytes", ctypes.c_double),
        ("solve_ms", ctypes.c_double),
        ("host_svd_ms", ctypes.c_double),
        ("restarts", ctypes.c_int32),
        ("refills", ctypes.c_int32),
        ("forward_timed", ctypes.c_int64),
        ("adjoint_timed", ctypes.c_int64),
    ]


class Layout(ctypes.Structure):
    _fields_ = [
        ("nnz", ctypes.c_int64),
        ("slots", ctypes.c_int64),
        ("index_bytes", ctypes.c_int64),
        ("tile_len", ctypes.c_int32),
        ("n_tiles", ctypes.c_int32),
        ("vec", ctypes.c_int32),
        ("lanes_per_line", ctypes.c_int32),
        ("delta", ctypes.c_int32),
        ("blocks", ctypes.c_int32),
        ("dictionary", ctypes.c_int32),
        ("dictionary_run", ctypes.c_int32),
    ]


MATVEC_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_double))

_VP = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_I64 = ctypes.POINTER(ctypes.c_int64)
_i = ctypes.c_int
_i64 = ctypes.c_int64

SIGNATURES = {
    "irlba_hip_options_default": (None, [ctypes.POINTER(Options)]),
    "irlba_hip_last_error": (ctypes.c_char_p, []),
    "irlba_hip_context_create": (_i, [_i, _PP]),
    "irlba_hip_context_destroy": (_i, [_VP]),
    "irlba_hip_context_synchronize": (_i, [_VP]),
    "irlba_hip_get_stats": (_i, [_VP, ctypes.POINTER(Stats)]),
    "irlba_hip_context_set_profiling": (_i, [_VP, _i]),
    "irlba_hip_context_stream": (_VP, [_VP]),
    "irlba_hip_context_wait_stream": (_i, [_VP, _VP]),
    "irlba_hip_stream_wait_context": (_i, [_VP, _VP]),
    "irlba_hip_sparse_create": (_i, [_VP, _i64, _i64, _i, _VP, _VP, _VP, _i, _i, _i, _i, _PP]),
    "irlba_hip_dense_create": (_i, [_VP, _i64, _i64, _i, _VP, _i, _i, _i, _PP]),
    "irlba_hip_centered_create": (_i, [_VP, _VP, _i, _PP]),
    "irlba_hip_scaled_create": (_i, [_VP, _VP, _i, _i, _i, _PP]),
    "irlba_hip_matrix_destroy": (_i, [_VP]),
    "irlba_hip_matrix_dims": (_i, [_VP, _I64, _I64]),
    "irlba_hip_multiply": (_i, [_VP, _i, _VP, _VP, _i]),
    "irlba_hip_realize": (_i, [_VP, _VP, _i]),
    "irlba_hip_compute": (_i, [_VP, _i64, ctypes.POINTER(Options), _VP, _VP, _VP, _I64, _I32, _I32, _i]),
    "irlba_hip_compute_callback": (_i, [_VP, _i64, _i64, MATVEC_FN, _VP, _i64, ctypes.POINTER(Options), _VP, _VP,
                                        _VP, _I64, _I32, _I32]),
    "irlba_hip_column_stats": (_i, [_VP, _i, _i, _VP, _VP, _i]),
    "irlba_hip_pca": (_i, [_VP, _i, _i, _i64, ctypes.POINTER(Options), _VP, _VP, _VP, _I64, _I32, _I32, _i]),
    "irlba_hip_generate_sparse": (_i, [_VP, _i64, _i64, ctypes.c_double, ctypes.c_uint64, _i, _i, _PP]),
    "irlba_hip_generate_csc": (_i, [_VP, _i64, _i64, ctypes.c_double, ctypes.c_uint64, _i, _VP, _VP, _VP, _I64]),
}


class HipError(RuntimeError):
    pass


def lib():
    """Load libirlba_hip.so (raises if it was never built)."""
    global _lib
    if _lib is None:
        # torch ships its own HIP runtime under the same soname (libamdhip64.so.7,
        # libhsa-runtime64.so.1).  Loading it first makes this library bind to that
        # copy, so a process never holds two HIP runtimes (the second one to open
        # the GPU finds no device).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "(python cppirlba_amd/build.py)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise HipError(lib().irlba_hip_last_error().decode())


SIGNATURES.update({
    "irlba_hip_lanczos": (_i, [_VP, _i64, _VP, _VP, _VP, ctypes.c_uint64, _i64, ctypes.POINTER(Options), _VP]),
    "irlba_hip_orthogonalize": (_i, [_VP, _VP, _i64, _i64, _VP, _i64, _VP]),
    "irlba_hip_rotate": (_i, [_VP, _VP, _i64, _i64, _VP, _i64, _VP]),
    "irlba_hip_svd_backward": (_i, [_VP, _i64, _VP, _VP, _VP, _VP, _VP, _VP, _i, ctypes.c_int32, ctypes.c_double,
                                    _VP, _VP, _VP, _I32, _i]),
    "irlba_hip_sddmm": (_i, [_VP, _i64, _i64, _VP, _VP, _i, _i64, _VP, _VP, _VP, _i, _i]),
    "irlba_hip_sddmm_noalloc": (_i, [_VP, _i64, _i64, _VP, _VP, _i, _i64, _VP, _VP, _VP, _i]),
    "irlba_hip_svd_backward_async": (_i, [_VP, _i64, _VP, _VP, _VP, _VP, _VP, _VP, _i, ctypes.c_int32,
                                          ctypes.c_double, _VP, _VP, _VP, _I32]),
    "irlba_hip_sddmm_async": (_i, [_VP, _i64, _i64, _VP, _VP, _i, _i64, _VP, _VP, _VP, _i]),
    "irlba_hip_dense_gradient": (_i, [_VP, _i64, _i64, _i64, _VP, _VP, _VP, _i, _i, _i]),
    "irlba_hip_dense_gradient_async": (_i, [_VP, _i64, _i64, _i64, _VP, _VP, _VP, _i, _i]),
    "irlba_hip_multiply_block": (_i, [_VP, _i, _i64, _VP, _VP, _i]),
    "irlba_hip_sparse_layout": (_i, [_VP, _i, ctypes.POINTER(Layout)]),
    "irlba_hip_sparse_value_dictionary": (_i, [_VP, _i, ctypes.POINTER(ctypes.c_int32)]),
})
