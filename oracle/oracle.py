"""ctypes binding of the CPU oracle (oracle/irlba_oracle.cpp).

ORACLE — test infrastructure only.  Imported by tests/, by
``__graft_entry__.smoke()`` (as the checker) and by ``bench.py``'s
``cpu_baseline`` leg; never by the ``cppirlba_amd`` product path.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libirlba_oracle.so")
_lib = None


class OracleOptions(ctypes.Structure):
    _fields_ = [
        ("invariant_subspace_tolerance", ctypes.c_double),
        ("convergence_tolerance", ctypes.c_double),
        ("singular_value_ratio_tolerance", ctypes.c_double),
        ("extra_work", ctypes.c_int32),
        ("max_iterations", ctypes.c_int32),
        ("exact_for_small_matrix", ctypes.c_int32),
        ("exact_for_large_number", ctypes.c_int32),
        ("cap_number", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("initial", ctypes.POINTER(ctypes.c_double)),
        ("initial_len", ctypes.c_int64),
    ]


_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_I64 = ctypes.POINTER(ctypes.c_int64)
_VP = ctypes.c_void_p


def build():
    """Compile the oracle with its Makefile (g++)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "oracle_last_error": (ctypes.c_char_p, []),
            "oracle_dense_new": (_VP, [ctypes.c_int64, ctypes.c_int64, _D]),
            "oracle_sparse_new": (_VP, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, _VP, _I32, _I64,
                                        ctypes.c_int32, ctypes.c_int32]),
            "oracle_centered_new": (_VP, [_VP, _D]),
            "oracle_scaled_new": (_VP, [_VP, _D, ctypes.c_int32, ctypes.c_int32]),
            "oracle_free": (None, [_VP]),
            "oracle_dims": (None, [_VP, _I64, _I64]),
            "oracle_multiply": (ctypes.c_int, [_VP, ctypes.c_int32, _D, _D]),
            "oracle_realize": (ctypes.c_int, [_VP, _D]),
            "oracle_compute": (ctypes.c_int, [_VP, ctypes.c_int64, ctypes.POINTER(OracleOptions), _D, _D, _D, _I64,
                                              _I32, _I32]),
            "oracle_pca": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                          ctypes.POINTER(OracleOptions), _D, _D, _D, _I64, _I32, _I32]),
            "oracle_column_stats": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32, _D, _D]),
            "oracle_lanczos": (ctypes.c_int, [_VP, ctypes.c_int64, _D, _D, _D, ctypes.c_uint64, ctypes.c_int64,
                                              ctypes.POINTER(OracleOptions), _D]),
            "oracle_orthogonalize": (None, [_D, ctypes.c_int64, ctypes.c_int64, _D, ctypes.c_int64]),
            "oracle_update_k": (ctypes.c_int64, [ctypes.c_int64] * 4),
            "oracle_requested_ge_half": (ctypes.c_int32, [ctypes.c_int64, ctypes.c_int64]),
            "oracle_choose_work": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_int32, ctypes.c_int64]),
            "oracle_svd": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, _D, _D, _D, _D]),
            "oracle_partition": (ctypes.c_int, [_VP, _I64, _I64, _I64]),
            "oracle_normal_sampler": (None, [ctypes.c_int64, ctypes.c_int64, _D]),
            "oracle_simulate_compressed_sparse": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_int64, _D, _I32, _I64]),
            "oracle_mt19937_64_nth": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_int64]),
            "oracle_fill_random_normals": (None, [ctypes.c_uint64, ctypes.c_int64, _D]),
            "oracle_gen_nb_sparse": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                                                      ctypes.c_uint64, _I64, ctypes.c_int32, _VP, _I32]),
            "oracle_nb_draw": (ctypes.c_int32, [ctypes.c_uint32]),
            "oracle_density_threshold": (ctypes.c_uint32, [ctypes.c_double]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())


def make_options(**kw):
    """Options with the defaults of Options.hpp:27-82; ``initial`` may be an array."""
    o = OracleOptions()
    o.invariant_subspace_tolerance = kw.get("invariant_subspace_tolerance", -1.0)
    o.convergence_tolerance = kw.get("convergence_tolerance", 1e-5)
    o.singular_value_ratio_tolerance = kw.get("singular_value_ratio_tolerance", -1.0)
    o.extra_work = kw.get("extra_work", 7)
    o.max_iterations = kw.get("max_iterations", 1000)
    o.exact_for_small_matrix = int(kw.get("exact_for_small_matrix", True))
    o.exact_for_large_number = int(kw.get("exact_for_large_number", True))
    o.cap_number = int(kw.get("cap_number", False))
    o.seed = kw.get("seed", 5489)
    init = kw.get("initial")
    keep = None
    if init is not None:
        keep = np.ascontiguousarray(init, dtype=np.float64)
        o.initial = _p(keep)
        o.initial_len = keep.size
    return o, keep


class OMatrix:
    """Handle on an oracle operator (dense, sparse, centred, scaled)."""

    def __init__(self, handle, keep=()):
        self.h = handle
        self._keep = list(keep)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_free(self.h)
            self.h = None

    @property
    def shape(self):
        nr = ctypes.c_int64()
        nc = ctypes.c_int64()
        lib().oracle_dims(self.h, ctypes.byref(nr), ctypes.byref(nc))
        return nr.value, nc.value

    def multiply(self, x, adjoint=False):
        nr, nc = self.shape
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(nc if adjoint else nr)
        _check(lib().oracle_multiply(self.h, int(adjoint), _p(x), _p(out)))
        return out

    def realize(self):
        nr, nc = self.shape
        out = np.empty(nr * nc)
        _check(lib().oracle_realize(self.h, _p(out)))
        return out.reshape(nc, nr).T.copy()

    def partition(self, nthreads):
        nr, nc = self.shape
        # primary dimension is inferred by caller; allocate generously
        prim = max(nr, nc)
        pb = np.zeros(nthreads + 1, dtype=np.int64)
        sb = np.zeros(nthreads + 1, dtype=np.int64)
        snz = np.zeros((nthreads + 1) * prim, dtype=np.int64)
        _check(lib().oracle_partition(self.h, _p(pb, ctypes.c_int64), _p(sb, ctypes.c_int64),
                                      _p(snz, ctypes.c_int64)))
        return pb, sb, snz


def dense(A):
    A = np.asarray(A, dtype=np.float64)
    buf = np.ascontiguousarray(A.T).ravel()  # column-major
    return OMatrix(lib().oracle_dense_new(A.shape[0], A.shape[1], _p(buf)))


def sparse(nrow, ncol, values, indices, pointers, column_major=True, nthreads=1):
    values = np.ascontiguousarray(values)
    f32 = values.dtype == np.float32
    if not f32:
        values = values.astype(np.float64)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    pointers = np.ascontiguousarray(pointers, dtype=np.int64)
    h = lib().oracle_sparse_new(nrow, ncol, int(f32), values.ctypes.data_as(_VP), _p(indices, ctypes.c_int32),
                                _p(pointers, ctypes.c_int64), int(column_major), nthreads)
    return OMatrix(h)


def centered(inner, center):
    c = np.ascontiguousarray(center, dtype=np.float64)
    return OMatrix(lib().oracle_centered_new(inner.h, _p(c)), keep=[inner])


def scaled(inner, scale, column=True, divide=True):
    s = np.ascontiguousarray(scale, dtype=np.float64)
    return OMatrix(lib().oracle_scaled_new(inner.h, _p(s), int(column), int(divide)), keep=[inner])


def compute(mat, number, **opts):
    """compute.hpp:141 -> (U, V, D, converged, iterations)."""
    nr, nc = mat.shape
    o, keep = make_options(**opts)
    n = max(int(number), 0)
    U = np.zeros(nr * n)
    V = np.zeros(nc * n)
    D = np.zeros(max(n, 1))
    nout = ctypes.c_int64()
    conv = ctypes.c_int32()
    it = ctypes.c_int32()
    _check(lib().oracle_compute(mat.h, number, ctypes.byref(o), _p(U), _p(V), _p(D), ctypes.byref(nout),
                                ctypes.byref(conv), ctypes.byref(it)))
    k = nout.value
    return (U[: nr * k].reshape(k, nr).T.copy(), V[: nc * k].reshape(k, nc).T.copy(), D[:k].copy(),
            bool(conv.value), it.value)


def pca(mat, center, scale, number, **opts):
    """pca.hpp:55 -> (scores, rotation, variances, converged, iterations)."""
    nr, nc = mat.shape
    o, keep = make_options(**opts)
    n = max(int(number), 0)
    S = np.zeros(nr * n)
    R = np.zeros(nc * n)
    var = np.zeros(max(n, 1))
    nout = ctypes.c_int64()
    conv = ctypes.c_int32()
    it = ctypes.c_int32()
    _check(lib().oracle_pca(mat.h, int(center), int(scale), number, ctypes.byref(o), _p(S), _p(R), _p(var),
                            ctypes.byref(nout), ctypes.byref(conv), ctypes.byref(it)))
    k = nout.value
    return (S[: nr * k].reshape(k, nr).T.copy(), R[: nc * k].reshape(k, nc).T.copy(), var[:k].copy(),
            bool(conv.value), it.value)


def column_stats(mat, center=True, scale=True):
    nr, nc = mat.shape
    c = np.zeros(nc)
    s = np.zeros(nc)
    _check(lib().oracle_column_stats(mat.h, int(center), int(scale), _p(c), _p(s)))
    return c, s


def lanczos(mat, W, V, B, seed, start, **opts):
    """One sweep of lanczos.hpp:54-143 on copies; returns (W, V, B, F)."""
    nr, nc = mat.shape
    work = W.shape[1]
    Wb = np.ascontiguousarray(W.T).ravel().copy()
    Vb = np.ascontiguousarray(V.T).ravel().copy()
    Bb = np.ascontiguousarray(B.T).ravel().copy()
    F = np.zeros(nc)
    o, keep = make_options(**opts)
    _check(lib().oracle_lanczos(mat.h, work, _p(Wb), _p(Vb), _p(Bb), seed, start, ctypes.byref(o), _p(F)))
    return (Wb.reshape(work, nr).T.copy(), Vb.reshape(work, nc).T.copy(), Bb.reshape(work, work).T.copy(), F)


def orthogonalize(M, vec, ncols):
    M = np.asarray(M, dtype=np.float64)
    buf = np.ascontiguousarray(M.T).ravel()
    v = np.ascontiguousarray(vec, dtype=np.float64).copy()
    lib().oracle_orthogonalize(_p(buf), M.shape[0], M.shape[1], _p(v), ncols)
    return v


def update_k(k, req, nconv, work):
    return lib().oracle_update_k(k, req, nconv, work)


def requested_ge_half(req, smaller):
    return bool(lib().oracle_requested_ge_half(req, smaller))


def choose_work(req, extra, smaller):
    return lib().oracle_choose_work(req, extra, smaller)


def svd(A):
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    p = min(m, n)
    buf = np.ascontiguousarray(A.T).ravel()
    U = np.zeros(m * p)
    S = np.zeros(p)
    V = np.zeros(n * p)
    _check(lib().oracle_svd(m, n, _p(buf), _p(U), _p(S), _p(V)))
    return U.reshape(p, m).T.copy(), S, V.reshape(p, n).T.copy()


def normal_sampler(seed, n):
    out = np.empty(n)
    lib().oracle_normal_sampler(seed, n, _p(out))
    return out


def create_random_matrix(nr, nc, seed=42):
    """tests/src/NormalSampler.h:15-24 (filled column by column)."""
    return normal_sampler(seed, nr * nc).reshape(nc, nr).T.copy()


def create_random_vector(n, seed=50):
    """tests/src/NormalSampler.h:26-32."""
    return normal_sampler(seed, n)


def simulate_compressed_sparse(nr, nc):
    """tests/src/sparse.h:19-44 -> (values, indices, pointers) in CSC."""
    n = lib().oracle_simulate_compressed_sparse(nr, nc, None, None, None)
    vals = np.zeros(n)
    idx = np.zeros(n, dtype=np.int32)
    ptr = np.zeros(nc + 1, dtype=np.int64)
    lib().oracle_simulate_compressed_sparse(nr, nc, _p(vals), _p(idx, ctypes.c_int32), _p(ptr, ctypes.c_int64))
    return vals, idx, ptr


def mt19937_64_nth(seed, n):
    return lib().oracle_mt19937_64_nth(seed, n)


def fill_random_normals(seed, n):
    out = np.zeros(n)
    lib().oracle_fill_random_normals(seed, n, _p(out))
    return out


def gen_nb_sparse(n_minor, n_major, density, seed, line0=0, line1=None, f32=True):
    """Synthetic NB CSC (rows = n_minor) for lines [line0, line1) -> (values, indices, pointers)."""
    if line1 is None:
        line1 = n_major
    nl = line1 - line0
    counts = np.zeros(nl, dtype=np.int64)
    L = lib()
    n = L.oracle_gen_nb_sparse(n_minor, line0, line1, density, seed, _p(counts, ctypes.c_int64), 0, None, None)
    vals = np.zeros(n, dtype=np.float32 if f32 else np.float64)
    idx = np.zeros(n, dtype=np.int32)
    L.oracle_gen_nb_sparse(n_minor, line0, line1, density, seed, _p(counts, ctypes.c_int64), int(f32),
                           vals.ctypes.data_as(_VP), _p(idx, ctypes.c_int32))
    ptr = np.zeros(nl + 1, dtype=np.int64)
    np.cumsum(counts, out=ptr[1:])
    return vals, idx, ptr


def nb_draw(u):
    return lib().oracle_nb_draw(u)


def density_threshold(d):
    return lib().oracle_density_threshold(d)
