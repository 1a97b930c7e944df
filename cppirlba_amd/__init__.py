"""MI355X-native IRLBA: Python mirror of the irlba:: API over the HIP engine.

Names follow the reference (include/irlba/*.hpp):
  Options            <- Options.hpp:19-83
  compute()/Results  <- compute.hpp:141-308, 357-410
  compute_simple()   <- compute.hpp:333-350, 426-433
  pca()/PcaResults   <- pca.hpp:55-204
  HipSparseMatrix    <- ParallelSparseMatrix (Matrix/sparse.hpp:438-589)
  HipDenseMatrix     <- SimpleMatrix (Matrix/simple.hpp:124-178)
  CenteredMatrix     <- Matrix/centered.hpp:139-204
  ScaledMatrix       <- Matrix/scaled.hpp:195-268
  Matrix             <- Matrix/interface.hpp:141-223 (caller-defined products)

Every product, Gram-Schmidt step and restart rotation runs on the GPU
through libirlba_hip.so; the package has no CPU fallback.
"""

import ctypes
import threading
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from ._lib import BF16, F32, F64, HipError, check, lib

__all__ = [
    "Options", "Results", "PcaResults", "Context", "Matrix", "HipSparseMatrix", "HipDenseMatrix", "CenteredMatrix",
    "ScaledMatrix", "compute", "compute_simple", "pca", "column_stats", "generate_sparse", "default_context",
    "HipError", "F64", "F32", "BF16",
]

_TYPES = {"f64": F64, "float64": F64, "fp64": F64, "f32": F32, "float32": F32, "fp32": F32, "bf16": BF16,
          "bfloat16": BF16}


def _type_code(t):
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return _TYPES[str(t).lower()]


@dataclass
class Options:
    """irlba::Options (Options.hpp:27-82), same names and defaults."""
    invariant_subspace_tolerance: float = -1.0
    convergence_tolerance: float = 1e-5
    singular_value_ratio_tolerance: float = -1.0
    extra_work: int = 7
    max_iterations: int = 1000
    exact_for_small_matrix: bool = True
    exact_for_large_number: bool = True
    cap_number: bool = False
    seed: int = 5489
    initial: Optional[np.ndarray] = None

    def _to_c(self):
        o = _lib.Options()
        o.invariant_subspace_tolerance = self.invariant_subspace_tolerance
        o.convergence_tolerance = self.convergence_tolerance
        o.singular_value_ratio_tolerance = self.singular_value_ratio_tolerance
        o.extra_work = self.extra_work
        o.max_iterations = self.max_iterations
        o.exact_for_small_matrix = int(self.exact_for_small_matrix)
        o.exact_for_large_number = int(self.exact_for_large_number)
        o.cap_number = int(self.cap_number)
        o.seed = self.seed
        keep = None
        if self.initial is not None:
            keep = np.ascontiguousarray(self.initial, dtype=np.float64)
            o.initial = keep.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            o.initial_length = keep.size
        return o, keep


@dataclass
class Results:
    """compute.hpp:357-388."""
    U: np.ndarray
    V: np.ndarray
    D: np.ndarray
    iterations: int
    converged: bool


@dataclass
class PcaResults:
    """pca.hpp:146-179."""
    scores: np.ndarray
    rotation: np.ndarray
    variances: np.ndarray
    iterations: int
    converged: bool


class Context:
    """One GPU and one HIP stream."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().irlba_hip_context_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().irlba_hip_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(lib().irlba_hip_context_synchronize(self.handle))

    def stats(self):
        s = _lib.Stats()
        check(lib().irlba_hip_get_stats(self.handle, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in _lib.Stats._fields_}

    def set_profiling(self, period):
        """Time one product in `period` (1 = all, 0 = none) with HIP events."""
        check(lib().irlba_hip_context_set_profiling(self.handle, int(period)))

    @property
    def stream(self):
        return lib().irlba_hip_context_stream(self.handle)


_default = None
_default_lock = threading.Lock()


def default_context():
    """The shared context on device 0, created once even when several threads ask first."""
    global _default
    if _default is None:
        with _default_lock:
            if _default is None:
                _default = Context(0)
    return _default


def _is_tensor(a):
    return hasattr(a, "data_ptr") and hasattr(a, "is_cuda")


def _ptr(a):
    """(pointer, is_device, keepalive) for a numpy array or a torch tensor."""
    if _is_tensor(a):
        if a.is_cuda:
            import torch

            torch.cuda.synchronize(a.device)  # the engine runs on its own stream
            return ctypes.c_void_p(a.data_ptr()), True, a
        a = a.numpy()
    return ctypes.c_void_p(a.ctypes.data), False, a


def _dtype_name(a):
    return str(a.dtype).replace("torch.", "")


def _sparse_values(x):
    """The values of a compressed matrix as float64 or float32 (the engine's two
    input types): numpy arrays and torch tensors, host or device.  Half and
    bfloat16 widen to float32 and integer types to float64, both exactly;
    anything else (complex, bool) is rejected."""
    name = _dtype_name(x) if hasattr(x, "dtype") else "float64"
    if "complex" in name or name == "bool":
        raise TypeError(f"sparse values must be real numbers, not {name}")
    if _is_tensor(x):
        import torch

        if name in ("float16", "bfloat16"):
            x = x.to(torch.float32)
        elif name not in ("float32", "float64"):
            x = x.to(torch.float64)
        return x.contiguous(), (F32 if x.dtype == torch.float32 else F64)
    x = np.asarray(x)
    if x.dtype == np.float32:
        return np.ascontiguousarray(x), F32
    if x.dtype == np.float16:
        return np.ascontiguousarray(x, dtype=np.float32), F32
    return np.ascontiguousarray(x, dtype=np.float64), F64


def _int_array(a, want, what):
    """A contiguous int32 (indices) or int64 (pointers) array from a numpy array
    or torch tensor; integer inputs of another width are converted after a
    range check, anything else raises."""
    name = _dtype_name(a) if hasattr(a, "dtype") else "int64"
    if "int" not in name or name == "bool":
        raise TypeError(f"sparse {what} must be integers, not {name}")
    lim = 2 ** 31 if want == "int32" else 2 ** 63
    if _is_tensor(a):
        import torch

        tdt = torch.int32 if want == "int32" else torch.int64
        if a.dtype != tdt:
            if a.numel() and (int(a.max()) >= lim or int(a.min()) < -lim):
                raise ValueError(f"sparse {what} do not fit in {want}")
            a = a.to(tdt)
        return a.contiguous()
    a = np.asarray(a)
    if a.dtype != np.dtype(want):
        if a.size and (int(a.max()) >= lim or int(a.min()) < -lim):
            raise ValueError(f"sparse {what} do not fit in {want}")
    return np.ascontiguousarray(a, dtype=want)


class _Handle:
    """A device-resident matrix (opaque irlba_hip_matrix)."""

    def __init__(self, ctx, handle, keep=()):
        self.ctx = ctx
        self.handle = handle
        self._keep = list(keep)

    def __del__(self):
        try:
            if getattr(self, "handle", None) is not None and self.handle.value:
                lib().irlba_hip_matrix_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def rows(self):
        r = ctypes.c_int64()
        check(lib().irlba_hip_matrix_dims(self.handle, ctypes.byref(r), None))
        return r.value

    def cols(self):
        c = ctypes.c_int64()
        check(lib().irlba_hip_matrix_dims(self.handle, None, ctypes.byref(c)))
        return c.value

    @property
    def shape(self):
        return self.rows(), self.cols()

    def multiply(self, right, adjoint=False):
        """Workspace::multiply / AdjointWorkspace::multiply (interface.hpp:49,87)."""
        nr, nc = self.shape
        right = np.ascontiguousarray(right, dtype=np.float64)
        if right.size != (nr if adjoint else nc):
            raise ValueError("vector length does not match the matrix")
        out = np.empty(nc if adjoint else nr)
        check(lib().irlba_hip_multiply(self.handle, int(adjoint), right.ctypes.data, out.ctypes.data, 0))
        return out

    def adjoint_multiply(self, right):
        return self.multiply(right, adjoint=True)

    def multiply_block(self, right, adjoint=False):
        """M X (or M' X) for the columns of X at once (irlba_hip_multiply_block)."""
        nr, nc = self.shape
        X = np.asfortranarray(np.asarray(right, dtype=np.float64))
        if X.ndim != 2 or X.shape[0] != (nr if adjoint else nc):
            raise ValueError("block rows do not match the matrix")
        k = X.shape[1]
        out = np.zeros((nc if adjoint else nr, k), order="F")
        check(lib().irlba_hip_multiply_block(self.handle, int(adjoint), k, X.ctypes.data, out.ctypes.data, 0))
        return out

    def layout(self, adjoint=False):
        """Sparse handles: the device layout that A.v (adjoint=False) or A'.u streams, as a dict
        (irlba_hip_sparse_layout)."""
        from ._lib import Layout

        out = Layout()
        check(lib().irlba_hip_sparse_layout(self.handle, int(adjoint), ctypes.byref(out)))
        return {name: getattr(out, name) for name, _ in Layout._fields_}

    def value_dictionary(self, enable=True):
        """Sparse handles: stream one code byte per stored value, looked up in a table of the distinct
        values, instead of the FP32/FP64 value (lossless; needs at most 256 distinct values and the
        8-bit delta layout).  Acts on the stored matrix, so on every handle that shares it; returns the
        table size, 0 after ``enable=False`` (irlba_hip_sparse_value_dictionary)."""
        n = ctypes.c_int32()
        check(lib().irlba_hip_sparse_value_dictionary(self.handle, int(bool(enable)), ctypes.byref(n)))
        return n.value

    def realize(self):
        """RealizeWorkspace::realize_copy (interface.hpp:125-130)."""
        nr, nc = self.shape
        out = np.empty(nr * nc)
        check(lib().irlba_hip_realize(self.handle, out.ctypes.data, 0))
        return out.reshape(nc, nr).T.copy()


class HipSparseMatrix(_Handle):
    """Device drop-in for ParallelSparseMatrix(nrow, ncol, x, i, p, column_major, num_threads)
    (Matrix/sparse.hpp:438-472), same positional arguments.

    ``num_threads`` is accepted for the reference's signature; the products run
    on the GPU, whose grid replaces the thread partition.  ``x`` (values), ``i``
    (indices) and ``p`` (pointers) are numpy arrays or torch tensors, all on the
    host or all on the device; values are taken as float64 or float32 (half
    precision widens to float32, integers to float64), indices as int32 and
    pointers as int64 (other integer widths are converted after a range check).
    ``storage``: value type kept in HBM ("f64" or "f32"; default: the input's).
    ``accumulate``: "f64" or "f32" (mixed precision); default "f64" for FP64
    storage and "f32" for FP32 storage.
    """

    def __init__(self, nrow, ncol, x, i, p, column_major=True, num_threads=1, *, storage=None, accumulate=None,
                 ctx=None):
        if int(num_threads) < 1:
            raise ValueError("num_threads must be positive")
        x, in_t = _sparse_values(x)
        i = _int_array(i, "int32", "indices")
        p = _int_array(p, "int64", "pointers")
        primary = ncol if column_major else nrow
        if p.ndim != 1 or p.shape[0] != primary + 1:
            raise ValueError("pointers must have one entry per column (or row) plus one")
        xp, xdev, xk = _ptr(x)
        ip, idev, ik = _ptr(i)
        pp, pdev, pk = _ptr(p)
        if not (xdev == idev == pdev):
            raise ValueError("values, indices and pointers must all be on the host or all on the device")
        if xk.shape[0] != ik.shape[0]:
            raise ValueError("values and indices must have the same length")
        st = _type_code(storage) if storage is not None else in_t
        acc = _type_code(accumulate) if accumulate is not None else (F64 if st == F64 else F32)
        ctx = ctx or default_context()
        h = ctypes.c_void_p()
        check(lib().irlba_hip_sparse_create(ctx.handle, nrow, ncol, in_t, xp, ip, pp, int(column_major), st,
                                            int(acc == F64), int(xdev), ctypes.byref(h)))
        super().__init__(ctx, h)
        self._x, self._i, self._p = xk, ik, pk
        self.column_major = bool(column_major)
        self.num_threads = int(num_threads)

    def get_values(self):
        """The values as given to the constructor (sparse.hpp:495-501)."""
        return self._x

    def get_indices(self):
        """The indices as given to the constructor (sparse.hpp:503-510)."""
        return self._i

    def get_pointers(self):
        """The pointers as given to the constructor (sparse.hpp:512-517)."""
        return self._p



class HipDenseMatrix(_Handle):
    """Device drop-in for SimpleMatrix over a dense matrix (simple.hpp:124-178).

    ``storage`` "f64" (FP64 products), "f32" or "bf16" (FP32 accumulation).
    """

    def __init__(self, array, storage="f64", ctx=None):
        if _is_tensor(array):
            import torch

            dt = _dtype_name(array)
            if "complex" in dt or dt == "bool":
                raise TypeError(f"dense values must be real numbers, not {dt}")
            if dt == "float16" or (dt == "bfloat16" and not array.is_cuda):
                array = array.to(torch.float32)  # exact
            elif dt not in ("float64", "float32", "bfloat16"):
                array = array.to(torch.float64)  # integers, exact up to 2^53
        if _is_tensor(array) and array.is_cuda:
            nr, nc = array.shape
            col_major = not array.is_contiguous()
            if col_major and not array.t().is_contiguous():
                array = array.contiguous()
                col_major = False
            in_t = {"float64": F64, "float32": F32, "bfloat16": BF16}[_dtype_name(array)]
            ptr, dev, keep = ctypes.c_void_p(array.data_ptr()), 1, array
        else:
            a = array.numpy() if _is_tensor(array) else np.asarray(array)
            if a.dtype.kind not in "fiu":
                raise TypeError(f"dense values must be real numbers, not {a.dtype}")
            nr, nc = a.shape
            if a.flags.f_contiguous and not a.flags.c_contiguous:
                col_major = True
            else:
                a = np.ascontiguousarray(a)
                col_major = False
            if a.dtype == np.float32:
                in_t = F32
            else:
                a = a.astype(np.float64, copy=False)
                in_t = F64
            ptr, dev, keep = ctypes.c_void_p(a.ctypes.data), 0, a
        ctx = ctx or default_context()
        h = ctypes.c_void_p()
        check(lib().irlba_hip_dense_create(ctx.handle, nr, nc, in_t, ptr, int(col_major), _type_code(storage), dev,
                                           ctypes.byref(h)))
        super().__init__(ctx, h)


class CenteredMatrix(_Handle):
    """CenteredMatrix(matrix, center) (centered.hpp:139-159): A - 1 c'."""

    def __init__(self, matrix, center):
        c = np.ascontiguousarray(center, dtype=np.float64)
        h = ctypes.c_void_p()
        check(lib().irlba_hip_centered_create(matrix.handle, c.ctypes.data, 0, ctypes.byref(h)))
        super().__init__(matrix.ctx, h, keep=[matrix])


class ScaledMatrix(_Handle):
    """ScaledMatrix(matrix, scale, column, divide) (scaled.hpp:195-223)."""

    def __init__(self, matrix, scale, column=True, divide=True):
        s = np.ascontiguousarray(scale, dtype=np.float64)
        h = ctypes.c_void_p()
        check(lib().irlba_hip_scaled_create(matrix.handle, s.ctypes.data, int(column), int(divide), 0,
                                            ctypes.byref(h)))
        super().__init__(matrix.ctx, h, keep=[matrix])


class Matrix:
    """Caller-defined matrix (interface.hpp:141-223): subclass and implement
    rows(), cols(), multiply(x) and adjoint_multiply(x) on host vectors; the
    rest of IRLBA runs on the GPU."""

    def rows(self):
        raise NotImplementedError

    def cols(self):
        raise NotImplementedError

    def multiply(self, right):
        raise NotImplementedError

    def adjoint_multiply(self, right):
        raise NotImplementedError


def _run_compute(matrix, number, options, ctx=None):
    options = options or Options()
    o, keep = options._to_c()
    n = max(int(number), 0)
    nout = ctypes.c_int64()
    conv = ctypes.c_int32()
    it = ctypes.c_int32()
    if isinstance(matrix, _Handle):
        nr, nc = matrix.shape
        U = np.zeros(nr * n)
        V = np.zeros(nc * n)
        D = np.zeros(max(n, 1))
        check(lib().irlba_hip_compute(matrix.handle, number, ctypes.byref(o), U.ctypes.data, V.ctypes.data,
                                      D.ctypes.data, ctypes.byref(nout), ctypes.byref(conv), ctypes.byref(it), 0))
    elif isinstance(matrix, Matrix):
        ctx = ctx or default_context()
        nr, nc = matrix.rows(), matrix.cols()
        err = []

        def cb(user, adjoint, right, output):
            try:
                nin = nr if adjoint else nc
                nout_ = nc if adjoint else nr
                x = np.ctypeslib.as_array(right, shape=(nin,))
                y = matrix.adjoint_multiply(x.copy()) if adjoint else matrix.multiply(x.copy())
                np.ctypeslib.as_array(output, shape=(nout_,))[:] = y
                return 0
            except Exception as e:  # reported after the call
                err.append(e)
                return 1

        fn = _lib.MATVEC_FN(cb)
        U = np.zeros(nr * n)
        V = np.zeros(nc * n)
        D = np.zeros(max(n, 1))
        rc = lib().irlba_hip_compute_callback(ctx.handle, nr, nc, fn, None, number, ctypes.byref(o), U.ctypes.data,
                                              V.ctypes.data, D.ctypes.data, ctypes.byref(nout), ctypes.byref(conv),
                                              ctypes.byref(it))
        if err:
            raise err[0]
        check(rc)
    else:
        raise TypeError("compute() takes a device matrix handle or an irlba Matrix subclass")
    k = nout.value
    # Column-major (Fortran-ordered) views of the engine's output buffers, as the
    # reference's Eigen matrices are laid out; no host transpose.
    return Results(U[: nr * k].reshape(k, nr).T, V[: nc * k].reshape(k, nc).T, D[:k], it.value, bool(conv.value))


def compute(matrix, number, options=None):
    """compute(matrix, number, options) -> Results (compute.hpp:141-308, 403-410)."""
    return _run_compute(matrix, number, options)


def compute_simple(array, number, options=None, storage="f64", ctx=None):
    """compute_simple(matrix, number, options) (compute.hpp:333-350, 426-433)."""
    return _run_compute(HipDenseMatrix(array, storage=storage, ctx=ctx), number, options)


def column_stats(matrix, center=True, scale=True):
    """Column means and standard deviations used by pca() (pca.hpp:66-109)."""
    nc = matrix.cols()
    c = np.zeros(nc)
    s = np.zeros(nc)
    check(lib().irlba_hip_column_stats(matrix.handle, int(center), int(scale), c.ctypes.data, s.ctypes.data, 0))
    return (c if center else None), (s if scale else None)


def pca(matrix, center, scale, number, options=None, storage="f64", ctx=None, device=False):
    """pca(matrix, center, scale, number, options) -> PcaResults (pca.hpp:55-139, 197-204).

    ``matrix`` is a device handle, or a dense array (wrapped like SimpleMatrix).
    The results are column-major numpy arrays, or with ``device=True`` float64
    torch tensors left in the GPU's memory (no transfer over PCIe)."""
    if not isinstance(matrix, _Handle):
        matrix = HipDenseMatrix(matrix, storage=storage, ctx=ctx)
    options = options or Options()
    o, keep = options._to_c()
    nr, nc = matrix.shape
    n = max(int(number), 0)
    if device:
        import torch

        dev = torch.device("cuda", matrix.ctx.device)
        S = torch.empty(max(n, 1) * nr, dtype=torch.float64, device=dev)
        R = torch.empty(max(n, 1) * nc, dtype=torch.float64, device=dev)
        var = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
        torch.cuda.synchronize(dev)  # the engine writes them on its own stream
        ptrs = (S.data_ptr(), R.data_ptr(), var.data_ptr())
    else:
        S = np.zeros(nr * n)
        R = np.zeros(nc * n)
        var = np.zeros(max(n, 1))
        ptrs = (S.ctypes.data, R.ctypes.data, var.ctypes.data)
    nout = ctypes.c_int64()
    conv = ctypes.c_int32()
    it = ctypes.c_int32()
    check(lib().irlba_hip_pca(matrix.handle, int(center), int(scale), number, ctypes.byref(o), *ptrs,
                              ctypes.byref(nout), ctypes.byref(conv), ctypes.byref(it), int(bool(device))))
    k = nout.value
    # Column-major views (pca.hpp:154-160 lays them out as Eigen matrices): no transpose copy.
    return PcaResults(S[: nr * k].reshape(k, nr).T, R[: nc * k].reshape(k, nc).T, var[:k], it.value,
                      bool(conv.value))


def generate_sparse(nrow, ncol, density, seed, storage="f32", accumulate=None, ctx=None):
    """Synthetic CSC built on the device: Bernoulli(density) positions, values
    log1p(NB(mu=2, phi=0.5)); identical to the CPU generator for the same seed."""
    ctx = ctx or default_context()
    st = _type_code(storage)
    acc = _type_code(accumulate) if accumulate is not None else (F64 if st == F64 else F32)
    h = ctypes.c_void_p()
    check(lib().irlba_hip_generate_sparse(ctx.handle, nrow, ncol, density, seed, st, int(acc == F64),
                                          ctypes.byref(h)))
    return _Handle(ctx, h)


def generate_csc_host(nrow, ncol, density, seed, value_type="f32", ctx=None):
    """The device generator's CSC copied to host arrays (values, indices, pointers)."""
    import torch

    ctx = ctx or default_context()
    vt = _type_code(value_type)
    nnz = ctypes.c_int64()
    check(lib().irlba_hip_generate_csc(ctx.handle, nrow, ncol, density, seed, vt, None, None, None,
                                       ctypes.byref(nnz)))
    dev = torch.device("cuda", ctx.device)
    vals = torch.empty(nnz.value, dtype=torch.float32 if vt == F32 else torch.float64, device=dev)
    idx = torch.empty(nnz.value, dtype=torch.int32, device=dev)
    ptr = torch.empty(ncol + 1, dtype=torch.int64, device=dev)
    check(lib().irlba_hip_generate_csc(ctx.handle, nrow, ncol, density, seed, vt, ctypes.c_void_p(vals.data_ptr()),
                                       ctypes.c_void_p(idx.data_ptr()), ctypes.c_void_p(ptr.data_ptr()),
                                       ctypes.byref(nnz)))
    ctx.synchronize()
    return vals.cpu().numpy(), idx.cpu().numpy(), ptr.cpu().numpy()


def _colmajor(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def run_lanczos_bidiagonalization(matrix, W, V, B, seed, start, options=None):
    """One sweep of lanczos.hpp:54-143 on the device; returns (W, V, B, F)."""
    options = options or Options()
    o, keep = options._to_c()
    W = _colmajor(W).copy(order="F")
    V = _colmajor(V).copy(order="F")
    B = _colmajor(B).copy(order="F")
    F = np.zeros(matrix.cols())
    check(lib().irlba_hip_lanczos(matrix.handle, W.shape[1], W.ctypes.data, V.ctypes.data, B.ctypes.data, seed, start,
                                  ctypes.byref(o), F.ctypes.data))
    return W, V, B, F


def orthogonalize_vector(M, vec, ncols, ctx=None):
    """lanczos.hpp:17-21 on the device: vec - M[:, :ncols] M[:, :ncols]' vec; returns (vec, norm)."""
    ctx = ctx or default_context()
    M = _colmajor(M)
    v = np.ascontiguousarray(vec, dtype=np.float64).copy()
    nrm = ctypes.c_double()
    check(lib().irlba_hip_orthogonalize(ctx.handle, M.ctypes.data, M.shape[0], M.shape[1], v.ctypes.data, ncols,
                                        ctypes.byref(nrm)))
    return v, nrm.value


def rotate(M, Y, ctx=None):
    """The restart rotation M Y (compute.hpp:268,279) on the matrix cores."""
    ctx = ctx or default_context()
    M = _colmajor(M)
    Y = _colmajor(Y)
    out = np.zeros((M.shape[0], Y.shape[1]), order="F")
    check(lib().irlba_hip_rotate(ctx.handle, M.ctypes.data, M.shape[0], M.shape[1], Y.ctypes.data, Y.shape[1],
                                 out.ctypes.data))
    return out


class SvdBackward:
    """Factored gradient dL/dA = X Y' of a top-k SVD (see svd_backward)."""

    def __init__(self, X, Y, grad_center, cg_iterations):
        self.X = X
        self.Y = Y
        self.grad_center = grad_center
        self.cg_iterations = cg_iterations

    def dense(self):
        return self.X @ self.Y.T


_COMPLEMENT = {"none": 0, "rank_k": 1, "exact": 2}


def svd_backward(matrix, U, D, V, grad_D=None, grad_U=None, grad_V=None, complement="exact",
                 max_cg_iterations=1000, cg_tolerance=1e-12, grad_center=False):
    """Backward of the top-k SVD of ``matrix`` (U rows x k, D k, V cols x k).

    Returns dL/dA for L with gradients grad_D, grad_U, grad_V (None = zero) as
    ``SvdBackward`` (X rows x 2k, Y cols x 2k).  ``complement``: "none",
    "rank_k" (exact when the matrix has rank k) or "exact" (CG solves through the
    matrix's own products).  ``grad_center``: also return -Y (X' 1), the gradient
    with respect to a column-centre vector subtracted from the matrix.
    """
    nr, nc = matrix.shape
    U = _colmajor(U)
    V = _colmajor(V)
    D = np.ascontiguousarray(D, dtype=np.float64)
    k = D.size
    if U.shape != (nr, k) or V.shape != (nc, k):
        raise ValueError("U, D, V do not match the matrix")

    def opt(a, shape):
        if a is None:
            return None
        a = _colmajor(a)
        if a.shape != shape and not (len(shape) == 1 and a.size == shape[0]):
            raise ValueError("gradient shape does not match")
        return a

    gD, gU, gV = opt(grad_D, (k,)), opt(grad_U, (nr, k)), opt(grad_V, (nc, k))
    X = np.zeros((nr, 2 * k), order="F")
    Y = np.zeros((nc, 2 * k), order="F")
    gc = np.zeros(nc) if grad_center else None
    its = ctypes.c_int32()
    p = lambda a: None if a is None else a.ctypes.data  # noqa: E731
    check(lib().irlba_hip_svd_backward(matrix.handle, k, p(U), p(D), p(V), p(gD), p(gU), p(gV),
                                       _COMPLEMENT[complement], int(max_cg_iterations), float(cg_tolerance),
                                       p(X), p(Y), p(gc), ctypes.byref(its), 0))
    return SvdBackward(X, Y, gc, its.value)


def sddmm(nrow, ncol, indices, pointers, X, Y, column_major=True, out_type="f64", ctx=None):
    """out[e] = X[row_e] . Y[col_e] over the stored entries of a compressed pattern.

    ``out_type`` "f64" or "f32" (FP32 factors, FP64 accumulation, FP32 output).
    """
    X = _colmajor(X)
    Y = _colmajor(Y)
    if X.shape[0] != nrow or Y.shape[0] != ncol or X.shape[1] != Y.shape[1]:
        raise ValueError("factor shapes do not match the pattern")
    idx = _int_array(indices, "int32", "indices")
    ptr = _int_array(pointers, "int64", "pointers")
    if _is_tensor(idx) or _is_tensor(ptr):
        idx, ptr = np.asarray(idx.cpu() if _is_tensor(idx) else idx), np.asarray(ptr.cpu() if _is_tensor(ptr) else ptr)
    if ptr.ndim != 1 or ptr.size != (ncol if column_major else nrow) + 1:
        raise ValueError("pointers have the wrong length")
    # The output has one slot per stored index: the pattern must address exactly
    # those, or the kernel would read past the indices and write past the output.
    if ptr[0] != 0 or ptr[-1] != idx.size or (ptr.size > 1 and np.any(np.diff(ptr) < 0)):
        raise ValueError("sparse pointers must start at 0, be non-decreasing and end at the number of indices")
    code = _type_code(out_type)
    out = np.zeros(idx.size, dtype=np.float32 if code == F32 else np.float64)
    ctx = ctx or default_context()
    check(lib().irlba_hip_sddmm(ctx.handle, nrow, ncol, idx.ctypes.data, ptr.ctypes.data, int(column_major),
                                X.shape[1], X.ctypes.data, Y.ctypes.data, out.ctypes.data, code, 0))
    return out


# ---------------------------------------------------------------- libtorch autograd binding
_torch_ops_loaded = False


def _torch_ops():
    """torch.ops.irlba_hip (lib/libirlba_torch.so: a torch::autograd::Function in C++)."""
    global _torch_ops_loaded
    import os

    import torch

    if not _torch_ops_loaded:
        lib()  # the engine first, bound to torch's HIP runtime
        path = os.path.join(os.path.dirname(_lib.LIB_PATH), "libirlba_torch.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with __graft_entry__.build()")
        torch.ops.load_library(path)
        _torch_ops_loaded = True
    return torch.ops.irlba_hip


def _op_args(options, storage, complement):
    o = options or Options()
    return (int(o.extra_work), float(o.convergence_tolerance), int(o.max_iterations), int(o.seed),
            _type_code(storage), _COMPLEMENT[complement])


def svd_sparse(values, indices, pointers, nrow, ncol, number, center=None, column_major=True, options=None,
               storage="f64", complement="exact"):
    """Differentiable top-``number`` SVD of a compressed sparse matrix (torch tensors).

    Returns (U, S, V) in float64; gradients flow to ``values`` (on the stored
    entries) and to ``center`` (the matrix is A - 1 center' when given).
    """
    return _torch_ops().svd_sparse(values, indices, pointers, int(nrow), int(ncol), bool(column_major), center,
                                   int(number), *_op_args(options, storage, complement))


def svd_dense(matrix, number, center=None, options=None, storage="f64", complement="exact"):
    """Differentiable top-``number`` SVD of a dense torch matrix; gradients to ``matrix`` and ``center``."""
    return _torch_ops().svd_dense(matrix, center, int(number), *_op_args(options, storage, complement))
