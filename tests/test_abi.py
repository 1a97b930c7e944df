"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every function that include/irlba_hip.h declares (no compute calls)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "irlba_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:[\w\*]+\s+)+\**(irlba_hip_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["irlba_hip_compute", "irlba_hip_pca", "irlba_hip_multiply", "irlba_hip_sparse_create",
                 "irlba_hip_dense_create", "irlba_hip_centered_create", "irlba_hip_scaled_create",
                 "irlba_hip_realize", "irlba_hip_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from cppirlba_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libirlba_hip.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # every symbol the Python binding uses is declared in the header
    assert set(_lib.SIGNATURES) <= set(declared_functions())


def test_options_defaults_match_reference():
    from cppirlba_amd import _lib
    import cppirlba_amd as irlba

    L = _lib.lib()
    o = _lib.Options()
    L.irlba_hip_options_default(ctypes.byref(o))
    py = irlba.Options()
    # Options.hpp:27-82
    assert o.invariant_subspace_tolerance == -1 == py.invariant_subspace_tolerance
    assert o.convergence_tolerance == 1e-5 == py.convergence_tolerance
    assert o.singular_value_ratio_tolerance == -1 == py.singular_value_ratio_tolerance
    assert o.extra_work == 7 == py.extra_work
    assert o.max_iterations == 1000 == py.max_iterations
    assert o.exact_for_small_matrix == 1 and o.exact_for_large_number == 1 and o.cap_number == 0
    assert o.seed == 5489 == py.seed
    assert ctypes.sizeof(o) == 72


def test_no_device_reports_error_not_crash():
    """Without a GPU, creating a context fails with a message (no fallback)."""
    import cppirlba_amd as irlba

    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(irlba.HipError):
        irlba.Context(0)


def test_null_handles_report_errors():
    """Entry points given a null context or matrix return an error status with a
    message instead of dereferencing it; the destroy calls accept null."""
    from cppirlba_amd._lib import Stats, lib

    L = lib()
    i64 = ctypes.c_int64
    r, c = i64(), i64()
    n32 = ctypes.c_int32()
    st = Stats()
    calls = {
        "context_synchronize": lambda: L.irlba_hip_context_synchronize(None),
        "get_stats": lambda: L.irlba_hip_get_stats(None, ctypes.byref(st)),
        "set_profiling": lambda: L.irlba_hip_context_set_profiling(None, 1),
        "matrix_dims": lambda: L.irlba_hip_matrix_dims(None, ctypes.byref(r), ctypes.byref(c)),
        "sparse_value_dictionary": lambda: L.irlba_hip_sparse_value_dictionary(None, 1, ctypes.byref(n32)),
    }
    for name, call in calls.items():
        assert call() != 0, name
        assert L.irlba_hip_last_error().decode().startswith("null"), name
    assert L.irlba_hip_context_destroy(None) == 0
    assert L.irlba_hip_matrix_destroy(None) == 0


def test_torch_binding_registers_operators():
    """libirlba_torch.so (the torch::autograd::Function binding) loads and
    registers its operators; without a GPU a call raises instead of falling back."""
    torch = pytest.importorskip("torch")
    import cppirlba_amd as irlba

    ops = irlba._torch_ops()
    assert "svd_sparse(Tensor values" in str(ops.svd_sparse.default._schema)
    assert "svd_dense(Tensor matrix" in str(ops.svd_dense.default._schema)
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError):
            irlba.svd_dense(torch.randn(6, 4, dtype=torch.float64, requires_grad=True), 2)


def test_negative_number_is_rejected_before_any_work():
    """compute()/pca() with a negative count fail with a message (the reference
    sizes its outputs from `number`, compute.hpp:141-160); checked before the
    matrix is touched, so no device is needed."""
    from cppirlba_amd._lib import lib

    L = lib()
    i64, i32 = ctypes.c_int64(), ctypes.c_int32()
    d = ctypes.c_double()
    assert L.irlba_hip_compute(None, -1, None, None, None, None, ctypes.byref(i64), ctypes.byref(i32),
                               ctypes.byref(i32), 0) != 0
    assert "non-negative" in L.irlba_hip_last_error().decode()
    assert L.irlba_hip_pca(None, 1, 1, -3, None, None, None, None, ctypes.byref(i64), ctypes.byref(i32),
                           ctypes.byref(i32), 0) != 0
    assert "non-negative" in L.irlba_hip_last_error().decode()
    assert d.value == 0


def test_sparse_inputs_are_validated_on_the_host():
    """HipSparseMatrix takes float64/float32 values, int32 indices and int64
    pointers; other types are converted when that is exact and rejected
    otherwise, before any device call."""
    import numpy as np

    import cppirlba_amd as irlba

    ptr = np.array([0, 2, 3])
    idx = np.array([0, 3, 1])
    vals = np.array([1.0, 2.0, 3.0])
    with pytest.raises(TypeError, match="real numbers"):
        irlba.HipSparseMatrix(4, 2, vals.astype(np.complex128), idx, ptr)
    with pytest.raises(TypeError, match="real numbers"):
        irlba.HipSparseMatrix(4, 2, vals > 0, idx, ptr)
    with pytest.raises(TypeError, match="integers"):
        irlba.HipSparseMatrix(4, 2, vals, idx.astype(np.float64), ptr)
    with pytest.raises(TypeError, match="integers"):
        irlba.HipSparseMatrix(4, 2, vals, idx, ptr.astype(np.float32))
    with pytest.raises(ValueError, match="do not fit in int32"):
        irlba.HipSparseMatrix(4, 2, vals, np.array([0, 2 ** 32 + 1, 1]), ptr)
    with pytest.raises(ValueError, match="one entry per column"):
        irlba.HipSparseMatrix(4, 3, vals, idx, ptr)
    with pytest.raises(ValueError, match="same length"):
        irlba.HipSparseMatrix(4, 2, vals[:2], idx, ptr)
    with pytest.raises(ValueError, match="num_threads"):
        irlba.HipSparseMatrix(4, 2, vals, idx, ptr, True, 0)
    # Exact conversions.
    assert irlba._int_array(idx, "int32", "indices").dtype == np.int32
    assert irlba._int_array(idx.astype(np.int32), "int64", "pointers").dtype == np.int64
    v, t = irlba._sparse_values(vals.astype(np.float16))
    assert v.dtype == np.float32 and t == irlba.F32
    v, t = irlba._sparse_values(np.array([1, 2, 3], dtype=np.int16))
    assert v.dtype == np.float64 and t == irlba.F64 and v.tolist() == [1.0, 2.0, 3.0]


def test_torch_inputs_are_validated_on_the_host():
    torch = pytest.importorskip("torch")
    import cppirlba_amd as irlba

    i64 = torch.tensor([0, 3, 1])
    out = irlba._int_array(i64, "int32", "indices")
    assert out.dtype == torch.int32 and out.tolist() == [0, 3, 1]
    out = irlba._int_array(torch.tensor([0, 2, 3], dtype=torch.int32), "int64", "pointers")
    assert out.dtype == torch.int64
    with pytest.raises(ValueError, match="do not fit in int32"):
        irlba._int_array(torch.tensor([0, 2 ** 31]), "int32", "indices")
    with pytest.raises(TypeError, match="integers"):
        irlba._int_array(torch.tensor([0.0, 1.0]), "int32", "indices")
    v, t = irlba._sparse_values(torch.tensor([1.0, 2.5], dtype=torch.bfloat16))
    assert v.dtype == torch.float32 and t == irlba.F32 and v.tolist() == [1.0, 2.5]
    v, t = irlba._sparse_values(torch.tensor([1, 7], dtype=torch.int32))
    assert v.dtype == torch.float64 and t == irlba.F64
    with pytest.raises(TypeError, match="real numbers"):
        irlba._sparse_values(torch.tensor([1 + 1j]))
    with pytest.raises(TypeError, match="real numbers"):
        irlba.HipDenseMatrix(torch.ones(3, 2, dtype=torch.bool))


def test_dense_inputs_are_validated_on_the_host():
    import numpy as np

    import cppirlba_amd as irlba

    with pytest.raises(TypeError, match="real numbers"):
        irlba.HipDenseMatrix(np.ones((3, 2), dtype=np.complex64))
    with pytest.raises(TypeError, match="real numbers"):
        irlba.HipDenseMatrix(np.ones((3, 2), dtype=bool))


def test_sddmm_pattern_is_checked_on_the_host():
    """The SDDMM writes one value per stored index: pointers that do not end at
    the index count (or do not start at 0, or decrease) are rejected before the
    kernel could read past the indices or write past the output."""
    import numpy as np

    import cppirlba_amd as irlba

    X, Y = np.ones((4, 2)), np.ones((3, 2))
    idx = np.array([0, 1, 2, 3])
    for ptr in ([0, 2, 3, 5], [1, 2, 3, 4], [0, 3, 2, 4], [0, 1, 2, 3]):
        with pytest.raises(ValueError, match="pointers"):
            irlba.sddmm(4, 3, idx, np.array(ptr), X, Y)
