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
    st = Stats()
    calls = {
        "context_synchronize": lambda: L.irlba_hip_context_synchronize(None),
        "get_stats": lambda: L.irlba_hip_get_stats(None, ctypes.byref(st)),
        "set_profiling": lambda: L.irlba_hip_context_set_profiling(None, 1),
        "matrix_dims": lambda: L.irlba_hip_matrix_dims(None, ctypes.byref(r), ctypes.byref(c)),
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
