"""GPU parity of the value dictionary (irlba_hip_sparse_value_dictionary): the
products stream one code byte per stored value and read the value from a
table of the distinct stored values.  Each product term is the same number as
with the value stream; the order of the sums depends on the code stream's runs.

  * Runs of 4 slots per lane (orientations whose segments are not all whole
    runs of 16) with FP32 storage: the terms are added in the value stream's
    order, so at the same lanes per line the outputs are equal bit for bit (and
    so is a whole solve).
  * Runs of 16, or FP64 storage (whose value stream loads lane pairs from both
    halves of a step): the same terms in another order, relative 1e-14 of the
    output with FP64 accumulation, 2e-6 with FP32 accumulation.
  * Against the oracle's ParallelSparseMatrix restatement (include/irlba/Matrix/
    sparse.hpp:185-267) the value stream's tolerances hold: FP64 accumulation
    1e-12, FP32 accumulation 2e-6 of the output norm.
"""

import contextlib
import os

import numpy as np
import pytest

from compare import assert_equal_vectors, subspace_angle
from oracle import oracle as O

pytestmark = pytest.mark.gpu

irlba = pytest.importorskip("cppirlba_amd")


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@contextlib.contextmanager
def fixed_lanes(lpl):
    """Lanes per line fixed for matrices built (and dictionaries switched on)
    inside the block: the value stream and the codes then sum in one grouping."""
    os.environ["IRLBA_HIP_LPL"] = str(lpl)
    try:
        yield
    finally:
        del os.environ["IRLBA_HIP_LPL"]


@pytest.mark.parametrize("storage,acc", [("f32", "f32"), ("f32", "f64"), ("f64", "f64")])
@pytest.mark.parametrize("nr,nc", [(2000, 30000), (500, 100000), (30000, 2000)])
@pytest.mark.parametrize("lpl", [4, 16])
def test_dictionary_products_match_value_stream_and_oracle(storage, acc, nr, nc, lpl):
    # log1p of negative binomial counts: a few dozen distinct values.  The shapes
    # give one and several LDS tiles in each orientation.
    f32 = storage == "f32"
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 101 + nr, f32=f32)
    with fixed_lanes(lpl):
        H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, storage=storage, accumulate=acc)
        v = O.create_random_vector(nc, 3)
        u = O.create_random_vector(nr, 4)
        y0, z0 = H.multiply(v), H.adjoint_multiply(u)
        n = H.value_dictionary(True)
    distinct = set(np.unique(vals.view(np.uint32 if f32 else np.uint64)).tolist()) | {0}
    assert n == len(distinct)
    runs = []
    for adjoint in (False, True):
        lay = H.layout(adjoint)
        assert lay["dictionary"] == n and lay["delta"] == 1 and lay["lanes_per_line"] == lpl
        assert lay["dictionary_run"] in (4, 16)
        runs.append(lay["dictionary_run"])
    y1, z1 = H.multiply(v), H.adjoint_multiply(u)
    close = 1e-14 if acc == "f64" else 2e-6
    for out, ref, run in ((y1, y0, runs[0]), (z1, z0, runs[1])):
        if f32 and run == 4:
            np.testing.assert_array_equal(out, ref)
        else:
            assert _rel(out, ref) < close
    OM = O.sparse(nr, nc, vals, idx, ptr, True, 1)
    tol = 1e-12 if acc == "f64" else 2e-6
    assert _rel(y1, OM.multiply(v)) < tol
    assert _rel(z1, OM.multiply(u, adjoint=True)) < tol
    # Switched off: the value stream again, the same outputs as before.
    assert H.value_dictionary(False) == 0
    assert H.layout(False)["dictionary"] == 0 and H.layout(True)["dictionary"] == 0
    np.testing.assert_array_equal(H.multiply(v), y0)
    np.testing.assert_array_equal(H.adjoint_multiply(u), z0)
    # On again: the codes are kept, same table.
    assert H.value_dictionary(True) == n


@pytest.mark.parametrize("storage,acc", [("f32", "f32"), ("f64", "f64")])
@pytest.mark.parametrize("lpl", [8, 32, 64])
def test_dictionary_other_group_widths(storage, acc, lpl):
    # The widths the test above leaves out: the step totals of the code
    # stream are broadcast per width (half-row mirror, lane reads).
    f32 = storage == "f32"
    nr, nc = 1500, 40000
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 55 + lpl, f32=f32)
    with fixed_lanes(lpl):
        H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, storage=storage, accumulate=acc)
        v = O.create_random_vector(nc, 13)
        u = O.create_random_vector(nr, 14)
        y0, z0 = H.multiply(v), H.adjoint_multiply(u)
        H.value_dictionary(True)
    assert H.layout(False)["lanes_per_line"] == lpl and H.layout(False)["dictionary"] > 0
    close = 1e-14 if acc == "f64" else 2e-6
    y1, z1 = H.multiply(v), H.adjoint_multiply(u)
    assert _rel(y1, y0) < close and _rel(z1, z0) < close
    OM = O.sparse(nr, nc, vals, idx, ptr, True, 1)
    tol = 1e-12 if acc == "f64" else 2e-6
    assert _rel(y1, OM.multiply(v)) < tol
    assert _rel(z1, OM.multiply(u, adjoint=True)) < tol


def test_dictionary_signed_zero_and_negative_values():
    # -0.0 and +0.0 are distinct bit patterns (two codes); negative values sort first.
    rng = np.random.default_rng(7)
    nr, nc = 3000, 4000
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.04, 17, f32=True)
    vals = vals.copy()
    pick = rng.random(vals.size)
    vals[pick < 0.1] = -0.0
    vals[(pick >= 0.1) & (pick < 0.3)] *= -1.5
    with fixed_lanes(8):
        H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, accumulate="f32")
        v = O.create_random_vector(nc, 5)
        u = O.create_random_vector(nr, 6)
        y0, z0 = H.multiply(v), H.adjoint_multiply(u)
        n = H.value_dictionary(True)
    assert n == len(set(np.unique(vals.view(np.uint32)).tolist()) | {0})
    for out, ref, adjoint in ((H.multiply(v), y0, False), (H.adjoint_multiply(u), z0, True)):
        if H.layout(adjoint)["dictionary_run"] == 4:
            np.testing.assert_array_equal(out, ref)
        else:
            assert _rel(out, ref) < 2e-6


def test_dictionary_pca_equals_value_stream_and_oracle():
    # tests/src/pca.cpp:218-240: the centred+scaled chain over a dictionary-coded
    # matrix.  FP32 storage with runs of 4 in both orientations (short lines):
    # the whole solve equals the value stream's bit for bit.
    nr, nc = 400, 1000
    vals32, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 23, f32=True)
    init = O.create_random_vector(nc, 8)
    opts = irlba.Options(extra_work=20, initial=init, convergence_tolerance=1e-8)
    with fixed_lanes(8):
        H = irlba.HipSparseMatrix(nr, nc, vals32, idx, ptr, accumulate="f32")
        r0 = irlba.pca(H, True, True, 8, opts)
        H.value_dictionary(True)
    assert H.layout(False)["dictionary_run"] == 4 and H.layout(True)["dictionary_run"] == 4
    r1 = irlba.pca(H, True, True, 8, opts)
    assert r1.iterations == r0.iterations
    np.testing.assert_array_equal(r1.variances, r0.variances)
    np.testing.assert_array_equal(r1.rotation, r0.rotation)
    np.testing.assert_array_equal(r1.scores, r0.scores)
    # Longer rows (runs of 16 on A.v): FP64 storage against the oracle with the
    # same start vector, and FP32 storage (FP32 products) within the mixed
    # solve's 1e-5 of it.
    nr, nc = 400, 6000
    vals32, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 23, f32=True)
    vals = vals32.astype(np.float64)
    init = O.create_random_vector(nc, 8)
    opts = irlba.Options(extra_work=20, initial=init, convergence_tolerance=1e-8)
    H64 = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr)
    assert H64.value_dictionary(True) > 0 and H64.layout(False)["dictionary_run"] == 16
    res = irlba.pca(H64, True, True, 8, opts)
    sc, rot, var, conv, it = O.pca(O.sparse(nr, nc, vals, idx, ptr, True, 1), True, True, 8, extra_work=20,
                                   initial=init, convergence_tolerance=1e-8)
    assert res.converged and conv
    assert_equal_vectors(res.variances, var, 1e-8)
    assert subspace_angle(res.rotation, rot) < 1e-5
    H32 = irlba.HipSparseMatrix(nr, nc, vals32, idx, ptr, accumulate="f32")
    H32.value_dictionary(True)
    r32 = irlba.pca(H32, True, True, 8, opts)
    assert_equal_vectors(r32.variances, var, 1e-5)


def test_dictionary_shared_by_decorators():
    # The dictionary belongs to the stored matrix: a CenteredMatrix / ScaledMatrix
    # built on it streams the codes too, with the same results.
    nr, nc = 2000, 5000
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 29, f32=False)
    H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr)
    c = np.linspace(0.1, 0.9, nc)
    s = np.linspace(1.0, 2.0, nc)
    M = irlba.ScaledMatrix(irlba.CenteredMatrix(H, c), s, column=True, divide=True)
    v = O.create_random_vector(nc, 1)
    u = O.create_random_vector(nr, 2)
    y0, z0 = M.multiply(v), M.adjoint_multiply(u)
    assert H.value_dictionary(True) > 0
    assert M.layout(False)["dictionary"] > 0
    assert _rel(M.multiply(v), y0) < 1e-14
    assert _rel(M.adjoint_multiply(u), z0) < 1e-14


def test_dictionary_refusals_leave_the_value_stream():
    # More than 256 distinct values: an error, and the products keep working.
    rng = np.random.default_rng(3)
    nr, nc = 1000, 4000
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 31, f32=False)
    vals = vals + rng.standard_normal(vals.size)
    H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr)
    v = O.create_random_vector(nc, 1)
    y0 = H.multiply(v)
    with pytest.raises(irlba.HipError, match="more than 256 distinct values"):
        H.value_dictionary(True)
    assert H.layout(False)["dictionary"] == 0
    np.testing.assert_array_equal(H.multiply(v), y0)
    # Exactly 256 distinct values (zero among them) are accepted.
    vals256 = (np.arange(vals.size) % 256).astype(np.float64)
    H256 = irlba.HipSparseMatrix(nr, nc, vals256, idx, ptr)
    y256 = H256.multiply(v)
    assert H256.value_dictionary(True) == 256
    assert _rel(H256.multiply(v), y256) < 1e-14
    # 256 non-zero values plus the zero of the padding slots: 257 entries, refused.
    H257 = irlba.HipSparseMatrix(nr, nc, vals256 + 1.0, idx, ptr)
    with pytest.raises(irlba.HipError, match="more than 256 distinct values"):
        H257.value_dictionary(True)
    # Short lines (no 8-bit delta layout) and dense handles are refused.
    svals, sidx, sptr = O.gen_nb_sparse(2000, 2000, 0.002, 5, f32=False)
    with pytest.raises(irlba.HipError, match="delta layout"):
        irlba.HipSparseMatrix(2000, 2000, svals, sidx, sptr).value_dictionary(True)
    with pytest.raises(irlba.HipError, match="not a sparse matrix handle"):
        irlba.HipDenseMatrix(np.ones((30, 20))).value_dictionary(True)


def test_dictionary_csr_input_vs_oracle():
    # Row-major (CSR) input: the same two tiled orientations, built from rows.
    nr, nc = 3000, 20000
    vals, idx, ptr = O.gen_nb_sparse(nc, nr, 0.05, 37, f32=False)  # CSC of the transpose = CSR of the matrix
    H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, column_major=False)
    OM = O.sparse(nr, nc, vals, idx, ptr, False, 1)
    v = O.create_random_vector(nc, 11)
    u = O.create_random_vector(nr, 12)
    assert H.value_dictionary(True) > 0
    assert _rel(H.multiply(v), OM.multiply(v)) < 1e-12
    assert _rel(H.adjoint_multiply(u), OM.multiply(u, adjoint=True)) < 1e-12


def test_dictionary_svd_backward_matches_value_stream():
    # The backward's forward solve streams the codes; its CG block products read
    # the value stream.  Same gradient factors as without the dictionary.
    nr, nc, k = 2000, 6000, 6
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 41, f32=False)
    H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr)
    opts = irlba.Options(extra_work=20, convergence_tolerance=1e-12, initial=O.create_random_vector(nc, 2))
    rng = np.random.default_rng(5)
    w, GU, GV = rng.standard_normal(k), rng.standard_normal((nr, k)), rng.standard_normal((nc, k))
    out = []
    for on in (False, True):
        H.value_dictionary(on)
        res = irlba.compute(H, k, opts)
        bw = irlba.svd_backward(H, res.U, res.D, res.V, w, GU, GV, complement="exact")
        out.append((res, bw.dense()))
    (r0, g0), (r1, g1) = out
    assert_equal_vectors(r1.D, r0.D, 1e-12)
    assert np.abs(g1 - g0).max() <= 1e-7 * np.abs(g0).max()
