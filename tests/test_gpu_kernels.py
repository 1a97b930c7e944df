"""GPU unit parity of the Lanczos building blocks against the oracle:
orthogonalize_vector (CGS), the restart rotation on MFMA, and one full
run_lanczos_bidiagonalization sweep (tests/src/lanczos.cpp on the device)."""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

irlba = pytest.importorskip("cppirlba_amd")


# (300007, 45, 45): long odd-length vectors take the 4-rows-per-thread dot and update kernels;
# (262144, 13, 11) and (600002, 37, 36) even ones the 16-byte (row pair) dots, and the latter
# also the row-pair update.
@pytest.mark.parametrize("rows,cols,n", [(5, 4, 4), (1000, 70, 33), (100003, 40, 40), (300, 9, 0), (300007, 45, 45),
                                         (262144, 13, 11), (600002, 37, 36)])
def test_orthogonalize_vs_oracle(rows, cols, n):
    rng = np.random.default_rng(rows + n)
    M, _ = np.linalg.qr(rng.standard_normal((rows, cols)))
    v = rng.standard_normal(rows)
    out, nrm = irlba.orthogonalize_vector(M, v, n)
    ref = O.orthogonalize(M, v, n)
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-13)
    assert abs(nrm - np.linalg.norm(ref)) <= 1e-12 * np.linalg.norm(ref)
    if n:
        assert np.abs(M[:, :n].T @ out).max() < 1e-12


# Bases of 65536+ rows with k <= 72 take the thread-per-row FMA rotation, the rest the MFMA tiles.
@pytest.mark.parametrize("rows,work,k", [(16, 4, 1), (64, 8, 5), (1000, 70, 67), (777, 27, 20), (50001, 33, 30),
                                         (300, 200, 150), (65536, 9, 3), (200001, 70, 66), (70001, 70, 67), (131073, 80, 72),
                                         (70000, 90, 80)])
def test_rotation_vs_numpy(rows, work, k):
    rng = np.random.default_rng(rows * 7 + k)
    M = rng.standard_normal((rows, work))
    Y = rng.standard_normal((work, k))
    out = irlba.rotate(M, Y)
    np.testing.assert_allclose(out, M @ Y, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("nr,nc,work", [(10, 10, 3), (30, 20, 7), (200, 150, 20), (1000, 300, 40)])
@pytest.mark.parametrize("start", [0, 1])
def test_lanczos_sweep_vs_oracle(nr, nc, work, start):
    A = O.create_random_matrix(nr, nc, 42)
    rng = np.random.default_rng(nr + nc)
    V = np.zeros((nc, work))
    W = np.zeros((nr, work))
    B = np.zeros((work, work))
    if start == 0:
        v0 = rng.standard_normal(nc)
        V[:, 0] = v0 / np.linalg.norm(v0)
    else:
        # a restart state: orthonormal W[:, :1], V[:, :2] from an oracle sweep
        V[:, 0] = rng.standard_normal(nc)
        V[:, 0] /= np.linalg.norm(V[:, 0])
        w2, v2, b2, f2 = O.lanczos(O.dense(A), W[:, :1], V[:, :1], B[:1, :1], 50, 0)
        W[:, :1] = w2
        V[:, :1] = v2
        V[:, 1] = f2 / np.linalg.norm(f2)
        B[0, 0] = b2[0, 0]
    oW, oV, oB, oF = O.lanczos(O.dense(A), W, V, B, 50, start)
    hW, hV, hB, hF = irlba.run_lanczos_bidiagonalization(irlba.HipDenseMatrix(A), W, V, B, 50, start)
    np.testing.assert_allclose(hB, oB, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(hW, oW, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(hV, oV, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(hF, oF, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(hW.T @ hW, np.eye(work), atol=1e-11)
    np.testing.assert_allclose(hV.T @ hV, np.eye(work), atol=1e-11)


@pytest.mark.parametrize("rank", [1, 2, 3, 5])
def test_lanczos_sweep_refills_vs_oracle(rank):
    # An exactly rank-deficient matrix: after `rank` steps the residual norms fall
    # to rounding level and every later step refills F and W_next with normals from
    # std::mt19937_64(seed) (lanczos.hpp:99-109,123-132).  The refilled columns of
    # V and W are those draws orthogonalised, so they pin the engine's draw order
    # and values to the oracle's.  The refill threshold eps^0.8 ~ 3e-13 is absolute
    # (lanczos.hpp:65-66), so the matrix is scaled to unit norm: its rounding-level
    # residuals (~1e-15) then fall clearly below it in both summation orders, where
    # at norm ~50 they sit at ~3e-13 and which side they land on is rounding.
    nr, nc, work = 60, 40, 9
    base = O.create_random_matrix(nr, rank, 4)
    A = np.empty((nr, nc))
    for i in range(nc):
        A[:, i] = base[:, i % rank]
    A /= np.linalg.norm(A, 2)
    V = np.zeros((nc, work))
    W = np.zeros((nr, work))
    B = np.zeros((work, work))
    v0 = np.random.default_rng(rank).standard_normal(nc)
    V[:, 0] = v0 / np.linalg.norm(v0)
    oW, oV, oB, oF = O.lanczos(O.dense(A), W, V, B, 123, 0)
    hW, hV, hB, hF = irlba.run_lanczos_bidiagonalization(irlba.HipDenseMatrix(A), W, V, B, 123, 0)
    assert np.sum(np.abs(np.diag(oB, 1)) < 1e-12) >= 1  # refills happened (R_F = 0)
    np.testing.assert_allclose(hB, oB, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(hV, oV, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(hW, oW, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(hV.T @ hV, np.eye(work), atol=1e-11)


def test_lanczos_inputs_land_after_the_zero_fill():
    # The standalone sweep zero-fills its new bases on the context's stream and
    # then uploads the caller's W and V; both go through that stream, so the
    # upload always lands after the fill.  A restart-state sweep on a 400000-row
    # basis, repeated: every run bit-identical to the first.
    nr, nc, work = 400000, 3000, 24
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.002, 31, f32=False)
    H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr)
    rng = np.random.default_rng(6)
    V = np.zeros((nc, work))
    W = np.zeros((nr, work))
    B = np.zeros((work, work))
    V[:, 0] = rng.standard_normal(nc)
    V[:, 0] /= np.linalg.norm(V[:, 0])
    w1, v1, b1, f1 = irlba.run_lanczos_bidiagonalization(H, W[:, :1], V[:, :1], B[:1, :1], 50, 0)
    W[:, :1], V[:, :1] = w1, v1
    V[:, 1] = f1 / np.linalg.norm(f1)
    B[0, 0] = b1[0, 0]
    first = irlba.run_lanczos_bidiagonalization(H, W, V, B, 50, 1)
    assert np.abs(first[2]).max() > 0
    for _ in range(8):
        again = irlba.run_lanczos_bidiagonalization(H, W, V, B, 50, 1)
        for a, b in zip(again, first):
            np.testing.assert_array_equal(a, b)


def _csc_dense(vals, idx, ptr, nr, nc):
    A = np.zeros((nr, nc))
    for c in range(nc):
        A[idx[ptr[c]:ptr[c + 1]], c] = vals[ptr[c]:ptr[c + 1]]
    return A


@pytest.mark.parametrize("storage,accumulate", [("f64", "f64"), ("f32", "f32"), ("f32", "f64")])
@pytest.mark.parametrize("k", [1, 3, 16, 33, 64, 128])
@pytest.mark.parametrize("decorated", [False, True])
def test_multiply_block_vs_numpy(storage, accumulate, k, decorated):
    # Block products on the tiled layout: 70000 columns span two FP32 tiles of
    # the row orientation; centring + scaling folded in as for the single product.
    nr, nc = 300, 70000
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.02, 3, f32=False)
    A = _csc_dense(vals, idx, ptr, nr, nc)
    H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, storage=storage, accumulate=accumulate)
    ref_A = A.astype(np.float32).astype(np.float64) if storage == "f32" else A
    M = H
    if decorated:
        c, s = irlba.column_stats(H)
        M = irlba.ScaledMatrix(irlba.CenteredMatrix(H, c), s)
        ref_A = (ref_A - c[None, :]) / s[None, :]
    rng = np.random.default_rng(k)
    X = rng.standard_normal((nc, k))
    U = rng.standard_normal((nr, k))
    tol = 1e-12 if storage == "f64" else (2e-6 if accumulate == "f32" else 1e-12)
    Y = M.multiply_block(X)
    Z = M.multiply_block(U, adjoint=True)
    assert np.linalg.norm(Y - ref_A @ X) <= tol * np.linalg.norm(ref_A @ X)
    assert np.linalg.norm(Z - ref_A.T @ U) <= tol * np.linalg.norm(ref_A.T @ U)
    # Each column equals the single-vector product to the accumulation's rounding.
    y0 = M.multiply(X[:, 0])
    assert np.linalg.norm(Y[:, 0] - y0) <= max(tol, 1e-12) * np.linalg.norm(y0)


@pytest.mark.parametrize("storage", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("order", ["C", "F"])
@pytest.mark.parametrize("k", [3, 16, 33])
def test_multiply_block_dense(storage, order, k):
    # Both orientations of both layouts (dot-lines and axpy-lines with chunk
    # partials); k = 33 exceeds one pass and falls back to single products.
    rng = np.random.default_rng(k)
    A = np.asarray(rng.standard_normal((700, 300)), order=order)
    H = irlba.HipDenseMatrix(A.astype(np.float32) if storage == "f32" else A, storage=storage)
    if storage == "bf16":
        import torch
        ref = torch.tensor(A).to(torch.bfloat16).double().numpy()
        tol = 1e-6  # the inputs are rounded to FP32, the stored values to BF16 (in ref)
    elif storage == "f32":
        ref = A.astype(np.float32).astype(np.float64)
        tol = 1e-6
    else:
        ref = A
        tol = 1e-12
    X = rng.standard_normal((300, k))
    U = rng.standard_normal((700, k))
    assert np.linalg.norm(H.multiply_block(X) - ref @ X) <= tol * np.linalg.norm(ref @ X)
    assert np.linalg.norm(H.multiply_block(U, adjoint=True) - ref.T @ U) <= tol * np.linalg.norm(ref.T @ U)


def test_multiply_block_dense_and_errors():
    A = np.random.default_rng(1).standard_normal((50, 40))
    H = irlba.HipDenseMatrix(A)
    X = np.random.default_rng(2).standard_normal((40, 5))
    np.testing.assert_allclose(H.multiply_block(X), A @ X, rtol=1e-12, atol=1e-12)
    c, s = irlba.column_stats(H)
    D = irlba.ScaledMatrix(irlba.CenteredMatrix(H, c), s)
    np.testing.assert_allclose(D.multiply_block(X), ((A - c) / s) @ X, rtol=1e-12, atol=1e-12)
    with pytest.raises(ValueError):
        H.multiply_block(np.ones((39, 2)))
    with pytest.raises(irlba.HipError, match="between 1 and 128"):
        H.multiply_block(np.ones((40, 129)))
