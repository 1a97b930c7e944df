"""Pins the CPU oracle (oracle/irlba_oracle.cpp) before it is trusted as the
parity checker: the reference's own known-answer tests, LAPACK (numpy), and
closed forms.  No GPU needed."""

import numpy as np
import pytest

from compare import (assert_equal_column_vectors, assert_equal_matrix, assert_equal_vectors, same_same)
from oracle import oracle as O


# ---------------------------------------------------------------- RNG facts
def test_mt19937_64_standard_value():
    # The C++ standard fixes the 10000th output of a default-seeded mt19937_64.
    assert O.mt19937_64_nth(5489, 10000) == 9981545732273789042


def test_normal_sampler_determinism():
    # tests/src/utils.cpp:7-19
    a = O.normal_sampler(10, 2)
    b = O.normal_sampler(10, 1)
    assert a[0] != a[1]
    assert a[0] == b[0]


def test_fill_random_normals_odd_even():
    # tests/src/utils.cpp:21-65: non-zero, distinct values for even and odd lengths.
    for n in (10, 11):
        v = O.fill_random_normals(1000, n)
        assert np.all(v != 0)
        assert len(np.unique(v)) == n
    # Box-Muller moments on a large draw (mean 0, variance 1).
    v = O.fill_random_normals(7, 200001)
    assert abs(v.mean()) < 0.01 and abs(v.var() - 1) < 0.01


# ---------------------------------------------------------------- CGS
def test_orthogonalize_known_answer():
    # tests/src/lanczos.cpp:13-51 (matrix from svd(matrix(rnorm(20), 5, 4))$u in R).
    m = np.array([
        [0.1186600, 0.9038716, 0.10194199, 0.2484399218],
        [-0.3090796, -0.1659502, 0.52915799, -0.3366910627],
        [-0.3107643, -0.2636561, -0.35950860, 0.7248074228],
        [-0.1683330, 0.1634341, -0.75523159, -0.5473305105],
        [0.8749166, -0.2434161, -0.09989224, -0.0004957394],
    ])
    v = np.array([-0.24054848, -0.04785069, -0.76491749, -0.65634291, 0.62815141])
    out = O.orthogonalize(m, v, 4)
    assert np.sum(out * out) > 0.1
    assert np.all(np.abs(m.T @ out) < 1e-6)  # the R matrix is only orthonormal to ~1e-7
    # exactly the CGS formula v - M (M' v)
    np.testing.assert_allclose(out, v - m @ (m.T @ v), rtol=0, atol=1e-15)
    out2 = O.orthogonalize(m, v, 2)
    assert not np.array_equal(out, out2)


# ---------------------------------------------------------------- Lanczos
@pytest.mark.parametrize("nr", [10, 20, 30])
@pytest.mark.parametrize("nc", [10, 20, 30])
@pytest.mark.parametrize("work", [3, 5, 7])
def test_lanczos_orthonormal_and_bidiagonal(nr, nc, work):
    # tests/src/lanczos.cpp:54-117 plus the Baglama-Reichel relations
    # A V_m = W_m B_m and A' W_m = V_m B_m' + F e_m'.
    A = O.create_random_matrix(nr, nc, 42)
    rest = O.normal_sampler(42, nr * nc + nc)[nr * nc:]
    V = np.zeros((nc, work))
    V[:, 0] = rest / np.linalg.norm(rest)
    W = np.zeros((nr, work))
    B = np.zeros((work, work))
    M = O.dense(A)
    W, V, B, F = O.lanczos(M, W, V, B, 50, 0)
    np.testing.assert_allclose(W.T @ W, np.eye(work), atol=1e-11)
    np.testing.assert_allclose(V.T @ V, np.eye(work), atol=1e-11)
    np.testing.assert_allclose(A @ V, W @ B, atol=1e-10)
    e = np.zeros(work)
    e[-1] = 1
    np.testing.assert_allclose(A.T @ W, V @ B.T + np.outer(F, e), atol=1e-10)
    assert np.allclose(np.tril(B, -1), 0) and np.allclose(np.triu(B, 2), 0)


@pytest.mark.parametrize("nr,nc,work", [(20, 30, 7), (30, 10, 5), (10, 20, 6)])
def test_lanczos_restart_from_mid(nr, nc, work):
    # tests/src/lanczos.cpp:119-163: a sweep restarted at `mid` reproduces the full sweep.
    A = O.create_random_matrix(nr, nc, 42)
    v0 = O.normal_sampler(3, nc)
    V = np.zeros((nc, work))
    V[:, 0] = v0 / np.linalg.norm(v0)
    W = np.zeros((nr, work))
    B = np.zeros((work, work))
    M = O.dense(A)
    mid = work // 2
    sW, sV, sB, F = O.lanczos(M, W[:, :mid], V[:, :mid], B[:mid, :mid], 50, 0)
    cW = np.zeros((nr, work))
    cV = np.zeros((nc, work))
    cB = np.zeros((work, work))
    cW[:, :mid] = sW
    cV[:, :mid] = sV
    cB[:mid, :mid] = sB
    cV[:, mid] = F / np.linalg.norm(F)
    cW, cV, cB, _ = O.lanczos(M, cW, cV, cB, 50, mid)
    fW, fV, fB, _ = O.lanczos(M, W, V, B, 50, 0)
    np.testing.assert_allclose(cW, fW, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(cV, fV, rtol=1e-6, atol=1e-9)


def test_lanczos_null_space_start_throws():
    # lanczos.hpp:82-84 (no reference test covers it).
    A = np.zeros((10, 8))
    A[:, 0] = 1.0
    V = np.zeros((8, 4))
    V[1, 0] = 1.0  # A v = 0
    with pytest.raises(RuntimeError, match="starting vector near the null space"):
        O.lanczos(O.dense(A), np.zeros((10, 4)), V, np.zeros((4, 4)), 1, 0)


# ---------------------------------------------------------------- integer helpers
def test_requested_vs_half_smaller():
    # tests/src/compute.cpp:33-41
    assert O.requested_ge_half(10, 10)
    assert not O.requested_ge_half(10, 30)
    assert O.requested_ge_half(10, 19)
    assert O.requested_ge_half(10, 20)
    assert not O.requested_ge_half(10, 21)
    assert not O.requested_ge_half(10, 22)


def test_choose_requested_plus_extra_work():
    # tests/src/compute.cpp:43-50
    for smaller, want in [(5, 5), (10, 10), (15, 15), (25, 25), (30, 30), (35, 30)]:
        assert O.choose_work(10, 20, smaller) == want


def test_update_k_truth_table():
    # tests/src/compute.cpp:52-76
    def ref(k, req, conv, work):
        if k < req + conv:
            k = req + conv
        if k > work - 3:
            k = work - 3
        if k < 1:
            k = 1
        return k

    vals = [0, 5, 10, 25, 30]
    for k in vals:
        for r in vals:
            for c in vals:
                for w in vals:
                    assert O.update_k(k, r, c, w) == ref(k, r, c, w)


# ---------------------------------------------------------------- SVD
@pytest.mark.parametrize("m,n", [(20, 10), (10, 20), (70, 70), (7, 7), (100, 3)])
def test_jacobi_svd_vs_lapack(m, n):
    rng = np.random.default_rng(m * 100 + n)
    A = rng.standard_normal((m, n))
    U, S, V = O.svd(A)
    Ul, Sl, Vlt = np.linalg.svd(A, full_matrices=False)
    np.testing.assert_allclose(S, Sl, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(U @ np.diag(S) @ V.T, A, atol=1e-12)
    np.testing.assert_allclose(U.T @ U, np.eye(min(m, n)), atol=1e-12)
    np.testing.assert_allclose(V.T @ V, np.eye(min(m, n)), atol=1e-12)
    assert_equal_column_vectors(U, Ul, 1e-8)


def test_jacobi_svd_rank_deficient_completion():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((12, 3)) @ rng.standard_normal((3, 8))
    U, S, V = O.svd(A)
    assert np.all(S[3:] < 1e-12)
    np.testing.assert_allclose(U.T @ U, np.eye(8), atol=1e-10)
    np.testing.assert_allclose(U @ np.diag(S) @ V.T, A, atol=1e-12)


def test_jacobi_svd_broken_arrow():
    # The shape of B after a restart (compute.hpp:282-293): diag + column k + bidiagonal tail.
    rng = np.random.default_rng(9)
    work, k = 12, 7
    B = np.zeros((work, work))
    B[np.arange(k), np.arange(k)] = np.sort(rng.random(k))[::-1] * 10
    B[:k, k] = rng.standard_normal(k) * 1e-3
    for j in range(k, work):
        B[j, j] = rng.random()
        if j + 1 < work:
            B[j, j + 1] = rng.random()
    U, S, V = O.svd(B)
    np.testing.assert_allclose(S, np.linalg.svd(B, compute_uv=False), rtol=1e-12)
    np.testing.assert_allclose(U @ np.diag(S) @ V.T, B, atol=1e-13)


# ---------------------------------------------------------------- compute
def test_compute_compare_to_exact():
    # tests/src/compute.cpp:15-31 (work == smaller gives a near-exact match).
    A = O.create_random_matrix(20, 10)
    U, V, D, conv, it = O.compute(O.dense(A), 5, exact_for_large_number=False)
    Ul, Sl, Vlt = np.linalg.svd(A, full_matrices=False)
    assert_equal_vectors(D, Sl[:5], 1e-8)
    assert_equal_column_vectors(U, Ul[:, :5], 1e-8)
    assert_equal_column_vectors(V, Vlt.T[:, :5], 1e-8)


@pytest.mark.parametrize("nr", [20, 50, 100])
@pytest.mark.parametrize("nc", [20, 50, 100])
@pytest.mark.parametrize("rank", [2, 5, 10])
def test_compute_sweep(nr, nc, rank):
    # tests/src/compute.cpp:78-123
    A = O.create_random_matrix(nr, nc)
    M = O.dense(A)
    U, V, D, conv, it = O.compute(M, rank, convergence_tolerance=1e-9)
    assert U.shape == (nr, rank) and V.shape == (nc, rank) and D.shape == (rank,)
    Ul, Sl, Vlt = np.linalg.svd(A, full_matrices=False)
    assert_equal_vectors(D, Sl[:rank], 1e-6)
    assert_equal_column_vectors(U, Ul[:, :rank], 1e-6)
    assert_equal_column_vectors(V, Vlt.T[:, :rank], 1e-6)
    init = O.create_random_vector(nc, 1239 + nr + nc + rank)
    _, _, D2, _, _ = O.compute(M, rank, convergence_tolerance=1e-9, initial=init)
    assert_equal_vectors(D, D2, 1e-6)


def test_compute_small_and_large_exact():
    # tests/src/compute.cpp:157-191 (exact fallback, cap_number).
    small = O.create_random_matrix(10, 3)
    U, V, D, conv, it = O.compute(O.dense(small), 2)
    assert conv and it == 0
    np.testing.assert_allclose(D, np.linalg.svd(small, compute_uv=False)[:2], rtol=1e-13)
    mat = O.create_random_matrix(20, 50)
    M = O.dense(mat)
    Sl = np.linalg.svd(mat, compute_uv=False)
    _, _, D, _, it = O.compute(M, 15)
    assert it == 0
    np.testing.assert_allclose(D, Sl[:15], rtol=1e-12)
    U, V, D, _, _ = O.compute(M, 20)
    np.testing.assert_allclose(D, Sl, rtol=1e-12)
    U, V, D, _, _ = O.compute(M, 50, cap_number=True)
    assert D.shape == (20,) and U.shape == (20, 20) and V.shape == (50, 20)


def test_compute_failures():
    # tests/src/compute.cpp:193-238
    A = O.create_random_matrix(20, 10)
    M = O.dense(A)
    with pytest.raises(RuntimeError, match="cannot be greater than"):
        O.compute(M, 100)
    with pytest.raises(RuntimeError, match="must be less than"):
        O.compute(M, 10, exact_for_large_number=False)
    with pytest.raises(RuntimeError, match="must be positive"):
        O.compute(M, 0, extra_work=0)
    with pytest.raises(RuntimeError) as e:
        O.compute(M, 2, initial=np.zeros(1))
    assert str(e.value).startswith("initialization")


@pytest.mark.parametrize("rank", [2, 5, 10])
def test_invariant_low_rank(rank):
    # tests/src/invariant.cpp:31-80 (duplicated columns give exact low rank).
    nr, nc = 100, 50
    base = O.create_random_matrix(nr, rank, 42)
    A = np.empty((nr, nc))
    for i in range(nc):
        A[:, i] = base[:, i % rank]
    M = O.dense(A)
    U, V, D, conv, it = O.compute(M, rank + 3)
    assert np.all(D[:rank] > 0.01)
    assert np.all(D[rank:] < 1e-8)
    assert same_same(np.sum(D ** 2), np.sum(A ** 2), 1e-8)
    U, V, D, conv, it = O.compute(M, rank + 2)
    np.testing.assert_allclose(U @ np.diag(D) @ V.T, A, atol=1e-8)


# ---------------------------------------------------------------- sparse
@pytest.mark.parametrize("nr", [10, 49, 97])
@pytest.mark.parametrize("nc", [10, 51, 88])
@pytest.mark.parametrize("nt", [1, 3])
def test_parallel_sparse(nr, nc, nt):
    # tests/src/Matrix/sparse.cpp:43-148
    vals, idx, ptr = O.simulate_compressed_sparse(nr, nc)
    dense = np.zeros((nr, nc))
    for c in range(nc):
        dense[idx[ptr[c]:ptr[c + 1]], c] = vals[ptr[c]:ptr[c + 1]]
    A = O.sparse(nr, nc, vals, idx, ptr, True, nt)
    A2 = O.sparse(nc, nr, vals, idx, ptr, False, nt)
    assert A.shape == (nr, nc) and A2.shape == (nc, nr)
    assert_equal_matrix(A.realize(), dense)
    assert_equal_matrix(A2.realize().T, dense)
    if nt > 1:
        pb, sb, snz = A.partition(nt)
        snz = snz[: (nt + 1) * nc].reshape(nt + 1, nc)
        assert pb[0] == 0 and sb[0] == 0 and pb[-1] == nc and sb[-1] == nr
        assert np.all(np.diff(pb) >= 0) and np.all(np.diff(sb) >= 0)
        np.testing.assert_array_equal(snz[0], ptr[:-1])
        np.testing.assert_array_equal(snz[-1], ptr[1:])
        for c in range(nc):
            for t in range(nt):
                a, b = snz[t, c], snz[t + 1, c]
                assert a <= b
                if a != b:
                    assert idx[a] >= sb[t] and idx[b - 1] < sb[t + 1]
    v = O.create_random_vector(nc, nr * nc * nt // 10)
    # CSC A.v against the column-ordered dense GEMV, exactly (sparse.cpp:114).
    ref = np.zeros(nr)
    for c in range(nc):
        ref += dense[:, c] * v[c]
    for _ in range(2):  # re-use of non-zeroed outputs (sparse.cpp:121-125)
        np.testing.assert_array_equal(A.multiply(v), ref)
        assert_equal_vectors(A2.multiply(v, adjoint=True), dense @ v)
    u = O.create_random_vector(nr, nr * nc * nt // 20)
    assert_equal_vectors(A.multiply(u, adjoint=True), dense.T @ u)
    assert_equal_vectors(A2.multiply(u), dense.T @ u)


def test_sparse_float_values_accumulate_in_double():
    vals, idx, ptr = O.simulate_compressed_sparse(40, 30)
    v32 = vals.astype(np.float32)
    A = O.sparse(40, 30, v32, idx, ptr, True, 1)
    dense = np.zeros((40, 30))
    for c in range(30):
        dense[idx[ptr[c]:ptr[c + 1]], c] = v32[ptr[c]:ptr[c + 1]].astype(np.float64)
    x = O.create_random_vector(30, 3)
    np.testing.assert_allclose(A.multiply(x), dense @ x, rtol=1e-13, atol=1e-13)


def test_compute_dense_vs_sparse():
    # tests/src/compute.cpp:240-285
    vals, idx, ptr = O.simulate_compressed_sparse(99, 64)
    dense = np.zeros((99, 64))
    for c in range(64):
        dense[idx[ptr[c]:ptr[c + 1]], c] = vals[ptr[c]:ptr[c + 1]]
    r1 = O.compute(O.dense(dense), 8, extra_work=7)
    r2 = O.compute(O.sparse(99, 64, vals, idx, ptr, True, 1), 8, extra_work=7)
    r3 = O.compute(O.sparse(99, 64, vals, idx, ptr, True, 3), 8, extra_work=7)
    for r in (r2, r3):
        assert_equal_vectors(r1[2], r[2])
        assert_equal_column_vectors(r1[0], r[0])
        assert_equal_column_vectors(r1[1], r[1])
    U, V, D, _, _ = O.compute(O.sparse(99, 64, vals, idx, ptr, True, 1), 13, extra_work=20)
    Ul, Sl, Vlt = np.linalg.svd(dense, full_matrices=False)
    assert_equal_vectors(D, Sl[:13], 1e-5)
    assert_equal_column_vectors(U, Ul[:, :13], 1e-5)
    assert_equal_column_vectors(V, Vlt.T[:, :13], 1e-5)


# ---------------------------------------------------------------- decorators
def test_centered_and_scaled_algebra():
    # tests/src/Matrix/centered.cpp and scaled.cpp: realize, multiply, adjoint vs dense.
    A = O.create_random_matrix(20, 10)
    c = O.create_random_vector(10)
    M = O.dense(A)
    C = O.centered(M, c)
    exp = A - c[None, :]
    assert_equal_matrix(C.realize(), exp)
    x = O.create_random_vector(10, 1234)
    u = O.create_random_vector(20, 1234)
    for _ in range(2):
        assert_equal_matrix(C.multiply(x), exp @ x)
        assert_equal_matrix(C.multiply(u, adjoint=True), exp.T @ u)
    for column in (True, False):
        for divide in (True, False):
            s = np.abs(O.create_random_vector(10 if column else 20, 77)) + 0.5
            S = O.scaled(C, s, column, divide)
            f = (1 / s if divide else s)
            expS = exp * f[None, :] if column else exp * f[:, None]
            assert_equal_matrix(S.realize(), expS)
            assert_equal_matrix(S.multiply(x), expS @ x)
            assert_equal_matrix(S.multiply(u, adjoint=True), expS.T @ u)


# ---------------------------------------------------------------- pca
def _spawn_center_scale(A):
    # tests/src/pca.cpp:27-48
    nr = A.shape[0]
    mean = A.mean(axis=0)
    centered = A - mean
    var1 = ((A - mean) ** 2).sum(axis=0)
    var2 = (A ** 2).sum(axis=0)
    both = centered / np.sqrt(var1 / (nr - 1))
    scaled = A / np.sqrt(var2 / (nr - 1))
    return centered, scaled, both


@pytest.mark.parametrize("nr", [20, 50, 100])
@pytest.mark.parametrize("nc", [20, 50, 100])
@pytest.mark.parametrize("rank", [2, 5, 10])
def test_pca_identities(nr, nc, rank):
    # tests/src/pca.cpp:66-141
    A = O.create_random_matrix(nr, nc)
    centered, scaled, both = _spawn_center_scale(A)
    M = O.dense(A)
    sc, rot, var, conv, it = O.pca(M, True, True, rank)
    assert sc.shape == (nr, rank) and rot.shape == (nc, rank) and var.shape == (rank,)
    U, V, D, _, _ = O.compute(O.dense(both), rank)
    assert_equal_vectors(var, D ** 2 / (nr - 1))
    assert_equal_column_vectors(sc, U * D[None, :], centered=True)
    assert_equal_column_vectors(rot, V)
    for i in range(rank):
        assert abs(sc[:, i].sum()) < 1e-8
        assert same_same((sc[:, i] ** 2).sum() / (nr - 1), var[i], 1e-8)
    for (cflag, sflag, ref) in [(True, False, centered), (False, True, None), (False, False, A)]:
        sc2, rot2, var2, _, _ = O.pca(M, cflag, sflag, rank)
        assert not np.array_equal(var, var2)
        if ref is None:
            # scale without centring divides by sqrt(sum((x - 0)^2)/(n-1)).
            ref = A / np.sqrt((A ** 2).sum(axis=0) / (nr - 1))
        U, V, D, _, _ = O.compute(O.dense(ref), rank)
        assert_equal_vectors(var2, D ** 2 / (nr - 1))
        assert_equal_column_vectors(sc2, U * D[None, :], centered=cflag)
        assert_equal_column_vectors(rot2, V)


def test_pca_sparse_vs_dense():
    # tests/src/pca.cpp:218-240
    vals, idx, ptr = O.simulate_compressed_sparse(76, 128)
    dense = np.zeros((76, 128))
    for c in range(128):
        dense[idx[ptr[c]:ptr[c + 1]], c] = vals[ptr[c]:ptr[c + 1]]
    r1 = O.pca(O.dense(dense), True, True, 8, extra_work=7)
    r2 = O.pca(O.sparse(76, 128, vals, idx, ptr, True, 1), True, True, 8, extra_work=7)
    assert_equal_vectors(r1[2], r2[2])
    assert_equal_column_vectors(r1[1], r2[1])
    assert np.all(same_same(np.abs(r1[0]), np.abs(r2[0]), 1e-8))
    assert np.all(np.abs(r2[0].sum(axis=0)) < 1e-8)


def test_pca_failures_and_zero_variance():
    # tests/src/pca.cpp:242-276
    with pytest.raises(RuntimeError, match="cannot center"):
        O.pca(O.dense(np.zeros((0, 10))), True, True, 1)
    with pytest.raises(RuntimeError, match="cannot scale"):
        O.pca(O.dense(O.create_random_matrix(1, 10)), True, True, 1)
    sc, rot, var, conv, it = O.pca(O.dense(np.zeros((10, 2))), True, True, 1)
    assert conv and np.isfinite(sc[0, 0]) and np.isfinite(rot[0, 0])


def test_column_stats_closed_form():
    vals, idx, ptr = O.simulate_compressed_sparse(50, 40)
    dense = np.zeros((50, 40))
    for c in range(40):
        dense[idx[ptr[c]:ptr[c + 1]], c] = vals[ptr[c]:ptr[c + 1]]
    c, s = O.column_stats(O.sparse(50, 40, vals, idx, ptr, True, 1))
    np.testing.assert_allclose(c, dense.mean(axis=0), rtol=1e-13)
    np.testing.assert_allclose(s, dense.std(axis=0, ddof=1), rtol=1e-12)


# ---------------------------------------------------------------- synthetic generator
def test_nb_draw_pmf_closed_form():
    # NB(mu=2, phi=0.5): P(K=k) = (k+1)/2^(k+2); the integer sampler's thresholds
    # reproduce that pmf exactly on the 2^32 grid.
    edges = [0]
    for k in range(31):
        edges.append((1 << 32) - (k + 3) * (1 << (30 - k)))
    for k in range(12):
        lo, hi = edges[k], edges[k + 1]
        assert (hi - lo) / 2 ** 32 == (k + 1) / 2 ** (k + 2)
        assert O.nb_draw(lo) == k and O.nb_draw(hi - 1) == k
    assert O.nb_draw(2 ** 32 - 1) == 31


def test_gen_nb_sparse_statistics_and_slicing():
    nr, nc, d = 3000, 400, 0.05
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, d, 11, f32=True)
    nnz = ptr[-1]
    assert abs(nnz / (nr * nc) - d) < 0.003
    for c in range(0, nc, 37):
        seg = idx[ptr[c]:ptr[c + 1]]
        assert np.all(np.diff(seg) > 0)
    K = np.round(np.expm1(vals.astype(np.float64))).astype(int)
    assert abs(K.mean() - 2.0) < 0.05 and abs(K.var() - 4.0) < 0.25
    assert abs((K == 0).mean() - 0.25) < 0.01
    # slices of lines reproduce the same entries (counter-based)
    v2, i2, p2 = O.gen_nb_sparse(nr, nc, d, 11, line0=100, line1=200, f32=True)
    np.testing.assert_array_equal(i2, idx[ptr[100]:ptr[200]])
    np.testing.assert_array_equal(v2, vals[ptr[100]:ptr[200]])
    assert O.density_threshold(0.03) == int(np.floor(0.03 * 2 ** 32))
