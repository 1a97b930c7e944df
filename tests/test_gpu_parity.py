"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle and
LAPACK on the same seeded inputs.  Tolerances:
  * FP64 products: relative 1e-12 of the output norm (different summation order).
  * FP32 and BF16 storage (FP32 accumulation): relative 2e-6 of the output norm
    against the FP64 product of the stored (rounded) matrix.
  * FP64 solves: singular values within 1e-8 relative of the oracle's run with the
    same explicit start vector (and of LAPACK at the reference's tolerances).
"""

import os

import numpy as np
import pytest

from compare import (align_signs, assert_equal_column_vectors, assert_equal_vectors, same_same, subspace_angle)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

irlba = pytest.importorskip("cppirlba_amd")


def _dense_of(vals, idx, ptr, nr, nc):
    A = np.zeros((nr, nc))
    for c in range(nc):
        A[idx[ptr[c]:ptr[c + 1]], c] = vals[ptr[c]:ptr[c + 1]]
    return A


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


# ---------------------------------------------------------------- sparse products
@pytest.mark.parametrize("nr,nc", [(10, 10), (49, 51), (97, 88), (300, 20), (20, 300)])
@pytest.mark.parametrize("column_major", [True, False])
def test_sparse_products_fp64(nr, nc, column_major):
    # tests/src/Matrix/sparse.cpp:91-148 on the device.
    if column_major:
        vals, idx, ptr = O.simulate_compressed_sparse(nr, nc)
        dense = _dense_of(vals, idx, ptr, nr, nc)
        A = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, column_major=True)
    else:
        vals, idx, ptr = O.simulate_compressed_sparse(nc, nr)  # CSC of the transpose = CSR of dense
        dense = _dense_of(vals, idx, ptr, nc, nr).T
        A = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, column_major=False)
    assert A.shape == (nr, nc)
    v = O.create_random_vector(nc, 17)
    u = O.create_random_vector(nr, 18)
    for _ in range(2):
        y = A.multiply(v)
        z = A.adjoint_multiply(u)
        assert _rel(y, dense @ v) < 1e-13
        assert _rel(z, dense.T @ u) < 1e-13
    np.testing.assert_array_equal(A.realize(), dense)


@pytest.mark.parametrize("storage,acc,nc", [("f32", "f32", 39937), ("f32", "f32", 3 * 39936 + 5),
                                            ("f64", "f64", 19969), ("f64", "f64", 3 * 19968 + 1),
                                            ("f32", "f64", 2 * 19968 + 63)])
def test_equal_width_tiles_vs_oracle(storage, acc, nc):
    # Input widths just past a whole number of the largest LDS tiles (39936 FP32 or
    # 19968 FP64-accumulation entries): the tiles are of equal width, a multiple
    # of 64 (the last one shorter by less than 64 entries per tile), none a
    # remainder of a few entries, and the products equal the oracle's.
    nr = 40
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 23, f32=storage == "f32")
    A = O.sparse(nr, nc, vals, idx, ptr, True, 1)
    H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, storage=storage, accumulate=acc)
    lay = H.layout(False)
    tmax = 39936 if acc == "f32" else 19968
    n_tiles = -(-nc // tmax)
    assert lay["n_tiles"] == n_tiles
    assert lay["tile_len"] % 64 == 0 and lay["tile_len"] <= tmax
    last = nc - (n_tiles - 1) * lay["tile_len"]
    assert lay["tile_len"] - 64 * n_tiles < last <= lay["tile_len"]
    v = O.create_random_vector(nc, 3)
    u = O.create_random_vector(nr, 4)
    # FP32 values (as stored) with FP64 accumulation in the oracle: 2e-6 for FP32 accumulation.
    tol = 2e-6 if acc == "f32" else 1e-12
    assert _rel(H.multiply(v), A.multiply(v)) < tol
    assert _rel(H.adjoint_multiply(u), A.multiply(u, adjoint=True)) < tol


def test_sparse_multitile_both_orientations():
    # 40000 columns exceed one FP64 LDS tile, so A.v runs with several tiles and partials.
    nr, nc = 64, 40000
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 7, f32=False)
    A = O.sparse(nr, nc, vals, idx, ptr, True, 1)
    H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr)
    v = O.create_random_vector(nc, 3)
    u = O.create_random_vector(nr, 4)
    assert _rel(H.multiply(v), A.multiply(v)) < 1e-12
    assert _rel(H.adjoint_multiply(u), A.multiply(u, adjoint=True)) < 1e-12
    # and the transpose shape: many row tiles on the adjoint side
    Ht = irlba.HipSparseMatrix(nc, nr, vals, idx, ptr, column_major=False)
    At = O.sparse(nc, nr, vals, idx, ptr, False, 1)
    assert _rel(Ht.multiply(u), At.multiply(u)) < 1e-12
    assert _rel(Ht.adjoint_multiply(v), At.multiply(v, adjoint=True)) < 1e-12


def test_sparse_mixed_precision_products():
    nr, nc = 3000, 5000
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.03, 11, f32=True)
    A = O.sparse(nr, nc, vals, idx, ptr, True, 1)  # float values, double accumulation (reference semantics)
    H32 = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, accumulate="f32")
    H64 = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, accumulate="f64")
    v = O.create_random_vector(nc, 5)
    u = O.create_random_vector(nr, 6)
    assert _rel(H64.multiply(v), A.multiply(v)) < 1e-13
    assert _rel(H64.adjoint_multiply(u), A.multiply(u, adjoint=True)) < 1e-13
    assert _rel(H32.multiply(v), A.multiply(v)) < 2e-6
    assert _rel(H32.adjoint_multiply(u), A.multiply(u, adjoint=True)) < 2e-6


@pytest.mark.parametrize("column_major", [True, False])
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_sparse_without_entries(column_major, storage):
    # A 50 x 40 matrix with no stored entries: zero products and realize,
    # column statistics (0, 1) (pca.hpp:85-109 with the zero-variance rule),
    # zeros from the exact path, and the reference's null-space error from the
    # Lanczos start (lanczos.hpp:83) in compute() and pca(), as the oracle raises.
    nr, nc = 50, 40
    ptr = np.zeros((nc if column_major else nr) + 1, dtype=np.int64)
    idx = np.zeros(0, dtype=np.int32)
    vals = np.zeros(0, dtype=np.float32 if storage == "f32" else np.float64)
    H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, column_major=column_major, storage=storage)
    A = O.sparse(nr, nc, np.zeros(0), idx, ptr, column_major, 1)
    assert not np.any(H.multiply(np.ones(nc))) and not np.any(H.adjoint_multiply(np.ones(nr)))
    assert not np.any(H.realize())
    c, s = irlba.column_stats(H)
    assert not np.any(c) and np.all(s == 1)
    init = O.create_random_vector(nc, 3)
    with pytest.raises(RuntimeError, match="null space"):
        O.compute(A, 3, initial=init)
    with pytest.raises(irlba.HipError, match="starting vector near the null space"):
        irlba.compute(H, 3, irlba.Options(initial=init))
    with pytest.raises(irlba.HipError, match="starting vector near the null space"):
        irlba.pca(H, True, True, 2, irlba.Options(initial=init))
    r = irlba.compute(H, 25)
    assert r.converged and r.iterations == 0 and r.D.shape == (25,) and not np.any(r.D)


def _ragged_matrix():
    # 300 x 200 with 287 empty rows and 157 empty columns: a half-filled 10 x 40
    # block plus three isolated entries (first row/column, last row/column).
    rng = np.random.default_rng(7)
    M = np.zeros((300, 200))
    M[5:15, 150:190] = (rng.random((10, 40)) < 0.5) * rng.standard_normal((10, 40))
    M[250, 3], M[299, 199], M[0, 0] = 2.0, -1.5, 0.5
    return M


def _compressed(M, column_major):
    lines = M if not column_major else M.T
    nz = [np.nonzero(line)[0] for line in lines]
    ptr = np.zeros(len(nz) + 1, dtype=np.int64)
    ptr[1:] = np.cumsum([len(z) for z in nz])
    idx = np.concatenate(nz).astype(np.int32)
    vals = np.concatenate([line[z] for line, z in zip(lines, nz)])
    return vals, idx, ptr


@pytest.mark.parametrize("column_major", [True, False])
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_sparse_ragged_vs_oracle(column_major, storage):
    # Empty rows and columns everywhere: products against the dense GEMV, and
    # compute() / pca() against the oracle from the same start vector.  FP32
    # storage is compared with the oracle on the FP32-rounded values.
    M = _ragged_matrix()
    if storage == "f32":
        M = M.astype(np.float32).astype(np.float64)
    vals, idx, ptr = _compressed(M, column_major)
    H = irlba.HipSparseMatrix(300, 200, vals.astype(np.float32) if storage == "f32" else vals, idx, ptr,
                              column_major=column_major, storage=storage)
    A = O.sparse(300, 200, vals, idx, ptr, column_major, 1)
    tp, ts = (1e-13, 1e-8) if storage == "f64" else (2e-6, 1e-5)
    v = O.create_random_vector(200, 11)
    u = O.create_random_vector(300, 12)
    assert _rel(H.multiply(v), M @ v) < tp
    assert _rel(H.adjoint_multiply(u), M.T @ u) < tp
    np.testing.assert_array_equal(H.realize(), M)
    init = O.create_random_vector(200, 5)
    res = irlba.compute(H, 4, irlba.Options(initial=init))
    oU, oV, oD, oconv, oit = O.compute(A, 4, initial=init)
    assert res.converged and oconv
    assert_equal_vectors(res.D, oD, ts)
    assert_equal_vectors(res.D, np.linalg.svd(M, compute_uv=False)[:4], ts)
    # Entries of the singular vectors on rows orthogonal to the top four sit at
    # rounding level (1e-17 .. 1e-13), so vectors compare with an absolute bound.
    tv = 1e-9 if storage == "f64" else 1e-5
    np.testing.assert_allclose(align_signs(res.U, oU), oU, rtol=0, atol=tv)
    np.testing.assert_allclose(align_signs(res.V, oV), oV, rtol=0, atol=tv)
    pres = irlba.pca(H, True, True, 3, irlba.Options(initial=init))
    oS, oR, ovar, pconv, _ = O.pca(A, True, True, 3, initial=init)
    assert pres.converged and pconv
    assert_equal_vectors(pres.variances, ovar, ts)


def test_sparse_rejects_bad_input():
    vals = np.ones(3)
    with pytest.raises(irlba.HipError, match="indices"):
        irlba.HipSparseMatrix(4, 2, vals, np.array([0, 5, 1]), np.array([0, 2, 3]))
    with pytest.raises(irlba.HipError, match="pointers"):
        irlba.HipSparseMatrix(4, 2, vals, np.array([0, 1, 1]), np.array([0, 3, 2]))
    with pytest.raises(irlba.HipError, match="more than once"):
        irlba.HipSparseMatrix(4, 2, vals, np.array([1, 1, 2]), np.array([0, 2, 3]))


@pytest.mark.parametrize("column_major", [True, False])
@pytest.mark.parametrize("f32", [False, True])
def test_unsorted_lines_match_sorted(column_major, f32):
    # The reference's products take the entries of a line in any order
    # (sparse.hpp:185-267 only sums them); the device layout sorts each segment,
    # so shuffled lines give bit-identical products to sorted ones.
    n_minor, n_major = 3000, 2000
    vals, idx, ptr = O.gen_nb_sparse(n_minor, n_major, 0.05, 12, f32=f32)
    rng = np.random.default_rng(4)
    perm = np.arange(idx.size)
    for c in range(n_major):
        rng.shuffle(perm[ptr[c]:ptr[c + 1]])
    nr, nc = (n_minor, n_major) if column_major else (n_major, n_minor)
    Hs = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, column_major=column_major)
    Hu = irlba.HipSparseMatrix(nr, nc, vals[perm], idx[perm], ptr, column_major=column_major)
    assert Hs.layout(False)["delta"] and Hu.layout(False)["delta"]
    v = rng.standard_normal(nc)
    u = rng.standard_normal(nr)
    np.testing.assert_array_equal(Hu.multiply(v), Hs.multiply(v))
    np.testing.assert_array_equal(Hu.adjoint_multiply(u), Hs.adjoint_multiply(u))
    np.testing.assert_array_equal(Hu.realize(), Hs.realize())


def test_generator_bit_identical_to_oracle():
    nr, nc, d, seed = 700, 300, 0.04, 99
    v, i, p = irlba.generate_csc_host(nr, nc, d, seed, "f32")
    ov, oi, op = O.gen_nb_sparse(nr, nc, d, seed, f32=True)
    np.testing.assert_array_equal(p, op)
    np.testing.assert_array_equal(i, oi)
    np.testing.assert_array_equal(v, ov)


def test_products_beyond_2e31_entries():
    # 30000 x 1,000,000 at 8%: ~2.4e9 stored entries, past the int32 range, so
    # every slot offset of the last columns (and of the last tiles of the row
    # layout) lies above 2^31.  FP32 values with FP64 accumulation, against the
    # oracle on the first and the last 2000 columns (the same generator on the
    # CPU, bit-identical), and the adjoint identity u'(A v) = (A'u)'v over all of it.
    nr, nc, d, seed, w = 30000, 1000000, 0.08, 3, 2000
    H = irlba.generate_sparse(nr, nc, d, seed, storage="f32", accumulate="f64")
    lay = H.layout(False)
    assert lay["nnz"] > 2 ** 31 and lay["slots"] > 2 ** 31
    rng = np.random.default_rng(5)
    u = rng.standard_normal(nr)
    z = H.adjoint_multiply(u)
    for c0 in (0, nc - w):
        vals, idx, ptr = O.gen_nb_sparse(nr, nc, d, seed, line0=c0, line1=c0 + w, f32=True)
        A = O.sparse(nr, w, vals, idx, ptr, True, 1)
        vt = rng.standard_normal(w)
        v = np.zeros(nc)
        v[c0:c0 + w] = vt
        assert _rel(H.multiply(v), A.multiply(vt)) < 1e-12
        assert _rel(z[c0:c0 + w], A.multiply(u, adjoint=True)) < 1e-12
    v = rng.standard_normal(nc)
    a, b = u @ H.multiply(v), z @ v
    assert abs(a - b) <= 1e-10 * (abs(a) + abs(b))
    del H


# ---------------------------------------------------------------- dense products
@pytest.mark.parametrize("nr,nc", [(20, 10), (1000, 200), (33, 1001), (4099, 17), (40000, 300)])
@pytest.mark.parametrize("order", ["F", "C"])
def test_dense_products(nr, nc, order):
    # (40000, 300) row-major: A'u splits over more than 128 line chunks and
    # reduces in two stages.
    rng = np.random.default_rng(nr + nc)
    A = np.asarray(rng.standard_normal((nr, nc)), order=order)
    v = rng.standard_normal(nc)
    u = rng.standard_normal(nr)
    H = irlba.HipDenseMatrix(A, storage="f64")
    assert _rel(H.multiply(v), A @ v) < 1e-13
    assert _rel(H.adjoint_multiply(u), A.T @ u) < 1e-13
    H32 = irlba.HipDenseMatrix(A.astype(np.float32), storage="f32")
    assert _rel(H32.multiply(v), A.astype(np.float32).astype(np.float64) @ v) < 2e-6
    assert _rel(H32.adjoint_multiply(u), A.astype(np.float32).astype(np.float64).T @ u) < 2e-6
    # BF16 storage (simple.hpp:42-44,70-72 over a bfloat16 matrix): the products
    # of the stored matrix, i.e. of A rounded to nearest even through FP32 as
    # torch rounds it, accumulated in FP32 (the same bound as FP32 storage); the
    # looser check against A itself is the rounding of the storage.
    import torch

    Ab = torch.from_numpy(np.ascontiguousarray(A)).float().to(torch.bfloat16).double().numpy()
    Hb = irlba.HipDenseMatrix(A, storage="bf16")
    yb, zb = Hb.multiply(v), Hb.adjoint_multiply(u)
    assert _rel(yb, Ab @ v) < 2e-6
    assert _rel(zb, Ab.T @ u) < 2e-6
    assert _rel(yb, A @ v) < 2e-2 and _rel(zb, A.T @ u) < 2e-2
    # The same from a device bfloat16 tensor in either layout.
    At = torch.from_numpy(np.ascontiguousarray(A)).float().to(torch.bfloat16).cuda()
    if order == "F":
        At = At.t().contiguous().t()
    Hd = irlba.HipDenseMatrix(At, storage="bf16")
    assert _rel(Hd.multiply(v), Ab @ v) < 2e-6
    assert _rel(Hd.adjoint_multiply(u), Ab.T @ u) < 2e-6


@pytest.mark.parametrize("kind", ["sparse", "sparse-multitile", "sparse-f32", "dense-C", "dense-F", "decorated"])
def test_products_overwrite_their_output(kind):
    # Workspace::multiply fully overwrites a pre-sized output (interface.hpp:49,87;
    # re-zeroing checked in sparse.cpp:119-124, centered.cpp:56-58): device
    # outputs filled with NaN, then products twice in a row, against the
    # products into fresh host buffers.
    import ctypes

    import torch
    from cppirlba_amd._lib import check, lib

    if kind.startswith("sparse") or kind == "decorated":
        nr, nc = (64, 40000) if kind == "sparse-multitile" else (300, 200)
        vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 4, f32=kind == "sparse-f32")
        H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr)
        if kind == "decorated":
            rng = np.random.default_rng(2)
            H = irlba.ScaledMatrix(irlba.CenteredMatrix(H, rng.standard_normal(nc)), rng.uniform(0.5, 2, nc))
    else:
        nr, nc = 300, 200
        A = np.asarray(np.random.default_rng(3).standard_normal((nr, nc)), order="F" if kind == "dense-F" else "C")
        H = irlba.HipDenseMatrix(A)
    for adjoint, (n_in, n_out) in ((0, (nc, nr)), (1, (nr, nc))):
        xh = np.random.default_rng(9 + adjoint).standard_normal(n_in)
        ref = H.multiply(xh, adjoint=bool(adjoint))
        x = torch.from_numpy(xh).cuda()
        out = torch.full((n_out,), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        for _ in range(2):
            check(lib().irlba_hip_multiply(H.handle, adjoint, ctypes.c_void_p(x.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()), 1))
            np.testing.assert_array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("storage", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("order", ["C", "F"])
def test_dense_product_into_its_own_input(storage, order):
    # A square dense product whose device output is its own input buffer gives
    # the product of the input as it was, bit-equal to the product into a
    # separate buffer: every read of the input ends before an output is written
    # (the dot orientation reads the prologue's copy while it writes finished
    # outputs, the axpy one reads the input itself and writes partials).
    import ctypes

    import torch
    from cppirlba_amd._lib import check, lib

    n = 3000
    A = np.asarray(np.random.default_rng(5).standard_normal((n, n)), order=order)
    H = irlba.HipDenseMatrix(A.astype(np.float32) if storage == "f32" else A, storage=storage)
    for adjoint in (0, 1):
        xh = np.random.default_rng(11 + adjoint).standard_normal(n)
        x = torch.from_numpy(xh).cuda()
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        same = x.clone()
        torch.cuda.synchronize()
        check(lib().irlba_hip_multiply(H.handle, adjoint, ctypes.c_void_p(x.data_ptr()),
                                       ctypes.c_void_p(out.data_ptr()), 1))
        check(lib().irlba_hip_multiply(H.handle, adjoint, ctypes.c_void_p(same.data_ptr()),
                                       ctypes.c_void_p(same.data_ptr()), 1))
        np.testing.assert_array_equal(same.cpu().numpy(), out.cpu().numpy())
        ref = H.multiply(xh, adjoint=bool(adjoint))
        np.testing.assert_array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("shape", [(5, 0), (0, 5)])
@pytest.mark.parametrize("kind", ["dense-C", "dense-F", "dense-f32", "sparse-csc", "sparse-csr"])
def test_empty_dimension_products(shape, kind):
    # A matrix with no rows or no columns: products are zero vectors of the
    # output length (Eigen's result for the reference's SimpleMatrix), block
    # products likewise, and compute() raises the reference's error, as the
    # oracle does.
    nr, nc = shape
    if kind.startswith("dense"):
        A = np.zeros(shape, order="F" if kind == "dense-F" else "C")
        H = irlba.HipDenseMatrix(A.astype(np.float32) if kind == "dense-f32" else A,
                                 storage="f32" if kind == "dense-f32" else "f64")
    else:
        cm = kind == "sparse-csc"
        ptr = np.zeros((nc if cm else nr) + 1, dtype=np.int64)
        H = irlba.HipSparseMatrix(nr, nc, np.zeros(0), np.zeros(0, dtype=np.int32), ptr, column_major=cm)
    y = H.multiply(np.ones(nc))
    z = H.adjoint_multiply(np.ones(nr))
    assert y.shape == (nr,) and z.shape == (nc,) and not y.any() and not z.any()
    # The output is overwritten whatever it held (interface.hpp:49,87; the
    # reference's tests fill it with garbage first, sparse.cpp:121-125).
    import ctypes

    import torch
    from cppirlba_amd._lib import check, lib

    for adjoint, (n_in, n_out) in ((0, (nc, nr)), (1, (nr, nc))):
        x = torch.ones(max(n_in, 1), dtype=torch.float64, device="cuda")
        out = torch.full((max(n_out, 1),), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        check(lib().irlba_hip_multiply(H.handle, adjoint, ctypes.c_void_p(x.data_ptr()),
                                       ctypes.c_void_p(out.data_ptr()), 1))
        assert not out[:n_out].cpu().numpy().any()
    Yb = H.multiply_block(np.ones((nc, 3)))
    Zb = H.multiply_block(np.ones((nr, 3)), adjoint=True)
    assert Yb.shape == (nr, 3) and Zb.shape == (nc, 3) and not Yb.any() and not Zb.any()
    assert H.realize().shape == shape
    msg = "cannot be greater than the smaller matrix dimension"
    with pytest.raises(RuntimeError, match=msg):
        O.compute(O.dense(np.zeros(shape)), 1)
    with pytest.raises(irlba.HipError, match=msg):
        irlba.compute(H, 1)


def test_dense_storage_conversions_bit_exact():
    # Every input type to every storage type, host and device input: the stored
    # matrix (read back by realize) equals the round-to-nearest-even conversion
    # torch makes (FP64 -> BF16 through FP32), bit for bit.
    import torch

    rng = np.random.default_rng(21)
    A64 = rng.standard_normal((37, 23)) * np.exp(rng.uniform(-20, 20, (37, 23)))
    ref = {"f64": torch.from_numpy(A64)}
    ref["f32"] = ref["f64"].float()
    ref["bf16"] = ref["f32"].to(torch.bfloat16)
    for src in ("f64", "f32", "bf16"):
        for dst in ("f64", "f32", "bf16"):
            expect = ref[src].to({"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}[dst])
            expect = expect.double().numpy()
            inputs = [ref[src].cuda(), ref[src].cuda().t().contiguous().t()]
            if src != "bf16":
                inputs.append(ref[src].numpy())
            for x in inputs:
                H = irlba.HipDenseMatrix(x, storage=dst)
                np.testing.assert_array_equal(H.realize(), expect, err_msg=f"{src} -> {dst}")
    # and a product on converted storage
    H = irlba.HipDenseMatrix(ref["bf16"].cuda(), storage="f32")
    v = rng.standard_normal(23)
    assert _rel(H.multiply(v), ref["bf16"].double().numpy() @ v) < 2e-6


@pytest.mark.parametrize("order", ["C", "F"])
def test_dense_products_beyond_2e31_elements(order):
    # 200000 x 11000 FP32 (2.2e9 elements, past the int32 range) made on the
    # device with torch and handed over as a device pointer, row- or
    # column-major.  Both products against FP64 torch products on 20000-row slices.
    import torch

    nr, nc = 200000, 11000
    g = torch.Generator(device="cuda").manual_seed(17)
    T = torch.randn(nr, nc, dtype=torch.float32, device="cuda", generator=g)
    if order == "F":
        T = T.t().contiguous().t()  # column-major storage, same values
    H = irlba.HipDenseMatrix(T, storage="f32")
    rng = np.random.default_rng(8)
    v = rng.standard_normal(nc)
    u = rng.standard_normal(nr)
    y = H.multiply(v)
    z = H.adjoint_multiply(u)
    vd = torch.from_numpy(v).cuda()
    ud = torch.from_numpy(u).cuda()
    y_ref = torch.empty(nr, dtype=torch.float64, device="cuda")
    z_ref = torch.zeros(nc, dtype=torch.float64, device="cuda")
    for r0 in range(0, nr, 20000):
        blk = T[r0:r0 + 20000].double()
        y_ref[r0:r0 + 20000] = blk @ vd
        z_ref += blk.t() @ ud[r0:r0 + 20000]
        del blk
    assert _rel(y, y_ref.cpu().numpy()) < 2e-6
    assert _rel(z, z_ref.cpu().numpy()) < 2e-6
    del H, T
    torch.cuda.empty_cache()


# ---------------------------------------------------------------- decorators
@pytest.mark.parametrize("storage,acc", [("f64", "f64"), ("f32", "f32"), ("f32", "f64")])
@pytest.mark.parametrize("column,divide", [(True, True), (False, False)])
def test_centered_scaled_products_multitile(storage, acc, column, divide):
    # Decorated products whose input spans several LDS tiles (100000 columns: 3
    # FP32 or 6 FP64-accumulation tiles in A.v): each block scales the tiles it
    # stages, the block that starts a tile sums its centring dot products, and
    # the partial reduce adds the tiles' sums.  Against the oracle's decorators,
    # in both decorator orders and with two centring terms.
    nr, nc = 40, 100000
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 29, f32=storage == "f32")
    H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, storage=storage, accumulate=acc)
    assert H.layout(False)["n_tiles"] > 1
    OM = O.sparse(nr, nc, vals, idx, ptr, True, 1)
    c = O.create_random_vector(nc, 1)
    s = np.abs(O.create_random_vector(nc if column else nr, 2)) + 0.5
    tol = 2e-6 if acc == "f32" else 1e-12
    x = O.create_random_vector(nc, 3)
    u = O.create_random_vector(nr, 4)
    pairs = [(irlba.ScaledMatrix(irlba.CenteredMatrix(H, c), s, column, divide),
              O.scaled(O.centered(OM, c), s, column, divide)),
             (irlba.CenteredMatrix(irlba.ScaledMatrix(H, s, column, divide), c),
              O.centered(O.scaled(OM, s, column, divide), c)),
             (irlba.CenteredMatrix(irlba.CenteredMatrix(H, c), c * 0.5), O.centered(O.centered(OM, c), c * 0.5))]
    # A row scaling between two centrings keeps two rank-1 terms (different left vectors).
    sr = np.abs(O.create_random_vector(nr, 5)) + 0.5
    pairs.append((irlba.CenteredMatrix(irlba.ScaledMatrix(irlba.CenteredMatrix(H, c), sr, False, True), c * 0.5),
                  O.centered(O.scaled(O.centered(OM, c), sr, False, True), c * 0.5)))
    for HD, OD in pairs:
        for _ in range(2):
            assert _rel(HD.multiply(x), OD.multiply(x)) < tol
            assert _rel(HD.adjoint_multiply(u), OD.multiply(u, adjoint=True)) < tol


@pytest.mark.parametrize("column", [True, False])
@pytest.mark.parametrize("divide", [True, False])
@pytest.mark.parametrize("kind", ["dense", "sparse"])
def test_centered_scaled_products(column, divide, kind):
    # tests/src/Matrix/centered.cpp, scaled.cpp on the device.
    nr, nc = 60, 45
    if kind == "dense":
        A = O.create_random_matrix(nr, nc)
        H = irlba.HipDenseMatrix(A)
        OM = O.dense(A)
    else:
        vals, idx, ptr = O.simulate_compressed_sparse(nr, nc)
        H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr)
        OM = O.sparse(nr, nc, vals, idx, ptr, True, 1)
    c = O.create_random_vector(nc, 1)
    s = np.abs(O.create_random_vector(nc if column else nr, 2)) + 0.5
    HC = irlba.CenteredMatrix(H, c)
    HS = irlba.ScaledMatrix(HC, s, column, divide)
    OS = O.scaled(O.centered(OM, c), s, column, divide)
    x = O.create_random_vector(nc, 3)
    u = O.create_random_vector(nr, 4)
    for _ in range(2):
        assert _rel(HS.multiply(x), OS.multiply(x)) < 1e-13
        assert _rel(HS.adjoint_multiply(u), OS.multiply(u, adjoint=True)) < 1e-13
    np.testing.assert_allclose(HS.realize(), OS.realize(), rtol=1e-13, atol=1e-13)
    # the other decorator order, and two centring layers (merged on the device)
    HSC = irlba.CenteredMatrix(irlba.ScaledMatrix(H, s, column, divide), c)
    OSC = O.centered(O.scaled(OM, s, column, divide), c)
    assert _rel(HSC.multiply(x), OSC.multiply(x)) < 1e-13
    assert _rel(HSC.adjoint_multiply(u), OSC.multiply(u, adjoint=True)) < 1e-13
    HCC = irlba.CenteredMatrix(HC, c * 0.5)
    OCC = O.centered(O.centered(OM, c), c * 0.5)
    assert _rel(HCC.multiply(x), OCC.multiply(x)) < 1e-13


# ---------------------------------------------------------------- compute
def test_compute_compare_to_exact():
    # tests/src/compute.cpp:15-31
    A = O.create_random_matrix(20, 10)
    res = irlba.compute(irlba.HipDenseMatrix(A), 5, irlba.Options(exact_for_large_number=False))
    Ul, Sl, Vlt = np.linalg.svd(A, full_matrices=False)
    assert_equal_vectors(res.D, Sl[:5], 1e-8)
    assert_equal_column_vectors(res.U, Ul[:, :5], 1e-8)
    assert_equal_column_vectors(res.V, Vlt.T[:, :5], 1e-8)


@pytest.mark.parametrize("nr", [20, 50, 100])
@pytest.mark.parametrize("nc", [20, 50, 100])
@pytest.mark.parametrize("rank", [2, 5, 10])
def test_compute_sweep_vs_lapack_and_oracle(nr, nc, rank):
    # tests/src/compute.cpp:78-123 on the device, plus the oracle on the same start.
    A = O.create_random_matrix(nr, nc)
    init = O.create_random_vector(nc, 1239 + nr + nc + rank)
    H = irlba.HipDenseMatrix(A)
    res = irlba.compute(H, rank, irlba.Options(convergence_tolerance=1e-9, initial=init))
    Ul, Sl, Vlt = np.linalg.svd(A, full_matrices=False)
    assert_equal_vectors(res.D, Sl[:rank], 1e-6)
    # The reference checks vectors elementwise at 1e-6 from its own (aarand) start
    # vector; from this explicit start one tiny entry lands at 1.1e-6, so vectors
    # are held to 5e-6 elementwise and to 1e-6 rad in subspace angle.
    assert_equal_column_vectors(res.U, Ul[:, :rank], 5e-6)
    assert_equal_column_vectors(res.V, Vlt.T[:, :rank], 5e-6)
    assert subspace_angle(res.U, Ul[:, :rank]) < 1e-6 and subspace_angle(res.V, Vlt.T[:, :rank]) < 1e-6
    oU, oV, oD, oconv, oit = O.compute(O.dense(A), rank, convergence_tolerance=1e-9, initial=init)
    assert_equal_vectors(res.D, oD, 1e-8)
    assert res.converged == oconv


def test_compute_sparse_dense_parallel_agree():
    # tests/src/compute.cpp:240-285
    vals, idx, ptr = O.simulate_compressed_sparse(99, 64)
    dense = _dense_of(vals, idx, ptr, 99, 64)
    opt = irlba.Options(extra_work=7)
    r1 = irlba.compute(irlba.HipDenseMatrix(dense), 8, opt)
    r2 = irlba.compute(irlba.HipSparseMatrix(99, 64, vals, idx, ptr), 8, opt)
    assert_equal_vectors(r1.D, r2.D)
    assert_equal_column_vectors(r1.U, r2.U)
    assert_equal_column_vectors(r1.V, r2.V)
    r3 = irlba.compute(irlba.HipSparseMatrix(99, 64, vals, idx, ptr), 13, irlba.Options(extra_work=20))
    Ul, Sl, Vlt = np.linalg.svd(dense, full_matrices=False)
    assert_equal_vectors(r3.D, Sl[:13], 1e-5)
    assert_equal_column_vectors(r3.U, Ul[:, :13], 1e-5)
    assert_equal_column_vectors(r3.V, Vlt.T[:, :13], 1e-5)


def test_compute_small_large_exact_and_errors():
    # tests/src/compute.cpp:157-238
    small = O.create_random_matrix(10, 3)
    res = irlba.compute(irlba.HipDenseMatrix(small), 2)
    assert res.converged and res.iterations == 0
    np.testing.assert_allclose(res.D, np.linalg.svd(small, compute_uv=False)[:2], rtol=1e-12)
    mat = O.create_random_matrix(20, 50)
    H = irlba.HipDenseMatrix(mat)
    Sl = np.linalg.svd(mat, compute_uv=False)
    np.testing.assert_allclose(irlba.compute(H, 15).D, Sl[:15], rtol=1e-12)
    r = irlba.compute(H, 50, irlba.Options(cap_number=True))
    assert r.D.shape == (20,) and r.U.shape == (20, 20) and r.V.shape == (50, 20)
    A = irlba.HipDenseMatrix(O.create_random_matrix(20, 10))
    with pytest.raises(irlba.HipError, match="cannot be greater than"):
        irlba.compute(A, 100)
    with pytest.raises(irlba.HipError, match="must be less than"):
        irlba.compute(A, 10, irlba.Options(exact_for_large_number=False))
    with pytest.raises(irlba.HipError, match="must be positive"):
        irlba.compute(A, 0, irlba.Options(extra_work=0))
    with pytest.raises(irlba.HipError) as e:
        irlba.compute(A, 2, irlba.Options(initial=np.zeros(1)))
    assert str(e.value).startswith("initialization")


def test_negative_number_rejected():
    A = irlba.HipDenseMatrix(O.create_random_matrix(20, 10))
    with pytest.raises(irlba.HipError, match="non-negative"):
        irlba.compute(A, -1)
    with pytest.raises(irlba.HipError, match="non-negative"):
        irlba.pca(A, True, True, -2)


def test_wide_restart_rotation_vs_oracle_and_lapack():
    # number = 150 (work 157): every restart after the first keeps k >= 150 > 128
    # columns, so the in-place restart rotations of V and W go through a scratch
    # basis (compute.hpp:268,279 rotate into Vtmp / Wtmp the same way).
    rng = np.random.default_rng(150)
    m, n = 1000, 1500
    U0, _ = np.linalg.qr(rng.standard_normal((m, m)))
    V0, _ = np.linalg.qr(rng.standard_normal((n, m)))
    s = np.exp(-np.arange(m) / 60.0)
    A = (U0 * s) @ V0.T
    init = np.random.default_rng(151).standard_normal(n)
    opt = dict(initial=init, convergence_tolerance=1e-9)
    res = irlba.compute(irlba.HipDenseMatrix(A), 150, irlba.Options(**opt))
    oU, oV, oD, oconv, oit = O.compute(O.dense(A), 150, **opt)
    assert res.converged and oconv and res.iterations > 1
    assert_equal_vectors(res.D, oD, 1e-8)
    assert_equal_vectors(res.D, np.linalg.svd(A, compute_uv=False)[:150], 1e-6)
    assert np.abs(res.V.T @ res.V - np.eye(150)).max() < 1e-10
    assert np.abs(res.U.T @ res.U - np.eye(150)).max() < 1e-10
    np.testing.assert_allclose(A @ res.V[:, :20], res.U[:, :20] * res.D[:20], atol=1e-8)


@pytest.mark.parametrize("seed", [5489, 42, 2024])
def test_default_seed_start_matches_oracle(seed):
    # No Options::initial: the start vector is fill_with_random_normals on
    # std::mt19937_64(seed) (compute.hpp:179-190, utils.hpp:16-37).  The first
    # restart's B-SVD depends on it entirely, so one capped restart pins the
    # engine's draws to the oracle's (a different stream moves the values at
    # the 1e-2 level); the converged solve then matches to the parity bar.
    A = O.create_random_matrix(120, 81)  # odd column count: the last draw of a pair is dropped
    one = irlba.compute(irlba.HipDenseMatrix(A), 5, irlba.Options(seed=seed, max_iterations=1))
    _, _, oD1, oconv1, _ = O.compute(O.dense(A), 5, seed=seed, max_iterations=1)
    assert not one.converged and not oconv1
    assert_equal_vectors(one.D, oD1, 1e-12)
    _, _, other, _, _ = O.compute(O.dense(A), 5, seed=seed + 1, max_iterations=1)
    assert np.max(np.abs(other - oD1) / oD1) > 1e-6  # the check separates streams
    res = irlba.compute(irlba.HipDenseMatrix(A), 5, irlba.Options(seed=seed, convergence_tolerance=1e-9))
    oU, oV, oD, oconv, oit = O.compute(O.dense(A), 5, seed=seed, convergence_tolerance=1e-9)
    assert res.converged and oconv and res.iterations == oit
    assert_equal_vectors(res.D, oD, 1e-8)
    assert_equal_column_vectors(res.V, oV, 1e-6)
    assert_equal_column_vectors(res.U, oU, 1e-6)


@pytest.mark.parametrize("rank", [1, 2, 5])
def test_refill_draws_match_oracle(rank):
    # Invariant-subspace refills (lanczos.hpp:99-109,123-132) draw fresh normals
    # from the solve's engine: with the default seed and no initial vector, the
    # engine and the oracle take the same path (same restarts, same number of
    # singular values above zero) and agree on the non-zero values.
    # Unit norm, so that rounding-level residuals stay clearly below the absolute
    # refill threshold eps^0.8 in both summation orders.
    nr, nc = 100, 50
    base = O.create_random_matrix(nr, rank, 42)
    A = np.empty((nr, nc))
    for i in range(nc):
        A[:, i] = base[:, i % rank]
    A /= np.linalg.norm(A, 2)
    H = irlba.HipDenseMatrix(A)
    res = irlba.compute(H, rank + 3, irlba.Options(seed=77))
    assert H.ctx.stats()["refills"] > 0
    oU, oV, oD, oconv, oit = O.compute(O.dense(A), rank + 3, seed=77)
    assert res.converged == oconv and res.iterations == oit
    assert_equal_vectors(res.D[:rank], oD[:rank], 1e-8)
    assert np.all(res.D[rank:] < 1e-8) and np.all(oD[rank:] < 1e-8)
    assert_equal_column_vectors(res.V[:, :rank], oV[:, :rank], 1e-8)
    assert_equal_column_vectors(res.U[:, :rank], oU[:, :rank], 1e-8)


def test_sparse_reference_constructor_and_getters():
    # ParallelSparseMatrix(nrow, ncol, x, i, p, column_major, num_threads)
    # (sparse.hpp:469) positionally; the getters return the arrays given
    # (sparse.hpp:495-517); num_threads does not change the device products.
    nr, nc = 97, 88
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.1, 3, f32=False)
    H4 = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, True, 4)
    H1 = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr)
    assert H4.num_threads == 4 and H4.column_major
    np.testing.assert_array_equal(H4.get_values(), vals)
    np.testing.assert_array_equal(H4.get_indices(), idx)
    np.testing.assert_array_equal(H4.get_pointers(), ptr)
    v = np.random.default_rng(0).standard_normal(nc)
    np.testing.assert_array_equal(H4.multiply(v), H1.multiply(v))
    # torch inputs of torch's default integer type (int64 indices) are converted.
    import torch

    Ht = irlba.HipSparseMatrix(nr, nc, torch.tensor(vals).cuda(), torch.tensor(idx.astype(np.int64)).cuda(),
                               torch.tensor(ptr.astype(np.int32)).cuda(), True, 2)
    np.testing.assert_array_equal(Ht.multiply(v), H1.multiply(v))


def test_null_space_start_throws():
    A = np.zeros((30, 20))
    A[:, 0] = 1.0
    init = np.zeros(20)
    init[1] = 1.0
    with pytest.raises(irlba.HipError, match="starting vector near the null space"):
        irlba.compute(irlba.HipDenseMatrix(A), 2, irlba.Options(initial=init))


@pytest.mark.parametrize("rank", [2, 5, 10])
def test_invariant_subspace_refills(rank):
    # tests/src/invariant.cpp:31-80: exact low rank drives the refill path.
    nr, nc = 100, 50
    base = O.create_random_matrix(nr, rank, 42)
    A = np.empty((nr, nc))
    for i in range(nc):
        A[:, i] = base[:, i % rank]
    H = irlba.HipDenseMatrix(A)
    res = irlba.compute(H, rank + 3)
    assert np.all(res.D[:rank] > 0.01)
    assert np.all(res.D[rank:] < 1e-8)
    assert same_same(np.sum(res.D ** 2), np.sum(A ** 2), 1e-8)
    assert H.ctx.stats()["refills"] > 0
    res = irlba.compute(H, rank + 2)
    np.testing.assert_allclose(res.U @ np.diag(res.D) @ res.V.T, A, atol=1e-8)


def test_planted_dense_1000x200_vs_oracle():
    # Dense 1000x200 FP64 = U diag(10*0.8^i) V' + 0.01 N(0,1); top-5, extra_work=20,
    # max_iterations=50, tol=1e-5, explicit start vector (seed 1234).
    rng = np.random.default_rng(20240601)
    U0, _ = np.linalg.qr(rng.standard_normal((1000, 200)))
    V0, _ = np.linalg.qr(rng.standard_normal((200, 200)))
    s = 10 * 0.8 ** np.arange(200)
    A = (U0 * s) @ V0.T + 0.01 * rng.standard_normal((1000, 200))
    init = np.random.default_rng(1234).standard_normal(200)
    opt = irlba.Options(extra_work=20, max_iterations=50, convergence_tolerance=1e-5, initial=init)
    res = irlba.compute_simple(A, 5, opt)
    oU, oV, oD, oconv, oit = O.compute(O.dense(A), 5, extra_work=20, max_iterations=50, convergence_tolerance=1e-5,
                                       initial=init)
    assert res.converged and oconv
    assert_equal_vectors(res.D, oD, 1e-8)
    assert subspace_angle(res.U, oU) < 1e-6 and subspace_angle(res.V, oV) < 1e-6
    Sl = np.linalg.svd(A, compute_uv=False)
    assert_equal_vectors(res.D, Sl[:5], 1e-8)


# ---------------------------------------------------------------- pca
@pytest.mark.parametrize("center,scale", [(True, True), (True, False), (False, True), (False, False)])
def test_pca_dense_vs_oracle(center, scale):
    # tests/src/pca.cpp:66-141
    A = O.create_random_matrix(100, 50)
    init = O.create_random_vector(50, 3)
    res = irlba.pca(A, center, scale, 5, irlba.Options(initial=init, convergence_tolerance=1e-9))
    sc, rot, var, conv, it = O.pca(O.dense(A), center, scale, 5, initial=init, convergence_tolerance=1e-9)
    assert_equal_vectors(res.variances, var, 1e-8)
    assert_equal_column_vectors(res.rotation, rot, 1e-6)
    assert_equal_column_vectors(res.scores, sc, 1e-6, centered=center)
    if center:
        assert np.all(np.abs(res.scores.sum(axis=0)) < 1e-8)
        for i in range(5):
            assert same_same((res.scores[:, i] ** 2).sum() / 99, res.variances[i], 1e-8)


def test_pca_sparse_vs_oracle_and_stats():
    # tests/src/pca.cpp:218-240 with the NB generator.
    nr, nc = 400, 600
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 21, f32=False)
    H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr)
    OM = O.sparse(nr, nc, vals, idx, ptr, True, 1)
    c, s = irlba.column_stats(H)
    oc, os_ = O.column_stats(OM)
    np.testing.assert_allclose(c, oc, rtol=1e-13)
    np.testing.assert_allclose(s, os_, rtol=1e-12)
    init = O.create_random_vector(nc, 8)
    res = irlba.pca(H, True, True, 10, irlba.Options(extra_work=20, initial=init, convergence_tolerance=1e-8))
    sc, rot, var, conv, it = O.pca(OM, True, True, 10, extra_work=20, initial=init, convergence_tolerance=1e-8)
    assert_equal_vectors(res.variances, var, 1e-8)
    assert subspace_angle(res.rotation, rot) < 1e-5
    assert np.all(np.abs(res.scores.sum(axis=0)) < 1e-8)
    # Results left in device memory: the same run (fixed-order reductions), bit for bit.
    dres = irlba.pca(H, True, True, 10, irlba.Options(extra_work=20, initial=init, convergence_tolerance=1e-8),
                     device=True)
    assert dres.rotation.is_cuda and tuple(dres.rotation.shape) == (nc, 10) and tuple(dres.scores.shape) == (nr, 10)
    assert dres.iterations == res.iterations
    np.testing.assert_array_equal(dres.variances.cpu().numpy(), res.variances)
    np.testing.assert_array_equal(dres.rotation.cpu().numpy(), res.rotation)
    np.testing.assert_array_equal(dres.scores.cpu().numpy(), res.scores)


def test_pca_long_even_basis_vs_oracle():
    # 600002 columns: the V basis is long and of even length, so its Gram-Schmidt
    # runs on the 16-byte (row pair) dot and update kernels and its restart
    # rotation on the thread-per-row FMA kernel; same start vector as the oracle.
    nr, nc = 200, 600002
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.02, 31, f32=False)
    init = O.create_random_vector(nc, 9)
    opts = dict(extra_work=10, convergence_tolerance=1e-6)
    H = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr)
    res = irlba.pca(H, True, True, 5, irlba.Options(initial=init, **opts))
    sc, rot, var, conv, it = O.pca(O.sparse(nr, nc, vals, idx, ptr, True, 4), True, True, 5, initial=init, **opts)
    assert res.converged and conv
    assert_equal_vectors(res.variances, var, 1e-8)
    assert subspace_angle(res.rotation, rot) < 1e-5
    assert np.abs(res.rotation.T @ res.rotation - np.eye(5)).max() < 1e-10
    assert np.all(np.abs(res.scores.sum(axis=0)) < 1e-8)


def test_pca_full_size_fp64_and_mixed():
    # 30000 x 100000 at 5% (1.5e8 nnz), FP64 storage, pca top-50 with centring and
    # scaling, extra_work 20: too large for the oracle, so size-independent checks.
    #  * the bases are orthonormal, the variances sorted;
    #  * scores = (A - 1c') D^-1 rotation through an independent product route
    #    (decorator handles built from the column statistics), to 1e-9;
    #  * the adjoint residual |A_c' u_i - s_i v_i| <= 1e-5 s_1 (the convergence test);
    #  * every column of the centred, scaled matrix has variance 1, so the variances
    #    are bounded by their total cols and each lies in (0, cols];
    #  * the mixed-precision run (FP32 values and products) gives singular values
    #    within 1e-5 relative of the FP64 run.
    nr, nc, k = 30000, 100000, 50
    A = irlba.generate_sparse(nr, nc, 0.05, 17, storage="f64", accumulate="f64")
    init = np.random.default_rng(5).standard_normal(nc)
    opts = irlba.Options(extra_work=20, initial=init)
    res = irlba.pca(A, True, True, k, opts)
    assert res.converged
    var = res.variances
    assert np.all(np.diff(var) <= 0) and var[-1] > 0 and var.sum() <= nc
    R = res.rotation
    assert np.abs(R.T @ R - np.eye(k)).max() < 1e-10
    c, s = irlba.column_stats(A)
    Ac = irlba.ScaledMatrix(irlba.CenteredMatrix(A, c), s, column=True, divide=True)
    sv = np.sqrt(var * (nr - 1))
    U = res.scores / sv[None, :]
    assert np.abs(U.T @ U - np.eye(k)).max() < 1e-10
    worst_fwd, worst_adj = 0.0, 0.0
    for i in range(0, k, 7):
        y = Ac.multiply(R[:, i])
        worst_fwd = max(worst_fwd, np.linalg.norm(y - res.scores[:, i]) / sv[i])
        z = Ac.adjoint_multiply(U[:, i])
        worst_adj = max(worst_adj, np.linalg.norm(z - sv[i] * R[:, i]) / sv[0])
    assert worst_fwd < 1e-9
    assert worst_adj < 1e-5
    del A, Ac
    M = irlba.generate_sparse(nr, nc, 0.05, 17, storage="f32", accumulate="f32")
    mixed = irlba.pca(M, True, True, k, opts)
    assert mixed.converged
    assert np.max(np.abs(np.sqrt(mixed.variances) - np.sqrt(var)) / np.sqrt(var)) < 1e-5


def test_pca_failures_and_zero_variance():
    with pytest.raises(irlba.HipError, match="cannot scale"):
        irlba.pca(O.create_random_matrix(1, 10), True, True, 1)
    out = irlba.pca(np.zeros((10, 2)), True, True, 1)
    assert out.converged and np.isfinite(out.scores[0, 0]) and np.isfinite(out.rotation[0, 0])


# ---------------------------------------------------------------- caller-defined matrices
class NumpyMatrix(irlba.Matrix):
    def __init__(self, A):
        self.A = A

    def rows(self):
        return self.A.shape[0]

    def cols(self):
        return self.A.shape[1]

    def multiply(self, x):
        return self.A @ x

    def adjoint_multiply(self, x):
        return self.A.T @ x


def test_callback_matrix_matches_device_matrix():
    A = O.create_random_matrix(80, 60)
    init = O.create_random_vector(60, 9)
    opt = irlba.Options(initial=init)
    r1 = irlba.compute(NumpyMatrix(A), 6, opt)
    r2 = irlba.compute(irlba.HipDenseMatrix(A), 6, opt)
    assert_equal_vectors(r1.D, r2.D, 1e-10)
    assert_equal_column_vectors(r1.U, r2.U, 1e-6)


class FailingMatrix(NumpyMatrix):
    """Raises from its product after `ok` successful calls."""

    def __init__(self, A, ok):
        super().__init__(A)
        self.left = ok

    def multiply(self, x):
        self.left -= 1
        if self.left < 0:
            raise ArithmeticError("product failed on purpose")
        return self.A @ x


def test_failing_callback_leaves_context_usable():
    # An exception in a caller-defined product ends the solve with that
    # exception, at the start and in the middle of a sweep; the same context then
    # runs a profiled device solve whose results match a fresh run.
    ctx = irlba.default_context()
    A = O.create_random_matrix(120, 90)
    init = O.create_random_vector(90, 6)
    opt = irlba.Options(initial=init)
    for ok in (0, 7):
        with pytest.raises(ArithmeticError, match="on purpose"):
            irlba.compute(FailingMatrix(A, ok), 5, opt)
    ref = irlba.compute(NumpyMatrix(A), 5, opt)
    before = ctx.stats()["forward_timed"]
    ctx.set_profiling(2)
    try:
        res = irlba.compute(irlba.HipDenseMatrix(A), 5, opt)
        assert ctx.stats()["forward_timed"] > before
    finally:
        ctx.set_profiling(1)  # the default: every product timed
    assert_equal_vectors(res.D, ref.D, 1e-10)


def test_concurrent_solves_on_one_matrix():
    # A Matrix is const and shareable across concurrent compute() calls
    # (interface.hpp:171-185): threads sharing one device handle (one context)
    # get the serial results bit for bit; the context serialises their calls.
    from concurrent.futures import ThreadPoolExecutor

    vals, idx, ptr = O.gen_nb_sparse(3000, 2000, 0.05, 13, f32=False)
    H = irlba.HipSparseMatrix(3000, 2000, vals, idx, ptr)
    D = irlba.HipDenseMatrix(O.create_random_matrix(400, 300))
    inits = [O.create_random_vector(2000, s) for s in range(4)]

    def job(i):
        r1 = irlba.compute(H, 6, irlba.Options(initial=inits[i % 4]))
        r2 = irlba.pca(D, True, True, 4, irlba.Options(initial=O.create_random_vector(300, i % 4)))
        return r1.D, r1.U, r2.variances, r2.scores

    serial = [job(i) for i in range(4)]
    with ThreadPoolExecutor(max_workers=4) as ex:
        concurrent = list(ex.map(job, range(12)))
    for i, got in enumerate(concurrent):
        for a, b in zip(got, serial[i % 4]):
            np.testing.assert_array_equal(a, b)


def test_sampled_product_timing_and_device_results():
    # Context.set_profiling(period): one product in `period` of each direction
    # carries HIP events; the counts and the results do not depend on it.
    ctx = irlba.default_context()
    A = O.create_random_matrix(300, 120)
    init = O.create_random_vector(120, 4)
    opt = irlba.Options(initial=init, extra_work=10)
    H = irlba.HipDenseMatrix(A)
    runs = {}
    try:
        for period in (1, 3, 0):
            ctx.set_profiling(period)
            r = irlba.compute(H, 4, opt)
            st = ctx.stats()
            runs[period] = (r, st)
            for side in ("forward", "adjoint"):
                n, timed = st[f"{side}_products"], st[f"{side}_timed"]
                expect = 0 if period == 0 else (n + period - 1) // period
                assert timed == expect, (period, side, n, timed)
                assert (st[f"{side}_kernel_ms"] > 0) == (period > 0)
        with pytest.raises(irlba.HipError, match="non-negative"):
            ctx.set_profiling(-1)
    finally:
        ctx.set_profiling(1)
    for period in (3, 0):
        np.testing.assert_array_equal(runs[period][0].D, runs[1][0].D)
        np.testing.assert_array_equal(runs[period][0].U, runs[1][0].U)
    # pca with results left in device memory, dense input and a single component.
    P = irlba.pca(H, True, False, 1, opt, device=True)
    Q = irlba.pca(H, True, False, 1, opt)
    assert tuple(P.scores.shape) == (300, 1) and tuple(P.rotation.shape) == (120, 1)
    np.testing.assert_array_equal(P.variances.cpu().numpy(), Q.variances)
    np.testing.assert_array_equal(P.rotation.cpu().numpy(), Q.rotation)
