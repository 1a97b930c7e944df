"""GPU parity of the top-k SVD backward (irlba_hip_svd_backward, irlba_hip_sddmm)
against the oracle's dense restatement and against finite differences.

Tolerances (FP64 throughout):
  * vs the oracle: max |X Y' - X_o Y_o'| <= 1e-9 of max |X_o Y_o'| (the engine's
    complement comes from CG solves to 1e-13, the oracle's from a full SVD);
  * vs central finite differences: 1e-6 of the gradient's max entry;
  * SDDMM: 1e-13 relative (a dot product of length 2k per entry).
"""

import numpy as np
import pytest

from oracle import oracle as O
from oracle import svd_backward as OB

pytestmark = pytest.mark.gpu

irlba = pytest.importorskip("cppirlba_amd")


def _triplets(A, k):
    Uf, sf, Vtf = np.linalg.svd(A, full_matrices=False)
    return Uf[:, :k].copy(), sf[:k].copy(), Vtf[:k].T.copy()


def _grads(m, n, k, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(k), rng.standard_normal((m, k)), rng.standard_normal((n, k))


def _dense_of(vals, idx, ptr, nr, nc):
    A = np.zeros((nr, nc))
    for c in range(nc):
        A[idx[ptr[c]:ptr[c + 1]], c] = vals[ptr[c]:ptr[c + 1]]
    return A


def _close(G, ref, tol):
    return np.abs(G - ref).max() <= tol * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("m,n", [(64, 32), (32, 64)])
@pytest.mark.parametrize("complement", ["none", "rank_k", "exact"])
def test_dense_backward_vs_oracle(m, n, complement):
    k = 5
    A = np.random.default_rng(1).standard_normal((m, n))
    U, s, V = _triplets(A, k)
    w, GU, GV = _grads(m, n, k, 2)
    h = irlba.HipDenseMatrix(A, storage="f64")
    res = irlba.svd_backward(h, U, s, V, w, GU, GV, complement=complement)
    Xo, Yo = OB.svd_backward(A, U, s, V, w, GU, GV, complement)
    assert res.X.shape == (m, 2 * k) and res.Y.shape == (n, 2 * k)
    assert _close(res.dense(), Xo @ Yo.T, 1e-9)
    if complement != "exact":
        assert res.cg_iterations == 0


@pytest.mark.parametrize("storage,k", [("f64", 20), ("f32", 20), ("f64", 12)])
def test_exact_backward_wider_k(storage, k):
    # The lock-step CG shares each pass over the matrix between its two sets of
    # systems when 2k vectors fit one pass (dense: 32 packed columns); k = 20
    # runs the sets one after the other, k = 12 shares the passes.  A matrix
    # with a gap under s_k keeps the systems well conditioned.
    m, n = 300, 180
    rng = np.random.default_rng(30 + k)
    Q1, _ = np.linalg.qr(rng.standard_normal((m, k)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, k)))
    A = (Q1 * (40.0 + np.arange(k)[::-1])) @ Q2.T + 0.3 * rng.standard_normal((m, n))
    if storage == "f32":
        A = A.astype(np.float32).astype(np.float64)  # the stored matrix
    U, s, V = _triplets(A, k)
    w, GU, GV = _grads(m, n, k, 31)
    res = irlba.svd_backward(irlba.HipDenseMatrix(A, storage=storage), U, s, V, w, GU, GV, complement="exact")
    Xo, Yo = OB.svd_backward(A, U, s, V, w, GU, GV, "exact")
    assert res.cg_iterations > 0
    # FP32 storage: the products see FP32-rounded inputs, so the CG converges to 1e-7 of the answer.
    assert _close(res.dense(), Xo @ Yo.T, 1e-9 if storage == "f64" else 1e-6)


def test_sparse_exact_backward_k40():
    # 40 systems per set: the fused CG update kernels past 32 columns, and the
    # shared passes over the matrix with 80 packed vectors.
    nr, nc, k = 700, 400, 40
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.08, 41, f32=False)
    A = _dense_of(vals, idx, ptr, nr, nc)
    h = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, storage="f64")
    U, s, V = _triplets(A, k)
    w, GU, GV = _grads(nr, nc, k, 42)
    res = irlba.svd_backward(h, U, s, V, w, GU, GV, complement="exact")
    Xo, Yo = OB.svd_backward(A, U, s, V, w, GU, GV, "exact")
    assert res.cg_iterations > 0
    assert _close(res.dense(), Xo @ Yo.T, 1e-9)


def test_dense_backward_vs_finite_differences():
    # 64 x 32 FP64, the whole gradient by central differences of a sign-fixed SVD.
    m, n, k = 64, 32, 5
    A = np.random.default_rng(3).standard_normal((m, n))
    U, s, V = _triplets(A, k)
    w, GU, GV = _grads(m, n, k, 4)
    res = irlba.svd_backward(irlba.HipDenseMatrix(A), U, s, V, w, GU, GV, complement="exact")
    fd = OB.fd_gradient(A, k, w, GU, GV)
    assert _close(res.dense(), fd, 1e-6)
    assert res.cg_iterations > 0


def test_rank_k_modes_agree_on_rank_k_matrix():
    rng = np.random.default_rng(5)
    k = 4
    A = rng.standard_normal((80, k)) @ rng.standard_normal((k, 60))
    U, s, V = _triplets(A, k)
    w, GU, GV = _grads(80, 60, k, 6)
    h = irlba.HipDenseMatrix(A)
    r1 = irlba.svd_backward(h, U, s, V, w, GU, GV, complement="rank_k")
    r2 = irlba.svd_backward(h, U, s, V, w, GU, GV, complement="exact")
    assert _close(r1.dense(), r2.dense(), 1e-8)
    Xo, Yo = OB.svd_backward(A, U, s, V, w, GU, GV, "rank_k")
    assert _close(r1.dense(), Xo @ Yo.T, 1e-10)


@pytest.mark.parametrize("nr,nc,column_major", [(500, 300, True), (300, 500, True), (400, 250, False)])
def test_sparse_backward_and_sddmm(nr, nc, column_major):
    k = 6
    if column_major:
        vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 11, f32=False)
        A = _dense_of(vals, idx, ptr, nr, nc)
    else:
        vals, idx, ptr = O.gen_nb_sparse(nc, nr, 0.05, 11, f32=False)  # CSC of A' = CSR of A
        A = _dense_of(vals, idx, ptr, nc, nr).T.copy()
    h = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, column_major=column_major, storage="f64")
    U, s, V = _triplets(A, k)
    w, GU, GV = _grads(nr, nc, k, 12)
    res = irlba.svd_backward(h, U, s, V, w, GU, GV, complement="exact")
    Xo, Yo = OB.svd_backward(A, U, s, V, w, GU, GV, "exact")
    G = Xo @ Yo.T
    assert _close(res.dense(), G, 1e-9)
    # Gradient of the stored values: the dense gradient on the pattern.
    gvals = irlba.sddmm(nr, nc, idx, ptr, res.X, res.Y, column_major=column_major)
    lines = nc if column_major else nr
    ref = np.empty(idx.size)
    for l in range(lines):
        sl = slice(ptr[l], ptr[l + 1])
        ref[sl] = G[idx[sl], l] if column_major else G[l, idx[sl]]
    assert np.abs(gvals - ref).max() <= 1e-9 * np.abs(ref).max()
    direct = np.einsum("ij,ij->i", res.X[(idx if column_major else np.repeat(np.arange(lines), np.diff(ptr)))],
                       res.Y[(np.repeat(np.arange(lines), np.diff(ptr)) if column_major else idx)])
    assert np.abs(gvals - direct).max() <= 1e-13 * np.abs(direct).max()


@pytest.mark.parametrize("R", [1, 3, 10, 32, 64, 128])
@pytest.mark.parametrize("column_major", [True, False])
def test_sddmm_ranks_and_precisions(R, column_major):
    # Every lanes-per-entry variant (R up to 128, FP32 and FP64 factors) against numpy.
    nr, nc = 700, 450
    vals, idx, ptr = O.gen_nb_sparse(nr if column_major else nc, nc if column_major else nr, 0.04, R, f32=False)
    rng = np.random.default_rng(R)
    X = rng.standard_normal((nr, R))
    Y = rng.standard_normal((nc, R))
    lines = nc if column_major else nr
    owner = np.repeat(np.arange(lines), np.diff(ptr))
    rows, cols = (idx, owner) if column_major else (owner, idx)
    ref = np.einsum("ij,ij->i", X[rows], Y[cols])
    out64 = irlba.sddmm(nr, nc, idx, ptr, X, Y, column_major=column_major)
    assert out64.dtype == np.float64
    assert np.abs(out64 - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    out32 = irlba.sddmm(nr, nc, idx, ptr, X, Y, column_major=column_major, out_type="f32")
    assert out32.dtype == np.float32
    # FP32 factors: each product carries ~2^-24 relative rounding per factor.
    assert np.abs(out32 - ref).max() <= 1e-6 * np.sqrt(R) * np.abs(X).max() * np.abs(Y).max() * 4


def test_sddmm_rank_bounds_and_bad_pattern():
    X = np.ones((4, 2))
    Y = np.ones((3, 2))
    idx = np.array([0, 5], dtype=np.int32)  # row 5 is out of range for 4 rows
    ptr = np.array([0, 1, 2, 2], dtype=np.int64)
    with pytest.raises(irlba.HipError, match="out of range"):
        irlba.sddmm(4, 3, idx, ptr, X, Y)
    with pytest.raises(irlba.HipError, match="rank"):
        irlba.sddmm(4, 3, np.array([0, 1], dtype=np.int32), ptr, np.ones((4, 129)), np.ones((3, 129)))
    out = irlba.sddmm(4, 3, np.zeros(0, dtype=np.int32), np.zeros(4, dtype=np.int64), X, Y)
    assert out.size == 0


@pytest.mark.parametrize("column_major", [True, False])
@pytest.mark.parametrize("out_type", ["f64", "f32"])
def test_sddmm_large_entry_factors(column_major, out_type):
    # Entry factors larger than an XCD's share of L2 (20000 x 32) and lines
    # whose indices are not sorted.
    R, n_other, lines = 32, 20000, 300
    nr, nc = (n_other, lines) if column_major else (lines, n_other)
    vals, idx, ptr = O.gen_nb_sparse(n_other, lines, 0.02, 3, f32=False)
    rng = np.random.default_rng(11)
    for c in range(0, lines, 7):  # shuffle some lines
        rng.shuffle(idx[ptr[c]:ptr[c + 1]])
    X = rng.standard_normal((nr, R))
    Y = rng.standard_normal((nc, R))
    owner = np.repeat(np.arange(lines), np.diff(ptr))
    rows, cols = (idx, owner) if column_major else (owner, idx)
    ref = np.einsum("ij,ij->i", X[rows], Y[cols])
    out = irlba.sddmm(nr, nc, idx, ptr, X, Y, column_major=column_major, out_type=out_type)
    if out_type == "f64":
        assert np.abs(out - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    else:
        assert np.abs(out - ref).max() <= 1e-6 * np.sqrt(R) * np.abs(X).max() * np.abs(Y).max() * 4
    bad = idx.copy()
    bad[len(bad) // 3] = n_other + 5
    with pytest.raises(irlba.HipError, match="out of range"):
        irlba.sddmm(nr, nc, bad, ptr, X, Y, column_major=column_major, out_type=out_type)


def test_sddmm_device_pointers_and_checked_layout():
    # Device arrays are read in place, their layout checked by the kernel as it
    # streams them: the same results as the host route, and the same errors.
    import ctypes

    import torch

    from cppirlba_amd import _lib

    nr, nc, R = 300, 200, 8
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 4, f32=False)
    rng = np.random.default_rng(9)
    X = rng.standard_normal((nr, R))
    Y = rng.standard_normal((nc, R))
    host = irlba.sddmm(nr, nc, idx, ptr, X, Y)
    ctx = irlba.default_context()

    def dev_sddmm(i, p):
        ti = torch.tensor(i, dtype=torch.int32, device="cuda")
        tp = torch.tensor(p, dtype=torch.int64, device="cuda")
        tX = torch.tensor(np.asfortranarray(X).T.copy().reshape(-1), device="cuda")  # column-major nr x R
        tY = torch.tensor(np.asfortranarray(Y).T.copy().reshape(-1), device="cuda")
        out = torch.zeros(len(i), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _lib.check(_lib.lib().irlba_hip_sddmm(ctx.handle, nr, nc, ctypes.c_void_p(ti.data_ptr()),
                                              ctypes.c_void_p(tp.data_ptr()), 1, R, ctypes.c_void_p(tX.data_ptr()),
                                              ctypes.c_void_p(tY.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                              _lib.F64, 1))
        return out.cpu().numpy()

    np.testing.assert_array_equal(dev_sddmm(idx, ptr), host)
    bad_idx = idx.copy()
    bad_idx[len(bad_idx) // 2] = nr  # one row past the end
    with pytest.raises(irlba.HipError, match="out of range"):
        dev_sddmm(bad_idx, ptr)
    bad_ptr = ptr.copy()
    bad_ptr[5], bad_ptr[6] = ptr[6], ptr[5] - 1  # a decreasing pair of bounds
    with pytest.raises(irlba.HipError, match="compressed layout"):
        dev_sddmm(idx, bad_ptr)


@pytest.mark.parametrize("complement", ["none", "rank_k", "exact"])
def test_async_backward_and_sddmm_match_synchronous(complement):
    # The _async entry points queue the same kernels without waiting for them;
    # ordered against torch's current stream by an event each way, their
    # outputs equal the synchronous entry points' bit for bit.  The inputs are
    # written on torch's stream behind a long matrix product, so a read that
    # did not wait for that stream would see them unwritten.
    import ctypes

    import scipy.sparse as sp
    import torch

    from cppirlba_amd import _lib

    L = _lib.lib()
    ctx = irlba.default_context()
    nr, nc, k = 500, 300, 6
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.05, 41, f32=False)
    A = _dense_of(vals, idx, ptr, nr, nc)
    h = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, storage="f64")
    U, s, V = _triplets(A, k)
    w, GU, GV = _grads(nr, nc, k, 42)

    def dev(a):  # column-major FP64 on the device
        a = np.asarray(a, dtype=np.float64)
        return torch.tensor(a.T.copy().reshape(-1) if a.ndim == 2 else a, device="cuda")

    def vp(t):
        return ctypes.c_void_p(t.data_ptr())

    ins = [dev(a) for a in (U, s, V, w, GU, GV)]
    X0 = torch.zeros(nr * 2 * k, dtype=torch.float64, device="cuda")
    Y0 = torch.zeros(nc * 2 * k, dtype=torch.float64, device="cuda")
    c0 = torch.zeros(nc, dtype=torch.float64, device="cuda")
    it0, it1 = ctypes.c_int32(), ctypes.c_int32()
    code = {"none": 0, "rank_k": 1, "exact": 2}[complement]
    torch.cuda.synchronize()
    _lib.check(L.irlba_hip_svd_backward(h.handle, k, vp(ins[0]), vp(ins[1]), vp(ins[2]), vp(ins[3]), vp(ins[4]),
                                        vp(ins[5]), code, 1000, 1e-12, vp(X0), vp(Y0), vp(c0), ctypes.byref(it0), 1))
    patterns = {"csc": (idx, ptr, 1)}
    csr = sp.csr_matrix(A)
    patterns["csr"] = (csr.indices.astype(np.int32), csr.indptr.astype(np.int64), 0)
    sync_out = {}
    for name, (pi, pp, cm) in patterns.items():
        ti = torch.tensor(pi, device="cuda")
        tp = torch.tensor(pp, device="cuda")
        o = torch.zeros(len(pi), dtype=torch.float64, device="cuda")
        _lib.check(L.irlba_hip_sddmm(ctx.handle, nr, nc, vp(ti), vp(tp), cm, 2 * k, vp(X0), vp(Y0), vp(o),
                                     _lib.F64, 1))
        sync_out[name] = o.cpu()

    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # Host-to-device copies wait for torch's stream, so they come first.
    pats = {name: (torch.tensor(pi, device="cuda"), torch.tensor(pp, device="cuda"), cm)
            for name, (pi, pp, cm) in patterns.items()}
    big = torch.randn(4096, 4096, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    slow = big @ big  # queued ahead of the input writes on torch's stream
    late = [torch.empty_like(t) for t in ins]
    for dst, src in zip(late, ins):
        dst.copy_(src)
    X1 = torch.full_like(X0, float("nan"))
    Y1 = torch.full_like(Y0, float("nan"))
    c1 = torch.full_like(c0, float("nan"))
    outs = {name: torch.full((len(pi),), float("nan"), dtype=torch.float64, device="cuda")
            for name, (pi, _, _) in patterns.items()}
    _lib.check(L.irlba_hip_context_wait_stream(ctx.handle, stream))
    _lib.check(L.irlba_hip_svd_backward_async(h.handle, k, vp(late[0]), vp(late[1]), vp(late[2]), vp(late[3]),
                                              vp(late[4]), vp(late[5]), code, 1000, 1e-12, vp(X1), vp(Y1), vp(c1),
                                              ctypes.byref(it1)))
    for name, (ti, tp, cm) in pats.items():
        _lib.check(L.irlba_hip_sddmm_async(ctx.handle, nr, nc, vp(ti), vp(tp), cm, 2 * k, vp(X1), vp(Y1),
                                           vp(outs[name]), _lib.F64))
    _lib.check(L.irlba_hip_stream_wait_context(ctx.handle, stream))
    # Read on torch's stream, which now waits for the engine's.
    assert torch.equal(X1.cpu(), X0.cpu()) and torch.equal(Y1.cpu(), Y0.cpu()) and torch.equal(c1.cpu(), c0.cpu())
    for name in patterns:
        assert torch.equal(outs[name].cpu(), sync_out[name]), name
    assert it1.value == it0.value
    del slow
    # Device arrays and all three gradients only.
    with pytest.raises(irlba.HipError, match="all three gradients"):
        _lib.check(L.irlba_hip_svd_backward_async(h.handle, k, vp(ins[0]), vp(ins[1]), vp(ins[2]), None,
                                                  vp(ins[4]), vp(ins[5]), code, 1000, 1e-12, vp(X1), vp(Y1), None,
                                                  None))
    ti, tp, cm = pats["csc"]
    with pytest.raises(irlba.HipError, match="device arrays"):
        _lib.check(L.irlba_hip_sddmm_async(ctx.handle, nr, nc, vp(ti), vp(tp), cm, 2 * k, vp(X1), vp(Y1), None,
                                           _lib.F64))
    ctx.synchronize()


@pytest.mark.parametrize("nr,nc,R", [(1, 1, 1), (70, 130, 3), (300, 65, 32), (129, 200, 33), (64, 64, 128)])
@pytest.mark.parametrize("column_major", [True, False])
def test_dense_gradient_vs_numpy(nr, nc, R, column_major):
    # X Y' written as a dense matrix in FP64, FP32 and BF16 (ragged tiles, rank
    # slices of 32 and their remainders), from host and from device arrays; the
    # queued form gives the same bits.  FP64 to 1e-13 relative; FP32 and BF16
    # within one rounding (2^-24, 2^-8 relative) of the FP64 product.
    import ctypes

    import torch

    from cppirlba_amd import _lib

    L = _lib.lib()
    ctx = irlba.default_context()
    rng = np.random.default_rng(nr + nc + R)
    X = np.asfortranarray(rng.standard_normal((nr, R)))
    Y = np.asfortranarray(rng.standard_normal((nc, R)))
    ref = X @ Y.T
    scale = np.abs(X).sum(axis=1)[:, None] * np.abs(Y).sum(axis=1)[None, :]  # bound of each sum's magnitudes
    order = "F" if column_major else "C"
    for code, dt, tol in ((_lib.F64, np.float64, 1e-15), (_lib.F32, np.float32, 2.0 ** -24),
                          (_lib.BF16, None, 2.0 ** -8)):
        nbytes = 2 if dt is None else np.dtype(dt).itemsize
        host = np.zeros(nr * nc * nbytes, dtype=np.uint8)
        _lib.check(L.irlba_hip_dense_gradient(ctx.handle, nr, nc, R, X.ctypes.data, Y.ctypes.data, host.ctypes.data,
                                              code, int(column_major), 0))
        if dt is None:  # bfloat16: the high half of an FP32
            got = (host.view(np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        else:
            got = host.view(dt).astype(np.float64)
        got = got.reshape((nr, nc), order=order)
        err = np.abs(got - ref)
        # Both FP64 sums (the kernel's, numpy's) lie within R u sum|x_r y_r| of the exact one.
        assert np.all(err <= 8 * tol * np.abs(ref) + 2.5e-16 * R * scale), (code, err.max())
        # Device arrays, synchronous and queued.
        tX = torch.tensor(X.T.copy().reshape(-1), device="cuda")
        tY = torch.tensor(Y.T.copy().reshape(-1), device="cuda")
        d0 = torch.zeros(nr * nc * nbytes, dtype=torch.uint8, device="cuda")
        d1 = torch.full((nr * nc * nbytes,), 255, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        args = (ctx.handle, nr, nc, R, ctypes.c_void_p(tX.data_ptr()), ctypes.c_void_p(tY.data_ptr()))
        _lib.check(L.irlba_hip_dense_gradient(*args, ctypes.c_void_p(d0.data_ptr()), code, int(column_major), 1))
        _lib.check(L.irlba_hip_dense_gradient_async(*args, ctypes.c_void_p(d1.data_ptr()), code,
                                                    int(column_major)))
        ctx.synchronize()
        assert np.array_equal(d0.cpu().numpy(), host) and np.array_equal(d1.cpu().numpy(), host), code
    with pytest.raises(irlba.HipError, match="rank"):
        _lib.check(L.irlba_hip_dense_gradient(ctx.handle, nr, nc, 129, X.ctypes.data, Y.ctypes.data,
                                              host.ctypes.data, _lib.F64, 0, 0))


def test_centered_backward_and_center_gradient():
    nr, nc, k = 300, 120, 5
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.1, 21, f32=False)
    Bm = _dense_of(vals, idx, ptr, nr, nc)
    c = Bm.mean(axis=0)
    A = Bm - c[None, :]
    base = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, storage="f64")
    h = irlba.CenteredMatrix(base, c)
    U, s, V = _triplets(A, k)
    w, GU, GV = _grads(nr, nc, k, 22)
    res = irlba.svd_backward(h, U, s, V, w, GU, GV, complement="exact", grad_center=True)
    Xo, Yo = OB.svd_backward(A, U, s, V, w, GU, GV, "exact")
    assert _close(res.dense(), Xo @ Yo.T, 1e-9)
    gc = OB.grad_center(Xo, Yo)
    assert np.abs(res.grad_center - gc).max() <= 1e-9 * max(1.0, np.abs(gc).max())


def test_backward_of_gpu_compute():
    # End to end: triplets from the device solver, then the backward on them.
    nr, nc, k = 600, 400, 8
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.08, 31, f32=False)
    A = _dense_of(vals, idx, ptr, nr, nc)
    h = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, storage="f64")
    r = irlba.compute(h, k, irlba.Options(convergence_tolerance=1e-12, extra_work=20, exact_for_small_matrix=False))
    assert r.converged
    w, GU, GV = _grads(nr, nc, k, 32)
    res = irlba.svd_backward(h, r.U, r.D, r.V, w, GU, GV, complement="exact")
    Xo, Yo = OB.svd_backward(A, r.U, r.D, r.V, w, GU, GV, "exact")
    assert _close(res.dense(), Xo @ Yo.T, 1e-6)


def test_backward_argument_errors():
    A = np.random.default_rng(0).standard_normal((20, 10))
    U, s, V = _triplets(A, 3)
    h = irlba.HipDenseMatrix(A)
    with pytest.raises(irlba.HipError, match="non-increasing"):
        irlba.svd_backward(h, U, s[::-1].copy(), V)
    with pytest.raises(ValueError):
        irlba.svd_backward(h, U[:, :2], s, V)
    with pytest.raises(KeyError):
        irlba.svd_backward(h, U, s, V, complement="bogus")
    # Values-only gradient: U diag(w) V'.
    w = np.array([1.0, -2.0, 0.5])
    res = irlba.svd_backward(h, U, s, V, grad_D=w, complement="none")
    np.testing.assert_allclose(res.dense(), U @ np.diag(w) @ V.T, atol=1e-13)


@pytest.mark.parametrize("name", ["tall", "wide"])
def test_engine_matches_torch_fixture(name):
    # The engine's exact backward against the committed torch.linalg.svd derivative.
    import os

    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "svd_backward_torch.npz"))
    g = {k[len(name) + 1:]: f[k] for k in f.files if k.startswith(name + "_")}
    res = irlba.svd_backward(irlba.HipDenseMatrix(g["A"]), g["U"], g["s"], g["V"], g["w"], g["GU"], g["GV"],
                             complement="exact")
    assert _close(res.dense(), g["dA"], 1e-9)
