"""The libtorch torch::autograd::Function binding (torch.ops.irlba_hip.svd_sparse /
svd_dense) on the GPU: gradients through loss.backward() against the oracle's
dense FP64 derivative, torch.linalg.svd autograd (CPU, float64), central
finite differences and torch.autograd.gradcheck.

Tolerances: FP64 storage 1e-8 of the gradient's max entry against the oracle
and torch (the forward triplets converge to 1e-12); 1e-6 against finite
differences; FP32 storage 1e-4.
"""

import numpy as np
import pytest

from oracle import oracle as O
from oracle import svd_backward as OB

pytestmark = pytest.mark.gpu

irlba = pytest.importorskip("cppirlba_amd")
torch = pytest.importorskip("torch")

def _opts():
    return irlba.Options(convergence_tolerance=1e-12, extra_work=20)


def _dense_of(vals, idx, ptr, nr, nc):
    A = np.zeros((nr, nc))
    for c in range(nc):
        A[idx[ptr[c]:ptr[c + 1]], c] = vals[ptr[c]:ptr[c + 1]]
    return A


def _loss(U, S, V, w, GU, GV):
    return (w * S).sum() + (GU * U).sum() + (GV * V).sum()


def _close(G, ref, tol):
    G, ref = np.asarray(G), np.asarray(ref)
    return np.abs(G - ref).max() <= tol * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("m,n", [(64, 32), (40, 90)])
def test_dense_autograd_vs_oracle_and_torch(m, n):
    k = 5
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(1)
    A = rng.standard_normal((m, n))
    w, GU, GV = rng.standard_normal(k), rng.standard_normal((m, k)), rng.standard_normal((n, k))
    At = torch.tensor(A, device=dev, requires_grad=True)
    U, S, V = irlba.svd_dense(At, k, options=_opts())
    assert U.shape == (m, k) and V.shape == (n, k) and S.shape == (k,)
    assert U.is_cuda and U.dtype == torch.float64
    L = _loss(U, S, V, torch.tensor(w, device=dev), torch.tensor(GU, device=dev), torch.tensor(GV, device=dev))
    L.backward()
    grad = At.grad.cpu().numpy()
    Un, Sn, Vn = U.detach().cpu().numpy(), S.detach().cpu().numpy(), V.detach().cpu().numpy()
    Xo, Yo = OB.svd_backward(A, Un, Sn, Vn, w, GU, GV, "exact")
    assert _close(grad, Xo @ Yo.T, 1e-8)
    # torch's own SVD derivative (LAPACK, float64, CPU) with the gradients sign-matched.
    Ac = torch.tensor(A, requires_grad=True)
    Ut, St, Vht = torch.linalg.svd(Ac, full_matrices=False)
    sg = np.sign(np.sum(Ut[:, :k].detach().numpy() * Un, axis=0))
    Lt = _loss(Ut[:, :k], St[:k], Vht[:k].T, torch.tensor(w), torch.tensor(GU * sg), torch.tensor(GV * sg))
    Lt.backward()
    assert _close(grad, Ac.grad.numpy(), 1e-8)


def test_dense_autograd_vs_finite_differences():
    m, n, k = 64, 32, 5
    rng = np.random.default_rng(2)
    A = rng.standard_normal((m, n))
    w, GU, GV = rng.standard_normal(k), rng.standard_normal((m, k)), rng.standard_normal((n, k))
    At = torch.tensor(A, device="cuda:0", requires_grad=True)
    U, S, V = irlba.svd_dense(At, k, options=_opts())
    _loss(U, S, V, torch.tensor(w, device="cuda:0"), torch.tensor(GU, device="cuda:0"),
          torch.tensor(GV, device="cuda:0")).backward()
    # The differences follow LAPACK's signs; <G, U> with U = U_lapack * sg is
    # <G * sg, U_lapack>, so difference that loss.
    sg = np.sign(np.sum(np.linalg.svd(A, full_matrices=False)[0][:, :k] * U.detach().cpu().numpy(), axis=0))
    fd = OB.fd_gradient(A, k, w, GU * sg, GV * sg)
    assert _close(At.grad.cpu().numpy(), fd, 1e-6)


def test_gradcheck_sign_invariant_outputs():
    # torch.autograd.gradcheck (finite differences inside torch) on outputs that do
    # not depend on the signs of the vectors: the values and sum_i c_i u_i v_i'.
    k = 3
    A = torch.tensor(np.random.default_rng(3).standard_normal((12, 8)), device="cuda:0", requires_grad=True)
    c = torch.tensor([1.0, -0.5, 2.0], device="cuda:0")

    def f(a):
        U, S, V = irlba.svd_dense(a, k, options=_opts())
        return S, (U * c) @ V.T

    assert torch.autograd.gradcheck(f, (A,), eps=1e-6, atol=1e-6, rtol=1e-5)


@pytest.mark.parametrize("column_major", [True, False])
def test_sparse_autograd_values_and_center(column_major):
    nr, nc, k = 300, 160, 6
    dev = torch.device("cuda:0")
    if column_major:
        vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.08, 5, f32=False)
        B = _dense_of(vals, idx, ptr, nr, nc)
    else:
        vals, idx, ptr = O.gen_nb_sparse(nc, nr, 0.08, 5, f32=False)
        B = _dense_of(vals, idx, ptr, nc, nr).T.copy()
    c = B.mean(axis=0)
    A = B - c[None, :]
    rng = np.random.default_rng(6)
    w, GU, GV = rng.standard_normal(k), rng.standard_normal((nr, k)), rng.standard_normal((nc, k))
    x = torch.tensor(vals, device=dev, requires_grad=True)
    ct = torch.tensor(c, device=dev, requires_grad=True)
    U, S, V = irlba.svd_sparse(x, torch.tensor(idx, device=dev), torch.tensor(ptr, device=dev), nr, nc, k,
                               center=ct, column_major=column_major, options=_opts())
    _loss(U, S, V, torch.tensor(w, device=dev), torch.tensor(GU, device=dev), torch.tensor(GV, device=dev)).backward()
    Un, Sn, Vn = (t.detach().cpu().numpy() for t in (U, S, V))
    np.testing.assert_allclose(Sn, np.linalg.svd(A, compute_uv=False)[:k], rtol=1e-10)
    Xo, Yo = OB.svd_backward(A, Un, Sn, Vn, w, GU, GV, "exact")
    G = Xo @ Yo.T
    lines = nc if column_major else nr
    ref = np.empty(idx.size)
    for l in range(lines):
        sl = slice(ptr[l], ptr[l + 1])
        ref[sl] = G[idx[sl], l] if column_major else G[l, idx[sl]]
    assert x.grad.shape == x.shape
    assert _close(x.grad.cpu().numpy(), ref, 1e-8)
    assert _close(ct.grad.cpu().numpy(), OB.grad_center(Xo, Yo), 1e-8)


def test_sparse_autograd_host_tensors_and_fp32_storage():
    nr, nc, k = 200, 120, 4
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.1, 8, f32=False)
    A = _dense_of(vals, idx, ptr, nr, nc)
    rng = np.random.default_rng(9)
    w, GU = rng.standard_normal(k), rng.standard_normal((nr, k))
    # Host tensors: the engine copies them in and the gradients out.
    x = torch.tensor(vals, requires_grad=True)
    U, S, V = irlba.svd_sparse(x, torch.tensor(idx), torch.tensor(ptr), nr, nc, k, options=_opts())
    assert not U.is_cuda
    ((torch.tensor(w) * S).sum() + (torch.tensor(GU) * U).sum()).backward()
    assert x.grad.dtype == torch.float64 and not x.grad.is_cuda
    Un, Sn, Vn = (t.detach().numpy() for t in (U, S, V))
    Xo, Yo = OB.svd_backward(A, Un, Sn, Vn, w, GU, None, "exact")
    G = Xo @ Yo.T
    ref = np.concatenate([G[idx[ptr[c]:ptr[c + 1]], c] for c in range(nc)])
    assert _close(x.grad.numpy(), ref, 1e-8)
    # FP32 storage (mixed precision): a sign-invariant loss, so the gradient is
    # U diag(w) V' on the pattern up to the FP32 rounding of the matrix.
    grads = []
    for storage, dtype in (("f64", torch.float64), ("f32", torch.float32)):
        x = torch.tensor(vals, dtype=dtype, requires_grad=True)
        U, S, V = irlba.svd_sparse(x, torch.tensor(idx), torch.tensor(ptr), nr, nc, k, options=_opts(),
                                   storage=storage)
        (torch.tensor(w) * S).sum().backward()
        assert x.grad.dtype == dtype
        grads.append(x.grad.double().numpy())
    assert _close(grads[1], grads[0], 1e-4)


@pytest.mark.parametrize("complement", ["rank_k", "exact"])
def test_queued_backward_checks_like_the_engine(complement):
    # The device backward is queued without reading back the engine's check of
    # the singular values (positive, non-increasing); the binding makes that
    # check when the forward returns them.  A rank-1 matrix asked for 3
    # triplets gives s_2 = s_3 = 0 in the reference algorithm (the engine too
    # returns exact zeros there, so both backwards fail); whatever values the
    # forward returns, the queued backward must fail or succeed as the
    # synchronous C ABI does on the same triplets, with the same message or
    # the same gradient.
    rng = np.random.default_rng(77)
    A0 = np.outer(rng.standard_normal(40), rng.standard_normal(30))
    A = torch.tensor(A0, device="cuda", requires_grad=True)
    opts = irlba.Options(extra_work=7)
    U, S, V = irlba.svd_dense(A, 3, options=opts, complement=complement)
    w = torch.tensor(rng.standard_normal(3), device="cuda")
    Ud, Sd, Vd = (t.detach().cpu().numpy() for t in (U, S, V))
    try:
        ref = irlba.svd_backward(irlba.HipDenseMatrix(A0), Ud, Sd, Vd, grad_D=w.cpu().numpy(),
                                 complement=complement)
        ref_error = None
    except irlba.HipError as e:
        ref, ref_error = None, str(e)
    loss = (w * S).sum()
    print(f"singular values {Sd}; synchronous backward: {ref_error or 'ran'}")
    if ref_error is not None:
        assert "positive and non-increasing" in ref_error
        with pytest.raises(RuntimeError, match="positive and non-increasing"):
            loss.backward()
    else:
        loss.backward()
        np.testing.assert_allclose(A.grad.cpu().numpy(), ref.dense(), rtol=1e-10, atol=1e-12)
