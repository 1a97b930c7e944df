"""Pins the oracle's top-k SVD backward (oracle/svd_backward.py) on the CPU.

The backward has no counterpart in the reference, so it is pinned by what the
mathematics fixes: central finite differences of a sign-fixed SVD (64 x 32 and
32 x 64, FP64), torch.linalg.svd's own autograd (CPU, float64), the closed form
U diag(w) V' for a loss on the singular values alone, and the agreement of the
"rank_k" and "exact" complements on a matrix of rank k.
"""

import numpy as np
import pytest

from oracle import svd_backward as B


def _problem(m, n, k, seed, rank=None):
    rng = np.random.default_rng(seed)
    if rank is None:
        A = rng.standard_normal((m, n))
    else:
        A = rng.standard_normal((m, rank)) @ rng.standard_normal((rank, n))
    Uf, sf, Vtf = np.linalg.svd(A, full_matrices=False)
    U, s, V = Uf[:, :k].copy(), sf[:k].copy(), Vtf[:k].T.copy()
    w = rng.standard_normal(k)
    GU = rng.standard_normal((m, k))
    GV = rng.standard_normal((n, k))
    return A, U, s, V, w, GU, GV


@pytest.mark.parametrize("m,n", [(64, 32), (32, 64)])
def test_exact_matches_finite_differences(m, n):
    k = 5
    A, U, s, V, w, GU, GV = _problem(m, n, k, 11)
    X, Y = B.svd_backward(A, U, s, V, w, GU, GV, "exact")
    fd = B.fd_gradient(A, k, w, GU, GV)
    scale = max(1.0, np.abs(fd).max())
    assert np.abs(X @ Y.T - fd).max() < 1e-6 * scale


def test_dropping_a_term_fails_finite_differences():
    # Each part of the formula is needed: without the complement, or with the
    # in-subspace antisymmetrisation transposed, the gradient is wrong.
    k = 4
    A, U, s, V, w, GU, GV = _problem(40, 24, k, 5)
    fd = B.fd_gradient(A, k, w, GU, GV)
    X, Y = B.svd_backward(A, U, s, V, w, GU, GV, "none")
    assert np.abs(X @ Y.T - fd).max() > 1e-2
    J, K = U.T @ GU, V.T @ GV
    bad = B.in_subspace_matrix(s, w, J.T, K)
    assert np.abs(bad - B.in_subspace_matrix(s, w, J, K)).max() > 1e-3


def test_rank_k_complement_is_exact_on_rank_k_matrices():
    k = 3
    A, U, s, V, w, GU, GV = _problem(50, 30, k, 7, rank=k)
    X1, Y1 = B.svd_backward(A, U, s, V, w, GU, GV, "rank_k")
    X2, Y2 = B.svd_backward(A, U, s, V, w, GU, GV, "exact")
    np.testing.assert_allclose(X1 @ Y1.T, X2 @ Y2.T, atol=1e-10)
    fd = B.fd_gradient(A, k, w, GU, GV)
    assert np.abs(X1 @ Y1.T - fd).max() < 1e-6 * max(1.0, np.abs(fd).max())


def test_values_only_closed_form():
    k = 6
    A, U, s, V, w, _, _ = _problem(30, 20, k, 3)
    X, Y = B.svd_backward(A, U, s, V, w, None, None, "exact")
    np.testing.assert_allclose(X @ Y.T, U @ np.diag(w) @ V.T, atol=1e-13)


@pytest.mark.parametrize("m,n,k", [(64, 32, 5), (32, 64, 8), (50, 50, 10)])
def test_exact_matches_torch_autograd(m, n, k):
    torch = pytest.importorskip("torch")
    A, U, s, V, w, GU, GV = _problem(m, n, k, 21)
    At = torch.tensor(A, dtype=torch.float64, requires_grad=True)
    Ut, St, Vht = torch.linalg.svd(At, full_matrices=False)
    # torch's vectors may differ in sign from U, V: flip the gradients to match.
    sg = np.sign(np.sum(Ut[:, :k].detach().numpy() * U, axis=0))
    L = (torch.tensor(w) * St[:k]).sum() + (torch.tensor(GU * sg) * Ut[:, :k]).sum() + \
        (torch.tensor(GV * sg) * Vht[:k].T).sum()
    L.backward()
    X, Y = B.svd_backward(A, U, s, V, w, GU, GV, "exact")
    np.testing.assert_allclose(X @ Y.T, At.grad.numpy(), atol=1e-10 * max(1.0, np.abs(At.grad.numpy()).max()))


def test_center_gradient_matches_finite_differences():
    k = 4
    rng = np.random.default_rng(9)
    Bm = rng.standard_normal((40, 25)) + 3.0
    c = rng.standard_normal(25)
    A = Bm - c[None, :]
    A_, U, s, V, w, GU, GV = _problem(40, 25, k, 9)
    Uf, sf, Vtf = np.linalg.svd(A, full_matrices=False)
    U, s, V = Uf[:, :k], sf[:k], Vtf[:k].T
    X, Y = B.svd_backward(A, U, s, V, w, GU, GV, "exact")
    g = B.grad_center(X, Y)
    eps = 1e-6
    fd = np.zeros(25)
    for j in range(25):
        cp, cm = c.copy(), c.copy()
        cp[j] += eps
        cm[j] -= eps
        fd[j] = (B.loss(Bm - cp[None, :], k, w, GU, GV, U, V) - B.loss(Bm - cm[None, :], k, w, GU, GV, U, V)) / (2 * eps)
    assert np.abs(g - fd).max() < 1e-6 * max(1.0, np.abs(fd).max())


@pytest.mark.parametrize("name", ["tall", "wide"])
def test_oracle_matches_torch_fixture(name):
    # Committed vectors from torch.linalg.svd's autograd (tests/golden/make_backward_fixtures.py).
    import os

    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "svd_backward_torch.npz"))
    g = {k[len(name) + 1:]: f[k] for k in f.files if k.startswith(name + "_")}
    X, Y = B.svd_backward(g["A"], g["U"], g["s"], g["V"], g["w"], g["GU"], g["GV"], "exact")
    np.testing.assert_allclose(X @ Y.T, g["dA"], atol=1e-10 * max(1.0, np.abs(g["dA"]).max()))
