"""GPU unit parity of the Lanczos building blocks against the oracle:
orthogonalize_vector (CGS), the restart rotation on MFMA, and one full
run_lanczos_bidiagonalization sweep (tests/src/lanczos.cpp on the device)."""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

irlba = pytest.importorskip("cppirlba_amd")


# (300007, 45, 45): long vectors take the 4-rows-per-thread dot and update kernels.
@pytest.mark.parametrize("rows,cols,n", [(5, 4, 4), (1000, 70, 33), (100003, 40, 40), (300, 9, 0), (300007, 45, 45),
                                         (262144, 13, 11)])
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


@pytest.mark.parametrize("rows,work,k", [(16, 4, 1), (64, 8, 5), (1000, 70, 67), (777, 27, 20), (50001, 33, 30),
                                         (300, 200, 150)])
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
