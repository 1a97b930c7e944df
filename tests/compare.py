"""Sign-insensitive comparators with the semantics of the reference's
tests/src/compare.h:7-49 (relative tolerance on |x|, plus |column sum|)."""

import numpy as np


def same_same(left, right, tol):
    left = np.asarray(left, dtype=np.float64)
    right = np.asarray(right, dtype=np.float64)
    return np.abs(left - right) <= (np.abs(left) + np.abs(right)) * tol


def assert_equal_vectors(left, right, tol=1e-8):
    left = np.asarray(left)
    right = np.asarray(right)
    assert left.shape == right.shape
    ok = same_same(left, right, tol)
    assert ok.all(), f"max rel diff {np.max(np.abs(left - right) / (np.abs(left) + np.abs(right) + 1e-300))}"


def assert_equal_column_vectors(left, right, tol=1e-8, centered=False):
    left = np.asarray(left)
    right = np.asarray(right)
    assert left.shape == right.shape
    ok = same_same(np.abs(left), np.abs(right), tol)
    assert ok.all(), f"max rel |x| diff {np.max(np.abs(np.abs(left) - np.abs(right)) / (np.abs(left) + np.abs(right) + 1e-300))}"
    for i in range(left.shape[1]):
        ls = abs(left[:, i].sum())
        rs = abs(right[:, i].sum())
        if centered:
            assert ls < tol and rs < tol
        else:
            assert same_same(ls, rs, tol)


def assert_equal_matrix(left, right, tol=1e-8):
    left = np.asarray(left)
    right = np.asarray(right)
    assert left.shape == right.shape
    assert same_same(left, right, tol).all()


def subspace_angle(A, B):
    """Largest principal angle (radians) between the column spaces of A and B."""
    qa, _ = np.linalg.qr(A)
    qb, _ = np.linalg.qr(B)
    s = np.linalg.svd(qa.T @ qb, compute_uv=False)
    return float(np.arccos(np.clip(s.min(), -1.0, 1.0)))


def align_signs(X, ref):
    """Flip columns of X to match the signs of ref (vectors are defined up to sign)."""
    X = np.array(X, copy=True)
    for j in range(X.shape[1]):
        if np.dot(X[:, j], ref[:, j]) < 0:
            X[:, j] = -X[:, j]
    return X
