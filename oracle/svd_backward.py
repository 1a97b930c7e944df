"""Dense numpy restatement of the top-k SVD backward.

ORACLE — test infrastructure only (imported by tests/; never by the
``cppirlba_amd`` product path).

No reference counterpart: the reference (include/irlba/compute.hpp:141-308)
computes the forward SVD only.  This restatement follows the published
derivative of the thin SVD (J. Townsend, "Differentiating the Singular Value
Decomposition", 2016, eq. for A-bar) generalised to a top-k truncation.  It
takes the part of dU, dV outside span(U), span(V) from the FULL SVD of A
(a sum over the other singular triplets plus the null-space projector), which
is a different route from the engine's projected CG solves; both must agree.

For L = sum_i w_i s_i + <G_U, U> + <G_V, V>:
    dL/dA = X Y',  X = [U M + Z' | U],  Y = [V | Y'],
    M = diag(w) + (F o (J - J')) S + S (F o (K - K')),  J = U'G_U, K = V'G_V,
    F_ij = 1 / (s_j^2 - s_i^2)  (0 on the diagonal),
    complement "exact":   z_i = sum_{j>k} u~_j (u~_j' g_i) / (s_i^2 - s~_j^2) + N_U g_i / s_i^2,
                          t_i = sum_{j>k} v~_j (v~_j' h_i) / (s_i^2 - s~_j^2) + N_V h_i / s_i^2,
                          Z'_i = s_i z_i + A t_i,  Y'_i = A' z_i + s_i t_i,
    complement "rank_k":  Z'_i = (I - UU') g_i / s_i,  Y'_i = (I - VV') h_i / s_i,
    complement "none":    Z' = Y' = 0,
with N_U = I - U~U~' and N_V = I - V~V~' the projectors on the null spaces left
out by the thin full SVD (U~, s~, V~).

Checked in tests/test_oracle_backward.py against central finite differences
of a sign-fixed SVD and against torch.linalg.svd's autograd on the CPU.
"""

import numpy as np


def in_subspace_matrix(s, w, J, K):
    """M = diag(w) + (F o (J - J')) S + S (F o (K - K'))."""
    k = s.size
    s2 = s * s
    den = s2[None, :] - s2[:, None]  # den[i, j] = s_j^2 - s_i^2
    F = np.zeros((k, k))
    off = ~np.eye(k, dtype=bool)
    nz = off & (den != 0)
    F[nz] = 1.0 / den[nz]
    return np.diag(w) + (F * (J - J.T)) * s[None, :] + s[:, None] * (F * (K - K.T))


def svd_backward(A, U, s, V, w=None, GU=None, GV=None, complement="exact"):
    """Return (X, Y) with dL/dA = X @ Y.T (see the module docstring)."""
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    k = s.size
    w = np.zeros(k) if w is None else np.asarray(w, dtype=np.float64)
    GU = np.zeros((m, k)) if GU is None else np.asarray(GU, dtype=np.float64)
    GV = np.zeros((n, k)) if GV is None else np.asarray(GV, dtype=np.float64)
    J = U.T @ GU
    K = V.T @ GV
    M = in_subspace_matrix(s, w, J, K)
    Zp = np.zeros((m, k))
    Yp = np.zeros((n, k))
    if complement == "rank_k":
        Zp = (GU - U @ J) / s[None, :]
        Yp = (GV - V @ K) / s[None, :]
    elif complement == "exact":
        Uf, sf, Vtf = np.linalg.svd(A, full_matrices=False)
        Vf = Vtf.T
        # Align the leading k full-SVD vectors with the given U (sign only); the
        # remaining ones span the complement whatever their signs.
        Ut, Vt = Uf[:, k:], Vf[:, k:]
        st = sf[k:]
        for i in range(k):
            si2 = s[i] * s[i]
            g, h = GU[:, i], GV[:, i]
            coef = 1.0 / (si2 - st * st)
            z = Ut @ (coef * (Ut.T @ g)) + (g - Uf @ (Uf.T @ g)) / si2
            t = Vt @ (coef * (Vt.T @ h)) + (h - Vf @ (Vf.T @ h)) / si2
            Zp[:, i] = s[i] * z + A @ t
            Yp[:, i] = A.T @ z + s[i] * t
    elif complement != "none":
        raise ValueError(complement)
    X = np.concatenate([U @ M + Zp, U], axis=1)
    Y = np.concatenate([V, Yp], axis=1)
    return X, Y


def grad_center(X, Y):
    """Gradient with respect to c of A = B - 1 c': -(X Y')' 1 = -Y (X' 1)."""
    return -Y @ X.sum(axis=0)


def loss(A, k, w, GU, GV, U0, V0):
    """L(A) with the top-k triplets' signs fixed by alignment with U0 (smooth near A0)."""
    Uf, sf, Vtf = np.linalg.svd(A, full_matrices=False)
    U = Uf[:, :k].copy()
    V = Vtf[:k].T.copy()
    sg = np.sign(np.sum(U * U0, axis=0))
    U *= sg
    V *= sg
    return float(w @ sf[:k] + np.sum(GU * U) + np.sum(GV * V))


def fd_gradient(A, k, w, GU, GV, eps=1e-6, entries=None):
    """Central finite differences of loss() over all entries (or the listed (r, c))."""
    Uf, sf, Vtf = np.linalg.svd(A, full_matrices=False)
    U0, V0 = Uf[:, :k], Vtf[:k].T
    out = np.zeros_like(A)
    coords = entries if entries is not None else [(r, c) for c in range(A.shape[1]) for r in range(A.shape[0])]
    for r, c in coords:
        Ap = A.copy()
        Ap[r, c] += eps
        Am = A.copy()
        Am[r, c] -= eps
        out[r, c] = (loss(Ap, k, w, GU, GV, U0, V0) - loss(Am, k, w, GU, GV, U0, V0)) / (2 * eps)
    return out
