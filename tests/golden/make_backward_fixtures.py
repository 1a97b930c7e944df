"""Writes tests/golden/svd_backward_torch.npz: the derivative of a top-k SVD loss
from torch.linalg.svd's autograd (LAPACK, float64, CPU) on two seeded cases.

    python tests/golden/make_backward_fixtures.py

Each case stores the matrix A, the loss weights w, G_U, G_V, the top-k triplets
(U, s, V, with the signs torch returned) and dL/dA for
L = sum_i w_i s_i + <G_U, U> + <G_V, V>.  The backward has no counterpart in the
reference, so these vectors come from torch, not from the reference.
"""

import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))


def case(m, n, k, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n))
    w = rng.standard_normal(k)
    GU = rng.standard_normal((m, k))
    GV = rng.standard_normal((n, k))
    At = torch.tensor(A, dtype=torch.float64, requires_grad=True)
    U, S, Vh = torch.linalg.svd(At, full_matrices=False)
    L = (torch.tensor(w) * S[:k]).sum() + (torch.tensor(GU) * U[:, :k]).sum() + (torch.tensor(GV) * Vh[:k].T).sum()
    L.backward()
    return dict(A=A, w=w, GU=GU, GV=GV, U=U[:, :k].detach().numpy(), s=S[:k].detach().numpy(),
                V=Vh[:k].T.detach().numpy(), dA=At.grad.numpy())


def main():
    out = {}
    for name, (m, n, k, seed) in {"tall": (64, 32, 5, 101), "wide": (24, 48, 6, 202)}.items():
        for key, val in case(m, n, k, seed).items():
            out[f"{name}_{key}"] = val
    np.savez_compressed(os.path.join(HERE, "svd_backward_torch.npz"), **out)


if __name__ == "__main__":
    main()
