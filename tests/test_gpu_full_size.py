"""GPU parity at the full sizes of the benchmark workloads, where the oracle is
too slow to run: size-independent properties, LAPACK on the stored matrix where
it fits, and perturbation bounds.

  * Sparse PCA, 30000 x 1,000,000 at 3% (~9e8 stored entries, log1p NB values),
    top-50 centred + scaled, extra_work 20: FP32 values with FP32 products
    (mixed) and FP64 values with FP64 products, as the reference's pca() chain
    Scaled(Centered(ParallelSparseMatrix)) (pca.hpp:113-128) runs them.
  * Dense 200000 x 2000 = rank-30 plant (s_i = 100 * 0.9^i) + N(0,1), top-20,
    stored as BF16, FP32 and FP64 (simple.hpp:42-44,70-72).
  * Differentiable top-16 SVD of a dense 4096 x 1024 FP32 Gaussian + rank-16
    matrix and of a sparse 50000 x 20000 1% N(0,1) matrix.

Tolerances are stated where they are asserted.
"""

import numpy as np
import pytest

from compare import subspace_angle
from oracle import svd_backward as OB

pytestmark = pytest.mark.gpu

irlba = pytest.importorskip("cppirlba_amd")
torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)


# ---------------------------------------------------------------- sparse pca, 30000 x 1e6
def _triplet_residuals(Ac, scores, rotation, var, nr, k):
    """Singular triplets of the pca (U = scores / s, s = sqrt(var (nr-1)), V =
    rotation) and the column norms of their residuals R = Ac V - U S and
    S = Ac' U - V S, from FP64 block products of the FP64 operator."""
    sv = np.sqrt(var[:k] * (nr - 1))
    V = np.asfortranarray(rotation[:, :k])
    U = np.asfortranarray(scores[:, :k] / sv[None, :])
    R = Ac.multiply_block(V) - U * sv[None, :]
    S = Ac.multiply_block(U, adjoint=True) - V * sv[None, :]
    return sv, np.linalg.norm(R, axis=0), np.linalg.norm(S, axis=0)


def test_sparse_pca_30000x1e6_mixed_and_fp64():
    # The benchmark's workload: 30000 x 1e6 CSC at 3% from the device generator
    # (seed 42), start vector N(0,1) with numpy seed 1234, extra_work 20, tol 1e-5.
    nr, nc, k, tol = 30000, 1000000, 50, 1e-5
    init = np.random.default_rng(1234).standard_normal(nc)
    opts = irlba.Options(extra_work=20, initial=init, convergence_tolerance=tol)
    _say("sparse pca 30000x1e6: mixed top-50")
    M = irlba.generate_sparse(nr, nc, 0.03, 42, storage="f32", accumulate="f32")
    mixed = irlba.pca(M, True, True, k, opts)
    del M
    _say(f"  mixed: converged={mixed.converged}, {mixed.iterations} restarts")
    _say("sparse pca 30000x1e6: FP64 top-51 (the 51st value bounds the gap below the 50th)")
    A = irlba.generate_sparse(nr, nc, 0.03, 42, storage="f64", accumulate="f64")
    # On this flat spectrum the restart count to tol depends on the rounding of
    # the products (648 to 1033 restarts across builds for the top 51), so this
    # run may take more than the default 1000.
    opts51 = irlba.Options(extra_work=20, initial=init, convergence_tolerance=tol, max_iterations=2000)
    f64 = irlba.pca(A, True, True, k + 1, opts51)
    _say(f"  fp64: converged={f64.converged}, {f64.iterations} restarts")
    assert mixed.converged and f64.converged
    for res, n in ((mixed, k), (f64, k + 1)):
        var = res.variances
        # Every column of the standardised matrix has variance 1: the variances
        # are sorted, positive and sum to at most the column count.
        assert np.all(np.diff(var) <= 0) and var[-1] > 0 and var.sum() <= nc
        R = res.rotation
        assert np.abs(R.T @ R - np.eye(n)).max() < 1e-10
        assert np.all(np.abs(res.scores.sum(axis=0)) < 1e-6 * np.sqrt(var * (nr - 1)))

    # Residuals of both runs' triplets measured with FP64 products of the FP64
    # matrix, through decorator handles built from its column statistics (an
    # independent route from pca()'s own chain).
    c, s = irlba.column_stats(A)
    Ac = irlba.ScaledMatrix(irlba.CenteredMatrix(A, c), s, column=True, divide=True)
    sm, fm, am = _triplet_residuals(Ac, mixed.scores, mixed.rotation, mixed.variances, nr, k)
    sd, fd, ad = _triplet_residuals(Ac, f64.scores, f64.rotation, f64.variances, nr, k + 1)
    _say(f"  residual column norms / s_1: mixed fwd {fm.max() / sd[0]:.2e} adj {am.max() / sd[0]:.2e}; "
         f"fp64 fwd {fd.max() / sd[0]:.2e} adj {ad.max() / sd[0]:.2e}")
    # The convergence test (compute.hpp:246-264) counts Ritz values whose
    # residual estimate is within tol * max s, over all `work` of them
    # (compute.hpp:251), until `number` pass: the returned ones are those, in
    # practice the leading ones.  Here all 50 of the mixed run and the leading 50
    # of the FP64 top-51 run (its 51st, the last to settle, need not be) are
    # within that bound; the mixed run's residuals, measured on the FP64 matrix,
    # also carry the FP32 rounding of the values and products (~1e-7 relative).
    assert ad[:k].max() <= tol * sd[0]
    assert am.max() <= 2 * tol * sd[0]
    # Forward residuals: U = A V S^-1 holds to rounding in both (W B = A V).
    assert fd.max() / sd[0] < 1e-9
    assert fm.max() / sd[0] < 1e-5
    # Singular values: mixed within 1e-5 relative of FP64.
    rel = np.max(np.abs(sm - sd[:k]) / sd[:k])
    _say(f"  max singular value rel. error mixed vs fp64: {rel:.3e}")
    assert rel < 1e-5
    # Subspaces (Wedin's theorem): for each j the top-j right subspace of a run is
    # within sqrt(||R_j||_F^2 + ||S_j||_F^2) / delta_j of the exact one, with
    # delta_j = s~_j - s_{j+1}.  s_{j+1} is bounded above by the FP64 run's
    # (j+1)-th value plus its residual (the symmetric embedding [0 A; A' 0] has an
    # eigenvalue within the residual of s~_{j+1}, and the solve returns the top
    # 51 in order).  The two runs' top-j subspaces are then within the sum of
    # their bounds, checked wherever that sum is below 1 (a gap wider than the
    # residuals); the spectrum of this noise matrix is flat, so near the 50th
    # value the gaps are narrower than the 1e-5 * s_1 tolerance allows to resolve.
    s_up = sd[1:] + np.hypot(fd[1:], ad[1:])  # upper bounds on s_2 .. s_51
    rm = np.sqrt(np.cumsum(fm ** 2 + am ** 2))
    rd = np.sqrt(np.cumsum(fd[:k] ** 2 + ad[:k] ** 2))
    C = mixed.rotation.T @ f64.rotation[:, :k]
    checked = 0
    for j in range(1, k + 1):
        dm, dd = sm[j - 1] - s_up[j - 1], sd[j - 1] - s_up[j - 1]
        if dm <= 0 or dd <= 0:
            continue
        bound = rm[j - 1] / dm + rd[j - 1] / dd
        if bound >= 1:
            continue
        smin = np.linalg.svd(C[:j, :j], compute_uv=False).min()
        sin_j = np.sqrt(max(0.0, 1.0 - smin ** 2))
        assert sin_j <= bound + 1e-7, (j, sin_j, bound)  # 1e-7: resolution of sin from the cosines
        checked += 1
    angle = subspace_angle(mixed.rotation, f64.rotation[:, :k])
    _say(f"  Wedin bounds held for {checked} of {k} leading subspaces; 50-dim rotation subspace angle "
         f"mixed vs fp64 {angle:.3e} rad")
    # On this matrix the tolerance resolves 49 of the 50 leading subspaces (all
    # but the 50-dim one, whose last vector mixes with the 51st: its angle is
    # reported, not bounded; it depends on the run, e.g. 1.1e-3 rad against an
    # FP64 top-50 run and 1.5e-2 against this top-51 one).
    assert checked >= 45


def test_sparse_pca_30000x1e6_value_dictionary():
    # The same workload with the products streaming value-dictionary codes
    # (irlba_hip_sparse_value_dictionary): the log1p NB values take a few dozen
    # distinct values.  Products against the value stream's, then both solves
    # checked with the property set above on the FP64 value-stream operator.
    nr, nc, k, tol = 30000, 1000000, 50, 1e-5
    init = np.random.default_rng(1234).standard_normal(nc)
    opts = irlba.Options(extra_work=20, initial=init, convergence_tolerance=tol)
    x = torch.randn(nc, dtype=torch.float64, device="cuda").cpu().numpy()
    u = torch.randn(nr, dtype=torch.float64, device="cuda").cpu().numpy()
    results = {}
    for storage, close in (("f32", 2e-6), ("f64", 1e-14)):
        _say(f"sparse pca 30000x1e6 with the value dictionary: {storage} storage")
        M = irlba.generate_sparse(nr, nc, 0.03, 42, storage=storage, accumulate=storage)
        y0, z0 = M.multiply(x), M.adjoint_multiply(u)
        n = M.value_dictionary(True)
        y1, z1 = M.multiply(x), M.adjoint_multiply(u)
        ry = np.linalg.norm(y1 - y0) / np.linalg.norm(y0)
        rz = np.linalg.norm(z1 - z0) / np.linalg.norm(z0)
        _say(f"  {n} codes, runs {M.layout(False)['dictionary_run']}/{M.layout(True)['dictionary_run']}; "
             f"products vs the value stream: {ry:.2e} / {rz:.2e}")
        assert 2 <= n <= 256 and ry < close and rz < close
        results[storage] = irlba.pca(M, True, True, k, opts)
        _say(f"  converged={results[storage].converged}, {results[storage].iterations} restarts")
        if storage == "f64":
            M.value_dictionary(False)
            A = M
        del M
    mixed, f64 = results["f32"], results["f64"]
    assert mixed.converged and f64.converged
    c, s = irlba.column_stats(A)
    Ac = irlba.ScaledMatrix(irlba.CenteredMatrix(A, c), s, column=True, divide=True)
    sm, fm, am = _triplet_residuals(Ac, mixed.scores, mixed.rotation, mixed.variances, nr, k)
    sd, fd, ad = _triplet_residuals(Ac, f64.scores, f64.rotation, f64.variances, nr, k)
    _say(f"  residual column norms / s_1: mixed fwd {fm.max() / sd[0]:.2e} adj {am.max() / sd[0]:.2e}; "
         f"fp64 fwd {fd.max() / sd[0]:.2e} adj {ad.max() / sd[0]:.2e}")
    for res in (mixed, f64):
        R = res.rotation
        assert np.abs(R.T @ R - np.eye(k)).max() < 1e-10
        assert np.all(np.abs(res.scores.sum(axis=0)) < 1e-6 * np.sqrt(res.variances * (nr - 1)))
    # Tolerances as for the value stream's runs above.
    assert ad.max() <= tol * sd[0] and am.max() <= 2 * tol * sd[0]
    assert fd.max() / sd[0] < 1e-9 and fm.max() / sd[0] < 1e-5
    rel = np.max(np.abs(sm - sd) / sd)
    _say(f"  max singular value rel. error mixed vs fp64 (both coded): {rel:.3e}")
    assert rel < 1e-5


# ---------------------------------------------------------------- dense 200000 x 2000
def _gram_fp64(X):
    """X' X in FP64 on the device, by 20000-row panels."""
    n = X.shape[1]
    G = torch.zeros(n, n, dtype=torch.float64, device=X.device)
    for r0 in range(0, X.shape[0], 20000):
        P = X[r0:r0 + 20000].double()
        G += P.t().contiguous() @ P
    return G


def test_dense_200000x2000_bf16_fp32_fp64_vs_lapack():
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(42)
    m, n, r, k = 200000, 2000, 30, 20
    Q1, _ = torch.linalg.qr(torch.randn(m, r, generator=gen, device=dev, dtype=torch.float64))
    Q2, _ = torch.linalg.qr(torch.randn(n, r, generator=gen, device=dev, dtype=torch.float64))
    s = 100.0 * 0.9 ** torch.arange(r, device=dev, dtype=torch.float64)
    A64 = torch.randn(m, n, generator=gen, device=dev, dtype=torch.float64)
    A64.addmm_(Q1 * s, Q2.T)
    del Q1, Q2
    init = np.random.default_rng(4321).standard_normal(n)
    tol = 1e-5
    opts = irlba.Options(initial=init, convergence_tolerance=tol)  # extra_work: the default 7
    # Singular values within this relative distance of the stored matrix's (FP64
    # eigenvalues of X'X on the host, LAPACK): FP64 storage the 1e-8 parity bar;
    # FP32 accumulation (FP32 and BF16 storage) 1e-6.
    bars = {"f64": 1e-8, "f32": 1e-6, "bf16": 1e-6}
    kinds = {"bf16": torch.bfloat16, "f32": torch.float32, "f64": torch.float64}
    vals = {}
    for name, dtype in kinds.items():
        X = A64.to(dtype)
        h = irlba.HipDenseMatrix(X, storage=name)
        res = irlba.compute(h, k, opts)
        del h
        ev = np.linalg.eigvalsh(_gram_fp64(X).cpu().numpy())[::-1][:k]
        ref = np.sqrt(ev)
        rel = np.max(np.abs(res.D - ref) / ref)
        # Residuals of the returned triplets on the stored matrix, in FP64.
        Xd = X.double()
        Ut = torch.from_numpy(np.ascontiguousarray(res.U)).to(dev)
        Vt = torch.from_numpy(np.ascontiguousarray(res.V)).to(dev)
        Dt = torch.from_numpy(res.D).to(dev)
        fwd = torch.linalg.norm(Xd @ Vt - Ut * Dt, dim=0).max().item()
        adj = torch.linalg.norm(Xd.t() @ Ut - Vt * Dt, dim=0).max().item()
        del Xd, Ut, Vt, X
        torch.cuda.empty_cache()
        _say(f"dense 200000x2000 {name}: {res.iterations} restarts, converged={res.converged}, max rel sigma err vs LAPACK on the "
             f"stored matrix {rel:.3e}, residuals fwd {fwd / ref[0]:.2e} adj {adj / ref[0]:.2e} of s_1")
        assert res.converged
        assert rel < bars[name]
        assert np.abs(res.V.T @ res.V - np.eye(k)).max() < 1e-10
        assert np.abs(res.U.T @ res.U - np.eye(k)).max() < 1e-10
        # The convergence test bounds the adjoint residual by tol * s_1; FP32
        # accumulation adds its rounding (~1e-7) on top.
        assert adj <= (tol if name == "f64" else 2 * tol) * ref[0]
        assert fwd <= (1e-9 if name == "f64" else 1e-5) * ref[0]
        vals[name] = res
    # Mixed precision against the FP64 run, the benchmark's report.
    for name in ("f32", "bf16"):
        _say(f"dense 200000x2000 {name} vs f64: sigma {np.max(np.abs(vals[name].D - vals['f64'].D) / vals['f64'].D):.3e}, "
             f"subspace angle {subspace_angle(vals[name].V, vals['f64'].V):.3e} rad")


# ---------------------------------------------------------------- backward, dense 4096 x 1024
def _dense_4096x1024():
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(42)
    m, n, k = 4096, 1024, 16
    Q1, _ = torch.linalg.qr(torch.randn(m, k, generator=gen, device=dev, dtype=torch.float64))
    Q2, _ = torch.linalg.qr(torch.randn(n, k, generator=gen, device=dev, dtype=torch.float64))
    s = 400.0 * 0.95 ** torch.arange(k, device=dev, dtype=torch.float64)
    A = ((Q1 * s) @ Q2.T + torch.randn(m, n, generator=gen, device=dev, dtype=torch.float64)).float()
    rng = np.random.default_rng(7)
    return A, rng.standard_normal(k), rng.standard_normal((m, k)), rng.standard_normal((n, k))


def _loss(U, S, V, w, GU, GV):
    return (w * S).sum() + (GU * U).sum() + (GV * V).sum()


def test_backward_dense_4096x1024_vs_oracle_and_fd():
    A32, w, GU, GV = _dense_4096x1024()
    m, n, k = A32.shape[0], A32.shape[1], w.size
    dev = A32.device
    A = A32.double()  # the FP32 values, exactly
    An = A.cpu().numpy()
    opts = irlba.Options(convergence_tolerance=1e-12, extra_work=20)
    tw, tGU, tGV = (torch.tensor(x, device=dev) for x in (w, GU, GV))
    grads = {}
    for comp in ("exact", "rank_k"):
        At = A.clone().requires_grad_(True)
        U, S, V = irlba.svd_dense(At, k, options=opts, storage="f64", complement=comp)
        _loss(U, S, V, tw, tGU, tGV).backward()
        Un, Sn, Vn = (t.detach().cpu().numpy() for t in (U, S, V))
        Xo, Yo = OB.svd_backward(An, Un, Sn, Vn, w, GU, GV, comp)
        ref = Xo @ Yo.T
        err = np.abs(At.grad.cpu().numpy() - ref).max() / np.abs(ref).max()
        _say(f"backward dense 4096x1024 {comp}: max rel gradient difference vs the oracle {err:.3e}")
        # FP64 storage: the oracle's full-SVD derivative to 1e-8 of the largest entry.
        assert err < 1e-8
        grads[comp] = (At.grad.cpu().numpy(), Un, Vn)
    # One directional central difference of the loss through LAPACK SVDs (signs
    # aligned with the engine's vectors): <dL/dA, D> for a seeded N(0,1) direction.
    G, Un, Vn = grads["exact"]
    D = np.random.default_rng(11).standard_normal((m, n))
    h = 2e-6
    Lp = OB.loss(An + h * D, k, w, GU, GV, Un, Vn)
    Lm = OB.loss(An - h * D, k, w, GU, GV, Un, Vn)
    fd = (Lp - Lm) / (2 * h)
    an = float(np.sum(G * D))
    _say(f"backward dense 4096x1024 directional derivative: analytic {an:.10e}, central difference {fd:.10e}")
    # h = 2e-6 on gaps >= ~9 between the planted values (>= ~89 to the noise):
    # the O(h^2) term and the LAPACK rounding over 2h stay below 1e-5 relative.
    assert abs(fd - an) <= 1e-5 * max(1.0, abs(an))
    # FP32 storage (FP32 products in the forward and the CG): within 1e-4.
    At = A32.clone().requires_grad_(True)
    U, S, V = irlba.svd_dense(At, k, options=opts, storage="f32", complement="exact")
    sg = np.sign(np.sum(U.detach().cpu().numpy() * Un, axis=0))  # align with the FP64 run's signs
    _loss(U, S, V, tw, tGU * torch.tensor(sg, device=dev), tGV * torch.tensor(sg, device=dev)).backward()
    g32 = At.grad.double().cpu().numpy()
    err = np.abs(g32 - G).max() / np.abs(G).max()
    _say(f"backward dense 4096x1024 FP32 storage vs FP64: {err:.3e}")
    assert err < 1e-4


# ---------------------------------------------------------------- backward, sparse 50000 x 20000
def _sparse_50000x20000(dev):
    gen = torch.Generator(device=dev)
    gen.manual_seed(42)
    m, n, density = 50000, 20000, 0.01
    rows, counts = [], []
    for c0 in range(0, n, 1000):
        c1 = min(n, c0 + 1000)
        mask = torch.rand((c1 - c0, m), generator=gen, device=dev) < density
        nzc, nzr = mask.nonzero(as_tuple=True)
        rows.append(nzr.to(torch.int32))
        counts.append(torch.bincount(nzc, minlength=c1 - c0))
    idx = torch.cat(rows)
    ptr = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    ptr[1:] = torch.cumsum(torch.cat(counts), 0)
    vals = torch.randn(idx.numel(), generator=gen, device=dev, dtype=torch.float32)
    return m, n, vals, idx, ptr


def _pattern_product(U, w, V, idx, ptr, chunk=2_000_000):
    """(U diag(w) V')[r_e, c_e] for every stored entry e of a CSC pattern (host, FP64)."""
    cols = np.repeat(np.arange(ptr.size - 1), np.diff(ptr))
    out = np.empty(idx.size)
    Uw = U * w[None, :]
    for e0 in range(0, idx.size, chunk):
        e1 = min(idx.size, e0 + chunk)
        out[e0:e1] = np.einsum("ij,ij->i", Uw[idx[e0:e1]], V[cols[e0:e1]])
    return out


def test_backward_sparse_50000x20000_values_loss():
    dev = torch.device("cuda:0")
    m, n, vals32, idx, ptr = _sparse_50000x20000(dev)
    k = 16
    w = np.random.default_rng(8).standard_normal(k)
    tw = torch.tensor(w, device=dev)
    idx_h, ptr_h = idx.cpu().numpy(), ptr.cpu().numpy()
    _say(f"backward sparse 50000x20000: {m}x{n}, {idx.numel()} stored entries")
    opts = irlba.Options(convergence_tolerance=1e-10, extra_work=30)
    # L = sum w_i s_i: dL/dA = U diag(w) V' (no complement terms), on the pattern.
    # FP64 storage to rounding (1e-12 of the largest entry); FP32 storage (FP32
    # values, the configuration's) to the FP32 rounding of the factors and output.
    for storage, vals, bar in (("f64", vals32.double(), 1e-12), ("f32", vals32, 1e-5)):
        x = vals.clone().requires_grad_(True)
        U, S, V = irlba.svd_sparse(x, idx, ptr, m, n, k, options=opts, storage=storage, complement="exact")
        (tw * S).sum().backward()
        ref = _pattern_product(U.detach().cpu().numpy(), w, V.detach().cpu().numpy(), idx_h, ptr_h)
        g = x.grad.double().cpu().numpy()
        err = np.abs(g - ref).max() / np.abs(ref).max()
        _say(f"  {storage}: gradient vs U diag(w) V' on the pattern: {err:.3e}")
        assert err < bar
        if storage == "f64":
            g64, S64 = g, S.detach().cpu().numpy()
    # One directional central difference of sum w_i s_i along a seeded N(0,1)
    # direction on the pattern, through two more FP64 solves.  The top values of
    # a noise matrix are spaced ~0.05 apart; h = 1e-7 keeps the O(h^2) term and
    # the solves' rounding over 2h (their values converge quadratically in the
    # 1e-10 residual) below 1e-3 of the derivative.
    d = np.random.default_rng(12).standard_normal(idx_h.size)
    h = 1e-7
    base = vals32.double().cpu().numpy()
    Ls = []
    for sgn in (1, -1):
        H = irlba.HipSparseMatrix(m, n, base + sgn * h * d, idx_h, ptr_h, storage="f64")
        Ls.append(float(w @ irlba.compute(H, k, opts).D))
        del H
    fd = (Ls[0] - Ls[1]) / (2 * h)
    an = float(g64 @ d)
    _say(f"  directional derivative: analytic {an:.8e}, central difference {fd:.8e}")
    assert abs(fd - an) <= 1e-3 * abs(an)
    assert np.all(np.diff(S64) <= 0)
