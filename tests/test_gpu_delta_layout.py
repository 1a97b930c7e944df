"""GPU parity of the 8-bit delta index layout of the sparse products.

The delta layout stores one byte per slot (the distance to the previous
in-tile index) and bridges gaps above 255 with value-0 slots; the kernels
recover the indices with lane scans.  The matrices here hold a dense block
(short gaps) and a sparse frame (gaps of hundreds to thousands, so every
segment kind -- bridged, plain, empty -- occurs), in one and in several
tiles, at every lanes-per-line width.

Tolerances: FP64 storage and accumulation 1e-12 of the output's largest
magnitude (the summation order differs from numpy's); FP32 storage and
accumulation 2e-5; realize is exact.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

irlba = pytest.importorskip("cppirlba_amd")


def _framed(nr, nc, seed):
    rng = np.random.default_rng(seed)
    A = np.zeros((nr, nc))
    blk = rng.random((nr // 2, nc // 2)) < 0.3
    A[: nr // 2, : nc // 2][blk] = rng.standard_normal(blk.sum())
    frame = rng.random((nr, nc)) < 0.002
    frame[: nr // 2, : nc // 2] = False
    A[frame] = rng.standard_normal(frame.sum())
    return A


def _csc(A):
    nr, nc = A.shape
    idx, vals, ptr = [], [], [0]
    for c in range(nc):
        r = np.nonzero(A[:, c])[0]
        idx.append(r)
        vals.append(A[r, c])
        ptr.append(ptr[-1] + r.size)
    return np.concatenate(vals), np.concatenate(idx).astype(np.int32), np.array(ptr, dtype=np.int64)


def _rel(y, ref):
    return np.abs(y - ref).max() / max(1e-300, np.abs(ref).max())


@pytest.mark.parametrize("tile_bytes", [None, "4096"])
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_forced_delta_matches_numpy(monkeypatch, tile_bytes, storage):
    A = _framed(3000, 2000, 1)
    vals, idx, ptr = _csc(A)
    monkeypatch.setenv("IRLBA_HIP_DELTA", "2")
    if tile_bytes:
        monkeypatch.setenv("IRLBA_HIP_TILE_BYTES", tile_bytes)
    h = irlba.HipSparseMatrix(3000, 2000, vals, idx, ptr, storage=storage)
    tol = 1e-12 if storage == "f64" else 2e-5
    for adjoint in (False, True):
        lay = h.layout(adjoint)
        assert lay["delta"] == 1 and lay["vec"] == 4
        assert lay["slots"] > lay["nnz"]  # bridging slots and padding
        if tile_bytes:
            assert lay["n_tiles"] > 1
    rng = np.random.default_rng(2)
    x, u = rng.standard_normal(2000), rng.standard_normal(3000)
    assert _rel(h.multiply(x), A @ x) < tol
    assert _rel(h.adjoint_multiply(u), A.T @ u) < tol
    X = rng.standard_normal((2000, 5))
    assert _rel(h.multiply_block(X), A @ X) < tol
    np.testing.assert_array_equal(h.realize(), A if storage == "f64" else A.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("storage,acc", [("f64", "f64"), ("f32", "f32"), ("f32", "f64")])
@pytest.mark.parametrize("lpl", ["4", "8", "16", "32", "64"])
def test_delta_every_group_width(monkeypatch, lpl, storage, acc):
    # The lane scan and the broadcast of each step's group total are written
    # per group width (DPP row shifts, quad permutes, half-row mirrors, row
    # broadcasts, lane reads); FP64 values take the lane-pair loads.
    A = _framed(1800, 1400, int(lpl))
    vals, idx, ptr = _csc(A)
    monkeypatch.setenv("IRLBA_HIP_DELTA", "2")
    monkeypatch.setenv("IRLBA_HIP_LPL", lpl)
    h = irlba.HipSparseMatrix(1800, 1400, vals, idx, ptr, storage=storage, accumulate=acc)
    assert h.layout(False)["lanes_per_line"] == int(lpl) and h.layout(False)["delta"] == 1
    if storage == "f32":
        A = A.astype(np.float32).astype(np.float64)  # the stored matrix
    tol = 1e-12 if acc == "f64" else 2e-6
    rng = np.random.default_rng(3)
    x, u = rng.standard_normal(1400), rng.standard_normal(1800)
    assert _rel(h.multiply(x), A @ x) < tol
    assert _rel(h.adjoint_multiply(u), A.T @ u) < tol


def test_layouts_agree_and_heuristic(monkeypatch):
    # Short gaps (3% density): the delta layout is taken by itself and agrees
    # with the 16-bit layout; a centered, scaled pca runs on it unchanged.
    from oracle import oracle as O

    nr, nc = 2500, 3000
    vals, idx, ptr = O.gen_nb_sparse(nr, nc, 0.03, 7, f32=False)
    h1 = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, storage="f32")
    assert h1.layout(False)["delta"] == 1 and h1.layout(True)["delta"] == 1
    assert h1.layout(False)["index_bytes"] == h1.layout(False)["slots"]
    monkeypatch.setenv("IRLBA_HIP_DELTA", "0")
    h0 = irlba.HipSparseMatrix(nr, nc, vals, idx, ptr, storage="f32")
    assert h0.layout(False)["delta"] == 0 and h0.layout(False)["index_bytes"] == 2 * h0.layout(False)["slots"]
    rng = np.random.default_rng(4)
    x, u = rng.standard_normal(nc), rng.standard_normal(nr)
    assert _rel(h1.multiply(x), h0.multiply(x)) < 2e-6
    assert _rel(h1.adjoint_multiply(u), h0.adjoint_multiply(u)) < 2e-6
    init = O.create_random_vector(nc, 9)
    opts = dict(extra_work=20, convergence_tolerance=1e-8)
    r1 = irlba.pca(h1, True, True, 8, irlba.Options(initial=init, **opts))
    ref = O.pca(O.sparse(nr, nc, vals, idx, ptr, True, 1), True, True, 8, initial=init, **opts)
    assert r1.converged
    assert np.max(np.abs(r1.variances - ref[2]) / ref[2]) < 1e-5
