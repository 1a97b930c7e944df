"""The allocation-free SDDMM entry point (irlba_hip_sddmm_noalloc) through the C
ABI: bit-identical to irlba_hip_sddmm on the same device inputs, on two calls
in a row (the second smaller, so it runs over scratch the first call left
behind), for FP64 and FP32 outputs and both orientations; a call too large
for the fixed scratch fails with an error before any work."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

irlba = pytest.importorskip("cppirlba_amd")
torch = pytest.importorskip("torch")

from cppirlba_amd import _lib  # noqa: E402


def _pattern(n_lines, n_other, density, seed, dev):
    """A random compressed pattern (sorted indices per line) as device tensors."""
    rng = np.random.default_rng(seed)
    mask = rng.random((n_lines, n_other)) < density
    lines, other = np.nonzero(mask)
    ptr = np.zeros(n_lines + 1, dtype=np.int64)
    np.cumsum(np.bincount(lines, minlength=n_lines), out=ptr[1:])
    return (torch.tensor(other.astype(np.int32), device=dev), torch.tensor(ptr, device=dev))


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _both(ctx, nrow, ncol, idx, ptr, column_major, X, Y, out_type):
    """(regular, no-alloc) outputs for the same device inputs."""
    dt = torch.float32 if out_type == irlba.F32 else torch.float64
    nnz = idx.numel()
    ref = torch.full((nnz,), float("nan"), dtype=dt, device=X.device)
    got = torch.full((nnz,), float("nan"), dtype=dt, device=X.device)
    torch.cuda.synchronize()
    L = _lib.lib()
    rank = X.shape[1]
    _lib.check(L.irlba_hip_sddmm(ctx.handle, nrow, ncol, _vp(idx), _vp(ptr), int(column_major), rank,
                                 _vp(X), _vp(Y), _vp(ref), out_type, 1))
    _lib.check(L.irlba_hip_sddmm_noalloc(ctx.handle, nrow, ncol, _vp(idx), _vp(ptr), int(column_major), rank,
                                         _vp(X), _vp(Y), _vp(got), out_type))
    return ref.cpu().numpy(), got.cpu().numpy()


@pytest.mark.parametrize("out_type", [irlba.F64, irlba.F32])
@pytest.mark.parametrize("column_major", [True, False])
def test_noalloc_sddmm_matches_regular_twice(out_type, column_major):
    ctx = irlba.default_context()
    dev = torch.device("cuda", ctx.device)
    # (nrow, ncol, rank, density): a larger call, then a smaller one with a lower
    # rank, so the second packs fewer, shorter rows over the first call's scratch.
    for nrow, ncol, rank, density, seed in ((3000, 2000, 32, 0.02, 1), (700, 300, 5, 0.05, 2)):
        n_lines, n_other = (ncol, nrow) if column_major else (nrow, ncol)
        idx, ptr = _pattern(n_lines, n_other, density, seed, dev)
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        # Column-major factors: X nrow x rank, Y ncol x rank (as the C ABI reads them).
        X = torch.randn(rank, nrow, generator=gen, device=dev, dtype=torch.float64).t()
        Y = torch.randn(rank, ncol, generator=gen, device=dev, dtype=torch.float64).t()
        ref, got = _both(ctx, nrow, ncol, idx, ptr, column_major, X, Y, out_type)
        assert ref.size == idx.numel() > 0
        assert not np.isnan(ref).any()
        np.testing.assert_array_equal(got, ref)
        # and the value itself: X[row] . Y[col] for a few entries, in FP64
        ph, ih = ptr.cpu().numpy(), idx.cpu().numpy()
        Xh, Yh = X.cpu().numpy(), Y.cpu().numpy()
        for e in (0, ref.size // 2, ref.size - 1):
            line = int(np.searchsorted(ph, e, side="right") - 1)
            r, c = (ih[e], line) if column_major else (line, ih[e])
            tol = 1e-6 if out_type == irlba.F32 else 1e-13
            assert abs(ref[e] - Xh[r] @ Yh[c]) <= tol * (1 + np.abs(Xh[r] * Yh[c]).sum())


def test_noalloc_sddmm_rejects_oversize():
    ctx = irlba.default_context()
    dev = torch.device("cuda", ctx.device)
    # 300000 rows + columns at rank 64 in FP64 need 300000 x 64 x 8 = 153.6 MB of
    # packed factors, above the 32 MiB scratch: the host check fails the call
    # before any kernel runs, so these small placeholder arrays are never read.
    idx = torch.zeros(4, dtype=torch.int32, device=dev)
    ptr = torch.zeros(5, dtype=torch.int64, device=dev)
    X = torch.zeros(8, dtype=torch.float64, device=dev)
    out = torch.zeros(4, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    L = _lib.lib()
    rc = L.irlba_hip_sddmm_noalloc(ctx.handle, 200000, 100000, _vp(idx), _vp(ptr), 1, 64, _vp(X), _vp(X),
                                   _vp(out), irlba.F64)
    assert rc != 0
    msg = L.irlba_hip_last_error().decode()
    assert "fixed" in msg and "bytes of scratch" in msg, msg
    # The context is still usable afterwards.
    idx, ptr = _pattern(50, 40, 0.1, 3, dev)
    X = torch.randn(40, 3, device=dev, dtype=torch.float64)
    Y = torch.randn(50, 3, device=dev, dtype=torch.float64)
    ref, got = _both(ctx, 40, 50, idx, ptr, True, X, Y, irlba.F64)
    np.testing.assert_array_equal(got, ref)
