"""Benchmark: top-50 centred+scaled PCA of a 30000 x 1,000,000, 3%-dense sparse
matrix (~9e8 nonzeros, log1p NB(mu=2, phi=0.5) values, FP32 values with int32
indices) on one MI355X, mixed precision (FP32 products, FP64 Lanczos), against
the FP64 GPU path on the same matrix.

One step = one full irlba pca() solve (compute.hpp:141-308 through pca.hpp:55-139)
with the matrix already resident in HBM.  Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--gpus 1]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "top-50 PCA solve s + HBM GB/s, mixed vs FP64 on GPU; SVD backward ms"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def solve(irlba, mat, number, opts, device=True):
    """One pca() top-`number`; with `device` the scores, rotation and variances
    stay in HBM (the timed steps), otherwise they cross PCIe into host arrays."""
    t0 = time.perf_counter()
    res = irlba.pca(mat, True, True, number, opts, device=device)
    return res, time.perf_counter() - t0


def host_copy(res):
    """PcaResults with numpy arrays (for the error reports, outside timing)."""
    import numpy as np

    conv = (lambda a: a.cpu().numpy()) if not isinstance(res.rotation, np.ndarray) else (lambda a: a)
    return type(res)(conv(res.scores), conv(res.rotation), conv(res.variances), res.iterations, res.converged)


def cpu_baseline(nr, ncol_full, density, seed, nnz_full, n_fwd, n_adj, restarts, number, extra_work, sample_cols):
    """The oracle's restatement of the reference (ParallelSparseMatrix with its
    nnz-balanced std::thread partition, float values and double vectors,
    sparse.hpp:185-267; compute() and pca() with plain-loop Gram-Schmidt and
    rotations, compute.hpp:141-308, pca.hpp:55-139) timed on a column slice of
    the same generated matrix:
      * per-product medians (>= 5 after warm-up) of A.v and A'.u;
      * the centred+scaled pca() of the slice capped at 1 and at 4 restarts, whose
        difference gives the time of one restart and whose first gives the
        initial sweep.
    `value` is the conservative figure: the products alone, scaled by nnz to the
    full matrix and to the GPU solve's product count (a lower bound on a CPU
    solve).  The capped solves give a second, extrapolated figure: the slice's
    restart cost scaled by the column ratio (products, Gram-Schmidt against V
    and the V rotation scale with the columns; the W-side terms do not, so it
    overstates) times the GPU run's restart count."""
    import numpy as np

    from oracle import oracle as O

    threads = max(1, len(os.sched_getaffinity(0)))  # every core this process may run on
    vals, idx, ptr = O.gen_nb_sparse(nr, sample_cols, density, seed, line0=0, line1=sample_cols, f32=True)
    A = O.sparse(nr, sample_cols, vals, idx, ptr, True, threads)
    x = np.random.default_rng(0).standard_normal(sample_cols)
    u = np.random.default_rng(1).standard_normal(nr)
    A.multiply(x)
    A.multiply(u, adjoint=True)
    tf, ta = [], []
    for _ in range(5):
        t0 = time.perf_counter()
        A.multiply(x)
        tf.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        A.multiply(u, adjoint=True)
        ta.append(time.perf_counter() - t0)
    scale = nnz_full / max(int(ptr[-1]), 1)
    fwd, adj = float(np.median(tf)) * scale, float(np.median(ta)) * scale
    init = np.random.default_rng(1234).standard_normal(sample_cols)
    t_pca = {}
    for cap in (1, 4):
        t0 = time.perf_counter()
        O.pca(A, True, True, number, extra_work=extra_work, max_iterations=cap, initial=init)
        t_pca[cap] = time.perf_counter() - t0
    per_restart = (t_pca[4] - t_pca[1]) / 3
    col_scale = ncol_full / sample_cols
    solve = col_scale * (t_pca[1] + max(0, restarts - 1) * per_restart)
    return {
        "value": round(fwd * n_fwd + adj * n_adj, 1),
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": (f"reference algorithm restated (no Eigen), {threads} threads (the cores this process may "
                   f"use) for the products: a "
                   f"{nr}x{sample_cols} column slice of the same generated matrix ({int(ptr[-1])} nnz); "
                   f"value = A.v {fwd * 1e3:.1f} ms + A'.u {adj * 1e3:.1f} ms per full-size product (median of 5, "
                   f"scaled by nnz) x the GPU solve's {n_fwd} + {n_adj} products (products only, a lower bound). "
                   f"pca top-{number} of the slice capped at 1 and 4 restarts took {t_pca[1]:.2f} s and "
                   f"{t_pca[4]:.2f} s ({per_restart * 1e3:.1f} ms per restart); scaled by the column ratio "
                   f"x{col_scale:.1f} and the GPU's {restarts} restarts that extrapolates to {solve:.0f} s"),
        "extrapolated_solve_s": round(solve, 1),
        "slice_restart_ms": round(per_restart * 1e3, 1),
    }


def product_pair_bytes(h, value_bytes, nr, nc):
    """Bytes one A.v and one A'.u of a sparse handle stream: every stored slot
    (value + delta or 16-bit index), the segment pointers and first indices,
    and the FP64 input and output vectors."""
    total = 0
    for adjoint in (False, True):
        lay = h.layout(adjoint)
        nseg = lay["n_tiles"] * (nc if adjoint else nr)
        total += (lay["slots"] * value_bytes + lay["index_bytes"] + (nseg + 1) * 8 + (nseg * 2 if lay["delta"] else 0)
                  + (nr + nc) * 8)
    return total


def spmv_mode(args):
    """Products only (for rocprofv3 counter runs): `reps` A.v and A'.u products of
    the mixed-precision matrix, then a device copy of a known size whose
    FETCH_SIZE/WRITE_SIZE calibrate the counters."""
    import ctypes

    import torch

    import cppirlba_amd as irlba
    from cppirlba_amd import _lib

    ctx = irlba.default_context()
    storage = args.storage
    mat = irlba.generate_sparse(args.rows, args.cols, args.density, args.seed, storage=storage,
                                accumulate="f64" if storage == "f64" else "f32")
    if args.dictionary:
        mat.value_dictionary(True)
    ctx.synchronize()
    L = _lib.lib()
    x = torch.randn(args.cols, dtype=torch.float64, device="cuda")
    y = torch.randn(args.rows, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    for _ in range(args.reps):
        _lib.check(L.irlba_hip_multiply(mat.handle, 0, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), 1))
        _lib.check(L.irlba_hip_multiply(mat.handle, 1, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()), 1))
    st = ctx.stats()
    n = 1 << 27  # 1 GiB of FP64
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"mode": "spmv", "storage": storage, "dictionary": bool(args.dictionary), "reps": args.reps,
                      "forward_ms": st["forward_kernel_ms"] / max(1, st["forward_timed"]),
                      "adjoint_ms": st["adjoint_kernel_ms"] / max(1, st["adjoint_timed"]),
                      "forward_bytes": st["forward_bytes"], "adjoint_bytes": st["adjoint_bytes"],
                      "copy_bytes": n * 8, "forward_layout": mat.layout(False),
                      "adjoint_layout": mat.layout(True)}), flush=True)


def _sparse_uniform_csc(torch, m, n, density, gen, dev, chunk=1000):
    """CSC of an m x n matrix with Bernoulli(density) positions and N(0,1) FP32 values (on the device)."""
    rows, counts = [], []
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        mask = torch.rand((c1 - c0, m), generator=gen, device=dev) < density  # one row per column
        nzc, nzr = mask.nonzero(as_tuple=True)  # column-major order: by column, then row
        rows.append(nzr.to(torch.int32))
        counts.append(torch.bincount(nzc, minlength=c1 - c0))
    idx = torch.cat(rows)
    ptr = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    ptr[1:] = torch.cumsum(torch.cat(counts), 0)
    vals = torch.randn(idx.numel(), generator=gen, device=dev, dtype=torch.float32)
    return vals, idx, ptr


def _time_backward(torch, fwd, loss_of, reps):
    """Forward once, then `reps` timed backward passes through the autograd graph."""
    out = fwd()
    L = loss_of(*out)
    times = []
    for r in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.backward(retain_graph=True)
        torch.cuda.synchronize()
        if r:  # the first pass pays one-time allocations
            times.append(1e3 * (time.perf_counter() - t0))
    return out, times


def backward_leg(args):
    """Differentiable top-16 SVD: dense 4096 x 1024 FP32 (Gaussian + planted rank 16,
    s_i = 400 * 0.95^i) and sparse 50000 x 20000 at 1% (uniform pattern, N(0,1)
    FP32 values); loss sum w_i s_i + <G_U, U> + <G_V, V> with seeded Gaussian w, G;
    backward through the torch::autograd::Function binding."""
    import numpy as np
    import torch

    import cppirlba_amd as irlba

    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed)
    k = 16
    res = {"number": k, "loss": "sum w_i s_i + <G_U,U> + <G_V,V>, seeded Gaussian w, G"}
    opts = irlba.Options(extra_work=args.extra_work, convergence_tolerance=1e-8)

    def loss_fn(w, GU, GV):
        return lambda U, S, V: (w * S).sum() + (GU * U).sum() + (GV * V).sum()

    # Dense.
    m, n = 4096, 1024
    Q1, _ = torch.linalg.qr(torch.randn(m, k, generator=gen, device=dev, dtype=torch.float64))
    Q2, _ = torch.linalg.qr(torch.randn(n, k, generator=gen, device=dev, dtype=torch.float64))
    s = 400.0 * 0.95 ** torch.arange(k, device=dev, dtype=torch.float64)
    A = ((Q1 * s) @ Q2.T + torch.randn(m, n, generator=gen, device=dev, dtype=torch.float64)).float()
    A.requires_grad_(True)
    w = torch.randn(k, generator=gen, device=dev, dtype=torch.float64)
    GU = torch.randn(m, k, generator=gen, device=dev, dtype=torch.float64)
    GV = torch.randn(n, k, generator=gen, device=dev, dtype=torch.float64)
    for comp in ("exact", "rank_k"):
        t0 = time.perf_counter()
        _, times = _time_backward(torch, lambda: irlba.svd_dense(A, k, options=opts, storage="f32",
                                                                 complement=comp),
                                  loss_fn(w, GU, GV), args.reps)
        res[f"dense_{m}x{n}_{comp}_ms"] = round(float(np.median(times)), 3)
        log(f"  backward dense {m}x{n} ({comp}): {np.median(times):.3f} ms "
            f"(forward + {args.reps + 1} backward {time.perf_counter() - t0:.2f} s)")
    del A, GU, GV

    # Sparse.
    m, n, density = 50000, 20000, 0.01
    vals, idx, ptr = _sparse_uniform_csc(torch, m, n, density, gen, dev)
    vals.requires_grad_(True)
    res["sparse_nnz"] = int(idx.numel())
    GU = torch.randn(m, k, generator=gen, device=dev, dtype=torch.float64)
    GV = torch.randn(n, k, generator=gen, device=dev, dtype=torch.float64)
    comps = ("rank_k",) if args.skip_backward_exact_sparse else ("rank_k", "exact")
    for comp in comps:
        t0 = time.perf_counter()
        out, times = _time_backward(torch, lambda: irlba.svd_sparse(vals, idx, ptr, m, n, k, options=opts,
                                                                    storage="f32", complement=comp),
                                    loss_fn(w, GU, GV), args.reps)
        res[f"sparse_{m}x{n}_{comp}_ms"] = round(float(np.median(times)), 3)
        log(f"  backward sparse {m}x{n} nnz={idx.numel()} ({comp}): {np.median(times):.3f} ms "
            f"(forward + {args.reps + 1} backward {time.perf_counter() - t0:.2f} s)")
        if comp == "exact":
            # The CG iteration count of the exact complement (the same backward
            # through the C ABI on a handle of the same values), and the gap that
            # sets it: the projected systems are shifted by s_16^2 against the
            # spectrum below, so they converge at a rate set by s_16 - s_17.
            U, S, V = (t.detach() for t in out)
            h = irlba.HipSparseMatrix(m, n, vals.detach(), idx, ptr, storage="f32")
            bw = irlba.svd_backward(h, U.cpu().numpy(), S.cpu().numpy(), V.cpu().numpy(), w.cpu().numpy(),
                                    GU.cpu().numpy(), GV.cpu().numpy(), complement="exact")
            s17 = irlba.compute(h, k + 1, opts).D
            res[f"sparse_{m}x{n}_exact_cg_iterations"] = int(bw.cg_iterations)
            res[f"sparse_{m}x{n}_s16_s17"] = [round(float(s17[k - 1]), 6), round(float(s17[k]), 6)]
            log(f"    exact complement: {bw.cg_iterations} lock-step CG iterations; s_16 {s17[k - 1]:.5f}, "
                f"s_17 {s17[k]:.5f}")
            del h
    return res


def dense_leg(args):
    """Dense mixed precision: 200000 x 2000 = rank-30 planted signal (s_i = 100 * 0.9^i,
    random orthonormal factors) + N(0,1) noise, stored as BF16 and as FP32 (FP32
    accumulation) and as FP64; top-20 triplets (compute, extra_work 7)."""
    import numpy as np
    import torch

    import cppirlba_amd as irlba

    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed)
    m, n, r, k = args.dense_rows, args.dense_cols, 30, 20
    t0 = time.perf_counter()
    Q1, _ = torch.linalg.qr(torch.randn(m, r, generator=gen, device=dev, dtype=torch.float64))
    Q2, _ = torch.linalg.qr(torch.randn(n, r, generator=gen, device=dev, dtype=torch.float64))
    s = 100.0 * 0.9 ** torch.arange(r, device=dev, dtype=torch.float64)
    A64 = torch.randn(m, n, generator=gen, device=dev, dtype=torch.float64)
    A64.addmm_(Q1 * s, Q2.T)
    del Q1, Q2
    torch.cuda.synchronize()
    log(f"  dense {m}x{n} generated in {time.perf_counter() - t0:.2f} s")
    init = np.random.default_rng(4321).standard_normal(n)
    opts = irlba.Options(initial=init)
    res = {"rows": m, "cols": n, "number": k, "planted_rank": r, "extra_work": opts.extra_work}
    ctx = irlba.default_context()
    runs = {}
    kinds = {"bf16": torch.bfloat16, "f32": torch.float32, "f64": torch.float64}
    for name in args.dense_storages.split(","):
        dtype = kinds[name]
        At = A64.to(dtype)
        h = irlba.HipDenseMatrix(At, storage=name)
        del At
        torch.cuda.empty_cache()
        times = []
        for _ in range(max(1, args.dense_reps)):
            t0 = time.perf_counter()
            out = irlba.compute(h, k, opts)
            times.append(time.perf_counter() - t0)
        st = ctx.stats()
        fwd_ms = st["forward_kernel_ms"] / max(1, st["forward_timed"])
        adj_ms = st["adjoint_kernel_ms"] / max(1, st["adjoint_timed"])
        runs[name] = out
        res[name] = {
            "solve_s": round(float(np.median(times)), 4), "iterations": int(out.iterations),
            "converged": bool(out.converged),
            "forward_ms": round(fwd_ms, 4), "adjoint_ms": round(adj_ms, 4),
            "forward_GBps": round(st["forward_bytes"] / (fwd_ms * 1e6), 1),
            "adjoint_GBps": round(st["adjoint_bytes"] / (adj_ms * 1e6), 1),
        }
        log(f"  dense {name}: {np.median(times):.3f} s, {out.iterations} restarts, fwd {fwd_ms:.3f} ms "
            f"({res[name]['forward_GBps']} GB/s), adj {adj_ms:.3f} ms ({res[name]['adjoint_GBps']} GB/s)")
        del h
        torch.cuda.empty_cache()
    ref = runs.get("f64")
    for name in [n for n in runs if n != "f64" and ref is not None]:
        o = runs[name]
        res[name]["sv_rel_err_vs_fp64"] = float(np.max(np.abs(o.D - ref.D) / ref.D))
        qa, _ = np.linalg.qr(o.V)
        qb, _ = np.linalg.qr(ref.V)
        cos = np.linalg.svd(qa.T @ qb, compute_uv=False)
        res[name]["max_subspace_angle_vs_fp64"] = float(np.arccos(np.clip(cos.min(), -1, 1)))
        res[name]["speedup_vs_fp64"] = round(res["f64"]["solve_s"] / res[name]["solve_s"], 3)
    del A64
    torch.cuda.empty_cache()
    return res


def _dict_leg(mat, res, dt, st, nr, nc):
    """Solve time, per-product times and the roofline of the A.v product for a
    matrix whose products stream the value-dictionary codes."""
    import numpy as np

    fwd = st["forward_kernel_ms"] / max(1, st["forward_timed"])
    adj = st["adjoint_kernel_ms"] / max(1, st["adjoint_timed"])
    lay = mat.layout(False)
    nseg = lay["n_tiles"] * nr
    # {code, delta} bytes per slot, segment pointers and first indices, FP64 vectors
    stream = int(lay["slots"] * 2 + (nseg + 1) * 8 + nseg * 2 + (nr + nc) * 8)
    return {
        "solve_s": round(dt, 4), "iterations": int(res.iterations), "converged": bool(res.converged),
        "ms_per_restart": round(1e3 * dt / max(1, res.iterations), 3),
        "forward_ms": round(fwd, 4), "adjoint_ms": round(adj, 4),
        "forward_lanes_per_line": lay["lanes_per_line"], "adjoint_lanes_per_line": mat.layout(True)["lanes_per_line"],
        "forward_stream_bytes": stream, "forward_achieved_GBps": round(stream / (fwd * 1e6), 1),
        "forward_frac": round(stream / (fwd * 1e6) / HBM_PEAK_GBS, 4),
        "_variances": np.asarray(res.variances),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", choices=["solve", "spmv", "backward", "dense"], default="solve")
    ap.add_argument("--reps", type=int, default=20)  # timed backward passes (median) and spmv products
    ap.add_argument("--storage", default="f32")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rows", type=int, default=30000)
    ap.add_argument("--cols", type=int, default=1000000)
    ap.add_argument("--density", type=float, default=0.03)
    ap.add_argument("--number", type=int, default=50)
    ap.add_argument("--extra-work", type=int, default=20)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-backward", action="store_true")
    ap.add_argument("--skip-dense", action="store_true")
    ap.add_argument("--dense-rows", type=int, default=200000)
    ap.add_argument("--dense-cols", type=int, default=2000)
    ap.add_argument("--dense-reps", type=int, default=1)
    ap.add_argument("--dense-storages", default="bf16,f32,f64")
    ap.add_argument("--dictionary", action="store_true",
                    help="--mode spmv: stream the value-dictionary codes instead of the values")
    ap.add_argument("--skip-dictionary", action="store_true",
                    help="skip the value-dictionary solves (codes in place of the stored values)")
    ap.add_argument("--skip-backward-exact-sparse", action="store_true",
                    help="time only the rank-k complement on the sparse backward case")
    ap.add_argument("--cpu-sample-cols", type=int, default=33334)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the IRLBA solve is not sharded across GPUs (a column-sharded "
              "layout with one all-reduce per forward product is the design, not built); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.mode == "spmv":
        spmv_mode(args)
        return
    if args.mode == "backward":
        print(json.dumps(backward_leg(args)), flush=True)
        return
    if args.mode == "dense":
        print(json.dumps(dense_leg(args)), flush=True)
        return

    import numpy as np

    import cppirlba_amd as irlba

    ctx = irlba.default_context()
    # Time one product in 8 with HIP events: each timed product idles the GPU a
    # few microseconds around its kernel, which over ~5000 products would add
    # ~1% to the solve; 1 in 8 still samples ~650 launches per direction.
    ctx.set_profiling(8)
    nr, nc = args.rows, args.cols
    log(f"generating {nr}x{nc} at density {args.density} on the device (mixed: FP32 storage)")
    t0 = time.perf_counter()
    mixed = irlba.generate_sparse(nr, nc, args.density, args.seed, storage="f32", accumulate="f32")
    ctx.synchronize()
    setup_s = time.perf_counter() - t0
    log(f"  generated + built both orientations in {setup_s:.2f} s")
    t0 = time.perf_counter()
    irlba.column_stats(mixed)
    stats_ms = 1e3 * (time.perf_counter() - t0)
    init = np.random.default_rng(1234).standard_normal(nc)
    opts = irlba.Options(extra_work=args.extra_work, initial=init)

    # Warmups return host arrays: their time is the PCIe-inclusive figure (the
    # 1e6 x 50 rotation is 400 MB); the timed steps keep the results in HBM.
    host_s = None
    for w in range(args.warmup):
        res, dt = solve(irlba, mixed, args.number, opts, device=False)
        host_s = dt
        log(f"  warmup {w} (results to host memory): {dt:.3f} s, {res.iterations} restarts, converged={res.converged}")
    ctx.synchronize()
    times, stats = [], []
    t_start = time.perf_counter()
    for k in range(args.steps):
        res, dt = solve(irlba, mixed, args.number, opts)
        times.append(dt)
        stats.append(ctx.stats())
        log(f"  step {k}: {dt:.3f} s, {res.iterations} restarts, converged={res.converged}")
    ctx.synchronize()
    total = time.perf_counter() - t_start
    mixed_res = host_copy(res)
    ms_per_step = 1e3 * total / args.steps
    st = stats[-1]
    nnz = None
    n_fwd, n_adj = st["forward_products"], st["adjoint_products"]
    fwd_ms = sum(s["forward_kernel_ms"] for s in stats) / max(1, sum(s["forward_timed"] for s in stats))
    adj_ms = sum(s["adjoint_kernel_ms"] for s in stats) / max(1, sum(s["adjoint_timed"] for s in stats))
    fwd_gbs = st["forward_bytes"] / (fwd_ms * 1e6)
    adj_gbs = st["adjoint_bytes"] / (adj_ms * 1e6)
    # nnz from the algorithmic bytes: nnz * (4 + 4) + (nc + 1) * 8 + (nr + nc) * 8
    nnz = int(round((st["adjoint_bytes"] - (nc + 1) * 8 - (nr + nc) * 8) / 8))
    dominant = "adjoint" if adj_ms * n_adj >= fwd_ms * n_fwd else "forward"
    dom_ms = adj_ms if dominant == "adjoint" else fwd_ms
    dom_bytes = st["adjoint_bytes"] if dominant == "adjoint" else st["forward_bytes"]
    # The bytes one launch of the dominant product must move in the layout it
    # streams: every stored slot (value + 1-byte delta index; entries, padding
    # and gap-bridging slots), the segment pointers and first indices, and the
    # FP64 input and output vectors.
    lay = mixed.layout(dominant == "adjoint")
    n_major = nc if dominant == "adjoint" else nr
    nseg = lay["n_tiles"] * n_major
    stored_bytes = int(lay["slots"] * 4 + lay["index_bytes"] + (nseg + 1) * 8 + (nseg * 2 if lay["delta"] else 0)
                       + (nr + nc) * 8)
    achieved = stored_bytes / (dom_ms * 1e6)

    out = {
        "metric": METRIC,
        "value": round(float(np.mean(times)), 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"top-{args.number} centred+scaled PCA (irlba pca, extra_work={args.extra_work}) of a "
                         f"{nr}x{nc} CSC matrix at density {args.density}, log1p NB(mu=2,phi=0.5) values, FP32 values "
                         f"+ int32 indices, mixed precision (FP32 products, FP64 Lanczos/CGS/B-SVD)"),
            "rows": nr, "cols": nc, "nnz": nnz, "number": args.number, "extra_work": args.extra_work,
            "tolerance": opts.convergence_tolerance, "seed": args.seed, "start_vector_seed": 1234,
            "parallelism": "single GPU",
        },
        "iterations": int(mixed_res.iterations),
        "converged": bool(mixed_res.converged),
        "results_in": "HBM (torch tensors); the PCIe-inclusive solve is solve_with_host_results_s",
        "solve_with_host_results_s": None if host_s is None else round(host_s, 4),
        "setup": {
            "generate_and_build_s": round(setup_s, 3),
            "column_stats_ms": round(stats_ms, 2),
            "note": "the matrix is generated on the device and both tiled orientations are built there (no "
                    "host upload); pca() recomputes the column statistics inside the timed solve",
        },
        # The whole solve's product bytes over its wall time (the products are
        # about 80% of it; Gram-Schmidt, rotations and the host B-SVD the rest).
        "solve_matvec_GBps": round((st["forward_bytes"] * n_fwd + st["adjoint_bytes"] * n_adj)
                                   / (float(np.mean(times)) * 1e9), 1),
        "spmv": {
            "forward_ms": round(fwd_ms, 4), "adjoint_ms": round(adj_ms, 4),
            "forward_GBps": round(fwd_gbs, 1), "adjoint_GBps": round(adj_gbs, 1),
            "forward_products": n_fwd, "adjoint_products": n_adj,
            "bytes_definition": "nnz*(4 value + 4 int32 index) + (cols+1)*8 pointers + (rows+cols)*8 FP64 vectors",
        },
        "host_svd_ms": round(st["host_svd_ms"], 2),
        "roofline": {
            "kernel": f"k_spmv_tiled ({dominant} product)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_launch": stored_bytes,
            "launch_ms": round(dom_ms, 4),
            "note": "achieved = bytes of the layout one launch streams (slots x (4-byte FP32 value + 1-byte delta "
                    "index) + segment pointers and first indices + FP64 input/output vectors) / mean launch time "
                    "(HIP events on the engine's stream); frac_int32_format_equiv divides the nnz x (4 + 4-byte "
                    "int32 index) format bytes of the workload definition instead, which no launch reads",
            "frac_int32_format_equiv": round(dom_bytes / (dom_ms * 1e6) / HBM_PEAK_GBS, 4),
            "layout": lay,
        },
    }
    prof = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            tr = json.load(f)
        key = f"{dominant}_hbm_bytes_per_launch"
        if key in tr:
            # Measured HBM bytes of one launch (calibrated FETCH_SIZE + WRITE_SIZE) from
            # the committed counter profile of this workload, over this run's launch time.
            out["roofline"]["traffic"] = tr[key]
            out["roofline"]["traffic_source"] = ("profiles/spmv_traffic.json (rocprofv3 --pmc FETCH_SIZE / "
                                                 "WRITE_SIZE runs of bench.py --mode spmv, calibrated by a copy "
                                                 "kernel); not measured in this run")
            out["roofline"]["traffic_frac"] = round(tr[key] / (dom_ms * 1e6) / HBM_PEAK_GBS, 4)

    dict_out = None
    if not args.skip_dictionary:
        # The same matrix with its values streamed as one code byte per slot
        # (a table of the distinct stored values; lossless, and with FP32
        # storage every product term is the same number).
        log("value dictionary: mixed solve with the values streamed as codes")
        t0 = time.perf_counter()
        n_codes = mixed.value_dictionary(True)
        ctx.synchronize()
        build_dict_s = time.perf_counter() - t0
        res_d, dt_d = solve(irlba, mixed, args.number, opts)
        st_d = ctx.stats()
        res_d = host_copy(res_d)
        log(f"  dictionary of {n_codes} values built in {build_dict_s:.3f} s; mixed solve {dt_d:.3f} s, "
            f"{res_d.iterations} restarts")
        dict_out = {"distinct_values": n_codes, "build_s": round(build_dict_s, 3),
                    "mixed": _dict_leg(mixed, res_d, dt_d, st_d, nr, nc)}
        prof = os.path.join(ROOT, "profiles", "spmv_traffic.json")
        if os.path.exists(prof):
            with open(prof) as f:
                tr = json.load(f)
            if "dictionary_forward_hbm_bytes_per_launch" in tr:
                # Measured HBM bytes of one coded A.v (committed counter profile), over this run's launch time.
                t_b = tr["dictionary_forward_hbm_bytes_per_launch"]
                dict_out["mixed"]["forward_traffic"] = t_b
                dict_out["mixed"]["forward_traffic_frac"] = round(
                    t_b / (dict_out["mixed"]["forward_ms"] * 1e6) / HBM_PEAK_GBS, 4)
                dict_out["traffic_source"] = ("profiles/spmv_traffic.json (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE "
                                              "runs of bench.py --mode spmv --dictionary); not measured in this run")
        mixed.value_dictionary(False)

    if not args.skip_fp64:
        mixed_pair_bytes = product_pair_bytes(mixed, 4, nr, nc)
        del mixed
        log("FP64 path on the same matrix (FP64 storage and accumulation)")
        fp64 = irlba.generate_sparse(nr, nc, args.density, args.seed, storage="f64", accumulate="f64")
        ctx.synchronize()
        res64, dt64 = solve(irlba, fp64, args.number, opts)
        st64 = ctx.stats()
        res64 = host_copy(res64)
        log(f"  fp64: {dt64:.3f} s, {res64.iterations} restarts")
        from numpy.linalg import norm

        sv_m = np.sqrt(mixed_res.variances)
        sv_d = np.sqrt(res64.variances)
        rel = float(np.max(np.abs(sv_m - sv_d) / sv_d))
        qa, _ = np.linalg.qr(mixed_res.rotation)
        qb, _ = np.linalg.qr(res64.rotation)
        cosines = np.linalg.svd(qa.T @ qb, compute_uv=False)
        angle = float(np.arccos(np.clip(cosines.min(), -1, 1)))
        out["fp64"] = {
            "solve_s": round(dt64, 4), "iterations": int(res64.iterations),
            "forward_ms": round(st64["forward_kernel_ms"] / max(1, st64["forward_timed"]), 4),
            "adjoint_ms": round(st64["adjoint_kernel_ms"] / max(1, st64["adjoint_timed"]), 4),
        }
        out["mixed_vs_fp64_speedup"] = round(dt64 / float(np.mean(times)), 3)
        # The two runs need different restart counts (rounding steers the
        # restarts on this flat spectrum), so also the time per restart.
        out["mixed_vs_fp64_per_restart_speedup"] = round(
            (dt64 / max(1, res64.iterations)) / (float(np.mean(times)) / max(1, mixed_res.iterations)), 3)
        # What bounds the per-restart ratio: the products (the only part whose
        # bytes depend on the value precision) against their streamed bytes,
        # and the FP64 work both runs share (Gram-Schmidt, rotations, norms).
        ms_m = out["spmv"]["forward_ms"] + out["spmv"]["adjoint_ms"]
        ms_d = out["fp64"]["forward_ms"] + out["fp64"]["adjoint_ms"]
        pairs_m = min(out["spmv"]["forward_products"], out["spmv"]["adjoint_products"]) / max(1, mixed_res.iterations)
        restart_m = float(np.mean(times)) * 1e3 / max(1, mixed_res.iterations)
        restart_d = dt64 * 1e3 / max(1, res64.iterations)
        b_m, b_d = mixed_pair_bytes, product_pair_bytes(fp64, 8, nr, nc)
        out["mixed_vs_fp64_breakdown"] = {
            "products_speedup": round(ms_d / ms_m, 3),
            "product_bytes_ratio": round(b_d / b_m, 3),
            "product_pair_bytes": {"mixed": int(b_m), "fp64": int(b_d)},
            "product_pairs_per_restart": round(pairs_m, 2),
            "shared_fp64_ms_per_restart": {"mixed": round(restart_m - pairs_m * ms_m, 3),
                                           "fp64": round(restart_d - pairs_m * ms_d, 3)},
            "note": "per restart both precisions run the same FP64 Gram-Schmidt, rotations and norms (FP64 by "
                    "design); only the products change, and they stream FP32 value + 8-bit delta (5 bytes per "
                    "slot) against FP64 value + delta (9 bytes), so the per-restart ratio stays below the products' "
                    "own ratio; 2x per restart would need the products to beat their byte ratio by more than the "
                    "shared work costs",
        }
        out["sv_rel_err_vs_fp64"] = rel
        out["max_subspace_angle_vs_fp64"] = angle
        if dict_out is not None:
            log("value dictionary: FP64 solve with the values streamed as codes")
            fp64.value_dictionary(True)
            res_d64, dt_d64 = solve(irlba, fp64, args.number, opts)
            st_d64 = ctx.stats()
            res_d64 = host_copy(res_d64)
            log(f"  fp64 with codes: {dt_d64:.3f} s, {res_d64.iterations} restarts")
            dict_out["fp64"] = _dict_leg(fp64, res_d64, dt_d64, st_d64, nr, nc)
            dm = dict_out["mixed"]
            dict_out["mixed_vs_fp64_per_restart_speedup"] = round(
                dict_out["fp64"]["ms_per_restart"] / dm["ms_per_restart"], 3)
            var_m = dm.pop("_variances")
            var_d = dict_out["fp64"].pop("_variances")
            # singular values of the coded solves against the value-stream runs
            dict_out["mixed_sv_rel_err_vs_fp64"] = float(np.max(np.abs(np.sqrt(var_m) - sv_d) / sv_d))
            dict_out["mixed_sv_rel_err_vs_value_stream_mixed"] = float(np.max(np.abs(np.sqrt(var_m) - sv_m) / sv_m))
            dict_out["fp64_sv_rel_err_vs_value_stream_fp64"] = float(np.max(np.abs(np.sqrt(var_d) - sv_d) / sv_d))
        del fp64
    if dict_out is not None:
        for leg in ("mixed", "fp64"):
            if leg in dict_out:
                dict_out[leg].pop("_variances", None)
        dict_out["note"] = ("irlba_hip_sparse_value_dictionary on the same matrices: the products stream a "
                            "{1-byte code, 1-byte delta} pair per slot and read the value from an LDS table of the "
                            "distinct stored values (lossless); roofline bytes are those of that stream")
        out["value_dictionary"] = dict_out

    if not args.skip_backward:
        log("SVD backward (top-16, dense and sparse)")
        out["svd_backward"] = backward_leg(args)

    if not args.skip_dense:
        log("dense mixed precision (BF16 / FP32 / FP64 storage)")
        out["dense"] = dense_leg(args)

    if not args.skip_cpu and nnz:
        log("CPU baseline (oracle restatement of ParallelSparseMatrix) on a column slice")
        out["cpu_baseline"] = cpu_baseline(nr, nc, args.density, args.seed, nnz, n_fwd, n_adj,
                                           int(mixed_res.iterations), args.number, args.extra_work,
                                           min(args.cpu_sample_cols, nc))

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
