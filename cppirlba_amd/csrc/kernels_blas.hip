// FP64 Lanczos-vector kernels for gfx950.
//
//  * Classical Gram-Schmidt (include/irlba/lanczos.hpp:17-21) as a tall-skinny
//    GEMV pair: t = M[:, :n]' f with per-block partials and a fixed-order
//    reduction (deterministic), then f -= M[:, :n] t fused with the per-block
//    sums of squares of the updated f (the next Lanczos norm).
//  * Norm finalisation + normalisation (lanczos.hpp:81-85, 99-109, 123-132).
//  * The restart rotation V[:, :k] <- V Y (compute.hpp:268-280, 302-305): on the
//    vector FP64 units, a thread per row, for long bases with k <= 72, else on
//    the FP64 matrix cores (v_mfma_f64_16x16x4f64) per 16-row strip; in place.
#include "common.hpp"

namespace irlba_hip {

int cgs_blocks(i64 L) {
    return static_cast<int>(std::max<i64>(1, std::min<i64>(1024, (L + 255) / 256)));
}

static inline i64 cgs_chunk(i64 L, int G) { return (L + G - 1) / G; }

// ------------------------------------------------------------------ t = M' f (partials)
// Thread per row (coalesced column reads), U rows per thread in flight, the
// columns in batches of C held in registers (C * U accumulators).  Each batch's
// per-lane sums are reduced across the wave by recursive halving (lanes
// exchange the half of the batch the partner keeps), which leaves column q of
// the batch spread over 64 / C lanes; a few xor-adds finish it and four waves
// combine through LDS.  The basis is streamed once per call, so its loads are
// nontemporal (they would only evict the vectors from the caches).  8-column
// batches, one per block (grid.y), with 4 rows per thread for long vectors and
// one for short ones: few registers, many loads in flight, and enough blocks
// for short vectors too.

// Long vectors: 8192-row chunks, each block looping over its chunk (fewer,
// longer-lived blocks and fewer partials: 90 us against 100 us for one-pass
// 1024-row blocks on a 1e6 x 67 basis), at most 1024 blocks.  Their count is
// rounded up to a multiple of the 8 XCDs: workgroups go to the XCDs round-robin
// by linear id (x + y * gridDim.x), so every column batch (y) of a row chunk (x)
// then runs on one XCD and the chunk of f comes from that XCD's L2 after the
// first batch instead of from the MALL nine times (96.4 -> 90.6 us on a
// 1e6 x 67 basis).  Short ones: 256-row chunks (U = 1) so that enough blocks
// share the work.
constexpr i64 kDotLong = 262144;
int cgs_dot_blocks(i64 L) {
    if (L >= kDotLong) return static_cast<int>(std::min<i64>(1024, ((L + 8191) / 8192 + 7) / 8 * 8));
    return static_cast<int>(std::max<i64>(1, std::min<i64>(1024, (L + 255) / 256)));
}

template <int C, int U>
__global__ __launch_bounds__(256) void k_cgs_dots(const double* __restrict__ M, i64 ld, int n,
                                                  const double* __restrict__ f, i64 L, i64 chunk,
                                                  double* __restrict__ partials) {
    static_assert(C == 8 || C == 16 || C == 32, "batch width");
    constexpr int span = 64 / C;  // lanes holding parts of one column after the halving
    __shared__ double red[4][C];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const i64 r0 = static_cast<i64>(blockIdx.x) * chunk;
    const i64 r1 = min(L, r0 + chunk);
    // Column batches are spread over gridDim.y: with one batch per block (long
    // vectors) the blocks resident on a CU touch only C columns, so the pages
    // of the basis they walk stay within the CU's translation cache.
    for (int cb = blockIdx.y * C; cb < n; cb += C * gridDim.y) {
        double v[C];
#pragma unroll
        for (int q = 0; q < C; ++q) v[q] = 0;
        const int nb = min(C, n - cb);
        for (i64 r = r0 + threadIdx.x; r < r1; r += 256 * U) {
            double fr[U];
#pragma unroll
            for (int u = 0; u < U; ++u) fr[u] = r + 256 * u < r1 ? f[r + 256 * u] : 0.0;
            const double* p = M + r + static_cast<i64>(cb) * ld;
#pragma unroll
            for (int q = 0; q < C; ++q) {
                if (q < nb) {
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (r + 256 * u < r1) v[q] += __builtin_nontemporal_load(p + 256 * u) * fr[u];
                }
                p += ld;
            }
        }
        // Halving over offsets 32, 16, ..., span leaves partial sums of column q
        // in the `span` lanes with lane / span == q; xor-adds total them.
#pragma unroll
        for (int half = C / 2, o = 32; half >= 1; half >>= 1, o >>= 1) {
            const bool upper = (lane & o) != 0;
#pragma unroll
            for (int j = 0; j < half; ++j) {
                const double send = upper ? v[j] : v[j + half];
                const double keep = upper ? v[j + half] : v[j];
                v[j] = keep + __shfl_xor(send, o, 64);
            }
        }
#pragma unroll
        for (int o = span / 2; o >= 1; o >>= 1) v[0] += __shfl_xor(v[0], o, 64);
        if ((lane & (span - 1)) == 0) red[w][lane / span] = v[0];
        __syncthreads();
        if (threadIdx.x < nb)
            partials[static_cast<i64>(blockIdx.x) * n + cb + threadIdx.x] =
                (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
        __syncthreads();
    }
}

// The same product with 16-byte loads: a lane holds two consecutive rows, U
// row pairs per thread in flight (512-row steps).  Needs an even leading
// dimension, an even chunk and 16-byte aligned M and f (checked by the
// launcher); with 8192-row chunks, measured 90 us against 140 us for the
// 8-byte form on a 1e6 x 67 basis.
template <int C, int U>
__global__ __launch_bounds__(256) void k_cgs_dots2(const double* __restrict__ M, i64 ld, int n,
                                                   const double* __restrict__ f, i64 L, i64 chunk,
                                                   double* __restrict__ partials) {
    constexpr int span = 64 / C;
    __shared__ double red[4][C];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const i64 r0 = static_cast<i64>(blockIdx.x) * chunk;
    const i64 r1 = min(L, r0 + chunk);
    for (int cb = blockIdx.y * C; cb < n; cb += C * gridDim.y) {
        double v[C];
#pragma unroll
        for (int q = 0; q < C; ++q) v[q] = 0;
        const int nb = min(C, n - cb);
        for (i64 r = r0 + 2 * threadIdx.x; r < r1; r += 512 * U) {
            // r is even and L is even, so r < r1 also covers row r + 1.
            f64x2 fr[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                fr[u] = r + 512 * u < r1 ? *reinterpret_cast<const f64x2*>(f + r + 512 * u) : f64x2{0, 0};
            const double* p = M + r + static_cast<i64>(cb) * ld;
#pragma unroll
            for (int q = 0; q < C; ++q) {
                if (q < nb) {
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (r + 512 * u < r1) {
                            const f64x2 m = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(p + 512 * u));
                            v[q] += m[0] * fr[u][0] + m[1] * fr[u][1];
                        }
                }
                p += ld;
            }
        }
#pragma unroll
        for (int half = C / 2, o = 32; half >= 1; half >>= 1, o >>= 1) {
            const bool upper = (lane & o) != 0;
#pragma unroll
            for (int j = 0; j < half; ++j) {
                const double send = upper ? v[j] : v[j + half];
                const double keep = upper ? v[j + half] : v[j];
                v[j] = keep + __shfl_xor(send, o, 64);
            }
        }
#pragma unroll
        for (int o = span / 2; o >= 1; o >>= 1) v[0] += __shfl_xor(v[0], o, 64);
        if ((lane & (span - 1)) == 0) red[w][lane / span] = v[0];
        __syncthreads();
        if (threadIdx.x < nb)
            partials[static_cast<i64>(blockIdx.x) * n + cb + threadIdx.x] =
                (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
        __syncthreads();
    }
}

static inline bool pairs_ok(const double* a, const double* b, i64 ld, i64 L) {
    return (ld % 2 == 0) && (L % 2 == 0) && (reinterpret_cast<std::uintptr_t>(a) % 16 == 0) &&
           (reinterpret_cast<std::uintptr_t>(b) % 16 == 0);
}

void launch_cgs_dots(const double* M, i64 ld, int n, const double* f, i64 L, double* partials, hipStream_t s) {
    if (n <= 0) return;
    const int G = cgs_dot_blocks(L);
    if (L >= kDotLong && pairs_ok(M, f, ld, L)) {
        const i64 chunk = (cgs_chunk(L, G) + 1) & ~static_cast<i64>(1);
        hipLaunchKernelGGL((k_cgs_dots2<8, 1>), dim3(G, (n + 7) / 8), dim3(256), 0, s, M, ld, n, f, L, chunk, partials);
    } else if (L >= kDotLong)
        hipLaunchKernelGGL((k_cgs_dots<8, 4>), dim3(G, (n + 7) / 8), dim3(256), 0, s, M, ld, n, f, L, cgs_chunk(L, G),
                           partials);
    else
        hipLaunchKernelGGL((k_cgs_dots<8, 1>), dim3(G, (n + 7) / 8), dim3(256), 0, s, M, ld, n, f, L, cgs_chunk(L, G),
                           partials);
    IRLBA_HIP_LAUNCH_CHECK();
}

// t[i] = sum over blocks of the partials of column i: one block per column, a
// fixed strided split and tree, so the result does not depend on timing.
__global__ __launch_bounds__(256) void k_cgs_reduce(const double* __restrict__ partials, int G, int n,
                                                    double* __restrict__ t) {
    __shared__ double red[32];
    const int i = blockIdx.x;
    double acc = 0, dummy = 0;
    for (int b = threadIdx.x; b < G; b += blockDim.x) acc += partials[static_cast<i64>(b) * n + i];
    block_sum2(acc, dummy, red);
    if (threadIdx.x == 0) t[i] = acc;
}

void launch_cgs_reduce(const double* partials, i64 L, int n, double* t, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cgs_reduce, dim3(n), dim3(256), 0, s, partials, cgs_dot_blocks(L), n, t);
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ f -= M t, plus sum of squares
// Thread per row, U rows in flight per thread, columns CB at a time with
// nontemporal loads (U * CB = 16 loads in flight per thread); t staged in LDS.
template <int U, int CB>
__global__ __launch_bounds__(256) void k_cgs_update(const double* __restrict__ M, i64 ld, int n,
                                                    const double* __restrict__ t, double* __restrict__ f, i64 L,
                                                    i64 chunk, double* __restrict__ norm_part) {
    extern __shared__ double ts[];
    __shared__ double red[32];
    for (int i = threadIdx.x; i < n; i += blockDim.x) ts[i] = t[i];
    __syncthreads();
    const i64 r0 = static_cast<i64>(blockIdx.x) * chunk;
    const i64 r1 = min(L, r0 + chunk);
    double sq = 0;
    for (i64 r = r0 + threadIdx.x; r < r1; r += 256 * U) {
        bool in[U];
        double acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            in[u] = r + 256 * u < r1;
            acc[u] = 0;
        }
        const double* p = M + r;
        int i = 0;
        for (; i + CB <= n; i += CB) {
            double m[CB][U];
#pragma unroll
            for (int c = 0; c < CB; ++c)
#pragma unroll
                for (int u = 0; u < U; ++u) m[c][u] = in[u] ? __builtin_nontemporal_load(p + c * ld + 256 * u) : 0.0;
#pragma unroll
            for (int c = 0; c < CB; ++c)
#pragma unroll
                for (int u = 0; u < U; ++u) acc[u] += m[c][u] * ts[i + c];
            p += CB * ld;
        }
        for (; i < n; ++i) {
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (in[u]) acc[u] += __builtin_nontemporal_load(p + 256 * u) * ts[i];
            p += ld;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (in[u]) {
                const double v = f[r + 256 * u] - acc[u];
                f[r + 256 * u] = v;
                sq += v * v;
            }
    }
    double dummy = 0;
    block_sum2(sq, dummy, red);
    if (threadIdx.x == 0) norm_part[blockIdx.x] = sq;
}

constexpr int kMaxWideCols = 1024;

// The same update with 16-byte accesses (a lane holds two consecutive rows, U
// pairs per thread, CB columns per batch); the conditions of k_cgs_dots2 apply.
template <int U, int CB>
__global__ __launch_bounds__(256) void k_cgs_update2(const double* __restrict__ M, i64 ld, int n,
                                                     const double* __restrict__ t, double* __restrict__ f, i64 L,
                                                     i64 chunk, double* __restrict__ norm_part) {
    extern __shared__ double ts[];
    __shared__ double red[32];
    for (int i = threadIdx.x; i < n; i += blockDim.x) ts[i] = t[i];
    __syncthreads();
    const i64 r0 = static_cast<i64>(blockIdx.x) * chunk;
    const i64 r1 = min(L, r0 + chunk);
    double sq = 0;
    for (i64 r = r0 + 2 * threadIdx.x; r < r1; r += 512 * U) {
        bool in[U];
        f64x2 acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            in[u] = r + 512 * u < r1;
            acc[u] = f64x2{0, 0};
        }
        const double* p = M + r;
        int i = 0;
        for (; i + CB <= n; i += CB) {
            f64x2 m[CB][U];
#pragma unroll
            for (int c = 0; c < CB; ++c)
#pragma unroll
                for (int u = 0; u < U; ++u)
                    m[c][u] = in[u] ? __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(p + c * ld + 512 * u))
                                    : f64x2{0, 0};
#pragma unroll
            for (int c = 0; c < CB; ++c)
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    acc[u][0] += m[c][u][0] * ts[i + c];
                    acc[u][1] += m[c][u][1] * ts[i + c];
                }
            p += CB * ld;
        }
        for (; i < n; ++i) {
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (in[u]) {
                    const f64x2 m = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(p + 512 * u));
                    acc[u][0] += m[0] * ts[i];
                    acc[u][1] += m[1] * ts[i];
                }
            p += ld;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (in[u]) {
                f64x2 v = *reinterpret_cast<const f64x2*>(f + r + 512 * u);
                v[0] -= acc[u][0];
                v[1] -= acc[u][1];
                *reinterpret_cast<f64x2*>(f + r + 512 * u) = v;
                sq += v[0] * v[0] + v[1] * v[1];
            }
    }
    double dummy = 0;
    block_sum2(sq, dummy, red);
    if (threadIdx.x == 0) norm_part[blockIdx.x] = sq;
}

// The update for short vectors (at most 256 rows per block): 1024 threads, 4
// row groups of 64 x 4 column groups, so that a lane has about n/4 loads, all
// in flight at once, instead of n loads in dependent batches; the four column
// partials of a row are summed in a fixed order through LDS.
__global__ __launch_bounds__(1024) void k_cgs_update_wide(const double* __restrict__ M, i64 ld, int n,
                                                           const double* __restrict__ t, double* __restrict__ f,
                                                           i64 L, i64 chunk, double* __restrict__ norm_part) {
    __shared__ double ts[kMaxWideCols];
    __shared__ double part[4][256];
    __shared__ double red[32];
    for (int i = threadIdx.x; i < n; i += blockDim.x) ts[i] = t[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int rg = w & 3, cg = w >> 2;
    const i64 r0 = static_cast<i64>(blockIdx.x) * chunk;
    const i64 r1 = min(L, r0 + chunk);
    const int row_in = rg * 64 + lane;
    const i64 r = r0 + row_in;
    const int q0 = n * cg / 4, q1 = n * (cg + 1) / 4;
    double acc = 0;
    if (r < r1) {
        const double* p = M + r + static_cast<i64>(q0) * ld;
        int q = q0;
        for (; q + 8 <= q1; q += 8) {
            double m[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) m[c] = __builtin_nontemporal_load(p + c * ld);
#pragma unroll
            for (int c = 0; c < 8; ++c) acc += m[c] * ts[q + c];
            p += 8 * ld;
        }
        for (; q < q1; ++q) {
            acc += __builtin_nontemporal_load(p) * ts[q];
            p += ld;
        }
    }
    part[cg][row_in] = acc;
    __syncthreads();
    double sq = 0, dummy = 0;
    if (cg == 0 && r < r1) {
        const double v = f[r] - ((part[0][row_in] + part[1][row_in]) + (part[2][row_in] + part[3][row_in]));
        f[r] = v;
        sq = v * v;
    }
    block_sum2(sq, dummy, red);
    if (threadIdx.x == 0) norm_part[blockIdx.x] = sq;
}

void launch_cgs_update(const double* M, i64 ld, int n, const double* t, double* f, i64 L, double* norm_part,
                       hipStream_t s) {
    const int G = cgs_blocks(L);  // the norm finalisation sums exactly these partials
    const size_t lds = sizeof(double) * static_cast<size_t>(std::max(n, 1));
    if (lds > 60000) throw std::runtime_error("working subspace too large for the CGS update kernel");
    const i64 chunk = cgs_chunk(L, G);
    if (chunk <= 256 && n <= kMaxWideCols) {
        hipLaunchKernelGGL(k_cgs_update_wide, dim3(G), dim3(1024), 0, s, M, ld, n, t, f, L, chunk, norm_part);
    } else if (chunk >= 512 && pairs_ok(M, f, ld, L)) {
        const i64 chunk2 = (chunk + 1) & ~static_cast<i64>(1);  // G * chunk2 >= L; trailing blocks write 0
        hipLaunchKernelGGL((k_cgs_update2<2, 8>), dim3(G), dim3(256), lds, s, M, ld, n, t, f, L, chunk2, norm_part);
    } else if (chunk >= 512)
        hipLaunchKernelGGL((k_cgs_update<4, 4>), dim3(G), dim3(256), lds, s, M, ld, n, t, f, L, chunk, norm_part);
    else
        hipLaunchKernelGGL((k_cgs_update<1, 16>), dim3(G), dim3(256), lds, s, M, ld, n, t, f, L, chunk, norm_part);
    IRLBA_HIP_LAUNCH_CHECK();
}

__global__ __launch_bounds__(256) void k_sumsq(const double* __restrict__ f, i64 L, i64 chunk,
                                               double* __restrict__ norm_part) {
    __shared__ double red[32];
    const i64 r0 = static_cast<i64>(blockIdx.x) * chunk;
    const i64 r1 = min(L, r0 + chunk);
    double sq = 0, dummy = 0;
    for (i64 r = r0 + threadIdx.x; r < r1; r += blockDim.x) sq += f[r] * f[r];
    block_sum2(sq, dummy, red);
    if (threadIdx.x == 0) norm_part[blockIdx.x] = sq;
}

void launch_sumsq(const double* f, i64 L, double* norm_part, hipStream_t s) {
    const int G = cgs_blocks(L);
    hipLaunchKernelGGL(k_sumsq, dim3(G), dim3(256), 0, s, f, L, cgs_chunk(L, G), norm_part);
    IRLBA_HIP_LAUNCH_CHECK();
}

// norm = sqrt(sum of partials); optionally dst = f / norm.  Every block
// re-sums the G partials in the same order, so all agree bit for bit.
__global__ __launch_bounds__(256) void k_normalize(const double* __restrict__ f, i64 L, const double* __restrict__ np,
                                                   int G, double* __restrict__ norm_out, double* __restrict__ dst) {
    __shared__ double red[32];
    double acc = 0, dummy = 0;
    for (int b = threadIdx.x; b < G; b += blockDim.x) acc += np[b];
    block_sum2(acc, dummy, red);
    const double nrm = sqrt(acc);
    if (blockIdx.x == 0 && threadIdx.x == 0 && norm_out) *norm_out = nrm;
    if (dst) {
        for (i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; r < L;
             r += static_cast<i64>(gridDim.x) * blockDim.x)
            dst[r] = f[r] / nrm;
    }
}

void launch_normalize(const double* f, i64 L, const double* norm_part, int G, double* norm_out, double* dst,
                      hipStream_t s) {
    const i64 grid = dst ? std::max<i64>(1, std::min<i64>((L + 255) / 256, 2048)) : 1;
    hipLaunchKernelGGL(k_normalize, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, f, L, norm_part, G, norm_out,
                       dst);
    IRLBA_HIP_LAUNCH_CHECK();
}

__global__ void k_scale_copy(const double* __restrict__ src, i64 L, double denom, double* __restrict__ dst) {
    for (i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; r < L;
         r += static_cast<i64>(gridDim.x) * blockDim.x)
        dst[r] = src[r] / denom;
}

void launch_scale_copy(const double* src, i64 L, double denom, double* dst, hipStream_t s) {
    const i64 grid = std::max<i64>(1, std::min<i64>((L + 255) / 256, 4096));
    hipLaunchKernelGGL(k_scale_copy, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, src, L, denom, dst);
    IRLBA_HIP_LAUNCH_CHECK();
}

__global__ void k_scale_copy_dev(const double* __restrict__ src, i64 L, const double* __restrict__ denom,
                                 double* __restrict__ dst) {
    const double d = *denom;
    for (i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; r < L;
         r += static_cast<i64>(gridDim.x) * blockDim.x)
        dst[r] = src[r] / d;
}

void launch_scale_copy_dev(const double* src, i64 L, const double* denom, double* dst, hipStream_t s) {
    const i64 grid = std::max<i64>(1, std::min<i64>((L + 255) / 256, 4096));
    hipLaunchKernelGGL(k_scale_copy_dev, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, src, L, denom, dst);
    IRLBA_HIP_LAUNCH_CHECK();
}

// out = op(a, b): 0 a*b, 1 a/b, 2 a+b, 3 1/a, 4 a
__global__ void k_vec_combine(double* __restrict__ out, const double* __restrict__ a, const double* __restrict__ b,
                              i64 n, int op) {
    for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * blockDim.x) {
        const double x = a[i];
        double y;
        switch (op) {
            case 0: y = x * b[i]; break;
            case 1: y = x / b[i]; break;
            case 2: y = x + b[i]; break;
            case 3: y = 1.0 / x; break;
            default: y = x; break;
        }
        out[i] = y;
    }
}

void launch_vec_combine(double* out, const double* a, const double* b, i64 n, int op, hipStream_t s) {
    if (n <= 0) return;
    const i64 grid = std::max<i64>(1, std::min<i64>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(k_vec_combine, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, out, a, b, n, op);
    IRLBA_HIP_LAUNCH_CHECK();
}

// y -= (*coef) * src, the Lanczos axpy with the coefficient still on the device.
__global__ void k_axpy_dev(double* __restrict__ y, const double* __restrict__ src, const double* __restrict__ coef,
                           i64 n) {
    const double c = *coef;
    for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * blockDim.x)
        y[i] -= c * src[i];
}

void launch_axpy_dev(double* y, const double* src, const double* coef, i64 n, hipStream_t s) {
    if (n <= 0) return;
    const i64 grid = std::max<i64>(1, std::min<i64>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(k_axpy_dev, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, y, src, coef, n);
    IRLBA_HIP_LAUNCH_CHECK();
}

// M[:, j] *= d[j] for j < k (scores = U diag(d), pca.hpp:130).
__global__ void k_scale_columns(double* __restrict__ M, i64 ld, i64 L, int k, const double* __restrict__ d) {
    const i64 n = L * k;
    for (i64 e = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<i64>(gridDim.x) * blockDim.x) {
        const i64 j = e / L, r = e - j * L;
        M[r + j * ld] *= d[j];
    }
}

void launch_scale_columns(double* M, i64 ld, i64 L, int k, const double* d, hipStream_t s) {
    if (L <= 0 || k <= 0) return;
    const i64 grid = std::max<i64>(1, std::min<i64>((L * k + 255) / 256, 8192));
    hipLaunchKernelGGL(k_scale_columns, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, M, ld, L, k, d);
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ restart rotation on MFMA
// out[:, :k] = M[:, :work] * Y, Y given row-major as Yt[work][ldy] (ldy a multiple
// of 16, zero padded), k <= 128 per pass.  A wave owns S strips of 16 rows and
// all NCT = ceil(k/16) column tiles: per K step of 4 it loads S A fragments
// straight from HBM (each a 4-column x 16-row block = four full 128-byte
// lines) and NCT B fragments (L1/L2-resident Yt), then issues S*NCT
// v_mfma_f64_16x16x4f64; the next step's fragments are loaded before the
// current MFMAs.  The finished tiles go through a per-wave LDS transpose so that
// each store writes whole column segments.
// MFMA operand layout (lane l, checked on gfx950):
//   A operand: lane l holds M[row0 + l%16, kk + l/16]
//   B operand: lane l holds Y[kk + l/16, col0 + l%16]
//   D:         lane l, element i is out[row0 + l/16 + 4*i, col0 + l%16]
// A wave reads all of its rows before it writes any, and waves own disjoint
// rows, so out == M (the in-place restart update) is safe.
typedef double double4_t __attribute__((ext_vector_type(4)));

template <int NCT, int S>
__global__ __launch_bounds__(256) void k_rotate(const double* M, i64 ld, int work, const double* __restrict__ Yt,
                                                int ldy, int k, int col_begin, double* out, i64 ld_out, i64 L) {
    constexpr int ROWS = 16 * S;         // rows per wave
    constexpr int LDR = ROWS + 1;        // padded LDS column stride (bank spread)
    __shared__ double stage[4][16 * LDR];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const i64 row0 = (static_cast<i64>(blockIdx.x) * 4 + w) * ROWS;
    if (row0 >= L) return;
    const double* Yc = Yt + col_begin;

    double4_t acc[S][NCT];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int c = 0; c < NCT; ++c) acc[s][c] = double4_t{0, 0, 0, 0};

    auto load_a = [&](int kidx, double (&a)[S]) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const i64 row = row0 + s * 16 + lr;
            a[s] = (kidx < work && row < L) ? __builtin_nontemporal_load(M + row + static_cast<i64>(kidx) * ld) : 0.0;
        }
    };
    auto load_b = [&](int kidx, double (&b)[NCT]) {
#pragma unroll
        for (int c = 0; c < NCT; ++c) b[c] = kidx < work ? Yc[static_cast<i64>(kidx) * ldy + c * 16 + lr] : 0.0;
    };

    // A fragments are loaded PD steps ahead (a ring of PD register sets), so a
    // wave keeps PD*S*512 bytes of HBM reads in flight at 1-2 waves per SIMD.
    constexpr int PD = 16 / S;
    double abuf[PD][S];
#pragma unroll
    for (int p = 0; p < PD; ++p) load_a(4 * p + lk, abuf[p]);
    for (int kk0 = 0; kk0 < work; kk0 += 4 * PD) {
#pragma unroll
        for (int p = 0; p < PD; ++p) {
            const int kk = kk0 + 4 * p;
            if (kk < work) {
                double b[NCT];
                load_b(kk + lk, b);
#pragma unroll
                for (int s = 0; s < S; ++s)
#pragma unroll
                    for (int c = 0; c < NCT; ++c)
                        acc[s][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(abuf[p][s], b[c], acc[s][c], 0, 0, 0);
                load_a(kk + 4 * PD + lk, abuf[p]);
            }
        }
    }

    double* st = stage[w];
    constexpr int COLS_PER_STORE = 64 / ROWS;  // columns one store instruction covers
#pragma unroll
    for (int c = 0; c < NCT; ++c) {
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int i = 0; i < 4; ++i) st[lr * LDR + s * 16 + lk + 4 * i] = acc[s][c][i];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the LDS writes have landed
        const int sub = lane / ROWS, rr = lane - sub * ROWS;
        const i64 row = row0 + rr;
#pragma unroll
        for (int cc = 0; cc < 16; cc += COLS_PER_STORE) {
            const int col = c * 16 + cc + sub;
            if (col_begin + col < k && row < L)
                __builtin_nontemporal_store(st[(cc + sub) * LDR + rr], out + row + static_cast<i64>(col_begin + col) * ld_out);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_s_waitcnt(0xc07f);
    }
}

template <int NCT>
static void rotate_launch(const double* M, i64 ld, int work, const double* Yt, int ldy, int k, int col_begin,
                          double* out, i64 ld_out, i64 L, hipStream_t s) {
    constexpr int S = NCT <= 4 ? 4 : 2;
    const i64 rows_per_block = 4 * 16 * S;
    const i64 grid = (L + rows_per_block - 1) / rows_per_block;
    hipLaunchKernelGGL((k_rotate<NCT, S>), dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, M, ld, work, Yt, ldy, k,
                       col_begin, out, ld_out, L);
    IRLBA_HIP_LAUNCH_CHECK();
}

// The same rotation on the vector FP64 units: a thread per row keeps NK
// accumulators, streams its row of M (CB columns of loads in flight; lanes read
// consecutive rows, so every load is coalesced) and takes Y row by row as
// wave-uniform scalar loads.  It computes exactly k (rounded up to 8) columns
// instead of the 16-column MFMA tiles, and measured 303 us against 426 us for
// the MFMA form on a 1e6 x 70 basis with k = 66 (that shape sits at the FP64
// ridge point, ~8.5 flop per byte).  A thread reads its whole row before it
// writes, so out == M is safe.
template <int NK, int CB>
__global__ __launch_bounds__(256) void k_rotate_rows(const double* M, i64 ld, int work, const double* __restrict__ Yt,
                                                     int ldy, int k, double* out, i64 ld_out, i64 L) {
    const i64 r = static_cast<i64>(blockIdx.x) * 256 + threadIdx.x;
    if (r >= L) return;
    double acc[NK];
#pragma unroll
    for (int j = 0; j < NK; ++j) acc[j] = 0;
    const double* p = M + r;
    int c = 0;
    for (; c + CB <= work; c += CB) {
        double m[CB];
#pragma unroll
        for (int q = 0; q < CB; ++q) m[q] = __builtin_nontemporal_load(p + static_cast<i64>(c + q) * ld);
#pragma unroll
        for (int q = 0; q < CB; ++q) {
            const double* y = Yt + static_cast<i64>(c + q) * ldy;
#pragma unroll
            for (int j = 0; j < NK; ++j) acc[j] = __builtin_fma(m[q], y[j], acc[j]);
        }
    }
    for (; c < work; ++c) {
        const double mv = __builtin_nontemporal_load(p + static_cast<i64>(c) * ld);
        const double* y = Yt + static_cast<i64>(c) * ldy;
#pragma unroll
        for (int j = 0; j < NK; ++j) acc[j] = __builtin_fma(mv, y[j], acc[j]);
    }
#pragma unroll
    for (int j = 0; j < NK; ++j)
        if (j < k) __builtin_nontemporal_store(acc[j], out + r + static_cast<i64>(j) * ld_out);
}

template <int NK>
static void rotate_rows_launch(const double* M, i64 ld, int work, const double* Yt, int ldy, int k, double* out,
                               i64 ld_out, i64 L, hipStream_t s) {
    const i64 grid = (L + 255) / 256;
    hipLaunchKernelGGL((k_rotate_rows<NK, 5>), dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, M, ld, work, Yt,
                       ldy, k, out, ld_out, L);
    IRLBA_HIP_LAUNCH_CHECK();
}

// Bases at least this long with k <= 72 take the vector-FMA rotation (enough
// threads to fill the GPU); shorter ones and wider k take the MFMA tiles.
constexpr i64 kRotateRowsMin = 65536;

void launch_rotate(const double* M, i64 ld, int work, const double* Yt, int ldy, int k, double* out, i64 ld_out,
                   i64 L, hipStream_t s) {
    if (k <= 0 || L <= 0) return;
    if (L >= kRotateRowsMin && k <= 72 && (k + 7) / 8 * 8 <= ldy) {
        // Y columns between k and the accumulator count (the next multiple of
        // 8, or of 2 from 65 to 68, where the restarts of a 70-vector subspace
        // sit) are read (ldy covers them) and their sums discarded.
        if (k > 64 && k <= 66) {
            rotate_rows_launch<66>(M, ld, work, Yt, ldy, k, out, ld_out, L, s);
            return;
        }
        if (k > 66 && k <= 68) {
            rotate_rows_launch<68>(M, ld, work, Yt, ldy, k, out, ld_out, L, s);
            return;
        }
        switch ((k + 7) / 8) {
            case 1: rotate_rows_launch<8>(M, ld, work, Yt, ldy, k, out, ld_out, L, s); break;
            case 2: rotate_rows_launch<16>(M, ld, work, Yt, ldy, k, out, ld_out, L, s); break;
            case 3: rotate_rows_launch<24>(M, ld, work, Yt, ldy, k, out, ld_out, L, s); break;
            case 4: rotate_rows_launch<32>(M, ld, work, Yt, ldy, k, out, ld_out, L, s); break;
            case 5: rotate_rows_launch<40>(M, ld, work, Yt, ldy, k, out, ld_out, L, s); break;
            case 6: rotate_rows_launch<48>(M, ld, work, Yt, ldy, k, out, ld_out, L, s); break;
            case 7: rotate_rows_launch<56>(M, ld, work, Yt, ldy, k, out, ld_out, L, s); break;
            case 8: rotate_rows_launch<64>(M, ld, work, Yt, ldy, k, out, ld_out, L, s); break;
            default: rotate_rows_launch<72>(M, ld, work, Yt, ldy, k, out, ld_out, L, s); break;
        }
        return;
    }
    if (out == M && k > 128) throw std::runtime_error("in-place rotation covers at most 128 columns per pass");
    for (int cb = 0; cb < k; cb += 128) {
        const int nct = std::min(8, (k - cb + 15) / 16);
        switch (nct) {
            case 1: rotate_launch<1>(M, ld, work, Yt, ldy, k, cb, out, ld_out, L, s); break;
            case 2: rotate_launch<2>(M, ld, work, Yt, ldy, k, cb, out, ld_out, L, s); break;
            case 3: rotate_launch<3>(M, ld, work, Yt, ldy, k, cb, out, ld_out, L, s); break;
            case 4: rotate_launch<4>(M, ld, work, Yt, ldy, k, cb, out, ld_out, L, s); break;
            case 5: rotate_launch<5>(M, ld, work, Yt, ldy, k, cb, out, ld_out, L, s); break;
            case 6: rotate_launch<6>(M, ld, work, Yt, ldy, k, cb, out, ld_out, L, s); break;
            case 7: rotate_launch<7>(M, ld, work, Yt, ldy, k, cb, out, ld_out, L, s); break;
            default: rotate_launch<8>(M, ld, work, Yt, ldy, k, cb, out, ld_out, L, s); break;
        }
    }
}

}  // namespace irlba_hip
