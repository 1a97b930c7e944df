// FP64 Lanczos-vector kernels for gfx950.
//
//  * Classical Gram-Schmidt (include/irlba/lanczos.hpp:17-21) as a tall-skinny
//    GEMV pair: t = M[:, :n]' f with per-block partials and a fixed-order
//    reduction (deterministic), then f -= M[:, :n] t fused with the per-block
//    sums of squares of the updated f (the next Lanczos norm).
//  * Norm finalisation + normalisation (lanczos.hpp:81-85, 99-109, 123-132).
//  * The restart rotation V[:, :k] <- V Y (compute.hpp:268-280, 302-305) on the
//    FP64 matrix cores (v_mfma_f64_16x16x4f64), in place per 16-row strip.
#include "common.hpp"

namespace irlba_hip {

int cgs_blocks(i64 L) {
    return static_cast<int>(std::max<i64>(1, std::min<i64>(1024, (L + 255) / 256)));
}

static inline i64 cgs_chunk(i64 L, int G) { return (L + G - 1) / G; }

// ------------------------------------------------------------------ t = M' f (partials)
constexpr int kDotBatch = 8;

__global__ __launch_bounds__(256) void k_cgs_dots(const double* __restrict__ M, i64 ld, int n,
                                                  const double* __restrict__ f, i64 L, i64 chunk,
                                                  double* __restrict__ partials) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const i64 r0 = static_cast<i64>(blockIdx.x) * chunk;
    const i64 r1 = min(L, r0 + chunk);
    // wave w owns columns w, w+4, w+8, ...; batches of kDotBatch columns share each f load.
    for (int q0 = 0; w + 4 * q0 < n; q0 += kDotBatch) {
        double acc[kDotBatch];
#pragma unroll
        for (int q = 0; q < kDotBatch; ++q) acc[q] = 0;
        for (i64 r = r0 + lane; r < r1; r += 64) {
            const double fr = f[r];
#pragma unroll
            for (int q = 0; q < kDotBatch; ++q) {
                const int i = w + 4 * (q0 + q);
                if (i < n) acc[q] += M[r + i * ld] * fr;
            }
        }
#pragma unroll
        for (int q = 0; q < kDotBatch; ++q) {
            const int i = w + 4 * (q0 + q);
            const double v = wave_sum(acc[q]);
            if (lane == 0 && i < n) partials[static_cast<i64>(blockIdx.x) * n + i] = v;
        }
    }
}

void launch_cgs_dots(const double* M, i64 ld, int n, const double* f, i64 L, double* partials, hipStream_t s) {
    if (n <= 0) return;
    const int G = cgs_blocks(L);
    hipLaunchKernelGGL(k_cgs_dots, dim3(G), dim3(256), 0, s, M, ld, n, f, L, cgs_chunk(L, G), partials);
    IRLBA_HIP_LAUNCH_CHECK();
}

__global__ void k_cgs_reduce(const double* __restrict__ partials, int G, int n, double* __restrict__ t) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double acc = 0;
        for (int b = 0; b < G; ++b) acc += partials[static_cast<i64>(b) * n + i];
        t[i] = acc;
    }
}

void launch_cgs_reduce(const double* partials, int G, int n, double* t, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cgs_reduce, dim3(1), dim3(256), 0, s, partials, G, n, t);
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ f -= M t, plus sum of squares
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
    for (i64 r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
        int i = 0;
        for (; i + 3 < n; i += 4) {
            a0 += M[r + i * ld] * ts[i];
            a1 += M[r + (i + 1) * ld] * ts[i + 1];
            a2 += M[r + (i + 2) * ld] * ts[i + 2];
            a3 += M[r + (i + 3) * ld] * ts[i + 3];
        }
        for (; i < n; ++i) a0 += M[r + i * ld] * ts[i];
        const double v = f[r] - ((a0 + a1) + (a2 + a3));
        f[r] = v;
        sq += v * v;
    }
    double dummy = 0;
    block_sum2(sq, dummy, red);
    if (threadIdx.x == 0) norm_part[blockIdx.x] = sq;
}

void launch_cgs_update(const double* M, i64 ld, int n, const double* t, double* f, i64 L, double* norm_part,
                       hipStream_t s) {
    const int G = cgs_blocks(L);
    const size_t lds = sizeof(double) * static_cast<size_t>(std::max(n, 1));
    if (lds > 60000) throw std::runtime_error("working subspace too large for the CGS update kernel");
    hipLaunchKernelGGL(k_cgs_update, dim3(G), dim3(256), lds, s, M, ld, n, t, f, L, cgs_chunk(L, G), norm_part);
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
// of 16, zero padded).  Each wave computes a 16-row strip for up to
// kRotTiles*16 columns with v_mfma_f64_16x16x4f64:
//   A operand: lane l holds M[row0 + l%16, kk + l/16]
//   B operand: lane l holds Y[kk + l/16, col0 + l%16]
//   D:         lane l, element i is out[row0 + l/16 + 4*i, col0 + l%16]  (checked on gfx950)
// All reads of a strip precede its writes, so out == M is safe when one pass
// covers every output column (k <= kRotTiles*16).
typedef double double4_t __attribute__((ext_vector_type(4)));
constexpr int kRotTiles = 8;

__global__ __launch_bounds__(256) void k_rotate(const double* M, i64 ld, int work,
                                                const double* __restrict__ Yt, int ldy, int k, int ct_begin,
                                                double* out, i64 ld_out, i64 L) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const i64 row0 = (static_cast<i64>(blockIdx.x) * 4 + w) * 16;
    if (row0 >= L) return;
    const int nct = min(kRotTiles, (k + 15) / 16 - ct_begin);
    double4_t acc[kRotTiles];
#pragma unroll
    for (int c = 0; c < kRotTiles; ++c) acc[c] = double4_t{0, 0, 0, 0};
    const int lr = lane & 15, lk = lane >> 4;
    const i64 arow = row0 + lr;
    for (int kk = 0; kk < work; kk += 4) {
        const int kidx = kk + lk;
        const double a = (arow < L && kidx < work) ? M[arow + static_cast<i64>(kidx) * ld] : 0.0;
#pragma unroll
        for (int c = 0; c < kRotTiles; ++c) {
            if (c < nct) {
                const int col = (ct_begin + c) * 16 + lr;
                const double b = kidx < work ? Yt[static_cast<i64>(kidx) * ldy + col] : 0.0;
                acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < kRotTiles; ++c) {
        if (c < nct) {
            const int col = (ct_begin + c) * 16 + lr;
            if (col < k) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const i64 row = row0 + lk + 4 * i;
                    if (row < L) out[row + static_cast<i64>(col) * ld_out] = acc[c][i];
                }
            }
        }
    }
}

void launch_rotate(const double* M, i64 ld, int work, const double* Yt, int ldy, int k, double* out, i64 ld_out,
                   i64 L, hipStream_t s) {
    if (k <= 0 || L <= 0) return;
    const int tiles = (k + 15) / 16;
    if (out == M && tiles > kRotTiles) throw std::runtime_error("in-place rotation limited to 128 columns");
    const i64 grid = (L + 63) / 64;
    for (int ct = 0; ct < tiles; ct += kRotTiles) {
        hipLaunchKernelGGL(k_rotate, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, M, ld, work, Yt, ldy, k, ct,
                           out, ld_out, L);
        IRLBA_HIP_LAUNCH_CHECK();
    }
}

}  // namespace irlba_hip
