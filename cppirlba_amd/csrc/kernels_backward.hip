// Kernels of the differentiable top-k SVD (no reference counterpart).
//
// The adjoint of L = sum_i w_i s_i + <G_U, U> + <G_V, V> with respect to the
// matrix is low rank, dL/dA = X Y' (X rows x R, Y cols x R, R = 2k).  For a
// sparse matrix only the stored entries need it: out[e] = X[row_e, :] . Y[col_e, :]
// (an SDDMM).  Both factors are first packed row-major (row stride Rp, zero
// padded, in the precision of the output), so that one stored entry needs one
// contiguous Rp-element run of the other factor: LPE lanes of a wavefront
// take one entry, each loading one 16-byte piece of the run, and reduce with
// shuffles.  A wavefront walks one line (column of a CSC) with the line's own
// factor row held in registers.
//
// The small dense products of the backward (U'G, the centre sums) are
// tall-skinny Gram products A'B with both operands L x k, k <= 64: blocks over
// ~64-row ranges stage 32-row tiles of both in LDS and write k x k partials
// that a second kernel sums in a fixed order.
#include "common.hpp"

namespace irlba_hip {

// ------------------------------------------------------------------ packing
// out[row * Rp + r] = X[row + r * ld] for r < R, 0 for R <= r < Rp.
template <typename FT>
__global__ void k_pack_rows(const double* __restrict__ X, i64 rows, i64 ld, int R, int Rp, FT* __restrict__ out) {
    const i64 n = rows * Rp;
    for (i64 e = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<i64>(gridDim.x) * blockDim.x) {
        const i64 row = e / Rp;
        const int r = static_cast<int>(e - row * Rp);
        out[e] = r < R ? static_cast<FT>(X[row + static_cast<i64>(r) * ld]) : FT(0);
    }
}

void launch_pack_rows(const double* X, i64 rows, i64 ld, int R, int Rp, bool f32, void* out, hipStream_t s) {
    if (rows <= 0 || Rp <= 0) return;
    const i64 grid = std::max<i64>(1, std::min<i64>((rows * Rp + 255) / 256, 16384));
    if (f32)
        hipLaunchKernelGGL(k_pack_rows<float>, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, X, rows, ld, R, Rp,
                           static_cast<float*>(out));
    else
        hipLaunchKernelGGL(k_pack_rows<double>, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, X, rows, ld, R, Rp,
                           static_cast<double*>(out));
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ SDDMM
template <typename FT>
struct Piece;  // one 16-byte load
template <>
struct Piece<float> {
    typedef f32x4 type;
    static constexpr int n = 4;
};
template <>
struct Piece<double> {
    typedef f64x2 type;
    static constexpr int n = 2;
};

// The compressed layout is checked as it is read (the checks of matrix
// creation): a line whose bounds are not monotone within [0, ptr[n_lines]], or
// ptr[0] != 0, sets bit 0 of *flag and is skipped; an index outside
// [0, n_other) sets bit 1 and its output is 0.
template <typename FT, int LPE>
__global__ __launch_bounds__(256) void k_sddmm(const std::int32_t* __restrict__ idx, const i64* __restrict__ ptr,
                                               i64 n_lines, i64 n_other, const FT* __restrict__ line_rm,
                                               const FT* __restrict__ entry_rm, int Rp, FT* __restrict__ out,
                                               int* __restrict__ flag) {
    typedef typename Piece<FT>::type PT;
    constexpr int VN = Piece<FT>::n;
    constexpr int G = kWave / LPE;  // entries per wavefront step
    constexpr int CH = LPE >= 16 ? 1 : 16 / LPE;
    constexpr int NB = CH * LPE < 4 ? CH * LPE : 4;
    const int lane = threadIdx.x & 63, q = lane % LPE, g = lane / LPE;
    const i64 wave = (static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const i64 nwaves = (static_cast<i64>(gridDim.x) * blockDim.x) >> 6;
    const i64 nnz = ptr[n_lines];
    if (wave == 0 && lane == 0 && ptr[0] != 0) atomicOr(flag, 1);
    for (i64 c = wave; c < n_lines; c += nwaves) {
        const PT yl = *reinterpret_cast<const PT*>(line_rm + c * Rp + q * VN);
        const i64 e0 = ptr[c], e1 = ptr[c + 1];
        if (!(0 <= e0 && e0 <= e1 && e1 <= nnz)) {  // uniform across the wavefront
            if (lane == 0) atomicOr(flag, 1);
            continue;
        }
        // Passes of CH chunks of 64 entries: each lane loads one entry's index
        // per chunk (coalesced, all chunks at once), the indices reach their
        // groups by shuffles, and the gathers follow in batches of NB, so the
        // HBM latency of the indices is paid once per pass, not once per G
        // entries.  Every lane runs the same trip count, so the group shuffles
        // stay converged.  Each entry's dot and group sum are formed as in a
        // walk of G entries at a time.
        for (i64 ps = e0; ps < e1; ps += 64 * CH) {
            int r[CH];
#pragma unroll
            for (int ch = 0; ch < CH; ++ch) {
                const i64 e = ps + 64 * ch + lane;
                r[ch] = e < e1 ? idx[e] : 0;
                if (e < e1 && !(r[ch] >= 0 && r[ch] < n_other)) atomicOr(flag, 2);
            }
#pragma unroll
            for (int n0 = 0; n0 < CH * LPE; n0 += NB) {
                double acc[NB];
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    const int ch = (n0 + i) / LPE, j = g + G * ((n0 + i) % LPE);
                    const i64 e = ps + 64 * ch + j;
                    const int row = __shfl(r[ch], j, kWave);
                    const bool ok = e < e1 && row >= 0 && row < n_other;
                    // A normal (cached) load: each factor row is shared by every entry of its row.
                    // Absent entries read a row that exists (the first of either factor).
                    const FT* src = ok ? entry_rm + static_cast<i64>(row) * Rp : (n_other > 0 ? entry_rm : line_rm);
                    const PT x = *reinterpret_cast<const PT*>(src + q * VN);
                    double a = 0;
#pragma unroll
                    for (int v = 0; v < VN; ++v) a += static_cast<double>(x[v]) * static_cast<double>(yl[v]);
                    acc[i] = ok ? a : 0.0;
                }
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    const double a = group_sum<LPE>(acc[i]);
                    const i64 e = ps + 64 * ((n0 + i) / LPE) + g + G * ((n0 + i) % LPE);
                    if (q == 0 && e < e1) out[e] = static_cast<FT>(a);
                }
            }
        }
    }
}

template <typename FT, int LPE>
static void sddmm_launch(const std::int32_t* idx, const i64* ptr, i64 n_lines, i64 n_other, const void* line_rm,
                         const void* entry_rm, int Rp, void* out, int* flag, hipStream_t s) {
    const i64 grid = std::max<i64>(1, std::min<i64>((n_lines + 3) / 4, 1 << 20));
    hipLaunchKernelGGL((k_sddmm<FT, LPE>), dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, idx, ptr, n_lines,
                       n_other, static_cast<const FT*>(line_rm), static_cast<const FT*>(entry_rm), Rp,
                       static_cast<FT*>(out), flag);
    IRLBA_HIP_LAUNCH_CHECK();
}

int sddmm_padded_rank(int R, bool f32, int* lpe_out) {
    const int vn = f32 ? 4 : 2;
    int lpe = 1;
    while (lpe * vn < R) lpe *= 2;
    if (lpe > 64) throw std::runtime_error("SDDMM rank too large");
    if (lpe_out) *lpe_out = lpe;
    return lpe * vn;
}

void launch_sddmm(const std::int32_t* idx, const i64* ptr, i64 n_lines, i64 n_other, const void* line_rm,
                  const void* entry_rm, int R, bool f32, void* out, int* flag, hipStream_t s) {
    if (n_lines <= 0) return;
    int lpe = 1;
    const int Rp = sddmm_padded_rank(R, f32, &lpe);
    if (f32) {
        switch (lpe) {
            case 1: sddmm_launch<float, 1>(idx, ptr, n_lines, n_other, line_rm, entry_rm, Rp, out, flag, s); break;
            case 2: sddmm_launch<float, 2>(idx, ptr, n_lines, n_other, line_rm, entry_rm, Rp, out, flag, s); break;
            case 4: sddmm_launch<float, 4>(idx, ptr, n_lines, n_other, line_rm, entry_rm, Rp, out, flag, s); break;
            case 8: sddmm_launch<float, 8>(idx, ptr, n_lines, n_other, line_rm, entry_rm, Rp, out, flag, s); break;
            case 16: sddmm_launch<float, 16>(idx, ptr, n_lines, n_other, line_rm, entry_rm, Rp, out, flag, s); break;
            case 32: sddmm_launch<float, 32>(idx, ptr, n_lines, n_other, line_rm, entry_rm, Rp, out, flag, s); break;
            default: sddmm_launch<float, 64>(idx, ptr, n_lines, n_other, line_rm, entry_rm, Rp, out, flag, s); break;
        }
    } else {
        switch (lpe) {
            case 1: sddmm_launch<double, 1>(idx, ptr, n_lines, n_other, line_rm, entry_rm, Rp, out, flag, s); break;
            case 2: sddmm_launch<double, 2>(idx, ptr, n_lines, n_other, line_rm, entry_rm, Rp, out, flag, s); break;
            case 4: sddmm_launch<double, 4>(idx, ptr, n_lines, n_other, line_rm, entry_rm, Rp, out, flag, s); break;
            case 8: sddmm_launch<double, 8>(idx, ptr, n_lines, n_other, line_rm, entry_rm, Rp, out, flag, s); break;
            case 16: sddmm_launch<double, 16>(idx, ptr, n_lines, n_other, line_rm, entry_rm, Rp, out, flag, s); break;
            case 32: sddmm_launch<double, 32>(idx, ptr, n_lines, n_other, line_rm, entry_rm, Rp, out, flag, s); break;
            default: sddmm_launch<double, 64>(idx, ptr, n_lines, n_other, line_rm, entry_rm, Rp, out, flag, s); break;
        }
    }
}

// ------------------------------------------------------------------ dense gradient
// out = X Y' written as a dense rows x cols matrix in the matrix's own type,
// row-major (out[i * cols + j]): the dense counterpart of the SDDMM, so that a
// dense matrix's gradient needs no FP64 intermediate and no separate
// conversion.  64 x 64 output tiles, rank slices of 32 staged in LDS, 4 x 4
// FP64 sums per thread (rows ty + 16u, columns tx + 16v, so a store writes 16
// consecutive columns of 4 rows); every sum runs over r in increasing order.
template <typename OT>
__device__ __forceinline__ OT from_double(double v);
template <>
__device__ __forceinline__ double from_double<double>(double v) {
    return v;
}
template <>
__device__ __forceinline__ float from_double<float>(double v) {
    return static_cast<float>(v);
}
template <>
__device__ __forceinline__ __hip_bfloat16 from_double<__hip_bfloat16>(double v) {
    return __float2bfloat16(static_cast<float>(v));  // through FP32, as torch converts
}

constexpr int kGradTile = 64, kGradSlice = 32;

template <typename OT>
__global__ __launch_bounds__(256) void k_dense_gradient(const double* __restrict__ X, i64 ldx,
                                                        const double* __restrict__ Y, i64 ldy, int R, i64 rows,
                                                        i64 cols, i64 tiles_j, OT* __restrict__ out) {
    __shared__ double xs[kGradSlice][kGradTile];
    __shared__ double ys[kGradSlice][kGradTile];
    const i64 b = blockIdx.x;
    const i64 i0 = (b / tiles_j) * kGradTile, j0 = (b % tiles_j) * kGradTile;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0;
    for (int r0 = 0; r0 < R; r0 += kGradSlice) {
        const int rc = min(kGradSlice, R - r0);
        // Factor columns are contiguous: consecutive threads read consecutive rows.
        for (int e = threadIdx.x; e < kGradSlice * kGradTile; e += 256) {
            const int r = e / kGradTile, c = e % kGradTile;
            const bool live = r < rc;
            xs[r][c] = live && i0 + c < rows ? X[i0 + c + static_cast<i64>(r0 + r) * ldx] : 0.0;
            ys[r][c] = live && j0 + c < cols ? Y[j0 + c + static_cast<i64>(r0 + r) * ldy] : 0.0;
        }
        __syncthreads();
        for (int r = 0; r < rc; ++r) {
            double a[4], bb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = xs[r][ty + 16 * u];
                bb[u] = ys[r][tx + 16 * u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = __builtin_fma(a[u], bb[v], acc[u][v]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const i64 i = i0 + ty + 16 * u;
        if (i >= rows) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const i64 j = j0 + tx + 16 * v;
            if (j < cols) out[i * cols + j] = from_double<OT>(acc[u][v]);
        }
    }
}

void launch_dense_gradient(const double* X, i64 ldx, const double* Y, i64 ldy, int R, i64 rows, i64 cols,
                           StorageType out_t, bool column_major, void* out, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return;
    // A column-major rows x cols output is the row-major output of Y X'.
    if (column_major) {
        std::swap(X, Y);
        std::swap(ldx, ldy);
        std::swap(rows, cols);
    }
    const i64 tiles_j = (cols + kGradTile - 1) / kGradTile;
    const i64 grid = ((rows + kGradTile - 1) / kGradTile) * tiles_j;
    if (grid > 0x7fffffff) throw std::runtime_error("dense gradient too large for one launch");
    const dim3 g(static_cast<unsigned>(grid)), blk(256);
    switch (out_t) {
        case ST_F64:
            hipLaunchKernelGGL(k_dense_gradient<double>, g, blk, 0, s, X, ldx, Y, ldy, R, rows, cols, tiles_j,
                               static_cast<double*>(out));
            break;
        case ST_F32:
            hipLaunchKernelGGL(k_dense_gradient<float>, g, blk, 0, s, X, ldx, Y, ldy, R, rows, cols, tiles_j,
                               static_cast<float*>(out));
            break;
        default:
            hipLaunchKernelGGL(k_dense_gradient<__hip_bfloat16>, g, blk, 0, s, X, ldx, Y, ldy, R, rows, cols,
                               tiles_j, static_cast<__hip_bfloat16*>(out));
            break;
    }
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ Gram products A'B
constexpr int kGramRows = 32;
constexpr int kGramMaxK = 64;
constexpr int kGramPad = kGramRows + 1;  // LDS row stride (breaks bank aliasing)

// About 64 rows (two LDS tiles) per block, at most 1024 blocks.
int gram_blocks(i64 L) { return static_cast<int>(std::max<i64>(1, std::min<i64>(1024, (L + 63) / 64))); }

__global__ __launch_bounds__(256) void k_gram(const double* __restrict__ A, i64 lda, const double* __restrict__ B,
                                              i64 ldb, int k, i64 L, i64 chunk, double* __restrict__ partials) {
    __shared__ double As[kGramMaxK * kGramPad];
    __shared__ double Bs[kGramMaxK * kGramPad];
    const int t = threadIdx.x;
    const int kk = k * k;
    double acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0;
    const i64 r0 = static_cast<i64>(blockIdx.x) * chunk, r1 = std::min<i64>(L, r0 + chunk);
    for (i64 base = r0; base < r1; base += kGramRows) {
        const int n = static_cast<int>(std::min<i64>(kGramRows, r1 - base));
        // Columns are contiguous in HBM: consecutive threads take consecutive rows.
        for (int e = t; e < kGramRows * k; e += 256) {
            const int r = e % kGramRows, c = e / kGramRows;
            As[c * kGramPad + r] = r < n ? A[base + r + static_cast<i64>(c) * lda] : 0.0;
            Bs[c * kGramPad + r] = r < n ? B[base + r + static_cast<i64>(c) * ldb] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int p = t + 256 * j;
            if (p < kk) {
                const int i = p / k, jj = p - i * k;
                const double* a = As + i * kGramPad;
                const double* b = Bs + jj * kGramPad;
                double s = 0;
                for (int r = 0; r < n; ++r) s += a[r] * b[r];
                acc[j] += s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int p = t + 256 * j;
        if (p < kk) partials[static_cast<i64>(blockIdx.x) * kk + p] = acc[j];
    }
}

// out[i * ld_out + j] = sum over blocks of partials[b][i * k + j] = (A'B)(i, j):
// row-major, the layout of the rotation coefficients.  One wavefront per output,
// lanes over the blocks in a fixed order.
__global__ __launch_bounds__(256) void k_gram_reduce(const double* __restrict__ partials, int nb, int k,
                                                     double* __restrict__ out, int ld_out) {
    const int kk = k * k;
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= kk) return;  // whole wavefronts only
    double s = 0;
    for (int b = lane; b < nb; b += kWave) s += partials[static_cast<i64>(b) * kk + p];
    s = wave_sum(s);
    if (lane == 0) {
        const int i = p / k, j = p - i * k;
        out[static_cast<i64>(i) * ld_out + j] = s;
    }
}

void launch_gram(const double* A, i64 lda, const double* B, i64 ldb, int k, i64 L, double* partials, double* out,
                 int ld_out, hipStream_t s) {
    if (k <= 0) return;
    if (k > kGramMaxK) throw std::runtime_error("Gram product wider than 64 columns");
    const int nb = gram_blocks(L);
    const i64 chunk = (L + nb - 1) / nb;
    hipLaunchKernelGGL(k_gram, dim3(nb), dim3(256), 0, s, A, lda, B, ldb, k, L, chunk, partials);
    IRLBA_HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gram_reduce, dim3((k * k + 3) / 4), dim3(256), 0, s, partials, nb, k, out, ld_out);
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ column-batched vector kernels
// out[i] = A[:, i] . B[:, i] for i < k: blocks (b, i) over row chunks, then a
// fixed-order sum of the gram_blocks(L) partials of each column.
__global__ __launch_bounds__(256) void k_coldots(const double* __restrict__ A, i64 lda, const double* __restrict__ B,
                                                 i64 ldb, i64 L, i64 chunk, double* __restrict__ partials) {
    __shared__ double red[8];
    const int i = blockIdx.y;
    const i64 r0 = static_cast<i64>(blockIdx.x) * chunk, r1 = std::min<i64>(L, r0 + chunk);
    const double* a = A + static_cast<i64>(i) * lda;
    const double* b = B + static_cast<i64>(i) * ldb;
    double s = 0, unused = 0;
    for (i64 r = r0 + threadIdx.x; r < r1; r += blockDim.x) s += a[r] * b[r];
    block_sum2(s, unused, red);
    if (threadIdx.x == 0) partials[static_cast<i64>(i) * gridDim.x + blockIdx.x] = s;
}

__global__ void k_coldots_reduce(const double* __restrict__ partials, int nb, int k, double* __restrict__ out) {
    __shared__ double red[8];
    const int i = blockIdx.x;
    double s = 0, unused = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += partials[static_cast<i64>(i) * nb + b];
    block_sum2(s, unused, red);
    if (threadIdx.x == 0) out[i] = s;
}

void launch_coldots(const double* A, i64 lda, const double* B, i64 ldb, int k, i64 L, double* partials, double* out,
                    hipStream_t s) {
    if (k <= 0) return;
    const int nb = gram_blocks(L);
    const i64 chunk = (L + nb - 1) / nb;
    hipLaunchKernelGGL(k_coldots, dim3(nb, k), dim3(256), 0, s, A, lda, B, ldb, L, chunk, partials);
    IRLBA_HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_coldots_reduce, dim3(k), dim3(256), 0, s, partials, nb, k, out);
    IRLBA_HIP_LAUNCH_CHECK();
}

// out[:, i] = out[:, i] * (a_scale * a_coef[i]) + x[:, i] * (b_scale * b_coef[i])
// for i < k (a null coefficient array reads as ones).  With a_coef null and
// a_scale 0 the old contents of out are not read (out may be uninitialised).
__global__ void k_cols_axpby(double* __restrict__ out, i64 ldo, const double* __restrict__ a_coef, double a_scale,
                             const double* __restrict__ x, i64 ldx, const double* __restrict__ b_coef,
                             double b_scale, int k, i64 L) {
    const i64 n = L * k;
    const bool overwrite = a_coef == nullptr && a_scale == 0.0;
    for (i64 e = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<i64>(gridDim.x) * blockDim.x) {
        const int i = static_cast<int>(e / L);
        const i64 l = e - static_cast<i64>(i) * L;
        double& o = out[l + static_cast<i64>(i) * ldo];
        const double bt = x[l + static_cast<i64>(i) * ldx] * (b_coef ? b_coef[i] * b_scale : b_scale);
        o = overwrite ? bt : o * (a_coef ? a_coef[i] * a_scale : a_scale) + bt;
    }
}

void launch_cols_axpby(double* out, i64 ldo, const double* a_coef, double a_scale, const double* x, i64 ldx,
                       const double* b_coef, double b_scale, int k, i64 L, hipStream_t s) {
    if (L <= 0 || k <= 0) return;
    const i64 grid = std::max<i64>(1, std::min<i64>((L * k + 255) / 256, 8192));
    hipLaunchKernelGGL(k_cols_axpby, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, out, ldo, a_coef, a_scale, x,
                       ldx, b_coef, b_scale, k, L);
    IRLBA_HIP_LAUNCH_CHECK();
}

// The k x k coefficients of the backward from J = U'G_U, K = V'G_V (rotation
// layout, J(i, j) at i * ldy + j), the singular values s and their weights w:
//   M_ij = w_i (i == j),  F_ij (J_ij - J_ji) s_j + s_i F_ij (K_ij - K_ji) else,
//   F_ij = 1 / (s_j^2 - s_i^2) (0 for repeated values: the term is dropped);
//   CX = M - J S^-1 and CY = -K S^-1 with the rank-k complement (mode 1), else
//   CX = M, CY = 0; d = 1 / s, sig2 = s^2.  Columns k..ldy-1 are zeroed.  Sets
//   bit 0 of *flag unless s is positive and non-increasing.  One block.
__global__ void k_backward_coefs(const double* __restrict__ J, const double* __restrict__ K, int ldy,
                                 const double* __restrict__ sv, const double* __restrict__ w, int k, int complement,
                                 double* __restrict__ CX, double* __restrict__ CY, double* __restrict__ d,
                                 double* __restrict__ sig2, int* __restrict__ flag) {
    for (int e = threadIdx.x; e < k * ldy; e += blockDim.x) {
        const int i = e / ldy, j = e - i * ldy;
        double cx = 0, cy = 0;
        if (j < k) {
            const double si = sv[i], sj = sv[j];
            double mij;
            if (i == j) {
                mij = w[i];
            } else {
                const double den = sj * sj - si * si;
                const double F = den != 0 ? 1.0 / den : 0.0;
                const double ju = J[i * ldy + j] - J[j * ldy + i], kv = K[i * ldy + j] - K[j * ldy + i];
                mij = F * ju * sj + si * F * kv;
            }
            cx = mij - (complement == 1 ? J[i * ldy + j] / sj : 0.0);
            cy = complement == 1 ? -K[i * ldy + j] / sj : 0.0;
        }
        CX[e] = cx;
        CY[e] = cy;
    }
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        const double si = sv[i];
        d[i] = 1.0 / si;
        sig2[i] = si * si;
        if (!(si > 0) || (i > 0 && !(si <= sv[i - 1]))) atomicOr(flag, 1);
    }
}

void launch_backward_coefs(const double* J, const double* K, int ldy, const double* sv, const double* w, int k,
                           int complement, double* CX, double* CY, double* d, double* sig2, int* flag,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_backward_coefs, dim3(1), dim3(256), 0, s, J, K, ldy, sv, w, k, complement, CX, CY, d, sig2,
                       flag);
    IRLBA_HIP_LAUNCH_CHECK();
}

// The scalar recurrences of the lock-step CG, one thread per system:
//   mode 0: out = a / b (0 unless b > 0)          (alpha = rr / p'Ap)
//   mode 1: out = a / b (0 unless b > 0); b = a    (beta = rr_new / rr; rr = rr_new)
//   mode 2: out = sqrt(a);  mode 3: out = 1 / a.
__global__ void k_cg_coefs(const double* a, double* b, double* out, int k, int mode) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    const double av = a[i];
    if (mode <= 1) {
        const double bv = b[i];
        out[i] = bv > 0 ? av / bv : 0.0;
        if (mode == 1) b[i] = av;
    } else {
        out[i] = mode == 2 ? sqrt(av) : 1.0 / av;
    }
}

void launch_cg_coefs(const double* a, double* b, double* out, int k, int mode, hipStream_t s) {
    if (k <= 0) return;
    hipLaunchKernelGGL(k_cg_coefs, dim3((k + 63) / 64), dim3(64), 0, s, a, b, out, k, mode);
    IRLBA_HIP_LAUNCH_CHECK();
}

// The CG operator's last step (the starting residual's; the iterations fuse
// it with the update, below):
//   ap[:, i] = (basis (basis' ap))[:, i] - ap[:, i] + sig2[i] in[:, i]
__global__ void k_cg_op_finish(double* __restrict__ ap, const double* __restrict__ uc, const double* __restrict__ in,
                               const double* __restrict__ sig2, i64 L, int k) {
    const i64 n = L * k;
    for (i64 e = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<i64>(gridDim.x) * blockDim.x) {
        const int i = static_cast<int>(e / L);
        const double t = ap[e] * -1.0 + uc[e] * 1.0;
        ap[e] = t * 1.0 + in[e] * (sig2[i] * 1.0);
    }
}

static i64 cols_grid(i64 L, int k) { return std::max<i64>(1, std::min<i64>((L * k + 255) / 256, 8192)); }

void launch_cg_op_finish(double* ap, const double* uc, const double* in, const double* sig2, i64 L, int k,
                         hipStream_t s) {
    if (L <= 0 || k <= 0) return;
    hipLaunchKernelGGL(k_cg_op_finish, dim3(static_cast<unsigned>(cols_grid(L, k))), dim3(256), 0, s, ap, uc, in, sig2,
                       L, k);
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ fused lock-step CG
// The CG's vector work after the Gram product proj = basis' (B p) and the
// rotation uc = basis proj in three kernels instead of six; each kernel's
// blocks sum the previous kernel's per-block partials themselves, in the same
// fixed order in every block, so no separate reduction kernels run.  All three
// use one grid: blocks (c, i) over row chunks c of the columns i.
//   k_cgf_finish_dot: ap = uc - ap + sig2 p (= sig2 p - P B p); partials of p'ap
//   k_cgf_xr:         alpha = rr / p'ap (0 unless p'ap > 0); x += alpha p;
//                     r -= alpha ap; partials of r'r
//   k_cgf_p:          beta = rr_new / rr (0 unless rr > 0); p = beta p + r;
//                     the first chunk's blocks store rr_new
// The element arithmetic is that of k_cg_op_finish and of the separate
// update kernels these replaced; the sums run over other partitions.
constexpr int kCgfMaxChunks = 64;  // row chunks per column
int cgf_chunks(i64 L) { return static_cast<int>(std::max<i64>(1, std::min<i64>(kCgfMaxChunks, (L + 1023) / 1024))); }

// The sum over b < nb of part[b] for one column: thread t adds b = t, t + 256,
// ..., then the block sum (wave sums, then the waves in order).
__device__ __forceinline__ double cgf_column_total(const double* __restrict__ part, int nb, double* red) {
    double v = 0, unused = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) v += part[b];
    block_sum2(v, unused, red);
    return v;
}

__global__ __launch_bounds__(256) void k_cgf_finish_dot(double* __restrict__ ap, const double* __restrict__ uc,
                                                        const double* __restrict__ p, const double* __restrict__ sig2,
                                                        i64 L, i64 chunk, double* __restrict__ part) {
    __shared__ double red[2 * 16];
    const int i = blockIdx.y;
    const double sg = sig2[i];
    const i64 r0 = static_cast<i64>(blockIdx.x) * chunk, r1 = std::min<i64>(L, r0 + chunk);
    const i64 off = static_cast<i64>(i) * L;
    double s = 0, unused = 0;
    for (i64 l = r0 + threadIdx.x; l < r1; l += blockDim.x) {
        const double t = ap[off + l] * -1.0 + uc[off + l] * 1.0;
        const double pe = p[off + l];
        const double a = t * 1.0 + pe * (sg * 1.0);
        ap[off + l] = a;
        s += pe * a;
    }
    block_sum2(s, unused, red);
    if (threadIdx.x == 0) part[static_cast<i64>(i) * gridDim.x + blockIdx.x] = s;
}

// Block (c, i): row chunk c of column i.
__global__ __launch_bounds__(256) void k_cgf_xr(double* __restrict__ x, double* __restrict__ r,
                                                const double* __restrict__ p, const double* __restrict__ ap,
                                                const double* __restrict__ rr, const double* __restrict__ part_in,
                                                i64 L, i64 chunk, double* __restrict__ part_out) {
    __shared__ double red[2 * 16];
    const int i = blockIdx.y;
    const double pv = cgf_column_total(part_in + static_cast<i64>(i) * gridDim.x, gridDim.x, red);
    const double alpha = pv > 0 ? rr[i] / pv : 0.0;
    const i64 r0 = static_cast<i64>(blockIdx.x) * chunk, r1 = std::min<i64>(L, r0 + chunk);
    const i64 off = static_cast<i64>(i) * L;
    double s = 0, unused = 0;
    for (i64 l = r0 + threadIdx.x; l < r1; l += blockDim.x) {
        x[off + l] = x[off + l] * 1.0 + p[off + l] * (alpha * 1.0);
        const double rv = r[off + l] * 1.0 + ap[off + l] * (alpha * -1.0);
        r[off + l] = rv;
        s += rv * rv;
    }
    block_sum2(s, unused, red);
    if (threadIdx.x == 0) part_out[static_cast<i64>(i) * gridDim.x + blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_cgf_p(double* __restrict__ p, const double* __restrict__ r,
                                               const double* __restrict__ part_in, const double* __restrict__ rr,
                                               double* __restrict__ rr_new, i64 L, i64 chunk) {
    __shared__ double red[2 * 16];
    const int i = blockIdx.y;
    const double rn = cgf_column_total(part_in + static_cast<i64>(i) * gridDim.x, gridDim.x, red);
    if (blockIdx.x == 0 && threadIdx.x == 0) rr_new[i] = rn;
    const double rv = rr[i];
    const double beta = rv > 0 ? rn / rv : 0.0;
    const i64 r0 = static_cast<i64>(blockIdx.x) * chunk, r1 = std::min<i64>(L, r0 + chunk);
    const i64 off = static_cast<i64>(i) * L;
    for (i64 l = r0 + threadIdx.x; l < r1; l += blockDim.x) p[off + l] = p[off + l] * (beta * 1.0) + r[off + l] * 1.0;
}

void launch_cg_fused(double* ap, const double* uc, double* p, double* x, double* r, const double* sig2,
                     const double* rr, double* rr_new, i64 L, int k, double* partials, hipStream_t s) {
    if (L <= 0 || k <= 0) return;
    const int nc = cgf_chunks(L);
    const i64 chunk = (L + nc - 1) / nc;
    double* part_a = partials;                             // p'ap per (column, chunk)
    double* part_b = partials + static_cast<i64>(nc) * k;  // r'r per (column, chunk)
    const dim3 grid(nc, k);
    hipLaunchKernelGGL(k_cgf_finish_dot, grid, dim3(256), 0, s, ap, uc, p, sig2, L, chunk, part_a);
    IRLBA_HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cgf_xr, grid, dim3(256), 0, s, x, r, p, ap, rr, part_a, L, chunk, part_b);
    IRLBA_HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cgf_p, grid, dim3(256), 0, s, p, r, part_b, rr, rr_new, L, chunk);
    IRLBA_HIP_LAUNCH_CHECK();
}

// out[i] = a[i] * alpha + b[i] * beta  (beta ignored when b is null); the CG and
// the adjoint assembly use it for their vector updates.
__global__ void k_axpby(double* __restrict__ out, const double* __restrict__ a, double alpha,
                        const double* __restrict__ b, double beta, i64 n) {
    for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * blockDim.x)
        out[i] = a[i] * alpha + (b ? b[i] * beta : 0.0);
}

void launch_axpby(double* out, const double* a, double alpha, const double* b, double beta, i64 n, hipStream_t s) {
    if (n <= 0) return;
    const i64 grid = std::max<i64>(1, std::min<i64>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(k_axpby, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, out, a, alpha, b, beta, n);
    IRLBA_HIP_LAUNCH_CHECK();
}

// out[i] = v for i < n.
__global__ void k_fill(double* __restrict__ out, double v, i64 n) {
    for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * blockDim.x)
        out[i] = v;
}

void launch_fill(double* out, double v, i64 n, hipStream_t s) {
    if (n <= 0) return;
    const i64 grid = std::max<i64>(1, std::min<i64>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(k_fill, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, out, v, n);
    IRLBA_HIP_LAUNCH_CHECK();
}

}  // namespace irlba_hip
