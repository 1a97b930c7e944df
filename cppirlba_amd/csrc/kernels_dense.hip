// Dense mat-vec kernels for gfx950 (the SimpleMatrix path,
// include/irlba/Matrix/simple.hpp:42-44 and :70-72).
//
// A dense matrix is a set of contiguous "lines" (columns if column-major).
// One product is either
//   * dot-lines : out[m]  = sum_j line_m[j] * xs[j]   (A'u column-major, A v row-major)
//   * axpy-lines: out[i]  = sum_l line_l[i] * xs[l]   (A v column-major, A'u row-major)
// Both stream the matrix once with 16-byte vector loads (2 x FP64, 4 x FP32,
// 8 x BF16), accumulate in FP32 (FP32/BF16 storage) or FP64, and write one
// partial per split; a reduce kernel sums the splits in a fixed order and
// applies the fused scaling/centring/axpy epilogue in FP64.  A prologue kernel
// scales the input vector into `xs` and computes the rank-1 dot products.  An
// unsplit dot product without rank-1 terms applies the epilogue itself (no
// reduce kernel), and an axpy product of an input without scaling or rank-1
// terms reads the FP64 input directly (no prologue kernel).
#include "common.hpp"

namespace irlba_hip {

namespace {

template <typename T>
struct Vec16;  // 16-byte vector of T
template <>
struct Vec16<double> {
    static constexpr int N = 2;
};
template <>
struct Vec16<float> {
    static constexpr int N = 4;
};
template <>
struct Vec16<__hip_bfloat16> {
    static constexpr int N = 8;
};

template <typename VT, typename AT, int N>
__device__ __forceinline__ void load_vec(const VT* p, AT (&out)[N]) {
    if constexpr (N == 1) {
        out[0] = static_cast<AT>(to_double_any(p[0]));
    } else if constexpr (sizeof(VT) == 8) {
        const f64x2 v = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(p));
        out[0] = v.x;
        out[1] = v.y;
    } else if constexpr (sizeof(VT) == 4) {
        const f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(p));
        out[0] = v.x;
        out[1] = v.y;
        out[2] = v.z;
        out[3] = v.w;
    } else {
        const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            out[2 * k] = __uint_as_float(w[k] << 16);
            out[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
        }
    }
}

// ------------------------------------------------------------------ prologue
template <typename AT>
__global__ void k_dense_prologue(Prologue pro, i64 n, AT* __restrict__ xs, double* __restrict__ gpart) {
    __shared__ double red[32];
    double g0 = 0, g1 = 0;
    for (i64 j = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; j < n;
         j += static_cast<i64>(gridDim.x) * blockDim.x) {
        const double xv = pro.x[j];
        xs[j] = static_cast<AT>(pro.xscale ? xv * pro.xscale[j] : xv);
        if (pro.n_terms > 0) g0 += (pro.tdot[0] ? pro.tdot[0][j] : 1.0) * xv;
        if (pro.n_terms > 1) g1 += (pro.tdot[1] ? pro.tdot[1][j] : 1.0) * xv;
    }
    block_sum2(g0, g1, red);
    if (threadIdx.x == 0) {
        gpart[2 * blockIdx.x] = g0;
        gpart[2 * blockIdx.x + 1] = g1;
    }
}

// Input element j as the axpy products use it: the prologue's scaled copy
// `xs`, or, for an input without scaling or rank-1 terms (no prologue), the
// FP64 input converted in registers (the value the prologue would have stored).
template <typename AT>
__device__ __forceinline__ AT input_at(const AT* __restrict__ xs, const double* __restrict__ x64, i64 j) {
    return x64 ? static_cast<AT>(x64[j]) : xs[j];
}

// ------------------------------------------------------------------ dot-lines
// Block: R lines x one segment of SEG elements; 256 threads, N elements each per step.
// With `fused` (one segment per line, no rank-1 terms) the block applies the
// epilogue and writes the finished outputs instead of partials.
template <typename VT, typename AT, int N, int R>
__global__ __launch_bounds__(256) void k_dot_lines(const VT* __restrict__ A, i64 n_lines, i64 Ln, i64 seg_len,
                                                   const AT* __restrict__ xs, AT* __restrict__ partial,
                                                   Epilogue epi, int fused) {
    __shared__ AT red[R][4];
    const i64 m0 = static_cast<i64>(blockIdx.x) * R;
    const i64 j0 = static_cast<i64>(blockIdx.y) * seg_len;
    const i64 j1 = min(Ln, j0 + seg_len);
    AT acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0;
    for (i64 j = j0 + static_cast<i64>(threadIdx.x) * N; j < j1; j += 256 * N) {
        AT x[N];
#pragma unroll
        for (int k = 0; k < N; ++k) x[k] = xs[j + k];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (m0 + r < n_lines) {
                AT a[N];
                load_vec<VT, AT, N>(A + (m0 + r) * Ln + j, a);
#pragma unroll
                for (int k = 0; k < N; ++k) acc[r] += a[k] * x[k];
            }
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const AT v = wave_sum(acc[r]);
        if (lane == 0) red[r][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < R && m0 + threadIdx.x < n_lines) {
        const int r = threadIdx.x;
        const AT sum = red[r][0] + red[r][1] + red[r][2] + red[r][3];
        if (fused) {
            const double gamma[2] = {0, 0};
            epi.out[m0 + r] = apply_epilogue(epi, m0 + r, static_cast<double>(sum), gamma);
        } else {
            partial[static_cast<i64>(blockIdx.y) * n_lines + m0 + r] = sum;
        }
    }
}

// ------------------------------------------------------------------ axpy-lines
// Block: 256 threads x N outputs each; grid.y splits the lines into chunks.
template <typename VT, typename AT, int N>
__global__ __launch_bounds__(256) void k_axpy_lines(const VT* __restrict__ A, i64 n_lines, i64 Ln, i64 chunk,
                                                    const AT* __restrict__ xs, const double* __restrict__ x64,
                                                    AT* __restrict__ partial) {
    const i64 i = (static_cast<i64>(blockIdx.x) * 256 + threadIdx.x) * N;
    const i64 l0 = static_cast<i64>(blockIdx.y) * chunk;
    const i64 l1 = min(n_lines, l0 + chunk);
    if (i >= Ln) return;
    AT acc[N];
#pragma unroll
    for (int k = 0; k < N; ++k) acc[k] = 0;
    if (i + N <= Ln) {
        i64 l = l0;
        for (; l + 3 < l1; l += 4) {
            AT a[4][N];
#pragma unroll
            for (int u = 0; u < 4; ++u) load_vec<VT, AT, N>(A + (l + u) * Ln + i, a[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const AT xv = input_at(xs, x64, l + u);
#pragma unroll
                for (int k = 0; k < N; ++k) acc[k] += a[u][k] * xv;
            }
        }
        for (; l < l1; ++l) {
            AT a[N];
            load_vec<VT, AT, N>(A + l * Ln + i, a);
            const AT xv = input_at(xs, x64, l);
#pragma unroll
            for (int k = 0; k < N; ++k) acc[k] += a[k] * xv;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) partial[static_cast<i64>(blockIdx.y) * Ln + i + k] = acc[k];
    } else {
        for (int k = 0; k < N && i + k < Ln; ++k) {
            AT s = 0;
            for (i64 l = l0; l < l1; ++l)
                s += static_cast<AT>(to_double_any(A[l * Ln + i + k])) * input_at(xs, x64, l);
            partial[static_cast<i64>(blockIdx.y) * Ln + i + k] = s;
        }
    }
}

// ------------------------------------------------------------------ reduce + epilogue
template <typename AT>
__global__ void k_dense_reduce(const AT* __restrict__ partial, int splits, i64 n_out, const double* __restrict__ gpart,
                               int n_gpart, Epilogue epi) {
    // The rank-1 dot products: the first wave of each block sums the prologue's
    // partials (in a fixed order) and shares them through LDS.
    __shared__ double gsh[2];
    double gamma[2] = {0, 0};
    if (epi.n_terms > 0) {
        if (threadIdx.x < kWave) {
            double g0 = 0, g1 = 0;
            for (int b = threadIdx.x; b < n_gpart; b += kWave) {
                g0 += gpart[2 * b];
                g1 += gpart[2 * b + 1];
            }
            g0 = wave_sum(g0);
            g1 = wave_sum(g1);
            if (threadIdx.x == 0) {
                gsh[0] = g0;
                gsh[1] = g1;
            }
        }
        __syncthreads();
        gamma[0] = gsh[0];
        gamma[1] = gsh[1];
    }
    for (i64 m = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; m < n_out;
         m += static_cast<i64>(gridDim.x) * blockDim.x) {
        double acc = 0;
        for (int s = 0; s < splits; ++s) acc += static_cast<double>(partial[static_cast<i64>(s) * n_out + m]);
        epi.out[m] = apply_epilogue(epi, m, acc, gamma);
    }
}

// First stage for many splits: block (x, g) sums splits [g*per, (g+1)*per) of
// 256 consecutive outputs into partial2[g][.], in a fixed order.
template <typename AT>
__global__ __launch_bounds__(256) void k_dense_fold(const AT* __restrict__ partial, int splits, int per, i64 n_out,
                                                    AT* __restrict__ partial2) {
    const i64 m = static_cast<i64>(blockIdx.x) * 256 + threadIdx.x;
    if (m >= n_out) return;
    const int s0 = blockIdx.y * per, s1 = min(splits, s0 + per);
    AT acc[4] = {0, 0, 0, 0};
    int s = s0;
    for (; s + 3 < s1; s += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] += partial[static_cast<i64>(s + u) * n_out + m];
    }
    for (; s < s1; ++s) acc[0] += partial[static_cast<i64>(s) * n_out + m];
    partial2[static_cast<i64>(blockIdx.y) * n_out + m] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

template <typename VT, typename AT>
void dense_matvec_typed(const DenseDesc& d, bool adjoint, const Prologue& pro, const Epilogue& epi, void* xs_scratch,
                        double* gamma_scratch, void* partial_v, i64 partial_cap, hipStream_t s, hipEvent_t ev0,
                        hipEvent_t ev1) {
    const VT* A = static_cast<const VT*>(d.a);
    AT* xs = static_cast<AT*>(xs_scratch);
    AT* partial = static_cast<AT*>(partial_v);
    const i64 n_lines = d.col_major ? d.nc : d.nr;
    const i64 Ln = d.col_major ? d.nr : d.nc;
    // Input length and whether the product is a dot over lines.
    const bool dot = d.col_major ? adjoint : !adjoint;
    const i64 n_in = dot ? Ln : n_lines;
    const i64 n_out = dot ? n_lines : Ln;

    // The axpy orientation of an input without scaling or rank-1 terms needs no
    // prologue: each block reads one input value per line, so it reads the FP64
    // input and converts it in registers.  The dot orientation re-reads its
    // input segment in every block and keeps the prologue's compact copy (read
    // directly in FP64 there, BF16 A.v of 200000 x 2000 took 0.22 ms instead of
    // 0.125).  Either way every input read ends before an output is written.
    const bool direct_x = !dot && pro.xscale == nullptr && pro.n_terms == 0;
    const double* x64 = direct_x ? pro.x : nullptr;
    const int pro_blocks = static_cast<int>(std::max<i64>(1, std::min<i64>((n_in + 255) / 256, 256)));
    if (!direct_x) {
        hipLaunchKernelGGL(k_dense_prologue<AT>, dim3(pro_blocks), dim3(256), 0, s, pro, n_in, xs, gamma_scratch);
        IRLBA_HIP_LAUNCH_CHECK();
    }

    constexpr int NV = Vec16<VT>::N;
    int splits = 1;
    if (n_lines == 0 || Ln == 0) {
        // No entries: the product is zero, and only the epilogue (centring,
        // scaling, Lanczos axpy) acts on it.  A timed product still records
        // its two events.
        if (ev0) IRLBA_HIP_CHECK(hipEventRecord(ev0, s));
        if (ev1) IRLBA_HIP_CHECK(hipEventRecord(ev1, s));
        if (n_out == 0) return;
        splits = 0;
    } else if (dot) {
        constexpr int R = 4;
        const i64 groups = (n_lines + R - 1) / R;
        // Segment the lines so that there are enough blocks to fill the chip.
        i64 nseg = std::max<i64>(1, std::min<i64>(1024 / std::max<i64>(groups, 1), (Ln + 8191) / 8192));
        i64 seg_len = (Ln + nseg - 1) / nseg;
        const bool vec_ok = (Ln % NV) == 0;
        if (vec_ok) seg_len = (seg_len + NV - 1) / NV * NV;
        nseg = (Ln + seg_len - 1) / seg_len;
        if (nseg * n_lines > partial_cap) throw std::runtime_error("dense partial buffer too small");
        // One segment per line and no rank-1 terms: the product kernel finishes
        // its outputs itself (no reduce kernel).
        const int fused = nseg == 1 && epi.n_terms == 0;
        const dim3 g(static_cast<unsigned>(groups), static_cast<unsigned>(nseg));
        if (ev0) IRLBA_HIP_CHECK(hipEventRecord(ev0, s));
        if (vec_ok)
            hipLaunchKernelGGL((k_dot_lines<VT, AT, NV, R>), g, dim3(256), 0, s, A, n_lines, Ln, seg_len, xs, partial,
                               epi, fused);
        else
            hipLaunchKernelGGL((k_dot_lines<VT, AT, 1, R>), g, dim3(256), 0, s, A, n_lines, Ln, seg_len, xs, partial,
                               epi, fused);
        IRLBA_HIP_LAUNCH_CHECK();
        if (ev1) IRLBA_HIP_CHECK(hipEventRecord(ev1, s));
        if (fused) return;
        splits = static_cast<int>(nseg);
    } else {
        const bool vec_ok = (Ln % NV) == 0;
        const int N = vec_ok ? NV : 1;
        const i64 xblocks = (Ln + 256 * N - 1) / (256 * N);
        i64 nchunk = std::max<i64>(1, std::min<i64>(std::max<i64>(1, 2048 / xblocks), (n_lines + 63) / 64));
        const i64 chunk = (n_lines + nchunk - 1) / nchunk;
        nchunk = (n_lines + chunk - 1) / chunk;
        if (nchunk * Ln > partial_cap) throw std::runtime_error("dense partial buffer too small");
        const dim3 g(static_cast<unsigned>(xblocks), static_cast<unsigned>(nchunk));
        if (ev0) IRLBA_HIP_CHECK(hipEventRecord(ev0, s));
        if (vec_ok)
            hipLaunchKernelGGL((k_axpy_lines<VT, AT, NV>), g, dim3(256), 0, s, A, n_lines, Ln, chunk, xs, x64,
                               partial);
        else
            hipLaunchKernelGGL((k_axpy_lines<VT, AT, 1>), g, dim3(256), 0, s, A, n_lines, Ln, chunk, xs, x64,
                               partial);
        IRLBA_HIP_LAUNCH_CHECK();
        if (ev1) IRLBA_HIP_CHECK(hipEventRecord(ev1, s));
        splits = static_cast<int>(nchunk);
    }
    // Many splits over few outputs (a tall row-major A'u): fold them in two
    // stages so that the serial chain per output stays short.  The second
    // stage's buffer is the tail of the partial buffer.
    const AT* red_in = partial;
    constexpr int kFold = 64;
    if (splits > 2 * kFold && (static_cast<i64>(splits) + kFold) * n_out <= partial_cap) {
        const int per = (splits + kFold - 1) / kFold;
        const int groups = (splits + per - 1) / per;
        AT* partial2 = partial + static_cast<i64>(splits) * n_out;
        const dim3 g(static_cast<unsigned>((n_out + 255) / 256), static_cast<unsigned>(groups));
        hipLaunchKernelGGL(k_dense_fold<AT>, g, dim3(256), 0, s, static_cast<const AT*>(partial), splits, per, n_out,
                           partial2);
        IRLBA_HIP_LAUNCH_CHECK();
        red_in = partial2;
        splits = groups;
    }
    const i64 rgrid = std::max<i64>(1, std::min<i64>((n_out + 255) / 256, 8192));
    hipLaunchKernelGGL(k_dense_reduce<AT>, dim3(static_cast<unsigned>(rgrid)), dim3(256), 0, s, red_in, splits, n_out,
                       gamma_scratch, pro_blocks, epi);
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ realize / column statistics
template <typename VT>
__global__ void k_dense_realize(const VT* __restrict__ A, i64 nr, i64 nc, int col_major, double* __restrict__ out) {
    const i64 n = nr * nc;
    for (i64 e = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<i64>(gridDim.x) * blockDim.x) {
        const i64 r = e % nr, c = e / nr;  // column-major output position
        out[e] = to_double_any(col_major ? A[r + c * nr] : A[r * nc + c]);
    }
}

// Column mean / SD as in pca.hpp:85-109 (two passes, all rows).
template <typename VT>
__global__ void k_dense_colstats(const VT* __restrict__ A, i64 nr, i64 nc, int col_major, double* center_out,
                                 double* scale_out) {
    __shared__ double red[32];
    for (i64 c = blockIdx.x; c < nc; c += gridDim.x) {
        double sum = 0, dummy = 0;
        for (i64 r = threadIdx.x; r < nr; r += blockDim.x)
            sum += to_double_any(col_major ? A[r + c * nr] : A[r * nc + c]);
        block_sum2(sum, dummy, red);
        const double mean = center_out ? sum / nr : 0.0;
        if (center_out && threadIdx.x == 0) center_out[c] = mean;
        if (scale_out) {
            double ss = 0;
            dummy = 0;
            for (i64 r = threadIdx.x; r < nr; r += blockDim.x) {
                const double dlt = to_double_any(col_major ? A[r + c * nr] : A[r * nc + c]) - mean;
                ss += dlt * dlt;
            }
            block_sum2(ss, dummy, red);
            if (threadIdx.x == 0) scale_out[c] = ss != 0 ? sqrt(ss / static_cast<double>(nr - 1)) : 1.0;
        }
    }
}

}  // namespace

void launch_dense_matvec(const DenseDesc& d, bool adjoint, const Prologue& pro, const Epilogue& epi, void* xs_scratch,
                         double* gamma_scratch, void* partial, i64 partial_cap, hipStream_t s, hipEvent_t ev0,
                         hipEvent_t ev1) {
    if (d.st == ST_F64)
        dense_matvec_typed<double, double>(d, adjoint, pro, epi, xs_scratch, gamma_scratch, partial, partial_cap, s,
                                           ev0, ev1);
    else if (d.st == ST_F32)
        dense_matvec_typed<float, float>(d, adjoint, pro, epi, xs_scratch, gamma_scratch, partial, partial_cap, s,
                                         ev0, ev1);
    else
        dense_matvec_typed<__hip_bfloat16, float>(d, adjoint, pro, epi, xs_scratch, gamma_scratch, partial,
                                                  partial_cap, s, ev0, ev1);
}

void launch_dense_realize(const DenseDesc& d, double* dense, hipStream_t s) {
    const i64 n = d.nr * d.nc;
    if (n == 0) return;
    const i64 grid = std::max<i64>(1, std::min<i64>((n + 255) / 256, 65536));
    const dim3 g(static_cast<unsigned>(grid)), b(256);
    const int cm = d.col_major ? 1 : 0;
    if (d.st == ST_F64)
        hipLaunchKernelGGL(k_dense_realize<double>, g, b, 0, s, static_cast<const double*>(d.a), d.nr, d.nc, cm, dense);
    else if (d.st == ST_F32)
        hipLaunchKernelGGL(k_dense_realize<float>, g, b, 0, s, static_cast<const float*>(d.a), d.nr, d.nc, cm, dense);
    else
        hipLaunchKernelGGL(k_dense_realize<__hip_bfloat16>, g, b, 0, s, static_cast<const __hip_bfloat16*>(d.a), d.nr,
                           d.nc, cm, dense);
    IRLBA_HIP_LAUNCH_CHECK();
}

void launch_dense_colstats(const DenseDesc& d, double* center_out, double* scale_out, hipStream_t s) {
    if (d.nc == 0) return;
    const i64 grid = std::max<i64>(1, std::min<i64>(d.nc, 65536));
    const dim3 g(static_cast<unsigned>(grid)), b(256);
    const int cm = d.col_major ? 1 : 0;
    if (d.st == ST_F64)
        hipLaunchKernelGGL(k_dense_colstats<double>, g, b, 0, s, static_cast<const double*>(d.a), d.nr, d.nc, cm,
                           center_out, scale_out);
    else if (d.st == ST_F32)
        hipLaunchKernelGGL(k_dense_colstats<float>, g, b, 0, s, static_cast<const float*>(d.a), d.nr, d.nc, cm,
                           center_out, scale_out);
    else
        hipLaunchKernelGGL(k_dense_colstats<__hip_bfloat16>, g, b, 0, s, static_cast<const __hip_bfloat16*>(d.a), d.nr,
                           d.nc, cm, center_out, scale_out);
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ storage conversion
// Element-wise conversion of a dense array between FP64, FP32 and BF16 on the
// device.  FP64 -> BF16 goes through FP32 (both steps round to nearest even);
// the BF16 rounding adds 0x7fff plus the lowest kept bit to the FP32 pattern
// and keeps the top half, and a NaN keeps its sign and top payload bits with
// the quiet bit set.
__device__ __forceinline__ std::uint16_t bf16_bits_rne(float f) {
    const std::uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<std::uint16_t>((u >> 16) | 0x40u);
    return static_cast<std::uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__global__ void k_convert_values(const void* __restrict__ in, int in_t, void* __restrict__ out, int out_t, i64 n) {
    for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * blockDim.x) {
        double v;
        if (in_t == ST_F64) v = static_cast<const double*>(in)[i];
        else if (in_t == ST_F32) v = static_cast<const float*>(in)[i];
        else v = __uint_as_float(static_cast<std::uint32_t>(static_cast<const std::uint16_t*>(in)[i]) << 16);
        if (out_t == ST_F64) static_cast<double*>(out)[i] = v;
        else if (out_t == ST_F32) static_cast<float*>(out)[i] = static_cast<float>(v);
        else static_cast<std::uint16_t*>(out)[i] = bf16_bits_rne(static_cast<float>(v));
    }
}

void launch_convert_values(const void* in, StorageType in_t, void* out, StorageType out_t, i64 n, hipStream_t s) {
    if (n <= 0) return;
    const i64 grid = std::max<i64>(1, std::min<i64>((n + 255) / 256, 65536));
    hipLaunchKernelGGL(k_convert_values, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, in, static_cast<int>(in_t),
                       out, static_cast<int>(out_t), n);
    IRLBA_HIP_LAUNCH_CHECK();
}

i64 dense_partial_elems(const DenseDesc& d) {
    // Upper bound on the partial buffer used by either direction.
    const i64 n_lines = d.col_major ? d.nc : d.nr;
    const i64 Ln = d.col_major ? d.nr : d.nc;
    const i64 dot_bound = n_lines * std::max<i64>(1, (Ln + 8191) / 8192 + 1);
    // axpy: at most min(2048, n_lines / 64) chunks, plus 64 rows for the second
    // stage of the split reduction.
    const i64 axpy_bound = Ln * (std::max<i64>(1, std::min<i64>(2048, (n_lines + 63) / 64)) + 1 + 64);
    return std::max(dot_bound, axpy_bound);
}

}  // namespace irlba_hip
