// Kernels of the differentiable top-k SVD (no reference counterpart).
//
// The adjoint of L = sum_i w_i s_i + <G_U, U> + <G_V, V> with respect to the
// matrix is low rank, dL/dA = X Y' (X rows x R, Y cols x R, R = 2k).  For a
// sparse matrix only the stored entries need it: out[e] = X[row_e, :] . Y[col_e, :]
// (an SDDMM).  One wavefront per line of the compressed pattern: the line's
// factor row is staged once in LDS, each lane takes entries of the line and
// reads the matching row of the other factor, stored row-major so that it is
// one contiguous R*8-byte run.
#include "common.hpp"

namespace irlba_hip {

constexpr int kSddmmMaxR = 128;

// X_rm[row * R + r] = X[row + r * ld]
__global__ void k_to_row_major(const double* __restrict__ X, i64 rows, i64 ld, int R, double* __restrict__ out) {
    const i64 n = rows * R;
    for (i64 e = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<i64>(gridDim.x) * blockDim.x) {
        const i64 row = e / R;
        const int r = static_cast<int>(e - row * R);
        out[e] = X[row + static_cast<i64>(r) * ld];
    }
}

void launch_to_row_major(const double* X, i64 rows, i64 ld, int R, double* out, hipStream_t s) {
    if (rows <= 0 || R <= 0) return;
    const i64 grid = std::max<i64>(1, std::min<i64>((rows * R + 255) / 256, 16384));
    hipLaunchKernelGGL(k_to_row_major, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, X, rows, ld, R, out);
    IRLBA_HIP_LAUNCH_CHECK();
}

// Lines are the primary dimension of the pattern (columns of a CSC); `line_f`
// is the factor indexed by the line (column-major, ld = n_lines), `entry_f`
// the factor indexed by the stored index, row-major (R per row).
__global__ __launch_bounds__(256) void k_sddmm(const std::int32_t* __restrict__ idx, const i64* __restrict__ ptr,
                                               i64 n_lines, const double* __restrict__ line_f,
                                               const double* __restrict__ entry_f_rm, int R,
                                               double* __restrict__ out) {
    __shared__ double yl[4][kSddmmMaxR];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const i64 wave = static_cast<i64>(blockIdx.x) * 4 + w;
    const i64 nwaves = static_cast<i64>(gridDim.x) * 4;
    for (i64 c = wave; c < n_lines; c += nwaves) {
        for (int r = lane; r < R; r += 64) yl[w][r] = line_f[c + static_cast<i64>(r) * n_lines];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the staged row is visible to the wave
        const i64 e0 = ptr[c], e1 = ptr[c + 1];
        for (i64 e = e0 + lane; e < e1; e += 64) {
            const double* xr = entry_f_rm + static_cast<i64>(idx[e]) * R;
            double a0 = 0, a1 = 0;
            int r = 0;
            for (; r + 1 < R; r += 2) {
                a0 += xr[r] * yl[w][r];
                a1 += xr[r + 1] * yl[w][r + 1];
            }
            if (r < R) a0 += xr[r] * yl[w][r];
            out[e] = a0 + a1;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_s_waitcnt(0xc07f);
    }
}

void launch_sddmm(const std::int32_t* idx, const i64* ptr, i64 n_lines, const double* line_f,
                  const double* entry_f_rm, int R, double* out, hipStream_t s) {
    if (n_lines <= 0) return;
    if (R > kSddmmMaxR) throw std::runtime_error("SDDMM rank above 128");
    const i64 grid = std::max<i64>(1, std::min<i64>((n_lines + 3) / 4, 65536));
    hipLaunchKernelGGL(k_sddmm, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, idx, ptr, n_lines, line_f,
                       entry_f_rm, R, out);
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

}  // namespace irlba_hip
