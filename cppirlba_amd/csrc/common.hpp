// Shared definitions for the MI355X (gfx950) IRLBA engine: error handling,
// wave64 reductions and the launcher declarations that the host engine
// (engine.cpp) calls.  Every kernel is written for 64-wide wavefronts.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace irlba_hip {

typedef std::int64_t i64;

#define IRLBA_HIP_CHECK(expr)                                                                         \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw std::runtime_error(std::string("HIP error in " #expr ": ") + hipGetErrorString(e_)); \
    } while (0)

#define IRLBA_HIP_LAUNCH_CHECK() IRLBA_HIP_CHECK(hipGetLastError())

enum StorageType { ST_F64 = 0, ST_F32 = 1, ST_BF16 = 2 };

constexpr int kWave = 64;

// Native clang vectors (the nontemporal load builtins take these, not HIP's struct vectors).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ device helpers
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum within aligned groups of `Width` lanes (Width a power of two <= 64).
template <int Width, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int off = Width / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Inclusive prefix sum within aligned groups of `Width` lanes.
template <int Width, typename T>
__device__ __forceinline__ T group_scan_incl(T v) {
    const int l = threadIdx.x & (Width - 1);
#pragma unroll
    for (int off = 1; off < Width; off <<= 1) {
        const T u = __shfl_up(v, off, Width);
        if (l >= off) v += u;
    }
    return v;
}

// The same scan for 32-bit integers with DPP lane moves (no LDS traffic): row
// shifts by 1, 2, 4, 8 inside 16-lane rows (a select keeps groups narrower
// than a row apart), then the row broadcasts of lane 15 and lane 31.
template <int Width>
__device__ __forceinline__ unsigned group_scan_incl_dpp(unsigned v) {
    static_assert(Width >= 2 && Width <= 64 && (Width & (Width - 1)) == 0, "group width");
    const int l = threadIdx.x & 15;
    int x = static_cast<int>(v);
    int u;
    u = __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);  // row_shr:1
    x += (Width >= 16 || (l & (Width - 1)) >= 1) ? u : 0;
    if (Width >= 4) {
        u = __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);  // row_shr:2
        x += (Width >= 16 || (l & (Width - 1)) >= 2) ? u : 0;
    }
    if (Width >= 8) {
        u = __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);  // row_shr:4
        x += (Width >= 16 || (l & (Width - 1)) >= 4) ? u : 0;
    }
    if (Width >= 16) {
        u = __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);  // row_shr:8
        x += u;
    }
    if (Width >= 32) {
        u = __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
        x += u;
    }
    if (Width >= 64) {
        u = __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
        x += u;
    }
    return static_cast<unsigned>(x);
}

// Value of the last lane of each aligned group of `Width` lanes.
template <int Width>
__device__ __forceinline__ unsigned group_last(unsigned v) {
    if constexpr (Width == 64) return static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
    else return static_cast<unsigned>(__shfl(static_cast<int>(v), Width - 1, Width));
}

// Block-wide sum of up to two doubles; every thread gets the totals.
// `scratch` needs 2 * (blockDim.x / 64) doubles of LDS.
__device__ __forceinline__ void block_sum2(double& a, double& b, double* scratch) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) {
        scratch[wid] = a;
        scratch[nw + wid] = b;
    }
    __syncthreads();
    double sa = 0, sb = 0;
    for (int w = 0; w < nw; ++w) {
        sa += scratch[w];
        sb += scratch[nw + w];
    }
    a = sa;
    b = sb;
    __syncthreads();
}

__device__ __forceinline__ float to_float(float v) { return v; }
__device__ __forceinline__ float to_float(__hip_bfloat16 v) { return __bfloat162float(v); }
__device__ __forceinline__ double to_double_any(double v) { return v; }
__device__ __forceinline__ double to_double_any(float v) { return v; }
__device__ __forceinline__ double to_double_any(__hip_bfloat16 v) { return __bfloat162float(v); }

// ------------------------------------------------------------------ fused affine terms
// A device operator is  M = D_r A D_c - sum_k a_k b_k'  (the closed form of any
// chain of CenteredMatrix / ScaledMatrix decorators).  For one product the
// "minor" side is the input and the "major" side the output.
constexpr int kMaxTerms = 2;

struct Prologue {                  // applied to the input vector while it is staged
    const double* x = nullptr;     // input (minor length)
    const double* xscale = nullptr;  // D on the input side (nullable)
    int n_terms = 0;
    const double* tdot[kMaxTerms] = {nullptr, nullptr};  // gamma_k = tdot_k . x (nullptr = ones)
};

struct Epilogue {                  // out[m] = oscale[m]*acc - sum_k tvec_k[m]*gamma_k - (*coef)*src[m]
    double* out = nullptr;
    const double* oscale = nullptr;
    int n_terms = 0;
    const double* tvec[kMaxTerms] = {nullptr, nullptr};  // nullptr = ones
    const double* axpy_src = nullptr;
    const double* axpy_coef = nullptr;  // device scalar
};

// The per-output operands of the epilogue, loaded before the output's
// reduction so their latency overlaps the stream of its nonzeros.
struct EpiOperands {
    double scale = 1.0, t0 = 1.0, t1 = 1.0, src = 0.0;
};

__device__ __forceinline__ EpiOperands load_epilogue(const Epilogue& e, i64 m) {
    EpiOperands o;
    if (e.oscale) o.scale = e.oscale[m];
    if (e.n_terms > 0 && e.tvec[0]) o.t0 = e.tvec[0][m];
    if (e.n_terms > 1 && e.tvec[1]) o.t1 = e.tvec[1][m];
    if (e.axpy_src) o.src = e.axpy_src[m];
    return o;
}

// `coef` is *e.axpy_coef, read once per kernel (0 when there is no axpy).
__device__ __forceinline__ double finish_epilogue(const Epilogue& e, const EpiOperands& p, double acc,
                                                  const double* gamma, double coef) {
    double o = acc * p.scale;
    if (e.n_terms > 0) o -= p.t0 * gamma[0];
    if (e.n_terms > 1) o -= p.t1 * gamma[1];
    if (e.axpy_src) o -= coef * p.src;
    return o;
}

__device__ __forceinline__ double apply_epilogue(const Epilogue& e, i64 m, double acc, const double* gamma) {
    return finish_epilogue(e, load_epilogue(e, m), acc, gamma, e.axpy_src ? *e.axpy_coef : 0.0);
}

// ------------------------------------------------------------------ tiled compressed lines
// One orientation of a sparse matrix: `n_major` output lines, the minor (input)
// dimension split into `n_tiles` tiles of `tile_len` entries so that one tile of
// the input vector fits in LDS.  Entries of tile t, line m live in
// [ptr[t*n_major + m], ptr[t*n_major + m + 1]) with 16-bit in-tile minor
// indices; every segment is padded with zeros to a multiple of `vec`.
struct TiledLines {
    i64 n_major = 0, n_minor = 0;
    int tile_len = 0, n_tiles = 0;
    i64 nnz = 0, slots = 0;     // true entries, stored slots (with padding)
    int vec = 1, lpl = 64;      // vector width per lane, lanes per line
    int B = 1;                  // blocks of one product (one resident wave)
    StorageType st = ST_F64;
    i64* ptr = nullptr;         // device, n_tiles*n_major + 1
    std::uint16_t* idx = nullptr;
    void* val = nullptr;
    i64* chunk = nullptr;       // device, B+1 segment boundaries (nnz quantiles)
    // Delta layout (vec == 4): one byte per slot instead of `idx` (kernels_sparse.hip,
    // "8-bit delta indices"); `idx` is then only filled on demand by launch_delta_expand.
    bool delta = false;
    std::uint8_t* dlt = nullptr;
    std::uint16_t* first = nullptr;  // device, n_tiles*n_major
    // Value dictionary (delta layout only; kernels_sparse.hip, "value
    // dictionary"): {codes, deltas} per run of 4 or 16 slots and the table of values
    // (storage type, n_dict entries).  The products read them when use_dict is
    // set; the value stream stays for the other kernels.
    const void* cd = nullptr;
    const void* dict = nullptr;
    int n_dict = 0;
    int dict_run = 4;                  // slots per code word: 4, or 16 when every segment holds whole runs of 16
    bool use_dict = false;
    int seg_align = 1;                 // every segment's slot count is a multiple of this
    int lpl_values = 0, lpl_dict = 0;  // lanes per line tuned for each stream
};

constexpr int kDictMaxCodes = 256;

// ------------------------------------------------------------------ launchers (kernels_*.hip)
// Sparse build.
void launch_count_entries(const std::int32_t* idx, const i64* ptr, i64 n_primary, bool transpose, int tile_len,
                          i64 n_major_out, unsigned long long* counts, hipStream_t s);
void launch_pad_counts(const unsigned long long* counts, i64* plen, i64 n, int vec, hipStream_t s);
void exclusive_scan_i64(const i64* in, i64* out, i64 n, hipStream_t s);  // out has n+1 entries
void launch_scatter_entries(const std::int32_t* idx, const void* val_in, StorageType in_t, const i64* ptr,
                            i64 n_primary, bool transpose, int tile_len, i64 n_major_out,
                            unsigned long long* cursor, std::uint16_t* out_idx, void* out_val, StorageType out_t,
                            hipStream_t s);
void launch_dict_collect(const void* val, StorageType st, i64 n, unsigned long long* gset, int* gcount, int* flag,
                         hipStream_t s);
void launch_dict_encode(const TiledLines& t, const unsigned long long* pats, int n_codes, void* cd, int* flag,
                        hipStream_t s);
void launch_sort_segments(const i64* ptr, const unsigned long long* counts, i64 n_seg, int tile_len,
                          const std::uint16_t* in_idx, const void* in_val, std::uint16_t* out_idx, void* out_val,
                          StorageType st, int* dup_flag, hipStream_t s);
void launch_chunk_bounds(const i64* ptr, i64 nseg, int B, i64* chunk, hipStream_t s);
// Delta layout: slot counts with the bridging slots, the conversion from the
// sorted 16-bit layout, and the expansion back to 16-bit indices.
void launch_delta_count(const i64* ptrA, const unsigned long long* counts, i64 nseg, const std::uint16_t* idxA,
                        unsigned long long* lenB, hipStream_t s);
void launch_delta_write(const i64* ptrA, const unsigned long long* counts, i64 nseg, const std::uint16_t* idxA,
                        const void* valA, StorageType st, const i64* ptrB, std::uint16_t* first, std::uint8_t* dltB,
                        void* valB, hipStream_t s);
void launch_delta_expand(const TiledLines& t, std::uint16_t* idx, hipStream_t s);
// Sparse products.
void launch_spmv_tiled(const TiledLines& t, const Prologue& pro, const Epilogue& epi, bool acc64, void* partial,
                       double* gamma_part, hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
void launch_validate_compressed(const std::int32_t* idx, const i64* ptr, i64 n_primary, i64 n_secondary, i64 nnz,
                                int* flag, hipStream_t s);
// Realize and column statistics.
void launch_sparse_realize(const TiledLines& t, bool major_is_col, i64 nr, double* dense, hipStream_t s);
void launch_sparse_colstats(const TiledLines& t, i64 nr, double* center_out, double* scale_out, hipStream_t s);

// Dense products.
struct DenseDesc {
    i64 nr = 0, nc = 0;
    bool col_major = true;
    StorageType st = ST_F64;
    const void* a = nullptr;
};
void launch_dense_matvec(const DenseDesc& d, bool adjoint, const Prologue& pro, const Epilogue& epi,
                         void* xs_scratch, double* gamma_scratch, void* partial, i64 partial_cap, hipStream_t s,
                         hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
void launch_dense_realize(const DenseDesc& d, double* dense, hipStream_t s);
void launch_dense_colstats(const DenseDesc& d, double* center_out, double* scale_out, hipStream_t s);
// Dense array conversion between FP64, FP32 and BF16 (round to nearest even).
void launch_convert_values(const void* in, StorageType in_t, void* out, StorageType out_t, i64 n, hipStream_t s);
i64 dense_partial_elems(const DenseDesc& d);

// Lanczos vector kernels (FP64).
int cgs_blocks(i64 L);
void launch_cgs_dots(const double* M, i64 ld, int n, const double* f, i64 L, double* partials, hipStream_t s);
int cgs_dot_blocks(i64 L);
void launch_cgs_reduce(const double* partials, i64 L, int n, double* t, hipStream_t s);
void launch_cgs_update(const double* M, i64 ld, int n, const double* t, double* f, i64 L, double* norm_part,
                       hipStream_t s);
void launch_sumsq(const double* f, i64 L, double* norm_part, hipStream_t s);
void launch_normalize(const double* f, i64 L, const double* norm_part, int G, double* norm_out, double* dst,
                      hipStream_t s);
void launch_rotate(const double* M, i64 ld, int work, const double* Yt, int ldy, int k, double* out, i64 ld_out,
                   i64 L, hipStream_t s);
void launch_scale_copy(const double* src, i64 L, double denom, double* dst, hipStream_t s);
// dst = src / *denom with the denominator read on the device.
void launch_scale_copy_dev(const double* src, i64 L, const double* denom, double* dst, hipStream_t s);
void launch_vec_combine(double* out, const double* a, const double* b, i64 n, int op, hipStream_t s);
void launch_axpy_dev(double* y, const double* src, const double* coef, i64 n, hipStream_t s);
void launch_scale_columns(double* M, i64 ld, i64 L, int k, const double* d, hipStream_t s);

// Differentiable SVD (kernels_backward.hip).
// Row stride of the packed SDDMM factors for rank R (FP32 or FP64 factors).
int sddmm_padded_rank(int R, bool f32, int* lanes_per_entry = nullptr);
void launch_pack_rows(const double* X, i64 rows, i64 ld, int R, int Rp, bool f32, void* out, hipStream_t s);
// out[e] = line_rm[line] . entry_rm[idx[e]] over a compressed pattern, checking
// the layout as it reads it (bit 0 of *flag: pointers, bit 1: indices).
void launch_sddmm(const std::int32_t* idx, const i64* ptr, i64 n_lines, i64 n_other, const void* line_rm,
                  const void* entry_rm, int R, bool f32, void* out, int* flag, hipStream_t s);
// out = X Y' (X rows x R, Y cols x R, column-major FP64) as a dense rows x cols
// matrix of type out_t, row-major or column-major.
void launch_dense_gradient(const double* X, i64 ldx, const double* Y, i64 ldy, int R, i64 rows, i64 cols,
                           StorageType out_t, bool column_major, void* out, hipStream_t s);
int gram_blocks(i64 L);
// out[i * ld_out + j] = A[:, i] . B[:, j] for L x k operands (k <= 64); partials
// holds gram_blocks(L) * k * k doubles.
void launch_gram(const double* A, i64 lda, const double* B, i64 ldb, int k, i64 L, double* partials, double* out,
                 int ld_out, hipStream_t s);
// out[i] = A[:, i] . B[:, i]; partials holds gram_blocks(L) * k doubles.
void launch_coldots(const double* A, i64 lda, const double* B, i64 ldb, int k, i64 L, double* partials, double* out,
                    hipStream_t s);
void launch_cols_axpby(double* out, i64 ldo, const double* a_coef, double a_scale, const double* x, i64 ldx,
                       const double* b_coef, double b_scale, int k, i64 L, hipStream_t s);
void launch_cg_coefs(const double* a, double* b, double* out, int k, int mode, hipStream_t s);
// ap = uc - ap + sig2 in over k columns of length L (column stride L).
void launch_cg_op_finish(double* ap, const double* uc, const double* in, const double* sig2, i64 L, int k,
                         hipStream_t s);
void launch_backward_coefs(const double* J, const double* K, int ldy, const double* sv, const double* w, int k,
                           int complement, double* CX, double* CY, double* d, double* sig2, int* flag,
                           hipStream_t s);
void launch_axpby(double* out, const double* a, double alpha, const double* b, double beta, i64 n, hipStream_t s);
// The lock-step CG's update after the projection terms (kernels_backward.hip):
// ap = uc - ap + sig2 p, alpha = rr / p'ap, x += alpha p, r -= alpha ap,
// rr_new = r'r, p = (rr_new / rr) p + r, for k columns of length L.
// `partials` holds 2 * cgf_chunks(L) * k doubles.
int cgf_chunks(i64 L);
void launch_cg_fused(double* ap, const double* uc, double* p, double* x, double* r, const double* sig2,
                     const double* rr, double* rr_new, i64 L, int k, double* partials, hipStream_t s);

void launch_fill(double* out, double v, i64 n, hipStream_t s);

// Multi-vector products on the tiled layout (kernels_spmm.hip).  X packed
// row-major with stride spmm_padded_k(n k, x_f32) (FP32 when x_f32), Y row-major
// FP64 with the same stride; the n blocks of k column-major vectors sit side
// by side, block b in packed columns [b k, (b + 1) k).  The stage fills every
// packed column (zeros after the blocks), the finish unpacks every block, one
// launch each.
constexpr int kMaxPackBlocks = 4;
struct StageBlocks {
    const double* X[kMaxPackBlocks];
    i64 ld[kMaxPackBlocks];
    int n;
};
struct FinishBlocks {
    double* Y[kMaxPackBlocks];
    i64 ld[kMaxPackBlocks];
    int n;
};
int spmm_padded_k(int k, bool x_f32, int* lanes_per_entry = nullptr);
void launch_spmm_stage(const StageBlocks& in, i64 L, int k, int kp, const double* xscale, bool x_f32, void* out,
                       hipStream_t s);
void launch_spmm_lines(const TiledLines& t, const void* X, int k, bool x_f32, double* Y, hipStream_t s);
// Y_b = oscale .* Y_rm[:, block b] - sum_j tvec_j gam[(2 b + j) k + :] for every block.
void launch_spmm_finish(const double* Yrm, i64 n, int k, int kp, const double* oscale, int n_terms,
                        const double* t0, const double* t1, const double* gam, const FinishBlocks& out,
                        hipStream_t s);
// Dense block products on packed inputs (kp <= 32); `scratch` (FP64) holds the
// per-chunk partials of the axpy orientation, up to dense_block_scratch(d, kp).
i64 dense_block_scratch(const DenseDesc& d, int kp);
void launch_dense_block(const DenseDesc& d, bool adjoint, int kp, bool x_f32, const void* X, double* Y,
                        double* scratch, i64 scratch_cap, hipStream_t s);

// Fixed device scratch of the allocation-free entry point (noalloc.hip): its
// size, its device address on `device` (resolved once per device, with that
// device current) and a kernel that clears an int flag in it.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
void* noalloc_scratch(int device);
void launch_clear_flag(int* flag, hipStream_t s);

// Synthetic data (counter-based, identical to the oracle's generator).
void launch_gen_nb_count(i64 n_minor, i64 n_major, std::uint32_t thr, std::uint64_t seed, i64* counts,
                         hipStream_t s);
void launch_gen_nb_fill(i64 n_minor, i64 n_major, std::uint32_t thr, std::uint64_t seed, const i64* ptr,
                        const double* table, bool f32, void* values, std::int32_t* indices, hipStream_t s);

}  // namespace irlba_hip
