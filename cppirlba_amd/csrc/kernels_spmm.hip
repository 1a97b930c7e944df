// Multi-vector products Y = M X (k right-hand sides) on the tiled sparse layout,
// for the lock-step CG of the SVD backward (and any block method).
//
// The single-vector product stages its input tile in LDS; with k vectors a tile
// of all k would not fit, so this kernel keeps the k inputs packed row-major in
// HBM (row r = k contiguous values, padded to kp) and lets the gathers hit L2 /
// the MALL: for the backward's k = 16 the packed block is a few MB.  The matrix
// is streamed once for all k.  One wavefront per output line walks the line's
// segments in every tile; LPE lanes take one stored entry (each lane a 16-byte
// piece of the k values), so a wavefront has 64 / LPE entries in flight, and
// the partial sums of the entry groups are combined by shuffles at the end of
// the line.  Accumulation is FP64 whatever the storage.
//
// The decorators of the handle (M = D_r A D_c - sum_j a_j b_j') are applied by
// the staging kernel (input scaling, packing) and the finishing kernel (output
// scaling, rank-1 terms with gamma_j = b_j' X computed beforehand).
#include "common.hpp"

namespace irlba_hip {

namespace {

template <typename FT>
struct Piece16;
template <>
struct Piece16<float> {
    typedef f32x4 type;
    static constexpr int n = 4;
};
template <>
struct Piece16<double> {
    typedef f64x2 type;
    static constexpr int n = 2;
};

template <typename VT, typename FT, int LPE>
__global__ __launch_bounds__(256) void k_spmm_lines(const i64* __restrict__ ptr, const std::uint16_t* __restrict__ idx,
                                                    const VT* __restrict__ val, i64 n_major, int n_tiles,
                                                    int tile_len, const FT* __restrict__ X, int kp,
                                                    double* __restrict__ Y) {
    typedef typename Piece16<FT>::type PT;
    constexpr int VN = Piece16<FT>::n;
    constexpr int G = kWave / LPE;
    const int lane = threadIdx.x & 63, q = lane % LPE, g = lane / LPE;
    // A pass covers CH chunks of 64 entries of a segment: each lane loads one
    // entry's value and index per chunk (coalesced, all chunks at once), the
    // entries reach their groups by shuffles, and the gathers follow in
    // batches of NB, so the HBM latency of the values and indices is paid once
    // per pass instead of once per few entries of each group.  Group g takes
    // entries g, g + G, g + 2G, ... of the segment, in that order.
    constexpr int CH = LPE >= 16 ? 1 : 16 / LPE;
    constexpr int NB = CH * LPE < 8 ? CH * LPE : 8;
    const i64 wave = (static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const i64 nwaves = (static_cast<i64>(gridDim.x) * blockDim.x) >> 6;
    for (i64 m = wave; m < n_major; m += nwaves) {
        double acc[VN];
#pragma unroll
        for (int v = 0; v < VN; ++v) acc[v] = 0;
        for (int t = 0; t < n_tiles; ++t) {
            const i64 seg = static_cast<i64>(t) * n_major + m;
            const i64 s0 = ptr[seg], s1 = ptr[seg + 1];
            const FT* __restrict__ Xt = X + static_cast<i64>(t) * tile_len * kp + q * VN;
            for (i64 ps = s0; ps < s1; ps += 64 * CH) {
                VT a[CH];
                int r[CH];
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    const i64 e = ps + 64 * c + lane;
                    a[c] = VT(0);
                    r[c] = 0;  // past the segment: value 0 at row `base`, a valid row
                    if (e < s1) {
                        a[c] = __builtin_nontemporal_load(val + e);
                        r[c] = idx[e];
                    }
                }
                // Entry n = c LPE + u of the pass is entry g + G u of chunk c;
                // batches of NB gathers, each batch's loads issued together.
#pragma unroll
                for (int n0 = 0; n0 < CH * LPE; n0 += NB) {
                    double aj[NB];
                    PT x[NB];
#pragma unroll
                    for (int i = 0; i < NB; ++i) {
                        const int c = (n0 + i) / LPE, j = g + G * ((n0 + i) % LPE);
                        aj[i] = to_double_any(static_cast<VT>(__shfl(a[c], j, kWave)));
                        const int rj = __shfl(r[c], j, kWave);
                        x[i] = *reinterpret_cast<const PT*>(Xt + static_cast<i64>(rj) * kp);
                    }
#pragma unroll
                    for (int i = 0; i < NB; ++i)
#pragma unroll
                        for (int v = 0; v < VN; ++v) acc[v] += aj[i] * static_cast<double>(x[i][v]);
                }
            }
        }
#pragma unroll
        for (int o = LPE; o < kWave; o <<= 1)
#pragma unroll
            for (int v = 0; v < VN; ++v) acc[v] += __shfl_xor(acc[v], o, kWave);
        if (g == 0) {
#pragma unroll
            for (int v = 0; v < VN; ++v) Y[m * kp + q * VN + v] = acc[v];
        }
    }
}

template <typename VT, typename FT, int LPE>
void spmm_launch(const TiledLines& t, const void* X, int kp, double* Y, hipStream_t s) {
    const i64 grid = std::max<i64>(1, std::min<i64>((t.n_major + 3) / 4, 1 << 20));
    hipLaunchKernelGGL((k_spmm_lines<VT, FT, LPE>), dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, t.ptr, t.idx,
                       static_cast<const VT*>(t.val), t.n_major, t.n_tiles, t.tile_len, static_cast<const FT*>(X), kp,
                       Y);
    IRLBA_HIP_LAUNCH_CHECK();
}

template <typename VT, typename FT>
void spmm_dispatch(const TiledLines& t, const void* X, int kp, int lpe, double* Y, hipStream_t s) {
    switch (lpe) {
        case 1: spmm_launch<VT, FT, 1>(t, X, kp, Y, s); break;
        case 2: spmm_launch<VT, FT, 2>(t, X, kp, Y, s); break;
        case 4: spmm_launch<VT, FT, 4>(t, X, kp, Y, s); break;
        case 8: spmm_launch<VT, FT, 8>(t, X, kp, Y, s); break;
        case 16: spmm_launch<VT, FT, 16>(t, X, kp, Y, s); break;
        case 32: spmm_launch<VT, FT, 32>(t, X, kp, Y, s); break;
        default: spmm_launch<VT, FT, 64>(t, X, kp, Y, s); break;
    }
}

// Entry b of a block list, chosen by comparisons (b varies across lanes, and
// the lists are kernel arguments).
template <typename T>
__device__ __forceinline__ T pick_block(const T (&a)[kMaxPackBlocks], int b) {
    static_assert(kMaxPackBlocks == 4, "pick_block covers four blocks");
    return b == 0 ? a[0] : b == 1 ? a[1] : b == 2 ? a[2] : a[3];
}

// X_rm[r * kp + v] = xscale[r] * X_b[r + c * ld_b] for v = b k + c inside the
// blocks, 0 in the padding columns.  Consecutive threads write consecutive
// packed values.
template <typename FT>
__global__ void k_spmm_stage(StageBlocks in, i64 L, int k, int kp, const double* __restrict__ xscale,
                             FT* __restrict__ out) {
    const int kin = in.n * k;
    const i64 n = L * kp;
    for (i64 e = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<i64>(gridDim.x) * blockDim.x) {
        const i64 r = e / kp;
        const int v = static_cast<int>(e - r * kp);
        double x = 0.0;
        if (v < kin) {
            const int b = v / k, c = v - b * k;
            x = pick_block(in.X, b)[r + static_cast<i64>(c) * pick_block(in.ld, b)];
            if (xscale) x *= xscale[r];
        }
        out[e] = static_cast<FT>(x);
    }
}

// The finishing transpose goes through LDS in tiles of `rows` rows: the
// row-major packed values are read a packed row at a time and the column-major
// output written `rows` consecutive rows at a time, both coalesced (the LDS
// row stride of k + 1 spreads a column over the banks).  Tiles hold 64 rows,
// or 32 from k = 64 on, so a tile takes at most ~33 KB of LDS at k = 128.
static int finish_rows(int k) { return k < 64 ? 64 : 32; }

// Y_b[m + c * ld_b] = oscale[m] * Y_rm[m * kp + v] - sum_j tvec_j[m] * gam[(2 b + j) k + c]
// for packed column v = b k + c of block b.
__global__ __launch_bounds__(256) void k_spmm_finish(const double* __restrict__ Yrm, i64 n, int k, int kp,
                                                     const double* __restrict__ oscale, int n_terms,
                                                     const double* __restrict__ t0, const double* __restrict__ t1,
                                                     const double* __restrict__ gam, FinishBlocks out, int rows) {
    extern __shared__ double tile[];
    const int kin = out.n * k;
    const int ts = kin + 1;
    for (i64 m0 = static_cast<i64>(blockIdx.x) * rows; m0 < n; m0 += static_cast<i64>(gridDim.x) * rows) {
        for (int e = threadIdx.x; e < rows * kin; e += blockDim.x) {
            const int rr = e / kin, v = e - rr * kin;
            const i64 m = m0 + rr;
            if (m < n) tile[rr * ts + v] = Yrm[m * kp + v];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < rows * kin; e += blockDim.x) {
            const int rr = e % rows, v = e / rows;
            const i64 m = m0 + rr;
            if (m < n) {
                const int b = v / k, c = v - b * k;
                double o = tile[rr * ts + v];
                if (oscale) o *= oscale[m];
                if (n_terms > 0) o -= (t0 ? t0[m] : 1.0) * gam[2 * b * k + c];
                if (n_terms > 1) o -= (t1 ? t1[m] : 1.0) * gam[2 * b * k + k + c];
                pick_block(out.Y, b)[m + static_cast<i64>(c) * pick_block(out.ld, b)] = o;
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ dense blocks
// Y = A X for a dense A (lines of length Ln, contiguous) and kp packed vectors,
// as a tiled product with FP64 accumulation: output row p, reduction index q,
//   dot orientation (the product along A's lines):  Y[p] = sum_q A[p * Ln + q] X[q]
//   axpy orientation (across A's lines):            Y[p] = sum_q A[q * Ln + p] X[q]
// A block owns TP = 256 output rows and one of NS slices of the reduction; per
// chunk of TQ reduction indices it stages the A tile and the X rows in LDS (as
// FP64; the next chunk is loaded into registers meanwhile), and each thread
// accumulates RPT rows x OPT consecutive outputs (4 x 8 for 32 vectors: 4 + 8
// LDS reads per 32 FMAs).  A thread's rows are r * (TP / RPT) + g, so the rows
// a wavefront reads together are consecutive and fall in distinct banks.  With
// NS > 1 the slices' partial sums land in `scratch` and k_dense_block_fold adds
// them in slice order.
constexpr int kDenseTQ = 16;

template <int KP>
struct DenseTile {
    static constexpr int OPT = KP >= 4 ? KP / 4 : 1;  // outputs per thread and row
    static constexpr int TPR = KP / OPT;              // threads per row group
    static constexpr int RPT = TPR;                   // rows per thread
    static constexpr int TP = 256;                    // output rows per block
    static constexpr int G = TP / RPT;                // row groups
    static_assert(G * TPR == 256, "a block is 256 threads");
};

template <typename VT, typename FT, int KP, bool DOT>
__global__ __launch_bounds__(256) void k_dense_block(const VT* __restrict__ A, i64 Ln, i64 P, i64 Q, i64 slice,
                                                     const FT* __restrict__ X, double* __restrict__ Y) {
    typedef DenseTile<KP> T;
    constexpr int AN = T::TP * kDenseTQ / 256;         // A tile elements per thread
    constexpr int XN = (kDenseTQ * KP + 255) / 256;    // X tile elements per thread
    __shared__ double As[T::TP * (kDenseTQ + 1)];
    __shared__ double Xs[kDenseTQ * (KP + 1)];
    const int g = threadIdx.x / T::TPR, tc = threadIdx.x % T::TPR;
    const i64 p0 = static_cast<i64>(blockIdx.x) * T::TP;
    const i64 q_begin = static_cast<i64>(blockIdx.y) * slice, q_end = min(Q, q_begin + slice);
    // Each load of a chunk is issued before the first use (absent elements
    // read element 0 and become 0), so a chunk waits for at most one memory
    // latency.  DOT: consecutive threads along a line (q); else along p, both
    // contiguous in A.
    double av[AN], xv[XN];
    auto fetch = [&](i64 q0) {
        const int nq = static_cast<int>(min(static_cast<i64>(kDenseTQ), q_end - q0));
#pragma unroll
        for (int u = 0; u < AN; ++u) {
            const int e = threadIdx.x + 256 * u;
            const int pr = DOT ? e / kDenseTQ : e % T::TP, qq = DOT ? e % kDenseTQ : e / T::TP;
            const i64 p = p0 + pr, q = q0 + qq;
            const bool in = p < P && qq < nq;
            const double a = to_double_any(A[in ? (DOT ? p * Ln + q : q * Ln + p) : 0]);
            av[u] = in ? a : 0.0;
        }
#pragma unroll
        for (int u = 0; u < XN; ++u) {
            const int e = threadIdx.x + 256 * u, qq = e / KP;
            const bool in = e < kDenseTQ * KP && qq < nq;
            const double x = static_cast<double>(X[in ? (q0 + qq) * KP + (e - qq * KP) : 0]);
            xv[u] = in ? x : 0.0;
        }
    };
    double acc[T::RPT][T::OPT];
#pragma unroll
    for (int r = 0; r < T::RPT; ++r)
#pragma unroll
        for (int o = 0; o < T::OPT; ++o) acc[r][o] = 0;
    if (q_begin < q_end) fetch(q_begin);
    for (i64 q0 = q_begin; q0 < q_end; q0 += kDenseTQ) {
        __syncthreads();  // the previous chunk's tiles are consumed
#pragma unroll
        for (int u = 0; u < AN; ++u) {
            const int e = threadIdx.x + 256 * u;
            const int pr = DOT ? e / kDenseTQ : e % T::TP, qq = DOT ? e % kDenseTQ : e / T::TP;
            As[pr * (kDenseTQ + 1) + qq] = av[u];
        }
#pragma unroll
        for (int u = 0; u < XN; ++u) {
            const int e = threadIdx.x + 256 * u, qq = e / KP;
            if (e < kDenseTQ * KP) Xs[qq * (KP + 1) + (e - qq * KP)] = xv[u];
        }
        __syncthreads();
        if (q0 + kDenseTQ < q_end) fetch(q0 + kDenseTQ);
        const double* xc = Xs + tc * T::OPT;
        // Absent reduction indices of the last chunk hold zeros in both tiles.
#pragma unroll 4
        for (int qq = 0; qq < kDenseTQ; ++qq) {
            double a[T::RPT], x[T::OPT];
#pragma unroll
            for (int r = 0; r < T::RPT; ++r) a[r] = As[(r * T::G + g) * (kDenseTQ + 1) + qq];
#pragma unroll
            for (int o = 0; o < T::OPT; ++o) x[o] = xc[qq * (KP + 1) + o];
#pragma unroll
            for (int r = 0; r < T::RPT; ++r)
#pragma unroll
                for (int o = 0; o < T::OPT; ++o) acc[r][o] += a[r] * x[o];
        }
    }
#pragma unroll
    for (int r = 0; r < T::RPT; ++r) {
        const i64 p = p0 + r * T::G + g;
        if (p < P) {
            double* out = Y + static_cast<i64>(blockIdx.y) * P * KP + p * KP + tc * T::OPT;
#pragma unroll
            for (int o = 0; o < T::OPT; ++o) out[o] = acc[r][o];
        }
    }
}

__global__ void k_dense_block_fold(const double* __restrict__ P, int chunks, i64 n, double* __restrict__ Y) {
    for (i64 e = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<i64>(gridDim.x) * blockDim.x) {
        double s = 0;
        for (int c = 0; c < chunks; ++c) s += P[static_cast<i64>(c) * n + e];
        Y[e] = s;
    }
}

// Output rows, reduction length and reduction slices of a dense block product.
struct DensePlan {
    i64 P, Q, tiles, slices, slice;
};

template <int KP>
DensePlan dense_plan(const DenseDesc& d, bool adjoint) {
    const i64 n_lines = d.col_major ? d.nc : d.nr;
    const i64 Ln = d.col_major ? d.nr : d.nc;
    const bool dot = d.col_major ? adjoint : !adjoint;
    DensePlan pl;
    pl.P = dot ? n_lines : Ln;
    pl.Q = dot ? Ln : n_lines;
    pl.tiles = std::max<i64>(1, (pl.P + DenseTile<KP>::TP - 1) / DenseTile<KP>::TP);
    if (pl.Q == 0) {  // no reduction: a zero product, written without the kernel
        pl.slices = 1;
        pl.slice = 0;
        return pl;
    }
    // Up to 512 blocks, slices of at least two chunks, and at most 2^20 doubles
    // of slice partials (the fold reads them all back).
    const i64 max_slices = std::max<i64>(1, pl.Q / (2 * kDenseTQ));
    const i64 fold_cap = std::max<i64>(1, (static_cast<i64>(1) << 20) / std::max<i64>(1, pl.P * KP));
    pl.slices = std::max<i64>(1, std::min<i64>(std::min(max_slices, fold_cap), 512 / pl.tiles));
    pl.slice = (pl.Q + pl.slices - 1) / pl.slices;
    pl.slice = (pl.slice + kDenseTQ - 1) / kDenseTQ * kDenseTQ;
    pl.slices = std::max<i64>(1, (pl.Q + pl.slice - 1) / pl.slice);
    return pl;
}

template <typename VT, typename FT, int KP>
void dense_block_typed(const DenseDesc& d, bool adjoint, const void* Xv, double* Y, double* scratch, i64 scratch_cap,
                       hipStream_t s) {
    const VT* A = static_cast<const VT*>(d.a);
    const FT* X = static_cast<const FT*>(Xv);
    const i64 Ln = d.col_major ? d.nr : d.nc;
    const bool dot = d.col_major ? adjoint : !adjoint;
    const DensePlan pl = dense_plan<KP>(d, adjoint);
    if (pl.P == 0) return;
    if (pl.Q == 0) {  // no entries: a zero product
        launch_fill(Y, 0.0, pl.P * KP, s);
        return;
    }
    if (pl.slices > 1 && pl.slices * pl.P * KP > scratch_cap)
        throw std::runtime_error("dense block product scratch too small");
    double* out = pl.slices > 1 ? scratch : Y;
    const dim3 grid(static_cast<unsigned>(pl.tiles), static_cast<unsigned>(pl.slices));
    if (dot)
        hipLaunchKernelGGL((k_dense_block<VT, FT, KP, true>), grid, dim3(256), 0, s, A, Ln, pl.P, pl.Q, pl.slice, X,
                           out);
    else
        hipLaunchKernelGGL((k_dense_block<VT, FT, KP, false>), grid, dim3(256), 0, s, A, Ln, pl.P, pl.Q, pl.slice, X,
                           out);
    IRLBA_HIP_LAUNCH_CHECK();
    if (pl.slices > 1) {
        const i64 n = pl.P * KP;
        hipLaunchKernelGGL(k_dense_block_fold, dim3(static_cast<unsigned>(std::min<i64>((n + 255) / 256, 4096))),
                           dim3(256), 0, s, scratch, static_cast<int>(pl.slices), n, Y);
        IRLBA_HIP_LAUNCH_CHECK();
    }
}

template <typename VT, typename FT>
void dense_block_kp(const DenseDesc& d, bool adjoint, int kp, const void* X, double* Y, double* scratch, i64 cap,
                    hipStream_t s) {
    switch (kp) {
        case 2: dense_block_typed<VT, FT, 2>(d, adjoint, X, Y, scratch, cap, s); break;
        case 4: dense_block_typed<VT, FT, 4>(d, adjoint, X, Y, scratch, cap, s); break;
        case 8: dense_block_typed<VT, FT, 8>(d, adjoint, X, Y, scratch, cap, s); break;
        case 16: dense_block_typed<VT, FT, 16>(d, adjoint, X, Y, scratch, cap, s); break;
        case 32: dense_block_typed<VT, FT, 32>(d, adjoint, X, Y, scratch, cap, s); break;
        default: throw std::runtime_error("dense block products take at most 32 columns per pass");
    }
}

}  // namespace

i64 dense_block_scratch(const DenseDesc& d, int kp) {
    // The larger of the two orientations' slice partials.
    auto need = [&](bool adjoint) -> i64 {
        DensePlan pl;
        switch (kp) {
            case 2: pl = dense_plan<2>(d, adjoint); break;
            case 4: pl = dense_plan<4>(d, adjoint); break;
            case 8: pl = dense_plan<8>(d, adjoint); break;
            case 16: pl = dense_plan<16>(d, adjoint); break;
            default: pl = dense_plan<32>(d, adjoint); break;
        }
        return pl.slices > 1 ? pl.slices * pl.P * kp : 0;
    };
    return std::max<i64>(1, std::max(need(false), need(true)));
}

void launch_dense_block(const DenseDesc& d, bool adjoint, int kp, bool x_f32, const void* X, double* Y,
                        double* scratch, i64 scratch_cap, hipStream_t s) {
    if (d.st == ST_F64) {
        if (x_f32) throw std::runtime_error("FP64 dense storage takes FP64 block inputs");
        dense_block_kp<double, double>(d, adjoint, kp, X, Y, scratch, scratch_cap, s);
    } else if (d.st == ST_F32) {
        if (!x_f32) throw std::runtime_error("FP32 dense storage takes FP32 block inputs");
        dense_block_kp<float, float>(d, adjoint, kp, X, Y, scratch, scratch_cap, s);
    } else {
        if (!x_f32) throw std::runtime_error("BF16 dense storage takes FP32 block inputs");
        dense_block_kp<__hip_bfloat16, float>(d, adjoint, kp, X, Y, scratch, scratch_cap, s);
    }
}

int spmm_padded_k(int k, bool f32, int* lanes_per_entry) {
    const int vn = f32 ? 4 : 2;
    int lpe = 1;
    while (lpe * vn < k) lpe *= 2;
    if (lpe > 64) throw std::runtime_error("multi-vector product wider than the wavefront");
    if (lanes_per_entry) *lanes_per_entry = lpe;
    return lpe * vn;
}

static void check_blocks(int n, int k, int kp) {
    if (n < 1 || n > kMaxPackBlocks || k <= 0 || static_cast<i64>(n) * k > kp)
        throw std::runtime_error("packed block columns out of range");
}

void launch_spmm_stage(const StageBlocks& in, i64 L, int k, int kp, const double* xscale, bool f32, void* out,
                       hipStream_t s) {
    check_blocks(in.n, k, kp);
    if (L <= 0) return;
    const i64 grid = std::max<i64>(1, std::min<i64>((L * kp + 255) / 256, 16384));
    if (f32)
        hipLaunchKernelGGL(k_spmm_stage<float>, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, in, L, k, kp,
                           xscale, static_cast<float*>(out));
    else
        hipLaunchKernelGGL(k_spmm_stage<double>, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, in, L, k, kp,
                           xscale, static_cast<double*>(out));
    IRLBA_HIP_LAUNCH_CHECK();
}

void launch_spmm_lines(const TiledLines& t, const void* X, int k, bool x_f32, double* Y, hipStream_t s) {
    if (t.n_major <= 0) return;
    int lpe = 1;
    const int kp = spmm_padded_k(k, x_f32, &lpe);
    if (t.st == ST_F64 && !x_f32) spmm_dispatch<double, double>(t, X, kp, lpe, Y, s);
    else if (t.st == ST_F32 && x_f32) spmm_dispatch<float, float>(t, X, kp, lpe, Y, s);
    else if (t.st == ST_F32) spmm_dispatch<float, double>(t, X, kp, lpe, Y, s);
    else throw std::runtime_error("multi-vector products need FP64 or FP32 sparse storage");
}

void launch_spmm_finish(const double* Yrm, i64 n, int k, int kp, const double* oscale, int n_terms,
                        const double* t0, const double* t1, const double* gam, const FinishBlocks& out,
                        hipStream_t s) {
    check_blocks(out.n, k, kp);
    if (n <= 0) return;
    const int kin = out.n * k;
    const int rows = finish_rows(kin);
    const i64 grid = std::max<i64>(1, std::min<i64>((n + rows - 1) / rows, 8192));
    const size_t lds = sizeof(double) * rows * (kin + 1);
    hipLaunchKernelGGL(k_spmm_finish, dim3(static_cast<unsigned>(grid)), dim3(256), lds, s, Yrm, n, k, kp, oscale,
                       n_terms, t0, t1, gam, out, rows);
    IRLBA_HIP_LAUNCH_CHECK();
}

}  // namespace irlba_hip
