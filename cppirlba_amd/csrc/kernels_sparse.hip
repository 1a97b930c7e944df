// Sparse kernels for gfx950: device-side construction of the tiled line
// layouts (both orientations of a CSC/CSR matrix) and the tiled SpMV that
// serves both A.v and A'.u.
//
// The reference's ParallelSparseMatrix (include/irlba/Matrix/sparse.hpp:185-267)
// computes CSC A.v as a scatter into thread-private row ranges and CSC A'.u as a
// per-column gather.  Here both directions are a gather over "lines": the
// matrix is kept twice in HBM, once per orientation, each split along its input
// (minor) dimension into tiles that fit LDS.  A workgroup stages one tile of
// the input vector in LDS (fusing the column/row scaling and the rank-1
// centring dot products into that pass), then streams values and 16-bit
// in-tile indices of its nnz-balanced chunk of lines with non-temporal vector
// loads, one group of lanes per line, and writes either the finished output
// (single tile: fused epilogue) or one FP partial per (tile, line).
#include "common.hpp"

namespace irlba_hip {

// ------------------------------------------------------------------ build: count / scatter
__global__ void k_count_entries(const std::int32_t* __restrict__ idx, const i64* __restrict__ ptr, i64 n_primary,
                                int transpose, int tile_len, i64 n_major_out, unsigned long long* counts) {
    const int lane = threadIdx.x & 63;
    const i64 wave = (static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const i64 nwaves = (static_cast<i64>(gridDim.x) * blockDim.x) >> 6;
    for (i64 p = wave; p < n_primary; p += nwaves) {
        const i64 e0 = ptr[p], e1 = ptr[p + 1];
        for (i64 e = e0 + lane; e < e1; e += 64) {
            const i64 s = idx[e];
            const i64 major = transpose ? s : p;
            const i64 minor = transpose ? p : s;
            const i64 t = minor / tile_len;
            atomicAdd(&counts[t * n_major_out + major], 1ULL);
        }
    }
}

void launch_count_entries(const std::int32_t* idx, const i64* ptr, i64 n_primary, bool transpose, int tile_len,
                          i64 n_major_out, unsigned long long* counts, hipStream_t s) {
    const int block = 256;
    const i64 waves_needed = n_primary;
    const i64 grid = std::max<i64>(1, std::min<i64>((waves_needed * 64 + block - 1) / block, 65536));
    hipLaunchKernelGGL(k_count_entries, dim3(static_cast<unsigned>(grid)), dim3(block), 0, s, idx, ptr, n_primary,
                       transpose ? 1 : 0, tile_len, n_major_out, counts);
    IRLBA_HIP_LAUNCH_CHECK();
}

__global__ void k_pad_counts(const unsigned long long* __restrict__ c, i64* __restrict__ plen, i64 n, int vec) {
    for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<i64>(gridDim.x) * blockDim.x) {
        const i64 v = static_cast<i64>(c[i]);
        plen[i] = (v + vec - 1) / vec * vec;
    }
}

void launch_pad_counts(const unsigned long long* counts, i64* plen, i64 n, int vec, hipStream_t s) {
    const i64 grid = std::max<i64>(1, std::min<i64>((n + 255) / 256, 65536));
    hipLaunchKernelGGL(k_pad_counts, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, counts, plen, n, vec);
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ exclusive scan (int64)
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ void k_scan_block_sums(const i64* __restrict__ in, i64 n, i64* __restrict__ sums) {
    __shared__ i64 red[kScanThreads / 64];
    const i64 base = static_cast<i64>(blockIdx.x) * kScanTile;
    i64 acc = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const i64 i = base + k * kScanThreads + threadIdx.x;
        if (i < n) acc += in[i];
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        i64 t = 0;
        for (int w = 0; w < kScanThreads / 64; ++w) t += red[w];
        sums[blockIdx.x] = t;
    }
}

__global__ void k_scan_apply(const i64* __restrict__ in, i64* __restrict__ out, i64 n,
                             const i64* __restrict__ block_offsets) {
    __shared__ i64 thread_tot[kScanThreads];
    const i64 base = static_cast<i64>(blockIdx.x) * kScanTile + static_cast<i64>(threadIdx.x) * kScanItems;
    i64 vals[kScanItems];
    i64 tot = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const i64 i = base + k;
        vals[k] = i < n ? in[i] : 0;
        tot += vals[k];
    }
    thread_tot[threadIdx.x] = tot;
    __syncthreads();
    // Hillis-Steele inclusive scan over the 256 thread totals.
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const i64 v = threadIdx.x >= static_cast<unsigned>(off) ? thread_tot[threadIdx.x - off] : 0;
        __syncthreads();
        thread_tot[threadIdx.x] += v;
        __syncthreads();
    }
    i64 run = block_offsets[blockIdx.x] + (threadIdx.x ? thread_tot[threadIdx.x - 1] : 0);
    for (int k = 0; k < kScanItems; ++k) {
        const i64 i = base + k;
        if (i < n) out[i] = run;
        run += vals[k];
        if (i == n - 1) out[n] = run;
    }
}

void exclusive_scan_i64(const i64* in, i64* out, i64 n, hipStream_t s) {
    if (n <= 0) {
        const i64 zero = 0;
        IRLBA_HIP_CHECK(hipMemcpyAsync(out, &zero, sizeof(i64), hipMemcpyHostToDevice, s));
        IRLBA_HIP_CHECK(hipStreamSynchronize(s));
        return;
    }
    const i64 nb = (n + kScanTile - 1) / kScanTile;
    i64* sums = nullptr;
    i64* offs = nullptr;
    IRLBA_HIP_CHECK(hipMalloc(&sums, sizeof(i64) * nb));
    IRLBA_HIP_CHECK(hipMalloc(&offs, sizeof(i64) * (nb + 1)));
    hipLaunchKernelGGL(k_scan_block_sums, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, s, in, n, sums);
    IRLBA_HIP_LAUNCH_CHECK();
    if (nb == 1) {
        const i64 zero = 0;
        IRLBA_HIP_CHECK(hipMemcpyAsync(offs, &zero, sizeof(i64), hipMemcpyHostToDevice, s));
    } else {
        exclusive_scan_i64(sums, offs, nb, s);
    }
    hipLaunchKernelGGL(k_scan_apply, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, s, in, out, n, offs);
    IRLBA_HIP_LAUNCH_CHECK();
    IRLBA_HIP_CHECK(hipStreamSynchronize(s));
    IRLBA_HIP_CHECK(hipFree(sums));
    IRLBA_HIP_CHECK(hipFree(offs));
}

// ------------------------------------------------------------------ scatter into (tile, line) buckets
template <typename TIn, typename TOut>
__device__ __forceinline__ TOut convert_value(TIn v) {
    return static_cast<TOut>(v);
}
template <>
__device__ __forceinline__ __hip_bfloat16 convert_value<double, __hip_bfloat16>(double v) {
    return __float2bfloat16(static_cast<float>(v));
}
template <>
__device__ __forceinline__ __hip_bfloat16 convert_value<float, __hip_bfloat16>(float v) {
    return __float2bfloat16(v);
}

template <typename TIn, typename TOut>
__global__ void k_scatter_entries(const std::int32_t* __restrict__ idx, const TIn* __restrict__ val,
                                  const i64* __restrict__ ptr, i64 n_primary, int transpose, int tile_len,
                                  i64 n_major_out, unsigned long long* cursor, std::uint16_t* __restrict__ out_idx,
                                  TOut* __restrict__ out_val) {
    const int lane = threadIdx.x & 63;
    const i64 wave = (static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const i64 nwaves = (static_cast<i64>(gridDim.x) * blockDim.x) >> 6;
    for (i64 p = wave; p < n_primary; p += nwaves) {
        const i64 e0 = ptr[p], e1 = ptr[p + 1];
        for (i64 e = e0 + lane; e < e1; e += 64) {
            const i64 sidx = idx[e];
            const i64 major = transpose ? sidx : p;
            const i64 minor = transpose ? p : sidx;
            const i64 t = minor / tile_len;
            const i64 pos = static_cast<i64>(atomicAdd(&cursor[t * n_major_out + major], 1ULL));
            out_idx[pos] = static_cast<std::uint16_t>(minor - t * tile_len);
            out_val[pos] = convert_value<TIn, TOut>(val[e]);
        }
    }
}

void launch_scatter_entries(const std::int32_t* idx, const void* val_in, StorageType in_t, const i64* ptr,
                            i64 n_primary, bool transpose, int tile_len, i64 n_major_out,
                            unsigned long long* cursor, std::uint16_t* out_idx, void* out_val, StorageType out_t,
                            hipStream_t s) {
    const int block = 256;
    const i64 grid = std::max<i64>(1, std::min<i64>((n_primary * 64 + block - 1) / block, 65536));
    const dim3 g(static_cast<unsigned>(grid)), b(block);
    const int tr = transpose ? 1 : 0;
#define IRLBA_SCATTER(TI, TO)                                                                                  \
    hipLaunchKernelGGL((k_scatter_entries<TI, TO>), g, b, 0, s, idx, static_cast<const TI*>(val_in), ptr,      \
                       n_primary, tr, tile_len, n_major_out, cursor, out_idx, static_cast<TO*>(out_val))
    if (in_t == ST_F64 && out_t == ST_F64) IRLBA_SCATTER(double, double);
    else if (in_t == ST_F64 && out_t == ST_F32) IRLBA_SCATTER(double, float);
    else if (in_t == ST_F32 && out_t == ST_F32) IRLBA_SCATTER(float, float);
    else if (in_t == ST_F32 && out_t == ST_F64) IRLBA_SCATTER(float, double);
    else if (in_t == ST_F64 && out_t == ST_BF16) IRLBA_SCATTER(double, __hip_bfloat16);
    else if (in_t == ST_F32 && out_t == ST_BF16) IRLBA_SCATTER(float, __hip_bfloat16);
    else throw std::runtime_error("unsupported sparse value conversion");
#undef IRLBA_SCATTER
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ segment sort (bitmap ranks)
// Entries of one (tile, line) segment have distinct 16-bit keys below tile_len.
// Each key sets one bit of an LDS bitmap; its rank is the popcount of the bits
// below it, so the sorted position needs no comparisons and no global order.
template <typename V>
__global__ __launch_bounds__(256) void k_sort_segments(const i64* __restrict__ ptr,
                                                       const unsigned long long* __restrict__ counts, i64 n_seg,
                                                       int words, const std::uint16_t* __restrict__ in_idx,
                                                       const V* __restrict__ in_val, std::uint16_t* __restrict__ out_idx,
                                                       V* __restrict__ out_val, int* dup_flag) {
    __shared__ unsigned int bitmap[2048];
    __shared__ int prefix[2048];
    __shared__ int thread_tot[256];
    for (i64 seg = blockIdx.x; seg < n_seg; seg += gridDim.x) {
        const i64 start = ptr[seg];
        const int n = static_cast<int>(counts[seg]);
        if (n <= 1) {
            if (n == 1 && threadIdx.x == 0) {
                out_idx[start] = in_idx[start];
                out_val[start] = in_val[start];
            }
            continue;
        }
        for (int w = threadIdx.x; w < words; w += blockDim.x) bitmap[w] = 0u;
        __syncthreads();
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const unsigned key = in_idx[start + e];
            atomicOr(&bitmap[key >> 5], 1u << (key & 31));
        }
        __syncthreads();
        const int per = (words + blockDim.x - 1) / blockDim.x;
        const int w0 = threadIdx.x * per;
        int tot = 0;
        for (int w = w0; w < w0 + per && w < words; ++w) tot += __popc(bitmap[w]);
        thread_tot[threadIdx.x] = tot;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int v = threadIdx.x >= static_cast<unsigned>(off) ? thread_tot[threadIdx.x - off] : 0;
            __syncthreads();
            thread_tot[threadIdx.x] += v;
            __syncthreads();
        }
        int run = threadIdx.x ? thread_tot[threadIdx.x - 1] : 0;
        for (int w = w0; w < w0 + per && w < words; ++w) {
            prefix[w] = run;
            run += __popc(bitmap[w]);
        }
        if (threadIdx.x == 255 && thread_tot[255] != n) atomicExch(dup_flag, 1);
        __syncthreads();
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const unsigned key = in_idx[start + e];
            const unsigned below = bitmap[key >> 5] & ((1u << (key & 31)) - 1u);
            const int rank = prefix[key >> 5] + __popc(below);
            if (rank < n) {
                out_idx[start + rank] = static_cast<std::uint16_t>(key);
                out_val[start + rank] = in_val[start + e];
            }
        }
        __syncthreads();
    }
}

void launch_sort_segments(const i64* ptr, const unsigned long long* counts, i64 n_seg, int tile_len,
                          const std::uint16_t* in_idx, const void* in_val, std::uint16_t* out_idx, void* out_val,
                          StorageType st, int* dup_flag, hipStream_t s) {
    const int words = (tile_len + 31) / 32;
    if (words > 2048) throw std::runtime_error("tile too long for the segment sort");
    const i64 grid = std::max<i64>(1, std::min<i64>(n_seg, 1 << 20));
    const dim3 g(static_cast<unsigned>(grid)), b(256);
    if (st == ST_F64)
        hipLaunchKernelGGL(k_sort_segments<double>, g, b, 0, s, ptr, counts, n_seg, words, in_idx,
                           static_cast<const double*>(in_val), out_idx, static_cast<double*>(out_val), dup_flag);
    else if (st == ST_F32)
        hipLaunchKernelGGL(k_sort_segments<float>, g, b, 0, s, ptr, counts, n_seg, words, in_idx,
                           static_cast<const float*>(in_val), out_idx, static_cast<float*>(out_val), dup_flag);
    else
        hipLaunchKernelGGL(k_sort_segments<__hip_bfloat16>, g, b, 0, s, ptr, counts, n_seg, words, in_idx,
                           static_cast<const __hip_bfloat16*>(in_val), out_idx,
                           static_cast<__hip_bfloat16*>(out_val), dup_flag);
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ nnz-balanced chunks
// The (tile, line) segments are one tile-major sequence whose offsets `ptr`
// increase; block b of a product owns segments [chunk[b], chunk[b+1]), cut at
// the nnz quantiles b*total/B.  B is one resident wave of blocks, so every CU
// gets the same number of nonzeros and no block waits for a second wave.
__global__ void k_chunk_bounds(const i64* __restrict__ ptr, i64 nseg, int B, i64* chunk) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > B) return;
    const i64 total = ptr[nseg];
    const i64 target = (total / B) * b + ((total % B) * b) / B;
    i64 lo = 0, hi = nseg;  // smallest s with ptr[s] >= target
    while (lo < hi) {
        const i64 mid = (lo + hi) / 2;
        if (ptr[mid] < target) lo = mid + 1; else hi = mid;
    }
    chunk[b] = b == B ? nseg : (b == 0 ? 0 : lo);
}

void launch_chunk_bounds(const i64* ptr, i64 nseg, int B, i64* chunk, hipStream_t s) {
    hipLaunchKernelGGL(k_chunk_bounds, dim3((B + 1 + 255) / 256), dim3(256), 0, s, ptr, nseg, B, chunk);
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ 8-bit delta indices
// A sorted segment of in-tile indices i_0 < i_1 < ... is stored as one 16-bit
// first index per segment and one byte per slot: 0 for the first entry, then
// i_j - i_{j-1}.  A gap above 255 is bridged by value-0 slots of delta 255
// placed before the entry ("dummies"); padding slots have delta 0.  The indices
// are recovered by prefix sums over lanes.

// lenB[seg] = entries + dummies of each sorted segment.  One wavefront per segment.
__global__ void k_delta_count(const i64* __restrict__ ptrA, const unsigned long long* __restrict__ counts, i64 nseg,
                              const std::uint16_t* __restrict__ idxA, unsigned long long* __restrict__ lenB) {
    const int lane = threadIdx.x & 63;
    const i64 wave = (static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const i64 nwaves = (static_cast<i64>(gridDim.x) * blockDim.x) >> 6;
    for (i64 seg = wave; seg < nseg; seg += nwaves) {
        const i64 cnt = static_cast<i64>(counts[seg]), e0 = ptrA[seg];
        i64 d = 0;
        for (i64 j = lane + 1; j < cnt; j += 64) d += (static_cast<i64>(idxA[e0 + j]) - idxA[e0 + j - 1] - 1) / 255;
        d = wave_sum(d);
        if (lane == 0) lenB[seg] = static_cast<unsigned long long>(cnt + d);
    }
}

template <typename VT>
__global__ void k_delta_write(const i64* __restrict__ ptrA, const unsigned long long* __restrict__ counts, i64 nseg,
                              const std::uint16_t* __restrict__ idxA, const VT* __restrict__ valA,
                              const i64* __restrict__ ptrB, std::uint16_t* __restrict__ first,
                              std::uint8_t* __restrict__ dltB, VT* __restrict__ valB) {
    const int lane = threadIdx.x & 63;
    const i64 wave = (static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const i64 nwaves = (static_cast<i64>(gridDim.x) * blockDim.x) >> 6;
    for (i64 seg = wave; seg < nseg; seg += nwaves) {
        const i64 cnt = static_cast<i64>(counts[seg]), e0 = ptrA[seg], o0 = ptrB[seg];
        if (lane == 0) first[seg] = cnt ? idxA[e0] : 0;
        i64 run = 0;  // output slots written so far (uniform over the wave)
        for (i64 jb = 0; jb < cnt; jb += 64) {
            const i64 j = jb + lane;
            const bool active = j < cnt;
            const i64 gap = (active && j > 0) ? static_cast<i64>(idxA[e0 + j]) - idxA[e0 + j - 1] : 0;
            const i64 d = gap > 0 ? (gap - 1) / 255 : 0;  // dummies before this entry
            const i64 w = active ? d + 1 : 0;
            const i64 incl = group_scan_incl<64>(w);
            if (active) {
                const i64 pos = o0 + run + incl - w;
                for (i64 q = 0; q < d; ++q) {
                    dltB[pos + q] = 255;
                    valB[pos + q] = VT(0);
                }
                dltB[pos + d] = static_cast<std::uint8_t>(gap - 255 * d);
                valB[pos + d] = valA[e0 + j];
            }
            run += __shfl(incl, 63, 64);
        }
    }
}

// idx[s] = first[seg] + (delta prefix through slot s): the 16-bit form, for the
// kernels that read indices out of order.
__global__ void k_delta_expand(const i64* __restrict__ ptr, i64 nseg, const std::uint16_t* __restrict__ first,
                               const std::uint8_t* __restrict__ dlt, std::uint16_t* __restrict__ idx) {
    const int lane = threadIdx.x & 63;
    const i64 wave = (static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const i64 nwaves = (static_cast<i64>(gridDim.x) * blockDim.x) >> 6;
    for (i64 seg = wave; seg < nseg; seg += nwaves) {
        const i64 s0 = ptr[seg], s1 = ptr[seg + 1];
        unsigned base = first[seg];
        for (i64 sb = s0; sb < s1; sb += 64) {
            const i64 s = sb + lane;
            const unsigned d = s < s1 ? dlt[s] : 0u;
            const unsigned incl = group_scan_incl<64>(d);
            if (s < s1) idx[s] = static_cast<std::uint16_t>(base + incl);
            base += __shfl(incl, 63, 64);
        }
    }
}

static i64 per_segment_grid(i64 nseg) { return std::max<i64>(1, std::min<i64>((nseg * 64 + 255) / 256, 65536)); }

void launch_delta_count(const i64* ptrA, const unsigned long long* counts, i64 nseg, const std::uint16_t* idxA,
                        unsigned long long* lenB, hipStream_t s) {
    hipLaunchKernelGGL(k_delta_count, dim3(static_cast<unsigned>(per_segment_grid(nseg))), dim3(256), 0, s, ptrA,
                       counts, nseg, idxA, lenB);
    IRLBA_HIP_LAUNCH_CHECK();
}

void launch_delta_write(const i64* ptrA, const unsigned long long* counts, i64 nseg, const std::uint16_t* idxA,
                        const void* valA, StorageType st, const i64* ptrB, std::uint16_t* first, std::uint8_t* dltB,
                        void* valB, hipStream_t s) {
    const dim3 g(static_cast<unsigned>(per_segment_grid(nseg))), b(256);
    if (st == ST_F64)
        hipLaunchKernelGGL(k_delta_write<double>, g, b, 0, s, ptrA, counts, nseg, idxA,
                           static_cast<const double*>(valA), ptrB, first, dltB, static_cast<double*>(valB));
    else if (st == ST_F32)
        hipLaunchKernelGGL(k_delta_write<float>, g, b, 0, s, ptrA, counts, nseg, idxA,
                           static_cast<const float*>(valA), ptrB, first, dltB, static_cast<float*>(valB));
    else
        throw std::runtime_error("the delta layout holds FP64 or FP32 values");
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ value dictionary
// Values that take at most 256 distinct bit patterns (count data: log1p of
// small integers, the integers themselves) can stream as one code byte per
// slot: the products then read {4 codes, 4 deltas} as one 8-byte word per lane
// step and look the values up in an LDS copy of the table.  The table holds
// the stored values themselves, so every product term is the same number as
// with the value stream.  Distinct values are found by bit pattern (so -0.0
// and +0.0 are two entries) in per-block LDS hash sets merged into a global one.

constexpr unsigned long long kDictEmpty = ~0ull;  // no FP32 pattern widens to it; an FP64 NaN payload could
constexpr int kDictLocalBits = 9, kDictGlobalBits = 10;

__device__ __forceinline__ unsigned dict_hash(unsigned long long p, int bits) {
    return static_cast<unsigned>((p * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}
__device__ __forceinline__ unsigned long long value_bits(float v) { return __float_as_uint(v); }
__device__ __forceinline__ unsigned long long value_bits(double v) {
    return static_cast<unsigned long long>(__double_as_longlong(v));
}

// Inserts p into a linear-probing set of 2^bits slots; returns 1 if p was new,
// 0 if present, -1 if the set is full.  Present entries are found by plain
// (relaxed atomic) reads, so the common case takes no atomic operation.
template <int BITS>
__device__ __forceinline__ int dict_insert(unsigned long long* set, unsigned long long p) {
    constexpr unsigned mask = (1u << BITS) - 1;
    unsigned h = dict_hash(p, BITS);
    for (int probe = 0; probe <= static_cast<int>(mask); ++probe, h = (h + 1) & mask) {
        unsigned long long cur = __hip_atomic_load(set + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == p) return 0;
        if (cur == kDictEmpty) {
            cur = atomicCAS(set + h, kDictEmpty, p);
            if (cur == kDictEmpty) return 1;
            if (cur == p) return 0;
        }
    }
    return -1;
}

// flag bit 1: more than 256 distinct patterns (or the reserved pattern) seen.
template <typename VT>
__global__ __launch_bounds__(256) void k_dict_collect(const VT* __restrict__ val, i64 n,
                                                      unsigned long long* __restrict__ gset, int* __restrict__ gcount,
                                                      int* __restrict__ flag) {
    __shared__ unsigned long long lset[1 << kDictLocalBits];
    __shared__ int lcount;
    for (int j = threadIdx.x; j < (1 << kDictLocalBits); j += blockDim.x) lset[j] = kDictEmpty;
    if (threadIdx.x == 0) lcount = 0;
    __syncthreads();
    unsigned long long last = kDictEmpty;  // this thread's previous pattern: runs of one value skip the set
    const i64 stride = static_cast<i64>(gridDim.x) * blockDim.x;
    for (i64 i = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const unsigned long long p = value_bits(__builtin_nontemporal_load(val + i));
        if (p == last) continue;
        last = p;
        if (p == kDictEmpty) {
            atomicOr(flag, 1);
            break;
        }
        const int r = dict_insert<kDictLocalBits>(lset, p);
        if (r < 0 || (r == 1 && atomicAdd(&lcount, 1) >= 256)) atomicOr(flag, 1);
        if (__hip_atomic_load(&lcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) > 256) break;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < (1 << kDictLocalBits); j += blockDim.x) {
        const unsigned long long p = lset[j];
        if (p == kDictEmpty) continue;
        const int r = dict_insert<kDictGlobalBits>(gset, p);
        if (r < 0 || (r == 1 && atomicAdd(gcount, 1) >= 256)) atomicOr(flag, 1);
    }
}

void launch_dict_collect(const void* val, StorageType st, i64 n, unsigned long long* gset, int* gcount, int* flag,
                         hipStream_t s) {
    if (n <= 0) return;
    const dim3 g(static_cast<unsigned>(std::max<i64>(1, std::min<i64>((n + 255) / 256, 4096)))), b(256);
    if (st == ST_F64)
        hipLaunchKernelGGL(k_dict_collect<double>, g, b, 0, s, static_cast<const double*>(val), n, gset, gcount, flag);
    else if (st == ST_F32)
        hipLaunchKernelGGL(k_dict_collect<float>, g, b, 0, s, static_cast<const float*>(val), n, gset, gcount, flag);
    else
        throw std::runtime_error("the value dictionary holds FP64 or FP32 values");
    IRLBA_HIP_LAUNCH_CHECK();
}

// Code stream, in runs of RUN = 4 or 16 slots: run r holds the RUN code bytes
// of slots RUN r .. RUN r + RUN - 1 (byte k = slot RUN r + k), then their RUN
// delta bytes, 2 RUN bytes in all.  pats[c] is the bit pattern of code c
// (n_codes distinct patterns).  flag bit 2: a slot value missing from the table.
template <typename VT, int RUN>
__global__ __launch_bounds__(256) void k_dict_encode(const VT* __restrict__ val, const unsigned* __restrict__ dlt4,
                                                     i64 nruns, const unsigned long long* __restrict__ pats,
                                                     int n_codes, unsigned* __restrict__ cd, int* __restrict__ flag) {
    constexpr int W = RUN / 4;  // words of codes (and of deltas) per run
    __shared__ unsigned long long kset[1 << kDictLocalBits];
    __shared__ unsigned char kcode[1 << kDictLocalBits];
    constexpr unsigned mask = (1u << kDictLocalBits) - 1;
    for (int j = threadIdx.x; j < (1 << kDictLocalBits); j += blockDim.x) kset[j] = kDictEmpty;
    __syncthreads();
    if (threadIdx.x == 0) {  // at most 256 patterns, distinct: a serial fill is simplest
        for (int c = 0; c < n_codes; ++c) {
            unsigned h = dict_hash(pats[c], kDictLocalBits);
            while (kset[h] != kDictEmpty) h = (h + 1) & mask;
            kset[h] = pats[c];
            kcode[h] = static_cast<unsigned char>(c);
        }
    }
    __syncthreads();
    const i64 stride = static_cast<i64>(gridDim.x) * blockDim.x;
    for (i64 r = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x; r < nruns; r += stride) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            unsigned codes = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const unsigned long long p = value_bits(val[RUN * r + 4 * w + k]);
                unsigned h = dict_hash(p, kDictLocalBits);
                int probe = 0;
                while (kset[h] != p && kset[h] != kDictEmpty && probe <= static_cast<int>(mask)) {
                    h = (h + 1) & mask;
                    ++probe;
                }
                if (kset[h] == p) codes |= static_cast<unsigned>(kcode[h]) << (8 * k);
                else atomicOr(flag, 2);
            }
            cd[2 * W * r + w] = codes;
            cd[2 * W * r + W + w] = dlt4[W * r + w];
        }
    }
}

void launch_dict_encode(const TiledLines& t, const unsigned long long* pats, int n_codes, void* cd, int* flag,
                        hipStream_t s) {
    const int run = t.dict_run;
    if (run != 4 && run != 16) throw std::runtime_error("value dictionary runs hold 4 or 16 slots");
    const i64 nruns = t.slots / run;
    if (nruns <= 0) return;
    const dim3 g(static_cast<unsigned>(std::max<i64>(1, std::min<i64>((nruns + 255) / 256, 8192)))), b(256);
    const unsigned* d4 = reinterpret_cast<const unsigned*>(t.dlt);
    unsigned* out = static_cast<unsigned*>(cd);
    if (t.st == ST_F64 && run == 16)
        hipLaunchKernelGGL((k_dict_encode<double, 16>), g, b, 0, s, static_cast<const double*>(t.val), d4, nruns,
                           pats, n_codes, out, flag);
    else if (t.st == ST_F64)
        hipLaunchKernelGGL((k_dict_encode<double, 4>), g, b, 0, s, static_cast<const double*>(t.val), d4, nruns,
                           pats, n_codes, out, flag);
    else if (t.st == ST_F32 && run == 16)
        hipLaunchKernelGGL((k_dict_encode<float, 16>), g, b, 0, s, static_cast<const float*>(t.val), d4, nruns,
                           pats, n_codes, out, flag);
    else if (t.st == ST_F32)
        hipLaunchKernelGGL((k_dict_encode<float, 4>), g, b, 0, s, static_cast<const float*>(t.val), d4, nruns,
                           pats, n_codes, out, flag);
    else
        throw std::runtime_error("the value dictionary holds FP64 or FP32 values");
    IRLBA_HIP_LAUNCH_CHECK();
}

void launch_delta_expand(const TiledLines& t, std::uint16_t* idx, hipStream_t s) {
    const i64 nseg = static_cast<i64>(t.n_tiles) * t.n_major;
    hipLaunchKernelGGL(k_delta_expand, dim3(static_cast<unsigned>(per_segment_grid(nseg))), dim3(256), 0, s, t.ptr,
                       nseg, t.first, t.dlt, idx);
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ tiled SpMV
template <typename AT>
struct SpmvParams {
    const i64* ptr;
    const std::uint16_t* idx;
    const void* val;
    const i64* chunk;
    const std::uint8_t* dlt;      // 8-bit delta indices (delta layout) ...
    const std::uint16_t* first;   // ... and the first index of each segment
    const void* cd;               // value dictionary: code stream in runs of DICT slots ...
    const void* dict;             // ... and the table (VT), n_dict entries
    int n_dict;
    i64 n_major, n_minor;
    int tile_len, n_tiles, B;
    Prologue pro;
    Epilogue epi;
    AT* partial;
    double* gamma_part;
};

template <typename VT, typename AT, int VEC>
struct Loader;

template <typename AT>
struct Loader<float, AT, 4> {
    __device__ __forceinline__ static void load(const float* v, const std::uint16_t* ix, i64 s, float (&vo)[4],
                                                unsigned (&io)[4]) {
        const f32x4 f = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(v + s));
        const u32x2 q = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(ix + s));
        vo[0] = f.x; vo[1] = f.y; vo[2] = f.z; vo[3] = f.w;
        io[0] = q.x & 0xffffu; io[1] = q.x >> 16; io[2] = q.y & 0xffffu; io[3] = q.y >> 16;
    }
};
template <typename AT>
struct Loader<double, AT, 4> {
    __device__ __forceinline__ static void load(const double* v, const std::uint16_t* ix, i64 s, double (&vo)[4],
                                                unsigned (&io)[4]) {
        const f64x2 a = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(v + s));
        const f64x2 b = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(v + s + 2));
        const u32x2 q = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(ix + s));
        vo[0] = a.x; vo[1] = a.y; vo[2] = b.x; vo[3] = b.y;
        io[0] = q.x & 0xffffu; io[1] = q.x >> 16; io[2] = q.y & 0xffffu; io[3] = q.y >> 16;
    }
};
template <typename VT, typename AT>
struct Loader<VT, AT, 1> {
    __device__ __forceinline__ static void load(const VT* v, const std::uint16_t* ix, i64 s, VT (&vo)[1],
                                                unsigned (&io)[1]) {
        vo[0] = __builtin_nontemporal_load(v + s);
        io[0] = __builtin_nontemporal_load(ix + s);
    }
};

// Four consecutive values (the delta layout loads its indices separately).
template <typename VT>
__device__ __forceinline__ void load_values4(const VT* v, i64 s, VT (&vo)[4]);
template <>
__device__ __forceinline__ void load_values4<float>(const float* v, i64 s, float (&vo)[4]) {
    const f32x4 f = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(v + s));
    vo[0] = f.x; vo[1] = f.y; vo[2] = f.z; vo[3] = f.w;
}
template <>
__device__ __forceinline__ void load_values4<double>(const double* v, i64 s, double (&vo)[4]) {
    const f64x2 a = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(v + s));
    const f64x2 b = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(v + s + 2));
    vo[0] = a.x; vo[1] = a.y; vo[2] = b.x; vo[3] = b.y;
}

// In-tile indices of a lane's four slots from their packed deltas: `carry` is
// the index reached before the group's step, `incl` the group's inclusive scan
// of the per-lane delta sums `tot`.
__device__ __forceinline__ void delta_indices(unsigned dw, unsigned carry, unsigned incl, unsigned tot,
                                              unsigned (&q)[4]) {
    q[0] = carry + incl - tot + (dw & 0xffu);
    q[1] = q[0] + ((dw >> 8) & 0xffu);
    q[2] = q[1] + ((dw >> 16) & 0xffu);
    q[3] = q[2] + (dw >> 24);
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b to XCD b % 8),
// and each XCD has its own L2.  Giving XCD x the consecutive chunks
// [x B/8, (x+1) B/8) keeps the blocks that stage the same input tile on one
// XCD, so a tile is fetched from HBM about once per product instead of once
// per block (A.v of the 31-tile 30000 x 1e6 matrix: ~190 MB less traffic).
__device__ __forceinline__ int xcd_chunk(int b, int nb) {
    constexpr int kXcds = 8;
    return (nb % kXcds == 0) ? (b % kXcds) * (nb / kXcds) + b / kXcds : b;
}

// Stages entries [base, base + cnt) of the scaled input into the LDS tile `xs`
// and adds the rank-1 dot products tdot_k . x to g0, g1: one loop per case
// (input scaling or not, number of terms) with the loads of four entries per
// thread (two with two terms) issued together.  A generic loop that checks the optional operands
// entry by entry waited for each entry's loads one after another (A.v of the
// centred and scaled 30000 x 1e6 matrix staged its 38464-entry tiles in ~35 us
// more than the plain matrix).  Each thread adds its entries in increasing
// order, as an entry-at-a-time loop does, so the sums are the same.
template <typename AT, bool SC, int NT>
__device__ __forceinline__ void stage_tile(const Prologue& pro, i64 base, int cnt, AT* __restrict__ xs, double& g0,
                                           double& g1) {
    const double* __restrict__ X = pro.x + base;
    const double* __restrict__ S = SC ? pro.xscale + base : nullptr;
    const double* __restrict__ T0 = (NT > 0 && pro.tdot[0]) ? pro.tdot[0] + base : nullptr;
    const double* __restrict__ T1 = (NT > 1 && pro.tdot[1]) ? pro.tdot[1] + base : nullptr;
    constexpr int U = NT == 2 ? 2 : 4;  // two terms at four entries would spill in the product loop
    const int stride = blockDim.x;
    for (int j0 = threadIdx.x; j0 < cnt; j0 += U * stride) {
        double xv[U], sv[U], t0[U], t1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u * stride;
            const bool ok = j < cnt;
            xv[u] = ok ? X[j] : 0.0;
            if constexpr (SC) sv[u] = ok ? S[j] : 0.0;
            if constexpr (NT > 0) t0[u] = (ok && T0) ? T0[j] : 1.0;
            if constexpr (NT > 1) t1[u] = (ok && T1) ? T1[j] : 1.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (j0 + u * stride < cnt) {
                xs[j0 + u * stride] = static_cast<AT>(SC ? xv[u] * sv[u] : xv[u]);
                if constexpr (NT > 0) g0 += t0[u] * xv[u];
                if constexpr (NT > 1) g1 += t1[u] * xv[u];
            }
        }
    }
}

// Decodes one run of 16 dictionary slots (four code words, four delta words)
// and accumulates its terms; q enters as the index before the run's first
// slot and leaves at its last.
template <typename AT>
__device__ __forceinline__ void dict_run16(const u32x4& c, const u32x4& d, unsigned& q, const AT* tab, const AT* xs,
                                           AT& acc0, AT& acc1) {
    const unsigned cw[4] = {c.x, c.y, c.z, c.w}, dw[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            q += (dw[j] >> (8 * k)) & 0xffu;
            const AT v = tab[(cw[j] >> (8 * k)) & 0xffu];
            if (j & 1) acc1 += v * xs[q];
            else acc0 += v * xs[q];
        }
    }
}

template <typename VT, typename AT, int LPL, int VEC, bool DELTA, int DICT>
__global__ __launch_bounds__(1024) void k_spmv_tiled(SpmvParams<AT> p) {
    extern __shared__ unsigned char smem_raw[];
    AT* xs = reinterpret_cast<AT*>(smem_raw);
    __shared__ double red[2 * 16];

    const int cid = xcd_chunk(blockIdx.x, gridDim.x);
    const i64 S0 = p.chunk[cid], S1 = p.chunk[cid + 1];
    if (S0 >= S1) return;  // uniform across the block
    const i64 t_first = S0 / p.n_major, t_last = (S1 - 1) / p.n_major;
    const VT* __restrict__ val = static_cast<const VT*>(p.val);
    const std::uint16_t* __restrict__ ix = p.idx;
    const int lane_in = threadIdx.x & (LPL - 1);
    const int grp = threadIdx.x / LPL, ngrp = blockDim.x / LPL;
    constexpr int STRIDE = LPL * VEC;
    // Steps in flight per lane: the tile in LDS allows one 16-wave block per
    // CU, so up to 128 VGPRs cost no occupancy; FP64 values or sums fit four steps.
    // Dictionary codes hold four registers per step (runs of 4; eight for
    // runs of 16) whatever the precision.
    constexpr int UNR = DICT == 16 ? 4 : DICT == 4 ? 8 : ((sizeof(VT) == 8 || sizeof(AT) == 8) ? 4 : 8);
    const Prologue& pro = p.pro;
    const bool single = p.n_tiles == 1;
    const double coef = (single && p.epi.axpy_src) ? *p.epi.axpy_coef : 0.0;
    // The value table sits after the input tile; the first tile's barrier
    // publishes it.
    [[maybe_unused]] AT* tab = xs + p.tile_len;
    if constexpr (DICT != 0) {
        static_assert(DELTA, "the value dictionary is a form of the delta layout");
        const VT* dict = static_cast<const VT*>(p.dict);
        for (int j = threadIdx.x; j < p.n_dict; j += blockDim.x) tab[j] = static_cast<AT>(dict[j]);
    }

    for (i64 t = t_first; t <= t_last; ++t) {
    if (t > t_first) __syncthreads();  // every group is done with the previous tile
    const i64 base = t * p.tile_len;
    const int cnt = static_cast<int>(min(static_cast<i64>(p.tile_len), p.n_minor - base));

    // Stage the input tile (scaled) in LDS; accumulate the rank-1 dot products.
    double g0 = 0, g1 = 0;
    if (pro.n_terms == 0)
        pro.xscale ? stage_tile<AT, true, 0>(pro, base, cnt, xs, g0, g1)
                   : stage_tile<AT, false, 0>(pro, base, cnt, xs, g0, g1);
    else if (pro.n_terms == 1)
        pro.xscale ? stage_tile<AT, true, 1>(pro, base, cnt, xs, g0, g1)
                   : stage_tile<AT, false, 1>(pro, base, cnt, xs, g0, g1);
    else
        pro.xscale ? stage_tile<AT, true, 2>(pro, base, cnt, xs, g0, g1)
                   : stage_tile<AT, false, 2>(pro, base, cnt, xs, g0, g1);
    if (pro.n_terms > 0) {
        block_sum2(g0, g1, red);
    } else {
        __syncthreads();
    }
    const i64 seg0 = t * p.n_major;  // the tile's first segment
    if (p.n_tiles > 1 && S0 <= seg0 && seg0 < S1 && threadIdx.x == 0) {
        p.gamma_part[2 * t] = g0;
        p.gamma_part[2 * t + 1] = g1;
    }
    const double gamma[2] = {g0, g1};

    const i64* __restrict__ P = p.ptr + seg0;
    const i64 L0 = max(S0, seg0) - seg0, L1 = min(S1, seg0 + p.n_major) - seg0;

    // The segment bounds of a group's next line are loaded while the current
    // line streams, so a line never starts with an exposed pointer load.
    const std::uint16_t* __restrict__ F = DELTA ? p.first + seg0 : nullptr;
    i64 nxt0 = 0, nxt1 = 0;
    unsigned nxtf = 0;
    if (L0 + grp < L1) {
        nxt0 = P[L0 + grp];
        nxt1 = P[L0 + grp + 1];
        if constexpr (DELTA) nxtf = F[L0 + grp];
    }
    for (i64 m = L0 + grp; m < L1; m += ngrp) {
        const i64 s0 = nxt0, s1 = nxt1;
        const unsigned first = nxtf;
        if (m + ngrp < L1) {
            nxt0 = P[m + ngrp];
            nxt1 = P[m + ngrp + 1];
            if constexpr (DELTA) nxtf = F[m + ngrp];
        }
        (void)first;
        EpiOperands eop;
        if (single && lane_in == 0) eop = load_epilogue(p.epi, m);  // in flight during the stream
        AT acc0 = 0, acc1 = 0;
        if constexpr (DELTA) {
            static_assert(VEC == 4, "the delta layout packs four indices per lane");
            // Group-uniform iterations of UNR steps: every lane of the group
            // takes part in each step's scan; lanes past the segment's end load
            // nothing and add nothing, so the tail costs one iteration, not one
            // exposed load per step.
            const std::uint8_t* __restrict__ dl = p.dlt;
            unsigned carry = first;
            // FP64 values: a lane's four slots are 32 bytes, so its own two
            // 16-byte loads would leave every load instruction of the group
            // covering only alternate halves of its cache lines (measured: 13%
            // more HBM fetch than the layout holds).  Instead each load
            // instruction of the group reads one contiguous run: lane l takes
            // slots 2l, 2l+1 of the step's first half and 2 LPL + 2l,
            // 2 LPL + 2l + 1 of its second half, while the deltas are decoded
            // as for FP32 (lane l: slots 4l .. 4l+3), and fetches its slots'
            // indices from the lanes that decode them (l/2 and LPL/2 + l/2) by
            // shuffles.  On a 9e8-entry matrix: A.v 1.79 -> 1.71 ms, A'.u
            // 1.50 -> 1.40 ms (with the lanes per line tuned for each form); a
            // form with each lane decoding its own two pairs (one scan of
            // packed pair sums, 16-bit delta loads) measured 1.77 / 1.46 ms.
            constexpr bool PAIRS = sizeof(VT) == 8 && DICT == 0;
            if constexpr (DICT == 16) {
                // Runs of 16 slots per lane step (a group step covers LPL runs,
                // 32 bytes each); one cross-lane scan per 16 slots.  Segments
                // hold whole runs.
                constexpr int RS = LPL * 16;  // slots per group step
                const u32x4* __restrict__ cp = static_cast<const u32x4*>(p.cd) + 2 * ((s0 >> 4) + lane_in);
                int left = static_cast<int>(s1 - s0);
                for (; left >= UNR * RS; left -= UNR * RS, cp += 2 * UNR * LPL) {
                    u32x4 c[UNR], d[UNR];
                    unsigned tot[UNR], inc[UNR];
#pragma unroll
                    for (int u = 0; u < UNR; ++u) {
                        c[u] = __builtin_nontemporal_load(cp + 2 * u * LPL);
                        d[u] = __builtin_nontemporal_load(cp + 2 * u * LPL + 1);
                    }
#pragma unroll
                    for (int u = 0; u < UNR; ++u) {
                        tot[u] = __builtin_amdgcn_sad_u8(
                            d[u].x, 0u, __builtin_amdgcn_sad_u8(
                                            d[u].y, 0u, __builtin_amdgcn_sad_u8(
                                                            d[u].z, 0u, __builtin_amdgcn_sad_u8(d[u].w, 0u, 0u))));
                        inc[u] = group_scan_incl_dpp<LPL>(tot[u]);
                    }
#pragma unroll
                    for (int u = 0; u < UNR; ++u) {
                        unsigned q = carry + inc[u] - tot[u];
                        carry += group_last<LPL>(inc[u]);
                        dict_run16(c[u], d[u], q, tab, xs, acc0, acc1);
                    }
                }
                if (left > 0) {
                    u32x4 c[UNR], d[UNR];
                    unsigned tot[UNR], inc[UNR];
                    bool ok[UNR];
#pragma unroll
                    for (int u = 0; u < UNR; ++u) {
                        ok[u] = lane_in * 16 + u * RS < left;
                        c[u] = u32x4{0, 0, 0, 0};
                        d[u] = u32x4{0, 0, 0, 0};
                        if (ok[u]) {
                            c[u] = __builtin_nontemporal_load(cp + 2 * u * LPL);
                            d[u] = __builtin_nontemporal_load(cp + 2 * u * LPL + 1);
                        }
                    }
#pragma unroll
                    for (int u = 0; u < UNR; ++u) {
                        tot[u] = __builtin_amdgcn_sad_u8(
                            d[u].x, 0u, __builtin_amdgcn_sad_u8(
                                            d[u].y, 0u, __builtin_amdgcn_sad_u8(
                                                            d[u].z, 0u, __builtin_amdgcn_sad_u8(d[u].w, 0u, 0u))));
                        inc[u] = group_scan_incl_dpp<LPL>(tot[u]);
                    }
#pragma unroll
                    for (int u = 0; u < UNR; ++u) {
                        unsigned q = carry + inc[u] - tot[u];
                        carry += group_last<LPL>(inc[u]);
                        if (ok[u]) dict_run16(c[u], d[u], q, tab, xs, acc0, acc1);
                    }
                }
            } else if constexpr (DICT == 4) {
                // One 8-byte word per lane step: the codes and deltas of its
                // four slots (lane l: slots 4l .. 4l+3 of the step).  Whole
                // iterations inside the segment run without predicates (no
                // compares, masks or branches: the product is bound by
                // instruction issue, not by its 2 bytes per slot), then one
                // predicated tail.
                const u32x2* __restrict__ cp = static_cast<const u32x2*>(p.cd) + (s0 >> 2) + lane_in;
                int left = static_cast<int>(s1 - s0);
                for (; left >= UNR * STRIDE; left -= UNR * STRIDE, cp += UNR * LPL) {
                    unsigned cw[UNR], dw[UNR], tot[UNR], inc[UNR];
#pragma unroll
                    for (int u = 0; u < UNR; ++u) {
                        const u32x2 w = __builtin_nontemporal_load(cp + u * LPL);
                        cw[u] = w.x;
                        dw[u] = w.y;
                    }
#pragma unroll
                    for (int u = 0; u < UNR; ++u) {
                        tot[u] = __builtin_amdgcn_sad_u8(dw[u], 0u, 0u);
                        inc[u] = group_scan_incl_dpp<LPL>(tot[u]);
                    }
#pragma unroll
                    for (int u = 0; u < UNR; ++u) {
                        unsigned q[4];
                        delta_indices(dw[u], carry, inc[u], tot[u], q);
                        carry += group_last<LPL>(inc[u]);
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const AT v = tab[(cw[u] >> (8 * k)) & 0xffu];
                            if (u & 1) acc1 += v * xs[q[k]];
                            else acc0 += v * xs[q[k]];
                        }
                    }
                }
                if (left > 0) {
                    unsigned cw[UNR], dw[UNR], tot[UNR], inc[UNR];
                    bool ok[UNR];
#pragma unroll
                    for (int u = 0; u < UNR; ++u) {
                        ok[u] = lane_in * 4 + u * STRIDE < left;
                        cw[u] = 0;
                        dw[u] = 0;
                        if (ok[u]) {
                            const u32x2 w = __builtin_nontemporal_load(cp + u * LPL);
                            cw[u] = w.x;
                            dw[u] = w.y;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < UNR; ++u) {
                        tot[u] = __builtin_amdgcn_sad_u8(dw[u], 0u, 0u);
                        inc[u] = group_scan_incl_dpp<LPL>(tot[u]);
                    }
#pragma unroll
                    for (int u = 0; u < UNR; ++u) {
                        unsigned q[4];
                        delta_indices(dw[u], carry, inc[u], tot[u], q);
                        carry += group_last<LPL>(inc[u]);
                        if (ok[u]) {
#pragma unroll
                            for (int k = 0; k < 4; ++k) {
                                const AT v = tab[(cw[u] >> (8 * k)) & 0xffu];
                                if (u & 1) acc1 += v * xs[q[k]];
                                else acc0 += v * xs[q[k]];
                            }
                        }
                    }
                }
            } else
            for (i64 sb = s0; sb < s1; sb += UNR * STRIDE) {
                VT v[UNR][4];
                unsigned dw[UNR], tot[UNR], inc[UNR];
                bool ok[UNR], oka[UNR], okb[UNR];
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    const i64 st = sb + u * STRIDE;
                    const i64 s = st + lane_in * 4;
                    ok[u] = s < s1;  // segments are padded to whole lanes (4 slots)
                    dw[u] = 0;
#pragma unroll
                    for (int k = 0; k < 4; ++k) v[u][k] = VT(0);
                    if constexpr (PAIRS) {
                        // Decoding lanes with slots in this step (group-uniform):
                        // pairs past the segment's end are neither loaded nor added.
                        const i64 nact = min(static_cast<i64>(LPL), max(static_cast<i64>(0), (s1 - st) / 4));
                        oka[u] = (lane_in >> 1) < nact;
                        okb[u] = (LPL / 2 + (lane_in >> 1)) < nact;
                        if (oka[u]) {
                            const f64x2 a =
                                __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(val + st + 2 * lane_in));
                            v[u][0] = a.x;
                            v[u][1] = a.y;
                        }
                        if (okb[u]) {
                            const f64x2 b = __builtin_nontemporal_load(
                                reinterpret_cast<const f64x2*>(val + st + 2 * LPL + 2 * lane_in));
                            v[u][2] = b.x;
                            v[u][3] = b.y;
                        }
                        if (ok[u]) dw[u] = __builtin_nontemporal_load(reinterpret_cast<const unsigned*>(dl + s));
                    } else if (ok[u]) {
                        load_values4<VT>(val, s, v[u]);
                        dw[u] = __builtin_nontemporal_load(reinterpret_cast<const unsigned*>(dl + s));
                    }
                }
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    tot[u] = __builtin_amdgcn_sad_u8(dw[u], 0u, 0u);  // sum of the four bytes
                    inc[u] = group_scan_incl_dpp<LPL>(tot[u]);
                }
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    unsigned q[4];
                    delta_indices(dw[u], carry, inc[u], tot[u], q);
                    carry += group_last<LPL>(inc[u]);
                    if constexpr (PAIRS) {
                        // In-tile indices fit 16 bits (tiles hold fewer than 65536 entries).
                        const int lo = static_cast<int>(q[0] | (q[1] << 16)), hi = static_cast<int>(q[2] | (q[3] << 16));
                        const int sa = lane_in >> 1, sbb = LPL / 2 + (lane_in >> 1);
                        const int alo = __shfl(lo, sa, LPL), ahi = __shfl(hi, sa, LPL);
                        const int blo = __shfl(lo, sbb, LPL), bhi = __shfl(hi, sbb, LPL);
                        const unsigned pa = static_cast<unsigned>((lane_in & 1) ? ahi : alo);
                        const unsigned pb = static_cast<unsigned>((lane_in & 1) ? bhi : blo);
                        AT& acc = (u & 1) ? acc1 : acc0;
                        if (oka[u]) {
                            acc += static_cast<AT>(v[u][0]) * xs[pa & 0xffffu];
                            acc += static_cast<AT>(v[u][1]) * xs[pa >> 16];
                        }
                        if (okb[u]) {
                            acc += static_cast<AT>(v[u][2]) * xs[pb & 0xffffu];
                            acc += static_cast<AT>(v[u][3]) * xs[pb >> 16];
                        }
                    } else if (ok[u]) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            if (u & 1) acc1 += static_cast<AT>(v[u][k]) * xs[q[k]];
                            else acc0 += static_cast<AT>(v[u][k]) * xs[q[k]];
                        }
                    }
                }
            }
        } else {
        i64 s = s0 + lane_in * VEC;
        for (; s + 3 * STRIDE < s1; s += 4 * STRIDE) {
            VT v[4][VEC];
            unsigned q[4][VEC];
#pragma unroll
            for (int u = 0; u < 4; ++u) Loader<VT, AT, VEC>::load(val, ix, s + u * STRIDE, v[u], q[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int k = 0; k < VEC; ++k) {
                    if (u & 1) acc1 += static_cast<AT>(v[u][k]) * xs[q[u][k]];
                    else acc0 += static_cast<AT>(v[u][k]) * xs[q[u][k]];
                }
        }
        for (; s < s1; s += STRIDE) {
            VT v[VEC];
            unsigned q[VEC];
            Loader<VT, AT, VEC>::load(val, ix, s, v, q);
#pragma unroll
            for (int k = 0; k < VEC; ++k) acc0 += static_cast<AT>(v[k]) * xs[q[k]];
        }
        }
        AT acc = group_sum<LPL>(acc0 + acc1);
        if (lane_in == 0) {
            if (single) p.epi.out[m] = finish_epilogue(p.epi, eop, static_cast<double>(acc), gamma, coef);
            else p.partial[seg0 + m] = acc;
        }
    }
    }
}

template <typename AT>
__global__ void k_spmv_reduce(const AT* __restrict__ partial, int n_tiles, i64 n_major,
                              const double* __restrict__ gamma_part, Epilogue epi) {
    // The rank-1 dot products: one wave sums the per-tile parts, shared through LDS.
    __shared__ double gsh[2];
    double gamma[2] = {0, 0};
    if (epi.n_terms > 0) {
        if (threadIdx.x < kWave) {
            double g0 = 0, g1 = 0;
            for (int t = threadIdx.x; t < n_tiles; t += kWave) {
                g0 += gamma_part[2 * t];
                g1 += gamma_part[2 * t + 1];
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
    // kSplit lanes per output line, each summing every kSplit-th tile, then a
    // fixed xor tree: more loads in flight than a lane per line (the per-tile
    // partials of a few tens of thousands of lines are a latency-bound read).
    // The loop bound is uniform across a wavefront (its first line), so every
    // lane reaches the shuffles.
    constexpr int kSplit = 4;
    const int sub = threadIdx.x & (kSplit - 1);
    const i64 g = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x;
    const i64 stride = static_cast<i64>(gridDim.x) * blockDim.x / kSplit;
    const i64 wave_first = (g & ~static_cast<i64>(kWave - 1)) / kSplit;
    for (i64 off = 0; wave_first + off < n_major; off += stride) {
        const i64 m = g / kSplit + off;
        const bool live = m < n_major;
        double acc = 0;
        if (live)
            for (int t = sub; t < n_tiles; t += kSplit) acc += static_cast<double>(partial[static_cast<i64>(t) * n_major + m]);
        acc += __shfl_xor(acc, 1, kWave);
        acc += __shfl_xor(acc, 2, kWave);
        if (live && sub == 0) epi.out[m] = apply_epilogue(epi, m, acc, gamma);
    }
}

template <typename VT, typename AT, int LPL, int VEC>
static void spmv_launch_one(const TiledLines& tl, const Prologue& pro, const Epilogue& epi, void* partial,
                            double* gamma_part, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    SpmvParams<AT> p;
    p.ptr = tl.ptr;
    p.idx = tl.idx;
    p.val = tl.val;
    p.chunk = tl.chunk;
    p.dlt = tl.dlt;
    p.first = tl.first;
    p.cd = tl.cd;
    p.dict = tl.dict;
    p.n_dict = tl.n_dict;
    p.n_major = tl.n_major;
    p.n_minor = tl.n_minor;
    p.tile_len = tl.tile_len;
    p.n_tiles = tl.n_tiles;
    p.B = tl.B;
    p.pro = pro;
    p.epi = epi;
    p.partial = static_cast<AT*>(partial);
    p.gamma_part = gamma_part;
    const bool dict = tl.use_dict && tl.cd && tl.delta;
    const size_t lds = (static_cast<size_t>(tl.tile_len) + (dict ? kDictMaxCodes : 0)) * sizeof(AT);
    if (ev0) IRLBA_HIP_CHECK(hipEventRecord(ev0, s));
    if constexpr (VEC == 4) {
        if (dict && tl.dict_run == 16)
            hipLaunchKernelGGL((k_spmv_tiled<VT, AT, LPL, 4, true, 16>), dim3(tl.B), dim3(1024), lds, s, p);
        else if (dict)
            hipLaunchKernelGGL((k_spmv_tiled<VT, AT, LPL, 4, true, 4>), dim3(tl.B), dim3(1024), lds, s, p);
        else if (tl.delta)
            hipLaunchKernelGGL((k_spmv_tiled<VT, AT, LPL, 4, true, 0>), dim3(tl.B), dim3(1024), lds, s, p);
        else
            hipLaunchKernelGGL((k_spmv_tiled<VT, AT, LPL, 4, false, 0>), dim3(tl.B), dim3(1024), lds, s, p);
    } else {
        hipLaunchKernelGGL((k_spmv_tiled<VT, AT, LPL, VEC, false, 0>), dim3(tl.B), dim3(1024), lds, s, p);
    }
    IRLBA_HIP_LAUNCH_CHECK();
    if (ev1) IRLBA_HIP_CHECK(hipEventRecord(ev1, s));
    if (tl.n_tiles > 1) {
        const i64 grid = std::max<i64>(1, std::min<i64>((tl.n_major * 4 + 255) / 256, 8192));  // 4 lanes per line
        hipLaunchKernelGGL(k_spmv_reduce<AT>, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s,
                           static_cast<const AT*>(partial), tl.n_tiles, tl.n_major, gamma_part, epi);
        IRLBA_HIP_LAUNCH_CHECK();
    }
}

template <typename VT, typename AT, int VEC>
static void spmv_dispatch_lpl(const TiledLines& tl, const Prologue& pro, const Epilogue& epi, void* partial,
                              double* gamma_part, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    switch (tl.lpl) {
        case 4: spmv_launch_one<VT, AT, 4, VEC>(tl, pro, epi, partial, gamma_part, s, ev0, ev1); break;
        case 8: spmv_launch_one<VT, AT, 8, VEC>(tl, pro, epi, partial, gamma_part, s, ev0, ev1); break;
        case 16: spmv_launch_one<VT, AT, 16, VEC>(tl, pro, epi, partial, gamma_part, s, ev0, ev1); break;
        case 32: spmv_launch_one<VT, AT, 32, VEC>(tl, pro, epi, partial, gamma_part, s, ev0, ev1); break;
        default: spmv_launch_one<VT, AT, 64, VEC>(tl, pro, epi, partial, gamma_part, s, ev0, ev1); break;
    }
}

template <typename VT, typename AT>
static void spmv_dispatch_vec(const TiledLines& tl, const Prologue& pro, const Epilogue& epi, void* partial,
                              double* gamma_part, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    if (tl.vec == 4) spmv_dispatch_lpl<VT, AT, 4>(tl, pro, epi, partial, gamma_part, s, ev0, ev1);
    else spmv_dispatch_lpl<VT, AT, 1>(tl, pro, epi, partial, gamma_part, s, ev0, ev1);
}

void launch_spmv_tiled(const TiledLines& tl, const Prologue& pro, const Epilogue& epi, bool acc64, void* partial,
                       double* gamma_part, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    if (tl.n_major == 0) return;
    if (tl.st == ST_F64) spmv_dispatch_vec<double, double>(tl, pro, epi, partial, gamma_part, s, ev0, ev1);
    else if (tl.st == ST_F32 && acc64) spmv_dispatch_vec<float, double>(tl, pro, epi, partial, gamma_part, s, ev0, ev1);
    else if (tl.st == ST_F32) spmv_dispatch_vec<float, float>(tl, pro, epi, partial, gamma_part, s, ev0, ev1);
    else throw std::runtime_error("unsupported sparse storage type");
}

// ------------------------------------------------------------------ input validation
// flag bits: 1 = pointers not starting at 0 / decreasing / not ending at nnz,
//            2 = an index outside [0, n_secondary).
__global__ void k_validate_compressed(const std::int32_t* __restrict__ idx, const i64* __restrict__ ptr,
                                      i64 n_primary, i64 n_secondary, i64 nnz, int* flag) {
    const i64 tid = static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x;
    const i64 nthreads = static_cast<i64>(gridDim.x) * blockDim.x;
    if (tid == 0 && (ptr[0] != 0 || ptr[n_primary] != nnz)) atomicOr(flag, 1);
    for (i64 p = tid; p < n_primary; p += nthreads)
        if (ptr[p + 1] < ptr[p]) atomicOr(flag, 1);
    for (i64 e = tid; e < nnz; e += nthreads) {
        const i64 v = idx[e];
        if (v < 0 || v >= n_secondary) atomicOr(flag, 2);
    }
}

void launch_validate_compressed(const std::int32_t* idx, const i64* ptr, i64 n_primary, i64 n_secondary, i64 nnz,
                                int* flag, hipStream_t s) {
    const i64 grid = std::max<i64>(1, std::min<i64>((std::max(nnz, n_primary) + 255) / 256, 16384));
    hipLaunchKernelGGL(k_validate_compressed, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, idx, ptr,
                       n_primary, n_secondary, nnz, flag);
    IRLBA_HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ realize / column statistics
template <typename VT>
__global__ void k_sparse_realize(const i64* __restrict__ ptr, const std::uint16_t* __restrict__ idx,
                                 const VT* __restrict__ val, i64 n_major, int n_tiles, int tile_len, int major_is_col,
                                 i64 nr, double* dense) {
    const int lane = threadIdx.x & 63;
    const i64 wave = (static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const i64 nwaves = (static_cast<i64>(gridDim.x) * blockDim.x) >> 6;
    const i64 nseg = static_cast<i64>(n_tiles) * n_major;
    for (i64 seg = wave; seg < nseg; seg += nwaves) {
        const i64 t = seg / n_major, m = seg - t * n_major;
        for (i64 e = ptr[seg] + lane; e < ptr[seg + 1]; e += 64) {
            const double v = to_double_any(val[e]);
            if (v == 0) continue;  // padding slots (stored zeros realize to zero anyway)
            const i64 minor = t * tile_len + idx[e];
            const i64 r = major_is_col ? minor : m;
            const i64 c = major_is_col ? m : minor;
            dense[r + c * nr] = v;
        }
    }
}

void launch_sparse_realize(const TiledLines& t, bool major_is_col, i64 nr, double* dense, hipStream_t s) {
    const i64 nseg = static_cast<i64>(t.n_tiles) * t.n_major;
    const i64 grid = std::max<i64>(1, std::min<i64>((nseg * 64 + 255) / 256, 65536));
    const dim3 g(static_cast<unsigned>(grid)), b(256);
    const int mc = major_is_col ? 1 : 0;
    if (t.st == ST_F64)
        hipLaunchKernelGGL(k_sparse_realize<double>, g, b, 0, s, t.ptr, t.idx, static_cast<const double*>(t.val),
                           t.n_major, t.n_tiles, t.tile_len, mc, nr, dense);
    else if (t.st == ST_F32)
        hipLaunchKernelGGL(k_sparse_realize<float>, g, b, 0, s, t.ptr, t.idx, static_cast<const float*>(t.val),
                           t.n_major, t.n_tiles, t.tile_len, mc, nr, dense);
    else
        hipLaunchKernelGGL(k_sparse_realize<__hip_bfloat16>, g, b, 0, s, t.ptr, t.idx,
                           static_cast<const __hip_bfloat16*>(t.val), t.n_major, t.n_tiles, t.tile_len, mc, nr,
                           dense);
    IRLBA_HIP_LAUNCH_CHECK();
}

// Column mean and standard deviation (pca.hpp:85-109) from the column-major
// lines: one wave per column, two passes over its slots.  Structural zeros and
// zero padding both enter as zeros: var = sum_slots (x-m)^2 + (nr - slots) m^2.
template <typename VT>
__global__ void k_sparse_colstats(const i64* __restrict__ ptr, const VT* __restrict__ val, i64 n_major, int n_tiles,
                                  i64 nr, double* center_out, double* scale_out) {
    const int lane = threadIdx.x & 63;
    const i64 wave = (static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const i64 nwaves = (static_cast<i64>(gridDim.x) * blockDim.x) >> 6;
    for (i64 c = wave; c < n_major; c += nwaves) {
        double sum = 0;
        i64 slots = 0;
        for (int t = 0; t < n_tiles; ++t) {
            const i64 seg = static_cast<i64>(t) * n_major + c;
            const i64 e0 = ptr[seg], e1 = ptr[seg + 1];
            slots += e1 - e0;
            for (i64 e = e0 + lane; e < e1; e += 64) sum += to_double_any(val[e]);
        }
        sum = wave_sum(sum);
        const double mean = center_out ? sum / nr : 0.0;
        if (center_out && lane == 0) center_out[c] = mean;
        if (scale_out) {
            double ss = 0;
            for (int t = 0; t < n_tiles; ++t) {
                const i64 seg = static_cast<i64>(t) * n_major + c;
                for (i64 e = ptr[seg] + lane; e < ptr[seg + 1]; e += 64) {
                    const double d = to_double_any(val[e]) - mean;
                    ss += d * d;
                }
            }
            ss = wave_sum(ss);
            if (lane == 0) {
                const double var = ss + static_cast<double>(nr - slots) * mean * mean;
                scale_out[c] = var != 0 ? sqrt(var / static_cast<double>(nr - 1)) : 1.0;
            }
        }
    }
}

void launch_sparse_colstats(const TiledLines& t, i64 nr, double* center_out, double* scale_out, hipStream_t s) {
    const i64 grid = std::max<i64>(1, std::min<i64>((t.n_major * 64 + 255) / 256, 65536));
    const dim3 g(static_cast<unsigned>(grid)), b(256);
    if (t.st == ST_F64)
        hipLaunchKernelGGL(k_sparse_colstats<double>, g, b, 0, s, t.ptr, static_cast<const double*>(t.val),
                           t.n_major, t.n_tiles, nr, center_out, scale_out);
    else if (t.st == ST_F32)
        hipLaunchKernelGGL(k_sparse_colstats<float>, g, b, 0, s, t.ptr, static_cast<const float*>(t.val), t.n_major,
                           t.n_tiles, nr, center_out, scale_out);
    else
        hipLaunchKernelGGL(k_sparse_colstats<__hip_bfloat16>, g, b, 0, s, t.ptr,
                           static_cast<const __hip_bfloat16*>(t.val), t.n_major, t.n_tiles, nr, center_out,
                           scale_out);
    IRLBA_HIP_LAUNCH_CHECK();
}

}  // namespace irlba_hip
