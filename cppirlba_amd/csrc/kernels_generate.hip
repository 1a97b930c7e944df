// Synthetic sparse input generated on the device.
//
// Counter-based: entry (r, c) exists when the top 32 bits of
// mix64(key_c ^ r*0xD1B54A32D192ED03) fall below floor(density * 2^32), and its
// value is log1p(K) with K ~ NB(mu = 2, phi = 0.5) drawn by inverting the dyadic
// CDF 1 - (k+3)/2^(k+2) on a second 32-bit draw.  The arithmetic is integer
// only (the log1p table comes from the host), so the device matrix is
// bit-identical to the one the CPU generator writes for any slice of columns.
// One wavefront per column: ballots give the positions, so entries are written
// in ascending row order without atomics.
#include "common.hpp"

namespace irlba_hip {

__device__ __forceinline__ std::uint64_t mix64(std::uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ std::uint64_t line_key(std::uint64_t seed, std::uint64_t line, std::uint64_t stream) {
    return mix64(seed ^ mix64(2 * line + stream));
}
__device__ __forceinline__ std::uint32_t draw32(std::uint64_t key, std::uint64_t pos) {
    return static_cast<std::uint32_t>(mix64(key ^ (pos * 0xD1B54A32D192ED03ULL)) >> 32);
}
__device__ __forceinline__ int nb_draw(std::uint32_t u) {
    for (int k = 0; k <= 30; ++k) {
        const std::uint64_t T = (1ULL << 32) - (static_cast<std::uint64_t>(k + 3) << (30 - k));
        if (u < T) return k;
    }
    return 31;
}

__global__ void k_gen_count(i64 n_minor, i64 n_major, std::uint32_t thr, std::uint64_t seed, i64* counts) {
    const int lane = threadIdx.x & 63;
    const i64 wave = (static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const i64 nwaves = (static_cast<i64>(gridDim.x) * blockDim.x) >> 6;
    for (i64 c = wave; c < n_major; c += nwaves) {
        const std::uint64_t kp = line_key(seed, c, 0);
        i64 cnt = 0;
        for (i64 r0 = 0; r0 < n_minor; r0 += 64) {
            const i64 r = r0 + lane;
            const bool in = r < n_minor && draw32(kp, r) < thr;
            cnt += __popcll(__ballot(in));
        }
        if (lane == 0) counts[c] = cnt;
    }
}

template <typename VT>
__global__ void k_gen_fill(i64 n_minor, i64 n_major, std::uint32_t thr, std::uint64_t seed, const i64* __restrict__ ptr,
                           const double* __restrict__ table, VT* __restrict__ values, std::int32_t* __restrict__ indices) {
    const int lane = threadIdx.x & 63;
    const i64 wave = (static_cast<i64>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const i64 nwaves = (static_cast<i64>(gridDim.x) * blockDim.x) >> 6;
    const unsigned long long below = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    for (i64 c = wave; c < n_major; c += nwaves) {
        const std::uint64_t kp = line_key(seed, c, 0), kv = line_key(seed, c, 1);
        i64 pos = ptr[c];
        for (i64 r0 = 0; r0 < n_minor; r0 += 64) {
            const i64 r = r0 + lane;
            const bool in = r < n_minor && draw32(kp, r) < thr;
            const unsigned long long mask = __ballot(in);
            if (in) {
                const i64 p = pos + __popcll(mask & below);
                values[p] = static_cast<VT>(table[nb_draw(draw32(kv, r))]);
                indices[p] = static_cast<std::int32_t>(r);
            }
            pos += __popcll(mask);
        }
    }
}

void launch_gen_nb_count(i64 n_minor, i64 n_major, std::uint32_t thr, std::uint64_t seed, i64* counts,
                         hipStream_t s) {
    const i64 grid = std::max<i64>(1, std::min<i64>((n_major * 64 + 255) / 256, 65536));
    hipLaunchKernelGGL(k_gen_count, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, n_minor, n_major, thr, seed,
                       counts);
    IRLBA_HIP_LAUNCH_CHECK();
}

void launch_gen_nb_fill(i64 n_minor, i64 n_major, std::uint32_t thr, std::uint64_t seed, const i64* ptr,
                        const double* table, bool f32, void* values, std::int32_t* indices, hipStream_t s) {
    const i64 grid = std::max<i64>(1, std::min<i64>((n_major * 64 + 255) / 256, 65536));
    if (f32)
        hipLaunchKernelGGL(k_gen_fill<float>, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, n_minor, n_major, thr,
                           seed, ptr, table, static_cast<float*>(values), indices);
    else
        hipLaunchKernelGGL(k_gen_fill<double>, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s, n_minor, n_major,
                           thr, seed, ptr, table, static_cast<double*>(values), indices);
    IRLBA_HIP_LAUNCH_CHECK();
}

}  // namespace irlba_hip
