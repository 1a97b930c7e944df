// Host FP64 SVD (one-sided Jacobi), compiled by the host compiler only.  The
// same code is built twice: for the baseline x86-64 ISA and with AVX2 + FMA,
// and the second runs where the CPU has them (__builtin_cpu_supports), about
// 30% faster on the 70 x 70 restart matrices.  The two builds round
// differently (fused multiply-adds), so results agree to rounding, not bit for bit.
#include "host_svd.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace irlba_hip {
namespace {

typedef std::int64_t i64;

// Dot product with four interleaved partial sums (a fixed order, so the
// result is reproducible) to break the serial add chain.
__attribute__((always_inline)) inline double dot4(const double* x, const double* y, i64 m) {
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    i64 r = 0;
    for (; r + 3 < m; r += 4) {
        s0 += x[r] * y[r];
        s1 += x[r + 1] * y[r + 1];
        s2 += x[r + 2] * y[r + 2];
        s3 += x[r + 3] * y[r + 3];
    }
    for (; r < m; ++r) s0 += x[r] * y[r];
    return (s0 + s1) + (s2 + s3);
}

// m >= n.
__attribute__((always_inline)) inline void jacobi_svd(const HostMat& A, HostMat& U, std::vector<double>& S,
                                                      HostMat& V) {
    const i64 m = A.nr, n = A.nc;
    HostMat G = A, Q(n, n);
    for (i64 i = 0; i < n; ++i) Q(i, i) = 1;
    std::vector<double> sq(n);
    const double eps = std::numeric_limits<double>::epsilon();
    for (int sweep = 0; sweep < 80; ++sweep) {
        // Exact squared norms at the start of each sweep; within the sweep they
        // follow the rotations in closed form (alpha -= t*gamma, beta += t*gamma).
        for (i64 j = 0; j < n; ++j) sq[j] = dot4(G.col(j), G.col(j), m);
        double worst = 0;
        for (i64 p = 0; p < n - 1; ++p) {
            for (i64 q = p + 1; q < n; ++q) {
                double* gp = G.col(p);
                double* gq = G.col(q);
                const double g = dot4(gp, gq, m);
                const double scale = std::sqrt(sq[p] * sq[q]);
                if (g == 0 || std::abs(g) <= eps * scale) continue;
                worst = std::max(worst, std::abs(g) / scale);
                const double zeta = (sq[q] - sq[p]) / (2 * g);
                const double t = std::copysign(1.0, zeta) / (std::abs(zeta) + std::sqrt(1 + zeta * zeta));
                const double c = 1 / std::sqrt(1 + t * t), s = c * t;
                for (i64 r = 0; r < m; ++r) {
                    const double x = gp[r], y = gq[r];
                    gp[r] = c * x - s * y;
                    gq[r] = s * x + c * y;
                }
                double* qp = Q.col(p);
                double* qq = Q.col(q);
                for (i64 r = 0; r < n; ++r) {
                    const double x = qp[r], y = qq[r];
                    qp[r] = c * x - s * y;
                    qq[r] = s * x + c * y;
                }
                sq[p] -= t * g;
                sq[q] += t * g;
            }
        }
        if (worst <= eps) break;
    }
    std::vector<i64> order(n);
    for (i64 j = 0; j < n; ++j) order[j] = j;
    std::vector<double> nrm(n);
    for (i64 j = 0; j < n; ++j) nrm[j] = std::sqrt(dot4(G.col(j), G.col(j), m));
    std::stable_sort(order.begin(), order.end(), [&](i64 x, i64 y) { return nrm[x] > nrm[y]; });
    U = HostMat(m, n);
    V = HostMat(n, n);
    S.assign(n, 0);
    std::vector<char> have(n, 0);
    for (i64 j = 0; j < n; ++j) {
        const i64 o = order[j];
        S[j] = nrm[o];
        std::copy(Q.col(o), Q.col(o) + n, V.col(j));
        if (nrm[o] > 0) {
            for (i64 r = 0; r < m; ++r) U(r, j) = G(r, o) / nrm[o];
            have[j] = 1;
        }
    }
    i64 probe = 0;
    for (i64 j = 0; j < n; ++j) {
        if (have[j]) continue;
        while (probe < m) {
            std::vector<double> e(m, 0.0);
            e[probe++] = 1;
            for (int pass = 0; pass < 2; ++pass)
                for (i64 k = 0; k < n; ++k) {
                    if (!have[k]) continue;
                    double d = 0;
                    for (i64 r = 0; r < m; ++r) d += U(r, k) * e[r];
                    for (i64 r = 0; r < m; ++r) e[r] -= d * U(r, k);
                }
            double nn = 0;
            for (double x : e) nn += x * x;
            nn = std::sqrt(nn);
            if (nn > 1e-8) {
                for (i64 r = 0; r < m; ++r) U(r, j) = e[r] / nn;
                have[j] = 1;
                break;
            }
        }
    }
}

__attribute__((target("avx2,fma"))) void jacobi_svd_avx2(const HostMat& A, HostMat& U, std::vector<double>& S,
                                                        HostMat& V) {
    jacobi_svd(A, U, S, V);
}

void jacobi_svd_base(const HostMat& A, HostMat& U, std::vector<double>& S, HostMat& V) { jacobi_svd(A, U, S, V); }

}  // namespace

namespace {

bool cpu_has_avx2_fma() {
    static const bool fast = [] {
        __builtin_cpu_init();
        return __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
    }();
    return fast;
}

void run_variant(const HostMat& A, HostMat& U, std::vector<double>& S, HostMat& V, bool avx2) {
    if (A.nr < A.nc) {
        HostMat T(A.nc, A.nr);
        for (i64 c = 0; c < A.nc; ++c)
            for (i64 r = 0; r < A.nr; ++r) T(c, r) = A(r, c);
        run_variant(T, V, S, U, avx2);
        return;
    }
    if (avx2) jacobi_svd_avx2(A, U, S, V);
    else jacobi_svd_base(A, U, S, V);
}

}  // namespace

void host_svd(const HostMat& A, HostMat& U, std::vector<double>& S, HostMat& V) {
    run_variant(A, U, S, V, cpu_has_avx2_fma());
}

bool host_svd_variant(const HostMat& A, HostMat& U, std::vector<double>& S, HostMat& V, int variant) {
    if (variant == 1 && !cpu_has_avx2_fma()) return false;
    run_variant(A, U, S, V, variant == 1);
    return true;
}

}  // namespace irlba_hip
