// Host FP64 SVD of the small projected matrix B (work x work) and the exact
// fallback: one-sided (Hestenes) Jacobi on the columns of an m x n matrix
// (m >= n is arranged by transposition), thin factors, values in decreasing
// order; left vectors of zero values are completed to an orthonormal set.
// Stands in for Eigen::JacobiSVD (include/irlba/compute.hpp:27-28).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace irlba_hip {

struct HostMat {
    std::int64_t nr = 0, nc = 0;
    std::vector<double> a;
    HostMat() = default;
    HostMat(std::int64_t r, std::int64_t c) : nr(r), nc(c), a(static_cast<std::size_t>(r * c), 0.0) {}
    double& operator()(std::int64_t r, std::int64_t c) { return a[r + c * nr]; }
    double operator()(std::int64_t r, std::int64_t c) const { return a[r + c * nr]; }
    double* col(std::int64_t c) { return a.data() + c * nr; }
    const double* col(std::int64_t c) const { return a.data() + c * nr; }
};

// A = U diag(S) V'.  Internal to the library.
__attribute__((visibility("hidden"))) void host_svd(const HostMat& A, HostMat& U, std::vector<double>& S, HostMat& V);

// The same with a chosen build: 0 the baseline x86-64 code, 1 the AVX2 + FMA
// code (only where the CPU has them; returns false otherwise).  For tests.
__attribute__((visibility("hidden"))) bool host_svd_variant(const HostMat& A, HostMat& U, std::vector<double>& S,
                                                            HostMat& V, int variant);

}  // namespace irlba_hip
