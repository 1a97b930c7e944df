// ORACLE — test infrastructure, not product code.
//
// A plain C++17 CPU restatement of the reference IRLBA path (LTLA/CppIrlba,
// include/irlba/*.hpp), without Eigen, used (a) as the parity checker for the
// HIP engine and (b) as the CPU baseline that bench.py times.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
//
// Each function names the reference file:line it restates.  What pins each
// part (checked in tests/test_oracle.py):
//   * orthogonalize_vector      — the R-generated 5x4 known-answer vectors of
//                                 tests/src/lanczos.cpp:13-51 (orthogonality).
//   * run_lanczos               — W'W = V'V = I to 1e-11 (lanczos.cpp:85-117),
//                                 restart-from-mid equivalence (lanczos.cpp:119-163),
//                                 and the bidiagonal relation A V = W B.
//   * update_k & helpers        — the integer truth tables of compute.cpp:33-76.
//   * jacobi_svd                — LAPACK (numpy.linalg.svd) on random and
//                                 rank-deficient inputs.
//   * compute / pca             — LAPACK SVD of the (centred/scaled) dense copy,
//                                 the 27-shape sweep of compute.cpp:78-123,
//                                 pca.cpp:66-151 identities, invariant.cpp ranks.
//   * sparse mat-vecs           — dense GEMV on the realized matrix, exact (tol 0)
//                                 for CSC A.v as in tests/src/Matrix/sparse.cpp:114,
//                                 and the partition invariants of sparse.cpp:56-89.
//   * synthetic NB generator    — the closed-form NB(mu=2, phi=0.5) pmf
//                                 (k+1)/2^(k+2) and Bernoulli density.
//   * standard_normal_pair      — parity unpinned: the reference draws refills and
//                                 default start vectors through aarand, which is not
//                                 vendored; every parity run passes Options::initial.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace oracle {

typedef std::int64_t Index;

// ---------------------------------------------------------------- dense helpers
struct Mat {
    Index nr = 0, nc = 0;
    std::vector<double> a;  // column-major
    Mat() = default;
    Mat(Index r, Index c) : nr(r), nc(c), a(static_cast<size_t>(r * c), 0.0) {}
    double& operator()(Index r, Index c) { return a[r + c * nr]; }
    double operator()(Index r, Index c) const { return a[r + c * nr]; }
    double* col(Index c) { return a.data() + c * nr; }
    const double* col(Index c) const { return a.data() + c * nr; }
};

static double dot(const double* x, const double* y, Index n) {
    double s = 0;
    for (Index i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
}
static double norm2(const double* x, Index n) { return std::sqrt(dot(x, x, n)); }

// ---------------------------------------------------------------- random normals
// Box-Muller pair from two uniforms in (0, 1], the published method that
// aarand::standard_normal implements (utils.hpp:27 calls it; aarand is not
// vendored, so exact draws are parity unpinned).
template <class Engine>
static double standard_uniform(Engine& eng) {
    const double range = static_cast<double>(Engine::max() - Engine::min()) + 1.0;
    double v;
    do {
        v = static_cast<double>(eng() - Engine::min()) / range;
    } while (v == 0);
    return v;
}
template <class Engine>
static std::pair<double, double> standard_normal_pair(Engine& eng) {
    const double pi = 3.14159265358979323846;
    const double r = std::sqrt(-2.0 * std::log(standard_uniform(eng)));
    const double theta = 2.0 * pi * standard_uniform(eng);
    return {r * std::sin(theta), r * std::cos(theta)};
}

// utils.hpp:16-37: pairs of normals; an odd length drops the second draw.
template <class Engine>
static void fill_with_random_normals(double* v, Index n, Engine& eng) {
    Index even = n - (n % 2);
    for (Index i = 0; i < even; i += 2) {
        auto p = standard_normal_pair(eng);
        v[i] = p.first;
        v[i + 1] = p.second;
    }
    if (n % 2) v[even] = standard_normal_pair(eng).first;
}

// ---------------------------------------------------------------- matrix operators
// Matrix/interface.hpp:141-223 collapsed to the three operations the algorithm uses.
struct OMat {
    virtual ~OMat() {}
    virtual Index rows() const = 0;
    virtual Index cols() const = 0;
    virtual void multiply(const double* in, double* out) const = 0;          // out (rows) = A in
    virtual void adjoint_multiply(const double* in, double* out) const = 0;  // out (cols) = A' in
    virtual void realize(Mat& out) const = 0;
    // matrix.col(i) of pca.hpp:88 (nr values).  The generic form multiplies by
    // the i-th unit vector, which is exact; storage types override it with a copy.
    virtual void column(Index i, double* out) const {
        std::vector<double> e(cols(), 0.0);
        e[i] = 1;
        multiply(e.data(), out);
    }
};

// Matrix/simple.hpp:42-44 (A v), :70-72 (A' u), :100-107 (realize).
struct ODense : OMat {
    Mat A;
    explicit ODense(Mat m) : A(std::move(m)) {}
    Index rows() const override { return A.nr; }
    Index cols() const override { return A.nc; }
    void multiply(const double* in, double* out) const override {
        for (Index r = 0; r < A.nr; ++r) out[r] = 0;
        for (Index c = 0; c < A.nc; ++c) {
            const double v = in[c];
            const double* col = A.col(c);
            for (Index r = 0; r < A.nr; ++r) out[r] += col[r] * v;
        }
    }
    void adjoint_multiply(const double* in, double* out) const override {
        for (Index c = 0; c < A.nc; ++c) out[c] = dot(A.col(c), in, A.nr);
    }
    void realize(Mat& out) const override { out = A; }
    void column(Index i, double* out) const override { std::copy(A.col(i), A.col(i) + A.nr, out); }
};

// std::thread fork/join, the shape of tests/src/custom_parallel.h:7-19 used as
// IRLBA_CUSTOM_PARALLEL (parallel.hpp:32-40).
template <class Fn>
static void parallelize(int nthreads, Fn fn) {
    std::vector<std::thread> jobs;
    jobs.reserve(nthreads);
    for (int t = 0; t < nthreads; ++t) jobs.emplace_back(fn, t);
    for (auto& j : jobs) j.join();
}

// Matrix/sparse.hpp:29-268 (ParallelSparseMatrixCore) and :283-409 (workspaces).
// Value_ is double or float; products and sums are carried in double, which is
// what the reference does with a float ValueArray_ and a double EigenVector_.
template <typename Value_>
struct OSparse : OMat {
    Index primary, secondary;
    int nthreads;
    bool column_major;
    std::vector<Value_> x;
    std::vector<std::int32_t> idx;
    std::vector<std::int64_t> ptr;
    std::vector<Index> primary_bounds, secondary_bounds;
    std::vector<std::vector<std::int64_t> > secondary_nz_bounds;

    OSparse(Index nrow, Index ncol, const Value_* vals, const std::int32_t* ii, const std::int64_t* pp,
            bool cm, int nt)
        : primary(cm ? ncol : nrow), secondary(cm ? nrow : ncol), nthreads(nt), column_major(cm) {
        const std::int64_t nnz = pp[primary];
        x.assign(vals, vals + nnz);
        idx.assign(ii, ii + nnz);
        ptr.assign(pp, pp + primary + 1);
        if (nthreads > 1) partition();
    }

    // sparse.hpp:52-114: nnz-balanced split of the primary dimension, of the
    // secondary dimension, and the per-(thread, primary) nonzero sub-ranges.
    void partition() {
        const std::int64_t total = ptr[primary];
        const std::int64_t floor_ = total / nthreads;
        const int extra = static_cast<int>(total % nthreads);
        primary_bounds.assign(nthreads + 1, 0);
        {
            Index counter = 0;
            std::int64_t sofar = 0;
            for (int t = 0; t < nthreads; ++t) {
                sofar += floor_ + (t < extra);
                while (counter < primary && ptr[counter + 1] <= sofar) ++counter;
                primary_bounds[t + 1] = counter;
            }
        }
        std::vector<std::int64_t> per_secondary(secondary, 0);
        for (std::int64_t s = 0; s < total; ++s) ++per_secondary[idx[s]];
        secondary_bounds.assign(nthreads + 1, 0);
        {
            Index counter = 0;
            std::int64_t sofar = 0, cum = 0;
            for (int t = 0; t < nthreads; ++t) {
                sofar += floor_ + (t < extra);
                while (counter < secondary && cum <= sofar) {
                    cum += per_secondary[counter];
                    ++counter;
                }
                secondary_bounds[t + 1] = counter;
            }
        }
        secondary_nz_bounds.assign(nthreads + 1, std::vector<std::int64_t>(primary));
        for (Index c = 0; c < primary; ++c) {
            std::int64_t s = ptr[c];
            const std::int64_t end = ptr[c + 1];
            secondary_nz_bounds[0][c] = s;
            for (int t = 0; t < nthreads; ++t) {
                const Index limit = secondary_bounds[t + 1];
                while (s < end && static_cast<Index>(idx[s]) < limit) ++s;
                secondary_nz_bounds[t + 1][c] = s;
            }
        }
    }

    Index rows() const override { return column_major ? secondary : primary; }
    Index cols() const override { return column_major ? primary : secondary; }

    // sparse.hpp:185-235: scatter along the primary dimension into the secondary.
    void indirect_multiply(const double* rhs, double* out) const {
        if (nthreads == 1) {
            std::fill(out, out + secondary, 0.0);
            for (Index c = 0; c < primary; ++c) {
                const double v = rhs[c];
                for (std::int64_t s = ptr[c]; s < ptr[c + 1]; ++s) out[idx[s]] += x[s] * v;
            }
            return;
        }
        parallelize(nthreads, [&](int t) {
            const Index s0 = secondary_bounds[t], s1 = secondary_bounds[t + 1];
            std::vector<double> buffer;
            double* o;
            if (t) {
                buffer.assign(s1 - s0, 0.0);
                o = buffer.data();
            } else {
                o = out + s0;
                std::fill(o, o + (s1 - s0), 0.0);
            }
            const auto& a = secondary_nz_bounds[t];
            const auto& b = secondary_nz_bounds[t + 1];
            for (Index c = 0; c < primary; ++c) {
                const double v = rhs[c];
                for (std::int64_t s = a[c]; s < b[c]; ++s) o[idx[s] - s0] += x[s] * v;
            }
            if (t) std::copy(buffer.begin(), buffer.end(), out + s0);
        });
    }

    // sparse.hpp:238-267: one dot product per primary element.
    double primary_dot(Index p, const double* rhs) const {
        double d = 0;
        for (std::int64_t s = ptr[p]; s < ptr[p + 1]; ++s) d += x[s] * rhs[idx[s]];
        return d;
    }
    void direct_multiply(const double* rhs, double* out) const {
        if (nthreads == 1) {
            for (Index p = 0; p < primary; ++p) out[p] = primary_dot(p, rhs);
            return;
        }
        parallelize(nthreads, [&](int t) {
            for (Index p = primary_bounds[t]; p < primary_bounds[t + 1]; ++p) out[p] = primary_dot(p, rhs);
        });
    }

    // sparse.hpp:305-311 and :346-352.
    void multiply(const double* in, double* out) const override {
        if (column_major) indirect_multiply(in, out); else direct_multiply(in, out);
    }
    void adjoint_multiply(const double* in, double* out) const override {
        if (column_major) direct_multiply(in, out); else indirect_multiply(in, out);
    }
    // sparse.hpp:382-409.
    void realize(Mat& out) const override {
        out = Mat(rows(), cols());
        for (Index p = 0; p < primary; ++p) {
            for (std::int64_t s = ptr[p]; s < ptr[p + 1]; ++s) {
                if (column_major) out(idx[s], p) = x[s]; else out(p, idx[s]) = x[s];
            }
        }
    }
    void column(Index i, double* out) const override {
        if (!column_major) return OMat::column(i, out);
        std::fill(out, out + secondary, 0.0);
        for (std::int64_t s = ptr[i]; s < ptr[i + 1]; ++s) out[idx[s]] = x[s];
    }
};

// Matrix/centered.hpp:46-52 (A v - (v.c) 1), :83-87 (A'u - (sum u) c), :118-122.
struct OCentered : OMat {
    std::shared_ptr<OMat> inner;
    std::vector<double> center;
    OCentered(std::shared_ptr<OMat> m, const double* c) : inner(std::move(m)), center(c, c + inner->cols()) {}
    Index rows() const override { return inner->rows(); }
    Index cols() const override { return inner->cols(); }
    void multiply(const double* in, double* out) const override {
        inner->multiply(in, out);
        const double beta = dot(in, center.data(), cols());
        for (Index r = 0; r < rows(); ++r) out[r] -= beta;
    }
    void adjoint_multiply(const double* in, double* out) const override {
        inner->adjoint_multiply(in, out);
        double beta = 0;
        for (Index r = 0; r < rows(); ++r) beta += in[r];
        for (Index c = 0; c < cols(); ++c) out[c] -= beta * center[c];
    }
    void realize(Mat& out) const override {
        inner->realize(out);
        for (Index c = 0; c < out.nc; ++c)
            for (Index r = 0; r < out.nr; ++r) out(r, c) -= center[c];
    }
};

// Matrix/scaled.hpp:52-69 (forward), :106-123 (adjoint), :158-177 (realize).
struct OScaled : OMat {
    std::shared_ptr<OMat> inner;
    std::vector<double> scale;
    bool column, divide;
    OScaled(std::shared_ptr<OMat> m, const double* s, bool col, bool div)
        : inner(std::move(m)), column(col), divide(div) {
        const Index n = column ? inner->cols() : inner->rows();
        scale.assign(s, s + n);
    }
    Index rows() const override { return inner->rows(); }
    Index cols() const override { return inner->cols(); }
    void apply(const double* in, double* out, Index n) const {
        for (Index i = 0; i < n; ++i) out[i] = divide ? in[i] / scale[i] : in[i] * scale[i];
    }
    void multiply(const double* in, double* out) const override {
        if (column) {
            std::vector<double> buf(cols());
            apply(in, buf.data(), cols());
            inner->multiply(buf.data(), out);
        } else {
            inner->multiply(in, out);
            apply(out, out, rows());
        }
    }
    void adjoint_multiply(const double* in, double* out) const override {
        if (column) {
            inner->adjoint_multiply(in, out);
            apply(out, out, cols());
        } else {
            std::vector<double> buf(rows());
            apply(in, buf.data(), rows());
            inner->adjoint_multiply(buf.data(), out);
        }
    }
    void realize(Mat& out) const override {
        inner->realize(out);
        for (Index c = 0; c < out.nc; ++c)
            for (Index r = 0; r < out.nr; ++r) {
                const double s = column ? scale[c] : scale[r];
                out(r, c) = divide ? out(r, c) / s : out(r, c) * s;
            }
    }
};

// ---------------------------------------------------------------- options
struct Opts {
    double invariant_subspace_tolerance = -1;   // Options.hpp:27
    double convergence_tolerance = 1e-5;        // :34
    double singular_value_ratio_tolerance = -1; // :41
    int extra_work = 7;                         // :47
    int max_iterations = 1000;                  // :53
    bool exact_for_small_matrix = true;         // :59
    bool exact_for_large_number = true;         // :65
    bool cap_number = false;                    // :71
    std::uint64_t seed = std::mt19937_64::default_seed;  // :76
    const double* initial = nullptr;            // :82
    Index initial_len = 0;
};

// ---------------------------------------------------------------- Lanczos
// lanczos.hpp:17-21: classical Gram-Schmidt, tmp = M[:, :n]' v; v -= M[:, :n] tmp.
static void orthogonalize_vector(const Mat& M, double* v, Index ncols, std::vector<double>& tmp) {
    if (static_cast<Index>(tmp.size()) < ncols) tmp.resize(ncols);
    for (Index j = 0; j < ncols; ++j) tmp[j] = dot(M.col(j), v, M.nr);
    for (Index r = 0; r < M.nr; ++r) {
        double acc = 0;
        for (Index j = 0; j < ncols; ++j) acc += M(r, j) * tmp[j];
        v[r] -= acc;
    }
}

struct LanczosWork {  // lanczos.hpp:23-39
    std::vector<double> F, W_next, tmp;
    explicit LanczosWork(const OMat& m) : F(m.cols()), W_next(m.rows()), tmp(m.cols()) {}
};

// lanczos.hpp:54-143.
static void run_lanczos(const OMat& mat, LanczosWork& lw, Mat& W, Mat& V, Mat& B, std::mt19937_64& eng,
                        Index start, const Opts& o) {
    const double raw = o.invariant_subspace_tolerance;
    const double eps = raw < 0 ? std::pow(std::numeric_limits<double>::epsilon(), 0.8) : raw;
    const Index work = W.nc, nr = mat.rows(), nc = mat.cols();
    auto& F = lw.F;
    auto& Wn = lw.W_next;

    std::copy(V.col(start), V.col(start) + nc, F.begin());
    mat.multiply(F.data(), Wn.data());
    if (start) orthogonalize_vector(W, Wn.data(), start, lw.tmp);
    double S = norm2(Wn.data(), nr);
    if (S < eps) throw std::runtime_error("starting vector near the null space of the input matrix");
    for (auto& w : Wn) w /= S;
    std::copy(Wn.begin(), Wn.end(), W.col(start));

    for (Index j = start; j < work; ++j) {
        mat.adjoint_multiply(Wn.data(), F.data());
        const double* vj = V.col(j);
        for (Index c = 0; c < nc; ++c) F[c] -= S * vj[c];
        orthogonalize_vector(V, F.data(), j + 1, lw.tmp);

        if (j + 1 < work) {
            double R_F = norm2(F.data(), nc);
            if (R_F < eps) {
                fill_with_random_normals(F.data(), nc, eng);
                orthogonalize_vector(V, F.data(), j + 1, lw.tmp);
                R_F = norm2(F.data(), nc);
                for (auto& f : F) f /= R_F;
                R_F = 0;
            } else {
                for (auto& f : F) f /= R_F;
            }
            std::copy(F.begin(), F.end(), V.col(j + 1));
            B(j, j) = S;
            B(j, j + 1) = R_F;

            mat.multiply(F.data(), Wn.data());
            const double* wj = W.col(j);
            for (Index r = 0; r < nr; ++r) Wn[r] -= R_F * wj[r];
            orthogonalize_vector(W, Wn.data(), j + 1, lw.tmp);

            S = norm2(Wn.data(), nr);
            if (S < eps) {
                fill_with_random_normals(Wn.data(), nr, eng);
                orthogonalize_vector(W, Wn.data(), j + 1, lw.tmp);
                S = norm2(Wn.data(), nr);
                for (auto& w : Wn) w /= S;
                S = 0;
            } else {
                for (auto& w : Wn) w /= S;
            }
            std::copy(Wn.begin(), Wn.end(), W.col(j + 1));
        } else {
            B(j, j) = S;
        }
    }
}

// ---------------------------------------------------------------- SVD
// One-sided (Hestenes) Jacobi SVD, a textbook method standing in for Eigen's
// JacobiSVD (compute.hpp:27-28).  Thin: U is m x p, V is n x p, p = min(m, n);
// values in decreasing order.  Left vectors of zero singular values are
// completed to an orthonormal set.
static void jacobi_svd(const Mat& A, Mat& U, std::vector<double>& S, Mat& V) {
    if (A.nr < A.nc) {
        Mat At(A.nc, A.nr);
        for (Index c = 0; c < A.nc; ++c)
            for (Index r = 0; r < A.nr; ++r) At(c, r) = A(r, c);
        jacobi_svd(At, V, S, U);
        return;
    }
    const Index m = A.nr, n = A.nc;
    Mat G = A;
    Mat Vf(n, n);
    for (Index i = 0; i < n; ++i) Vf(i, i) = 1;
    const double tol = 4 * std::numeric_limits<double>::epsilon();
    for (int sweep = 0; sweep < 100; ++sweep) {
        bool rotated = false;
        for (Index p = 0; p + 1 < n; ++p) {
            for (Index q = p + 1; q < n; ++q) {
                const double alpha = dot(G.col(p), G.col(p), m);
                const double beta = dot(G.col(q), G.col(q), m);
                const double gamma = dot(G.col(p), G.col(q), m);
                if (gamma == 0 || std::abs(gamma) <= tol * std::sqrt(alpha * beta)) continue;
                rotated = true;
                const double zeta = (beta - alpha) / (2 * gamma);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::abs(zeta) + std::sqrt(1 + zeta * zeta));
                const double c = 1 / std::sqrt(1 + t * t), s = c * t;
                double* gp = G.col(p);
                double* gq = G.col(q);
                for (Index r = 0; r < m; ++r) {
                    const double a = gp[r], b = gq[r];
                    gp[r] = c * a - s * b;
                    gq[r] = s * a + c * b;
                }
                double* vp = Vf.col(p);
                double* vq = Vf.col(q);
                for (Index r = 0; r < n; ++r) {
                    const double a = vp[r], b = vq[r];
                    vp[r] = c * a - s * b;
                    vq[r] = s * a + c * b;
                }
            }
        }
        if (!rotated) break;
    }
    std::vector<double> sv(n);
    for (Index j = 0; j < n; ++j) sv[j] = norm2(G.col(j), m);
    std::vector<Index> order(n);
    for (Index j = 0; j < n; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](Index a, Index b) { return sv[a] > sv[b]; });
    U = Mat(m, n);
    V = Mat(n, n);
    S.assign(n, 0);
    const double smax = n ? sv[order[0]] : 0;
    std::vector<bool> filled(n, false);
    for (Index j = 0; j < n; ++j) {
        const Index o = order[j];
        S[j] = sv[o];
        std::copy(Vf.col(o), Vf.col(o) + n, V.col(j));
        if (sv[o] > smax * 1e-300 && sv[o] > 0) {
            for (Index r = 0; r < m; ++r) U(r, j) = G(r, o) / sv[o];
            filled[j] = true;
        }
    }
    // Completion of left vectors for zero singular values (Gram-Schmidt on unit vectors).
    Index probe = 0;
    for (Index j = 0; j < n; ++j) {
        if (filled[j]) continue;
        while (probe < m) {
            std::vector<double> e(m, 0.0);
            e[probe++] = 1;
            for (int pass = 0; pass < 2; ++pass)
                for (Index k = 0; k < n; ++k)
                    if (filled[k]) {
                        const double d = dot(U.col(k), e.data(), m);
                        for (Index r = 0; r < m; ++r) e[r] -= d * U(r, k);
                    }
            const double nn = norm2(e.data(), m);
            if (nn > 1e-8) {
                for (Index r = 0; r < m; ++r) U(r, j) = e[r] / nn;
                filled[j] = true;
                break;
            }
        }
    }
}

// ---------------------------------------------------------------- compute helpers
// compute.hpp:49-56.
static bool requested_ge_half(Index requested, Index smaller) {
    const Index half = smaller / 2;
    if (requested == half) return smaller % 2 == 0;
    return requested > half;
}
// compute.hpp:59-67.
static Index choose_work(Index requested, int extra_work, Index smaller) {
    if (requested >= smaller) return smaller;
    return requested + std::min<Index>(smaller - requested, extra_work);
}
// compute.hpp:78-111.
static Index update_k(Index k, Index requested, Index n_converged, Index work) {
    if (work <= 3) return 1;
    const Index limit = work - 3;
    auto less = [&](Index val) { return val < requested || (val - requested) < n_converged; };
    if (less(k)) {
        if (less(limit)) return limit;
        return std::max<Index>(requested + n_converged, 1);
    }
    return std::max<Index>(std::min(k, limit), 1);
}

// compute.hpp:30-46.
static void exact(const OMat& mat, Index requested, Mat& U, Mat& V, std::vector<double>& D) {
    Mat A;
    mat.realize(A);
    Mat u, v;
    std::vector<double> s;
    jacobi_svd(A, u, s, v);
    D.assign(s.begin(), s.begin() + requested);
    U = Mat(u.nr, requested);
    V = Mat(v.nr, requested);
    std::copy(u.a.begin(), u.a.begin() + u.nr * requested, U.a.begin());
    std::copy(v.a.begin(), v.a.begin() + v.nr * requested, V.a.begin());
}

// dst[:, :k] = src[:, :work] * Y[:, :k] (Y is work x ?).
static void gemm_cols(const Mat& src, const Mat& Y, Index k, Mat& dst) {
    std::vector<double> rowbuf(k);
    for (Index r = 0; r < src.nr; ++r) {
        for (Index j = 0; j < k; ++j) {
            double acc = 0;
            for (Index l = 0; l < src.nc; ++l) acc += src(r, l) * Y(l, j);
            rowbuf[j] = acc;
        }
        for (Index j = 0; j < k; ++j) dst(r, j) = rowbuf[j];
    }
}

// compute.hpp:141-308.
static std::pair<bool, int> compute(const OMat& mat, Index number, Mat& outU, Mat& outV, std::vector<double>& outD,
                                    const Opts& o) {
    const Index smaller = std::min(mat.rows(), mat.cols());
    Index requested = number;
    if (requested > smaller) {
        if (o.cap_number) requested = smaller;
        else throw std::runtime_error("requested number of singular values cannot be greater than the smaller matrix dimension");
    } else if (requested == smaller && !o.exact_for_large_number) {
        throw std::runtime_error("requested number of singular values must be less than the smaller matrix dimension for IRLBA iterations");
    }
    if ((o.exact_for_small_matrix && smaller < 6) || (o.exact_for_large_number && requested_ge_half(requested, smaller))) {
        exact(mat, requested, outU, outV, outD);
        return {true, 0};
    }
    const Index work = choose_work(requested, o.extra_work, smaller);
    if (work == 0) throw std::runtime_error("number of requested dimensions must be positive");

    const Index nr = mat.rows(), nc = mat.cols();
    Mat V(nc, work);
    std::mt19937_64 eng(o.seed);
    if (o.initial) {
        if (o.initial_len != nc) throw std::runtime_error("initialization vector does not have expected number of rows");
        std::copy(o.initial, o.initial + nc, V.col(0));
    } else {
        fill_with_random_normals(V.col(0), nc, eng);
    }
    {
        const double n0 = norm2(V.col(0), nc);
        for (Index r = 0; r < nc; ++r) V(r, 0) /= n0;
    }

    bool converged = false;
    int iter = 0;
    Index k = 0;
    LanczosWork lw(mat);
    Mat W(nr, work), Wtmp(nr, work), Vtmp(nc, work), B(work, work);
    std::vector<double> res(work), prevS(work, 0.0);
    const double tol = o.convergence_tolerance;
    const double svtol = o.singular_value_ratio_tolerance >= 0 ? o.singular_value_ratio_tolerance : tol;
    Mat BU, BV;
    std::vector<double> BS;

    for (; iter < o.max_iterations; ++iter) {
        run_lanczos(mat, lw, W, V, B, eng, k, o);
        jacobi_svd(B, BU, BS, BV);
        if (B(work - 1, work - 1) == 0) {
            converged = true;
            break;
        }
        const double R_F = norm2(lw.F.data(), nc);
        for (Index j = 0; j < work; ++j) res[j] = R_F * BU(work - 1, j);

        Index n_conv = 0;
        if (iter > 0) {
            const double Smax = *std::max_element(BS.begin(), BS.end());
            const double thr = Smax * tol;
            for (Index j = 0; j < work; ++j) {
                if (std::abs(res[j]) <= thr) {
                    const double ratio = std::abs(BS[j] - prevS[j]) / BS[j];
                    if (ratio <= svtol) ++n_conv;
                }
            }
            if (n_conv >= requested) {
                converged = true;
                break;
            }
        }
        prevS = BS;

        k = update_k(k, requested, n_conv, work);
        gemm_cols(V, BV, k, Vtmp);
        for (Index j = 0; j < k; ++j) std::copy(Vtmp.col(j), Vtmp.col(j) + nc, V.col(j));
        for (Index r = 0; r < nc; ++r) V(r, k) = lw.F[r] / R_F;
        gemm_cols(W, BU, k, Wtmp);
        for (Index j = 0; j < k; ++j) std::copy(Wtmp.col(j), Wtmp.col(j) + nr, W.col(j));

        std::fill(B.a.begin(), B.a.end(), 0.0);
        for (Index l = 0; l < k; ++l) {
            B(l, l) = BS[l];
            B(l, k) = res[l];
        }
    }

    outD.assign(BS.begin(), BS.begin() + requested);
    outU = Mat(nr, requested);
    gemm_cols(W, BU, requested, outU);
    outV = Mat(nc, requested);
    gemm_cols(V, BV, requested, outV);
    return {converged, converged ? iter + 1 : iter};
}

// pca.hpp:66-109: column means and standard deviations over all rows
// (structural zeros included), SD := 1 when the variance is zero.
static void column_stats(const OMat& mat, bool center, bool scale, std::vector<double>& c, std::vector<double>& s) {
    const Index nr = mat.rows(), nc = mat.cols();
    if (center) {
        if (nr < 1) throw std::runtime_error("cannot center with no observations");
        c.assign(nc, 0);
    }
    if (scale) {
        if (nr < 2) throw std::runtime_error("cannot scale with fewer than two observations");
        s.assign(nc, 0);
    }
    if (!(center || scale)) return;
    std::vector<double> buffer(nr);
    for (Index i = 0; i < nc; ++i) {
        mat.column(i, buffer.data());
        double mean = 0;
        if (center) {
            double sum = 0;
            for (Index r = 0; r < nr; ++r) sum += buffer[r];
            mean = sum / nr;
            c[i] = mean;
        }
        if (scale) {
            double var = 0;
            for (Index r = 0; r < nr; ++r) var += (buffer[r] - mean) * (buffer[r] - mean);
            s[i] = var ? std::sqrt(var / (nr - 1)) : 1.0;
        }
    }
}

// pca.hpp:55-139, generalised from an Eigen input to any OMat.
static std::pair<bool, int> pca(const std::shared_ptr<OMat>& mat, bool center, bool scale, Index number, Mat& scores,
                                Mat& rotation, std::vector<double>& variances, const Opts& o) {
    const Index nr = mat->rows(), nc = mat->cols();
    std::vector<double> c, s;
    column_stats(*mat, center, scale, c, s);
    std::shared_ptr<OMat> ptr = mat;
    if (center) ptr = std::make_shared<OCentered>(ptr, c.data());
    if (scale) ptr = std::make_shared<OScaled>(ptr, s.data(), true, true);
    auto stats = compute(*ptr, number, scores, rotation, variances, o);
    for (Index j = 0; j < scores.nc; ++j)
        for (Index r = 0; r < scores.nr; ++r) scores(r, j) *= variances[j];
    if (nc > 1) {
        for (auto& v : variances) v = v * v / (nr - 1);
    }
    return stats;
}

// ---------------------------------------------------------------- synthetic data
// Counter-based generator shared, integer for integer, with the device
// generator: positions are Bernoulli(density) draws and values are
// log1p(K) with K ~ NB(mu = 2, phi = 0.5), i.e. size 2, prob 1/2, whose pmf
// (k+1)/2^(k+2) has the dyadic CDF 1 - (k+3)/2^(k+2).
static inline std::uint64_t mix64(std::uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline std::uint64_t line_key(std::uint64_t seed, std::uint64_t line, std::uint64_t stream) {
    return mix64(seed ^ mix64(2 * line + stream));
}
static inline std::uint32_t draw32(std::uint64_t key, std::uint64_t pos) {
    return static_cast<std::uint32_t>(mix64(key ^ (pos * 0xD1B54A32D192ED03ULL)) >> 32);
}
static inline std::uint32_t density_threshold(double density) {
    double t = std::floor(density * 4294967296.0);
    if (t < 0) t = 0;
    if (t > 4294967295.0) t = 4294967295.0;
    return static_cast<std::uint32_t>(t);
}
static inline int nb_draw(std::uint32_t u) {
    for (int k = 0; k <= 30; ++k) {
        const std::uint64_t T = (1ULL << 32) - (static_cast<std::uint64_t>(k + 3) << (30 - k));
        if (u < T) return k;
    }
    return 31;
}

}  // namespace oracle

// ===================================================================== C API
using namespace oracle;

namespace {
thread_local std::string last_error;

struct Handle {
    std::shared_ptr<OMat> m;
};

template <class Fn>
int guarded(Fn fn) {
    try {
        fn();
        return 0;
    } catch (const std::exception& e) {
        last_error = e.what();
        return 1;
    }
}
}  // namespace

extern "C" {

typedef struct {
    double invariant_subspace_tolerance;
    double convergence_tolerance;
    double singular_value_ratio_tolerance;
    int32_t extra_work;
    int32_t max_iterations;
    int32_t exact_for_small_matrix;
    int32_t exact_for_large_number;
    int32_t cap_number;
    int32_t pad_;
    uint64_t seed;
    const double* initial;
    int64_t initial_len;
} oracle_options;

static Opts to_opts(const oracle_options* p) {
    Opts o;
    if (!p) return o;
    o.invariant_subspace_tolerance = p->invariant_subspace_tolerance;
    o.convergence_tolerance = p->convergence_tolerance;
    o.singular_value_ratio_tolerance = p->singular_value_ratio_tolerance;
    o.extra_work = p->extra_work;
    o.max_iterations = p->max_iterations;
    o.exact_for_small_matrix = p->exact_for_small_matrix != 0;
    o.exact_for_large_number = p->exact_for_large_number != 0;
    o.cap_number = p->cap_number != 0;
    o.seed = p->seed;
    o.initial = p->initial;
    o.initial_len = p->initial_len;
    return o;
}

const char* oracle_last_error(void) { return last_error.c_str(); }

void* oracle_dense_new(int64_t nr, int64_t nc, const double* a) {
    Mat m(nr, nc);
    std::copy(a, a + nr * nc, m.a.begin());
    return new Handle{std::make_shared<ODense>(std::move(m))};
}

void* oracle_sparse_new(int64_t nr, int64_t nc, int32_t value_f32, const void* values, const int32_t* idx,
                        const int64_t* ptr, int32_t column_major, int32_t nthreads) {
    if (value_f32)
        return new Handle{std::make_shared<OSparse<float> >(nr, nc, static_cast<const float*>(values), idx, ptr,
                                                             column_major != 0, nthreads)};
    return new Handle{std::make_shared<OSparse<double> >(nr, nc, static_cast<const double*>(values), idx, ptr,
                                                          column_major != 0, nthreads)};
}

void* oracle_centered_new(void* inner, const double* center) {
    return new Handle{std::make_shared<OCentered>(static_cast<Handle*>(inner)->m, center)};
}

void* oracle_scaled_new(void* inner, const double* scale, int32_t column, int32_t divide) {
    return new Handle{std::make_shared<OScaled>(static_cast<Handle*>(inner)->m, scale, column != 0, divide != 0)};
}

void oracle_free(void* h) { delete static_cast<Handle*>(h); }

void oracle_dims(void* h, int64_t* nr, int64_t* nc) {
    *nr = static_cast<Handle*>(h)->m->rows();
    *nc = static_cast<Handle*>(h)->m->cols();
}

int oracle_multiply(void* h, int32_t adjoint, const double* in, double* out) {
    return guarded([&] {
        auto& m = static_cast<Handle*>(h)->m;
        if (adjoint) m->adjoint_multiply(in, out); else m->multiply(in, out);
    });
}

int oracle_realize(void* h, double* out) {
    return guarded([&] {
        Mat m;
        static_cast<Handle*>(h)->m->realize(m);
        std::copy(m.a.begin(), m.a.end(), out);
    });
}

// Outputs are sized for `number` columns; *nout receives the count written.
int oracle_compute(void* h, int64_t number, const oracle_options* opt, double* U, double* V, double* D, int64_t* nout,
                   int32_t* converged, int32_t* iterations) {
    return guarded([&] {
        Mat u, v;
        std::vector<double> d;
        auto st = compute(*static_cast<Handle*>(h)->m, number, u, v, d, to_opts(opt));
        std::copy(u.a.begin(), u.a.end(), U);
        std::copy(v.a.begin(), v.a.end(), V);
        std::copy(d.begin(), d.end(), D);
        *nout = static_cast<int64_t>(d.size());
        *converged = st.first;
        *iterations = st.second;
    });
}

int oracle_pca(void* h, int32_t center, int32_t scale, int64_t number, const oracle_options* opt, double* scores,
               double* rotation, double* variances, int64_t* nout, int32_t* converged, int32_t* iterations) {
    return guarded([&] {
        Mat sc, rot;
        std::vector<double> var;
        auto st = pca(static_cast<Handle*>(h)->m, center != 0, scale != 0, number, sc, rot, var, to_opts(opt));
        std::copy(sc.a.begin(), sc.a.end(), scores);
        std::copy(rot.a.begin(), rot.a.end(), rotation);
        std::copy(var.begin(), var.end(), variances);
        *nout = static_cast<int64_t>(var.size());
        *converged = st.first;
        *iterations = st.second;
    });
}

int oracle_column_stats(void* h, int32_t center, int32_t scale, double* c, double* s) {
    return guarded([&] {
        std::vector<double> cc, ss;
        column_stats(*static_cast<Handle*>(h)->m, center != 0, scale != 0, cc, ss);
        if (center) std::copy(cc.begin(), cc.end(), c);
        if (scale) std::copy(ss.begin(), ss.end(), s);
    });
}

// One Lanczos sweep (lanczos.hpp:54-143).  W (nr x work), V (nc x work) and B
// (work x work) are column-major and updated in place; F receives the residual.
int oracle_lanczos(void* h, int64_t work, double* W, double* V, double* B, uint64_t seed, int64_t start,
                   const oracle_options* opt, double* F) {
    return guarded([&] {
        const OMat& m = *static_cast<Handle*>(h)->m;
        Mat w(m.rows(), work), v(m.cols(), work), b(work, work);
        std::copy(W, W + w.a.size(), w.a.begin());
        std::copy(V, V + v.a.size(), v.a.begin());
        std::copy(B, B + b.a.size(), b.a.begin());
        LanczosWork lw(m);
        std::mt19937_64 eng(seed);
        run_lanczos(m, lw, w, v, b, eng, start, to_opts(opt));
        std::copy(w.a.begin(), w.a.end(), W);
        std::copy(v.a.begin(), v.a.end(), V);
        std::copy(b.a.begin(), b.a.end(), B);
        std::copy(lw.F.begin(), lw.F.end(), F);
    });
}

void oracle_orthogonalize(const double* M, int64_t nr, int64_t total_cols, double* vec, int64_t ncols) {
    Mat m(nr, total_cols);
    std::copy(M, M + nr * total_cols, m.a.begin());
    std::vector<double> tmp;
    orthogonalize_vector(m, vec, ncols, tmp);
}

int64_t oracle_update_k(int64_t k, int64_t req, int64_t nconv, int64_t work) { return update_k(k, req, nconv, work); }
int32_t oracle_requested_ge_half(int64_t req, int64_t smaller) { return requested_ge_half(req, smaller); }
int64_t oracle_choose_work(int64_t req, int32_t extra, int64_t smaller) { return choose_work(req, extra, smaller); }

int oracle_svd(int64_t m, int64_t n, const double* A, double* U, double* S, double* V) {
    return guarded([&] {
        Mat a(m, n);
        std::copy(A, A + m * n, a.a.begin());
        Mat u, v;
        std::vector<double> s;
        jacobi_svd(a, u, s, v);
        std::copy(u.a.begin(), u.a.end(), U);
        std::copy(s.begin(), s.end(), S);
        std::copy(v.a.begin(), v.a.end(), V);
    });
}

// Partition tables of sparse.hpp:52-114 (only when nthreads > 1).
int oracle_partition(void* h, int64_t* primary_bounds, int64_t* secondary_bounds, int64_t* secondary_nz) {
    return guarded([&] {
        auto* s64 = dynamic_cast<OSparse<double>*>(static_cast<Handle*>(h)->m.get());
        auto* s32 = dynamic_cast<OSparse<float>*>(static_cast<Handle*>(h)->m.get());
        auto fill = [&](auto* s) {
            std::copy(s->primary_bounds.begin(), s->primary_bounds.end(), primary_bounds);
            std::copy(s->secondary_bounds.begin(), s->secondary_bounds.end(), secondary_bounds);
            int64_t off = 0;
            for (auto& row : s->secondary_nz_bounds) {
                std::copy(row.begin(), row.end(), secondary_nz + off);
                off += static_cast<int64_t>(row.size());
            }
        };
        if (s64) fill(s64);
        else if (s32) fill(s32);
        else throw std::runtime_error("not a sparse matrix");
    });
}

// tests/src/NormalSampler.h:7-32: std::mt19937_64(seed) with std::normal_distribution<double>
// (libstdc++), column by column.
void oracle_normal_sampler(int64_t seed, int64_t n, double* out) {
    std::mt19937_64 eng(static_cast<std::mt19937_64::result_type>(seed));
    std::normal_distribution<> dist;
    for (int64_t i = 0; i < n; ++i) out[i] = dist(eng);
}

// tests/src/sparse.h:19-44 (simulate_compressed_sparse): seed nr*nc, Bernoulli(0.2)
// positions, N(0,1) values, CSC.  Call with values == NULL to get the nnz count.
int64_t oracle_simulate_compressed_sparse(int64_t nr, int64_t nc, double* values, int32_t* indices, int64_t* nzeros) {
    std::mt19937_64 rng(static_cast<std::uint64_t>(nr * nc));
    std::uniform_real_distribution<double> udist(0.0, 1.0);
    std::normal_distribution<double> ndist;
    int64_t count = 0;
    if (nzeros) nzeros[0] = 0;
    for (int64_t c = 0; c < nc; ++c) {
        for (int64_t r = 0; r < nr; ++r) {
            if (udist(rng) < 0.2) {
                const double v = ndist(rng);
                if (values) {
                    values[count] = v;
                    indices[count] = static_cast<int32_t>(r);
                }
                ++count;
            }
        }
        if (nzeros) nzeros[c + 1] = count;
    }
    return count;
}

uint64_t oracle_mt19937_64_nth(uint64_t seed, int64_t n) {
    std::mt19937_64 eng(seed);
    uint64_t v = 0;
    for (int64_t i = 0; i < n; ++i) v = eng();
    return v;
}

// fill_with_random_normals (utils.hpp:16-37) from mt19937_64(seed).
void oracle_fill_random_normals(uint64_t seed, int64_t n, double* out) {
    std::mt19937_64 eng(seed);
    fill_with_random_normals(out, n, eng);
}

// Synthetic CSC: lines [line0, line1) of a (n_minor x n_major) matrix.  counts
// receives the nnz per line; if values != NULL the entries are written too
// (values as float when value_f32, else double), rows ascending within a line.
int64_t oracle_gen_nb_sparse(int64_t n_minor, int64_t line0, int64_t line1, double density, uint64_t seed,
                             int64_t* counts, int32_t value_f32, void* values, int32_t* indices) {
    const std::uint32_t thr = density_threshold(density);
    double table[32];
    for (int k = 0; k < 32; ++k) table[k] = std::log1p(static_cast<double>(k));
    int64_t pos = 0;
    for (int64_t c = line0; c < line1; ++c) {
        const std::uint64_t kp = line_key(seed, c, 0), kv = line_key(seed, c, 1);
        int64_t cnt = 0;
        for (int64_t r = 0; r < n_minor; ++r) {
            if (draw32(kp, r) < thr) {
                if (values) {
                    const double v = table[nb_draw(draw32(kv, r))];
                    if (value_f32) static_cast<float*>(values)[pos] = static_cast<float>(v);
                    else static_cast<double*>(values)[pos] = v;
                    indices[pos] = static_cast<int32_t>(r);
                }
                ++pos;
                ++cnt;
            }
        }
        if (counts) counts[c - line0] = cnt;
    }
    return pos;
}

int32_t oracle_nb_draw(uint32_t u) { return nb_draw(u); }
uint32_t oracle_density_threshold(double d) { return density_threshold(d); }

}  // extern "C"
