// C++ test of the irlba:: header API over the GPU engine, in the style of the
// reference's tests (tests/src/compute.cpp, pca.cpp, Matrix/*.cpp): the same
// solve through different Matrix implementations must agree.  Exits non-zero
// on the first failed check.  Needs a GPU; built by __graft_entry__.build().

#include <cmath>
#include <cstdio>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "irlba/irlba.hpp"

namespace {

int failures = 0;

void check(bool ok, const char* what) {
    std::printf("%s %s\n", ok ? "ok  " : "FAIL", what);
    if (!ok) ++failures;
}

bool same_same(double a, double b, double tol) { return std::abs(a - b) <= (std::abs(a) + std::abs(b)) * tol; }

irlba::MatrixXd random_matrix(irlba::Index nr, irlba::Index nc, unsigned seed) {
    std::mt19937_64 eng(seed);
    std::normal_distribution<double> nd;
    irlba::MatrixXd A(nr, nc);
    for (irlba::Index c = 0; c < nc; ++c)
        for (irlba::Index r = 0; r < nr; ++r) A(r, c) = nd(eng);
    return A;
}

// A caller-defined Matrix whose products run on the host (the reference's
// plugin interface); compute() drives it from the GPU through callbacks.
class HostDense final : public irlba::Matrix<irlba::VectorXd, irlba::MatrixXd> {
public:
    explicit HostDense(const irlba::MatrixXd& a) : A(a) {}
    irlba::Index rows() const { return A.rows(); }
    irlba::Index cols() const { return A.cols(); }
    struct Fwd final : irlba::Workspace<irlba::VectorXd> {
        const irlba::MatrixXd& A;
        explicit Fwd(const irlba::MatrixXd& a) : A(a) {}
        void multiply(const irlba::VectorXd& x, irlba::VectorXd& y) {
            for (irlba::Index r = 0; r < A.rows(); ++r) y[r] = 0;
            for (irlba::Index c = 0; c < A.cols(); ++c)
                for (irlba::Index r = 0; r < A.rows(); ++r) y[r] += A(r, c) * x[c];
        }
    };
    struct Adj final : irlba::AdjointWorkspace<irlba::VectorXd> {
        const irlba::MatrixXd& A;
        explicit Adj(const irlba::MatrixXd& a) : A(a) {}
        void multiply(const irlba::VectorXd& x, irlba::VectorXd& y) {
            for (irlba::Index c = 0; c < A.cols(); ++c) {
                double s = 0;
                for (irlba::Index r = 0; r < A.rows(); ++r) s += A(r, c) * x[r];
                y[c] = s;
            }
        }
    };
    struct Real final : irlba::RealizeWorkspace<irlba::MatrixXd> {
        const irlba::MatrixXd& A;
        explicit Real(const irlba::MatrixXd& a) : A(a) {}
        const irlba::MatrixXd& realize(irlba::MatrixXd&) { return A; }
    };
    std::unique_ptr<irlba::Workspace<irlba::VectorXd> > new_workspace() const { return std::make_unique<Fwd>(A); }
    std::unique_ptr<irlba::AdjointWorkspace<irlba::VectorXd> > new_adjoint_workspace() const {
        return std::make_unique<Adj>(A);
    }
    std::unique_ptr<irlba::RealizeWorkspace<irlba::MatrixXd> > new_realize_workspace() const {
        return std::make_unique<Real>(A);
    }

private:
    irlba::MatrixXd A;
};

// A user exception type, to check that compute() rethrows the object itself.
struct UserError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Host dense matrix whose forward product throws on its `fail_at`-th call.
class ThrowingDense final : public irlba::Matrix<irlba::VectorXd, irlba::MatrixXd> {
public:
    ThrowingDense(const irlba::MatrixXd& a, int fail_at) : inner(a), fail(fail_at) {}
    irlba::Index rows() const { return inner.rows(); }
    irlba::Index cols() const { return inner.cols(); }
    struct Fwd final : irlba::Workspace<irlba::VectorXd> {
        std::unique_ptr<irlba::Workspace<irlba::VectorXd> > base;
        int fail, calls = 0;
        Fwd(std::unique_ptr<irlba::Workspace<irlba::VectorXd> > b, int f) : base(std::move(b)), fail(f) {}
        void multiply(const irlba::VectorXd& x, irlba::VectorXd& y) {
            if (++calls == fail) throw UserError("user product failed after " + std::to_string(fail) + " calls");
            base->multiply(x, y);
        }
    };
    std::unique_ptr<irlba::Workspace<irlba::VectorXd> > new_workspace() const {
        return std::make_unique<Fwd>(inner.new_workspace(), fail);
    }
    std::unique_ptr<irlba::AdjointWorkspace<irlba::VectorXd> > new_adjoint_workspace() const {
        return inner.new_adjoint_workspace();
    }
    std::unique_ptr<irlba::RealizeWorkspace<irlba::MatrixXd> > new_realize_workspace() const {
        return inner.new_realize_workspace();
    }

private:
    HostDense inner;
    int fail;
};

}  // namespace

int main() {
    const irlba::Index nr = 120, nc = 80;
    irlba::MatrixXd A = random_matrix(nr, nc, 42);
    irlba::Options<> opt;
    opt.convergence_tolerance = 1e-9;
    irlba::VectorXd init(nc);
    {
        std::mt19937_64 eng(7);
        std::normal_distribution<double> nd;
        for (auto& v : init) v = nd(eng);
    }
    opt.initial = init;

    // Device dense vs caller-defined host matrix (same algorithm, same start).
    irlba::HipDenseMatrix<> dense(A);
    auto r1 = irlba::compute(dense, 6, opt);
    HostDense host(A);
    auto r2 = irlba::compute(host, 6, opt);
    bool ok = r1.D.size() == 6 && r2.D.size() == 6;
    for (irlba::Index i = 0; ok && i < 6; ++i) ok = same_same(r1.D[i], r2.D[i], 1e-10);
    check(ok && r1.converged && r2.converged, "compute: HipDenseMatrix == host Matrix subclass");

    // Sparse (CSC) vs dense of the same values, with the reference's own array
    // types and constructor (tests/src/sparse.h:11-16; sparse.hpp:469).
    std::vector<double> x;
    std::vector<int> idx;
    std::vector<size_t> ptr(nc + 1, 0);
    for (irlba::Index c = 0; c < nc; ++c) {
        for (irlba::Index r = 0; r < nr; ++r)
            if ((r * 7 + c * 3) % 5 == 0) {
                x.push_back(A(r, c));
                idx.push_back(static_cast<int32_t>(r));
            }
        ptr[c + 1] = static_cast<int64_t>(x.size());
    }
    irlba::MatrixXd S(nr, nc);
    for (irlba::Index c = 0; c < nc; ++c)
        for (size_t e = ptr[c]; e < ptr[c + 1]; ++e) S(idx[e], c) = x[e];
    const int num_threads = 3;
    irlba::HipSparseMatrix<irlba::VectorXd, irlba::MatrixXd, std::vector<double>, std::vector<int>, std::vector<size_t> >
        sparse(nr, nc, x, idx, ptr, true, num_threads);
    irlba::HipDenseMatrix<> sdense(S);
    auto r3 = irlba::compute(sparse, 5, opt);
    auto r4 = irlba::compute(sdense, 5, opt);
    ok = true;
    for (irlba::Index i = 0; i < 5; ++i) ok = ok && same_same(r3.D[i], r4.D[i], 1e-8);
    check(ok, "compute: HipSparseMatrix == HipDenseMatrix");
    check(sparse.get_values() == x && sparse.get_indices() == idx && sparse.get_pointers() == ptr,
          "HipSparseMatrix getters return the constructor's arrays");

    // The thread partition getters (sparse.hpp:527-552), with the invariants of
    // tests/src/Matrix/sparse.cpp:56-89: boundaries run from 0 to the dimension,
    // each thread's primary range holds its share of the non-zeros, and the
    // per-primary non-zero sub-ranges cover exactly the entries whose secondary
    // index lies in the thread's secondary range.
    {
        const auto& pb = sparse.get_primary_boundaries();
        const auto& sb = sparse.get_secondary_boundaries();
        const auto& nzb = sparse.get_secondary_nonzero_boundaries();
        bool pok = pb.size() == num_threads + 1u && sb.size() == num_threads + 1u && nzb.size() == num_threads + 1u;
        pok = pok && pb.front() == 0 && pb.back() == nc && sb.front() == 0 && sb.back() == nr;
        const size_t total = ptr[nc];
        for (int t = 0; pok && t < num_threads; ++t) {
            pok = pb[t] <= pb[t + 1] && sb[t] <= sb[t + 1];
            // A primary boundary sits where the running count first reaches the
            // thread's cumulative share (floor(total/T), the first total%T one more).
            size_t share = 0;
            for (int u = 0; u <= t; ++u) share += total / num_threads + (static_cast<size_t>(u) < total % num_threads);
            pok = pok && ptr[pb[t + 1]] <= share;
            for (irlba::Index c = 0; pok && c < nc; ++c) {
                for (size_t e = nzb[t][c]; e < nzb[t + 1][c]; ++e) pok = pok && idx[e] >= sb[t] && idx[e] < sb[t + 1];
            }
        }
        for (irlba::Index c = 0; pok && c < nc; ++c) pok = nzb[0][c] == ptr[c] && nzb[num_threads][c] == ptr[c + 1];
        check(pok, "HipSparseMatrix partition getters follow the reference's nnz-balanced split");
        irlba::HipSparseMatrix<irlba::VectorXd, irlba::MatrixXd, std::vector<double>, std::vector<int>,
                               std::vector<size_t> > single(nr, nc, x, idx, ptr, true, 1);
        check(single.get_primary_boundaries().empty() && single.get_secondary_nonzero_boundaries().empty(),
              "HipSparseMatrix with one thread leaves the partition empty");
    }

    // Integer values (e.g. counts) and 64-bit indices are converted on the host.
    {
        std::vector<int> xi;
        std::vector<long> idx64(idx.begin(), idx.end());
        for (size_t e = 0; e < x.size(); ++e) xi.push_back(static_cast<int>(e % 7));
        irlba::MatrixXd Si(nr, nc);
        for (irlba::Index c = 0; c < nc; ++c)
            for (size_t e = ptr[c]; e < ptr[c + 1]; ++e) Si(idx[e], c) = xi[e];
        irlba::HipSparseMatrix<irlba::VectorXd, irlba::MatrixXd, std::vector<int>, std::vector<long>,
                               std::vector<size_t> > isparse(nr, nc, xi, idx64, ptr, true, 2);
        irlba::VectorXd vv(nc), y1(nr), y2(nr);
        for (irlba::Index i = 0; i < nc; ++i) vv[i] = std::cos(0.7 * i);
        isparse.new_workspace()->multiply(vv, y1);
        for (irlba::Index r = 0; r < nr; ++r) {
            double acc = 0;
            for (irlba::Index c = 0; c < nc; ++c) acc += Si(r, c) * vv[c];
            y2[r] = acc;
        }
        bool iok = true;
        for (irlba::Index r = 0; r < nr; ++r) iok = iok && std::abs(y1[r] - y2[r]) <= 1e-12 * (1 + std::abs(y2[r]));
        check(iok, "HipSparseMatrix over integer values and 64-bit indices");
        // Counts take few distinct values: the products can stream dictionary
        // codes instead (same terms; FP64 sums in another order).
        const int ncodes = isparse.use_value_dictionary();
        irlba::VectorXd y3(nr);
        isparse.new_workspace()->multiply(vv, y3);
        bool dok = ncodes == 7;
        for (irlba::Index r = 0; r < nr; ++r) dok = dok && std::abs(y3[r] - y1[r]) <= 1e-14 * (1 + std::abs(y1[r]));
        dok = dok && isparse.use_value_dictionary(false) == 0;
        check(dok, "HipSparseMatrix value dictionary: same products");
        std::string dmsg;
        try {
            sparse.use_value_dictionary();
        } catch (const std::exception& e) {
            dmsg = e.what();
        }
        check(dmsg.find("more than 256 distinct values") != std::string::npos,
              "HipSparseMatrix value dictionary refuses values with many distinct patterns");
        std::string emsg;
        std::vector<long> bad(idx64);
        bad[0] = static_cast<long>(nr) + (1L << 32);  // would wrap into range as int32
        try {
            irlba::HipSparseMatrix<irlba::VectorXd, irlba::MatrixXd, std::vector<int>, std::vector<long>,
                                   std::vector<size_t> > b(nr, nc, xi, bad, ptr, true, 1);
        } catch (const std::exception& e) {
            emsg = e.what();
        }
        check(emsg.find("within the secondary dimension") != std::string::npos,
              "HipSparseMatrix rejects a 64-bit index beyond the matrix");
    }

    // Workspaces (interface.hpp): products of decorators over device and host matrices agree.
    irlba::VectorXd c(nc), s(nc), v(nc), o1(nr), o2(nr), u(nr), z1(nc), z2(nc);
    for (irlba::Index i = 0; i < nc; ++i) {
        c[i] = 0.1 * i;
        s[i] = 1.0 + 0.01 * i;
        v[i] = std::sin(0.3 * i);
    }
    for (irlba::Index i = 0; i < nr; ++i) u[i] = std::cos(0.2 * i);
    irlba::CenteredMatrix<> cd(&dense, &c), ch(&host, &c);
    irlba::ScaledMatrix<> sd(&cd, &s, true, true), sh(&ch, &s, true, true);
    check(sd.device_handle() != nullptr && sh.device_handle() == nullptr, "decorators: device folding only over device matrices");
    sd.new_workspace()->multiply(v, o1);
    sh.new_workspace()->multiply(v, o2);
    sd.new_adjoint_workspace()->multiply(u, z1);
    sh.new_adjoint_workspace()->multiply(u, z2);
    ok = true;
    for (irlba::Index i = 0; i < nr; ++i) ok = ok && std::abs(o1[i] - o2[i]) <= 1e-12 * (1 + std::abs(o2[i]));
    for (irlba::Index i = 0; i < nc; ++i) ok = ok && std::abs(z1[i] - z2[i]) <= 1e-12 * (1 + std::abs(z2[i]));
    check(ok, "Scaled(Centered()) products: device == host composition");
    auto r5 = irlba::compute(sd, 4, opt);
    auto r6 = irlba::compute(sh, 4, opt);
    ok = true;
    for (irlba::Index i = 0; i < 4; ++i) ok = ok && same_same(r5.D[i], r6.D[i], 1e-10);
    check(ok, "compute on decorated device matrix == on decorated host matrix");

    // pca(): pca.cpp:66-141 identity with an explicitly centred and scaled copy.
    auto p = irlba::pca(A, true, true, 5, opt);
    irlba::MatrixXd B = A;
    for (irlba::Index j = 0; j < nc; ++j) {
        double m = 0;
        for (irlba::Index r = 0; r < nr; ++r) m += A(r, j);
        m /= nr;
        double ss = 0;
        for (irlba::Index r = 0; r < nr; ++r) ss += (A(r, j) - m) * (A(r, j) - m);
        const double sd_ = std::sqrt(ss / (nr - 1));
        for (irlba::Index r = 0; r < nr; ++r) B(r, j) = (A(r, j) - m) / sd_;
    }
    irlba::HipDenseMatrix<> bd(B);
    auto r7 = irlba::compute(bd, 5, opt);
    ok = p.variances.size() == 5;
    for (irlba::Index i = 0; ok && i < 5; ++i) ok = same_same(p.variances[i], r7.D[i] * r7.D[i] / (nr - 1), 1e-8);
    double colsum = 0;
    for (irlba::Index r = 0; r < nr; ++r) colsum += p.scores(r, 0);
    check(ok && std::abs(colsum) < 1e-8, "pca(center, scale) == compute on the standardised copy");
    auto ps = irlba::pca(sparse, true, true, 4, opt);
    auto pd = irlba::pca(S, true, true, 4, opt);
    ok = true;
    for (irlba::Index i = 0; i < 4; ++i) ok = ok && same_same(ps.variances[i], pd.variances[i], 1e-8);
    check(ok, "pca: HipSparseMatrix == dense host matrix");

    // Errors keep the reference's messages (compute.hpp:156, pca.hpp:80).
    std::string msg;
    try {
        irlba::compute(dense, 1000, irlba::Options<>());
    } catch (const std::exception& e) {
        msg = e.what();
    }
    check(msg.find("cannot be greater than") != std::string::npos, "error: requested number too large");
    msg.clear();
    try {
        irlba::MatrixXd one = random_matrix(1, 10, 3);
        irlba::pca(one, true, true, 1, irlba::Options<>());
    } catch (const std::exception& e) {
        msg = e.what();
    }
    check(msg.find("cannot scale") != std::string::npos, "error: cannot scale");

    // A caller-defined Matrix whose product throws: compute() ends with the
    // caller's own exception, as the reference's direct workspace calls give
    // (lanczos.hpp:74,93,115).
    {
        std::string emsg;
        bool right_type = false;
        ThrowingDense thrower(A, 3);
        try {
            irlba::compute(thrower, 4, opt);
        } catch (const UserError& e) {
            right_type = true;
            emsg = e.what();
        } catch (const std::exception& e) {
            emsg = e.what();
        }
        check(right_type && emsg == "user product failed after 3 calls",
              "compute: a caller-defined product's exception reaches the caller intact");
        auto again = irlba::compute(host, 6, opt);
        bool aok = again.D.size() == 6;
        for (irlba::Index i = 0; aok && i < 6; ++i) aok = same_same(again.D[i], r2.D[i], 1e-12);
        check(aok, "compute: the context works after a caller exception");
    }

    std::printf("%d failure(s)\n", failures);
    return failures == 0 ? 0 : 1;
}
