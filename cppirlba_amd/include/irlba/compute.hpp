#ifndef IRLBA_HIP_COMPUTE_HPP
#define IRLBA_HIP_COMPUTE_HPP

// irlba::compute() / compute_simple() / Results (include/irlba/compute.hpp:141-433
// of the reference).  A device-resident matrix (HipSparseMatrix, HipDenseMatrix
// and decorators over them) runs entirely on the GPU; any other Matrix subclass
// keeps its own products on the host while the Lanczos bases, Gram-Schmidt and
// restart rotations run on the GPU (irlba_hip_compute_callback).

#include <algorithm>
#include <exception>
#include <memory>
#include <stdexcept>
#include <utility>

#include "Matrix/hip.hpp"
#include "Matrix/interface.hpp"
#include "Options.hpp"

namespace irlba {

namespace detail {

template <class EigenVector_>
struct COptions {
    irlba_hip_options o;
    explicit COptions(const Options<EigenVector_>& opt) {
        irlba_hip_options_default(&o);
        o.invariant_subspace_tolerance = opt.invariant_subspace_tolerance;
        o.convergence_tolerance = opt.convergence_tolerance;
        o.singular_value_ratio_tolerance = opt.singular_value_ratio_tolerance;
        o.extra_work = opt.extra_work;
        o.max_iterations = opt.max_iterations;
        o.exact_for_small_matrix = opt.exact_for_small_matrix;
        o.exact_for_large_number = opt.exact_for_large_number;
        o.cap_number = opt.cap_number;
        o.seed = opt.seed;
        if (opt.initial.has_value()) {
            o.initial = opt.initial->data();
            o.initial_length = opt.initial->size();
        }
    }
};

// Host products of a caller-defined Matrix, called back from the GPU engine.
// An exception thrown by the caller's workspace cannot cross the C ABI, so it
// is kept here and rethrown by compute() once the engine has returned: the
// caller sees its own exception, as when the reference calls the workspace
// directly (lanczos.hpp:74,93,115).
template <class EigenVector_, class Matrix_>
struct HostProducts {
    const Matrix_& matrix;
    std::unique_ptr<Workspace<EigenVector_> > fwd;
    std::unique_ptr<AdjointWorkspace<EigenVector_> > adj;
    EigenVector_ in, out;
    std::exception_ptr error;
    explicit HostProducts(const Matrix_& m) : matrix(m), fwd(m.new_workspace()), adj(m.new_adjoint_workspace()) {}

    static int call(void* user, int adjoint, const double* right, double* output) {
        auto* self = static_cast<HostProducts*>(user);
        try {
            const Index nin = adjoint ? self->matrix.rows() : self->matrix.cols();
            const Index nout = adjoint ? self->matrix.cols() : self->matrix.rows();
            self->in.resize(nin);
            self->out.resize(nout);
            std::copy(right, right + nin, self->in.data());
            if (adjoint) self->adj->multiply(self->in, self->out);
            else self->fwd->multiply(self->in, self->out);
            std::copy(self->out.data(), self->out.data() + nout, output);
            return 0;
        } catch (...) {
            if (!self->error) self->error = std::current_exception();
            return 1;
        }
    }
};

}  // namespace detail

// compute(matrix, number, outU, outV, outD, options) (compute.hpp:141-308).
template <class Matrix_, class EigenMatrix_, class EigenVector_>
std::pair<bool, int> compute(const Matrix_& matrix, Index number, EigenMatrix_& outU, EigenMatrix_& outV,
                             EigenVector_& outD, const Options<EigenVector_>& options) {
    detail::COptions<EigenVector_> co(options);
    const Index nr = matrix.rows(), nc = matrix.cols();
    const Index alloc = std::max<Index>(0, std::min<Index>(number, std::min(nr, nc)));
    const Index n = options.cap_number ? alloc : std::max<Index>(number, 0);
    outU.resize(nr, n);
    outV.resize(nc, n);
    outD.resize(std::max<Index>(n, 1));
    std::int64_t nout = 0;
    std::int32_t conv = 0, iters = 0;
    const auto* dev = dynamic_cast<const DeviceResident*>(&matrix);
    if (dev && dev->device_handle()) {
        hip_check(irlba_hip_compute(dev->device_handle(), number, &co.o, outU.data(), outV.data(), outD.data(), &nout,
                                    &conv, &iters, 0));
    } else {
        detail::HostProducts<EigenVector_, Matrix_> hp(matrix);
        const int rc = irlba_hip_compute_callback(default_hip_context().get(), nr, nc,
                                                  &detail::HostProducts<EigenVector_, Matrix_>::call, &hp, number,
                                                  &co.o, outU.data(), outV.data(), outD.data(), &nout, &conv, &iters);
        if (hp.error) std::rethrow_exception(hp.error);
        hip_check(rc);
    }
    outU.resize(nr, nout);
    outV.resize(nc, nout);
    outD.resize(nout);
    return std::make_pair(conv != 0, static_cast<int>(iters));
}

// compute.hpp:357-388
template <class EigenMatrix_, class EigenVector_>
struct Results {
    EigenMatrix_ U;
    EigenMatrix_ V;
    EigenVector_ D;
    int iterations;
    bool converged;
};

// compute.hpp:403-410
template <class EigenMatrix_ = MatrixXd, class EigenVector_ = VectorXd, class Matrix_>
Results<EigenMatrix_, EigenVector_> compute(const Matrix_& matrix, Index number, const Options<EigenVector_>& options) {
    Results<EigenMatrix_, EigenVector_> out;
    const auto st = compute(matrix, number, out.U, out.V, out.D, options);
    out.converged = st.first;
    out.iterations = st.second;
    return out;
}

// compute_simple(matrix, number, ...) (compute.hpp:333-350, 426-433) for a
// dense column-major host matrix, run as an FP64 HipDenseMatrix.
template <class EigenMatrix_ = MatrixXd, class EigenVector_ = VectorXd>
std::pair<bool, int> compute_simple(const MatrixXd& matrix, Index number, EigenMatrix_& outU, EigenMatrix_& outV,
                                    EigenVector_& outD, const Options<EigenVector_>& options) {
    HipDenseMatrix<EigenVector_, EigenMatrix_> wrapped(matrix);
    return compute(wrapped, number, outU, outV, outD, options);
}

template <class EigenMatrix_ = MatrixXd, class EigenVector_ = VectorXd>
Results<EigenMatrix_, EigenVector_> compute_simple(const MatrixXd& matrix, Index number,
                                                   const Options<EigenVector_>& options) {
    Results<EigenMatrix_, EigenVector_> out;
    const auto st = compute_simple(matrix, number, out.U, out.V, out.D, options);
    out.converged = st.first;
    out.iterations = st.second;
    return out;
}

}  // namespace irlba

#endif
