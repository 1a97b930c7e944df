#ifndef IRLBA_HIP_MATRIX_CENTERED_HPP
#define IRLBA_HIP_MATRIX_CENTERED_HPP

// CenteredMatrix (include/irlba/Matrix/centered.hpp:139-204 of the reference):
// A - 1 c' for a column centre c, kept deferred.  Over a device-resident
// matrix the decorator is folded into the device operator (the centring term
// is applied inside the GPU products); over any other Matrix it composes the
// inner workspaces on the host as the reference does
// (centered.hpp:46-52, 83-87, 118-122).

#include <memory>
#include <type_traits>

#include "hip.hpp"
#include "interface.hpp"

namespace irlba {

namespace detail {

template <class EigenVector_, class Inner_, class Center_>
class CenteredWorkspace final : public Workspace<EigenVector_> {
public:
    CenteredWorkspace(const Inner_& m, const Center_& c) : my_work(m.new_known_workspace()), my_center(c) {}
    void multiply(const EigenVector_& right, EigenVector_& out) {
        my_work->multiply(right, out);
        double beta = 0;
        for (Index i = 0; i < right.size(); ++i) beta += right[i] * my_center[i];
        for (Index r = 0; r < out.size(); ++r) out[r] -= beta;
    }

private:
    decltype(std::declval<const Inner_&>().new_known_workspace()) my_work;
    const Center_& my_center;
};

template <class EigenVector_, class Inner_, class Center_>
class CenteredAdjointWorkspace final : public AdjointWorkspace<EigenVector_> {
public:
    CenteredAdjointWorkspace(const Inner_& m, const Center_& c) : my_work(m.new_known_adjoint_workspace()), my_center(c) {}
    void multiply(const EigenVector_& right, EigenVector_& out) {
        my_work->multiply(right, out);
        double beta = 0;
        for (Index r = 0; r < right.size(); ++r) beta += right[r];
        for (Index c = 0; c < out.size(); ++c) out[c] -= beta * my_center[c];
    }

private:
    decltype(std::declval<const Inner_&>().new_known_adjoint_workspace()) my_work;
    const Center_& my_center;
};

template <class EigenMatrix_, class Inner_, class Center_>
class CenteredRealizeWorkspace final : public RealizeWorkspace<EigenMatrix_> {
public:
    CenteredRealizeWorkspace(const Inner_& m, const Center_& c) : my_work(m.new_known_realize_workspace()), my_center(c) {}
    const EigenMatrix_& realize(EigenMatrix_& buffer) {
        my_work->realize_copy(buffer);
        for (Index c = 0; c < buffer.cols(); ++c)
            for (Index r = 0; r < buffer.rows(); ++r) buffer(r, c) -= my_center[c];
        return buffer;
    }

private:
    decltype(std::declval<const Inner_&>().new_known_realize_workspace()) my_work;
    const Center_& my_center;
};

}  // namespace detail

template <class EigenVector_ = VectorXd, class EigenMatrix_ = MatrixXd,
          class MatrixPointer_ = const Matrix<EigenVector_, EigenMatrix_>*, class CenterPointer_ = const EigenVector_*>
class CenteredMatrix final : public Matrix<EigenVector_, EigenMatrix_>, public DeviceResident {
public:
    CenteredMatrix(MatrixPointer_ matrix, CenterPointer_ center)
        : my_matrix(std::move(matrix)), my_center(std::move(center)) {
        const auto* dev = dynamic_cast<const DeviceResident*>(&(*my_matrix));
        const irlba_hip_matrix* inner = dev ? dev->device_handle() : nullptr;
        if (inner) {
            irlba_hip_matrix* h = nullptr;
            hip_check(irlba_hip_centered_create(inner, (*my_center).data(), 0, &h));
            my_handle = detail::own(h);
        }
    }

    Index rows() const { return my_matrix->rows(); }
    Index cols() const { return my_matrix->cols(); }
    const irlba_hip_matrix* device_handle() const { return my_handle.get(); }

    std::unique_ptr<Workspace<EigenVector_> > new_workspace() const {
        if (my_handle) return std::make_unique<detail::HandleWorkspace<EigenVector_> >(my_handle.get());
        return std::make_unique<detail::CenteredWorkspace<EigenVector_, Inner, Center> >(*my_matrix, *my_center);
    }
    std::unique_ptr<AdjointWorkspace<EigenVector_> > new_adjoint_workspace() const {
        if (my_handle) return std::make_unique<detail::HandleAdjointWorkspace<EigenVector_> >(my_handle.get());
        return std::make_unique<detail::CenteredAdjointWorkspace<EigenVector_, Inner, Center> >(*my_matrix, *my_center);
    }
    std::unique_ptr<RealizeWorkspace<EigenMatrix_> > new_realize_workspace() const {
        if (my_handle)
            return std::make_unique<detail::HandleRealizeWorkspace<EigenMatrix_> >(my_handle.get(), rows(), cols());
        return std::make_unique<detail::CenteredRealizeWorkspace<EigenMatrix_, Inner, Center> >(*my_matrix, *my_center);
    }

private:
    typedef typename std::decay<decltype(*std::declval<MatrixPointer_>())>::type Inner;
    typedef typename std::decay<decltype(*std::declval<CenterPointer_>())>::type Center;
    MatrixPointer_ my_matrix;
    CenterPointer_ my_center;
    detail::SharedHandle my_handle;
};

}  // namespace irlba

#endif
