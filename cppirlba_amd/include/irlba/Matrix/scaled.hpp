#ifndef IRLBA_HIP_MATRIX_SCALED_HPP
#define IRLBA_HIP_MATRIX_SCALED_HPP

// ScaledMatrix (include/irlba/Matrix/scaled.hpp:195-268 of the reference):
// columns (column = true) or rows scaled by `scale`, divided (divide = true)
// or multiplied.  Folded into the device operator over a device-resident
// matrix; composed on the host otherwise (scaled.hpp:52-69, 106-123, 158-177).

#include <memory>
#include <type_traits>

#include "hip.hpp"
#include "interface.hpp"

namespace irlba {

namespace detail {

template <class EigenVector_, class Scale_>
inline void scale_into(const EigenVector_& in, EigenVector_& out, const Scale_& s, bool divide) {
    out.resize(in.size());
    for (Index i = 0; i < in.size(); ++i) out[i] = divide ? in[i] / s[i] : in[i] * s[i];
}

template <class EigenVector_, class Inner_, class Scale_, bool Adjoint>
class ScaledWorkspaceT final
    : public std::conditional<Adjoint, AdjointWorkspace<EigenVector_>, Workspace<EigenVector_> >::type {
public:
    ScaledWorkspaceT(const Inner_& m, const Scale_& s, bool column, bool divide)
        : my_scale(s), my_column(column), my_divide(divide) {
        if constexpr (Adjoint) my_adj = m.new_known_adjoint_workspace();
        else my_fwd = m.new_known_workspace();
    }
    void multiply(const EigenVector_& right, EigenVector_& out) {
        // Forward: column scaling acts on the input, row scaling on the output;
        // the adjoint swaps the two.
        const bool on_input = Adjoint ? !my_column : my_column;
        if (on_input) {
            scale_into(right, my_buffer, my_scale, my_divide);
            run(my_buffer, out);
        } else {
            run(right, out);
            for (Index i = 0; i < out.size(); ++i) out[i] = my_divide ? out[i] / my_scale[i] : out[i] * my_scale[i];
        }
    }

private:
    void run(const EigenVector_& in, EigenVector_& out) {
        if constexpr (Adjoint) my_adj->multiply(in, out);
        else my_fwd->multiply(in, out);
    }
    decltype(std::declval<const Inner_&>().new_known_workspace()) my_fwd;
    decltype(std::declval<const Inner_&>().new_known_adjoint_workspace()) my_adj;
    const Scale_& my_scale;
    bool my_column, my_divide;
    EigenVector_ my_buffer;
};

template <class EigenMatrix_, class Inner_, class Scale_>
class ScaledRealizeWorkspace final : public RealizeWorkspace<EigenMatrix_> {
public:
    ScaledRealizeWorkspace(const Inner_& m, const Scale_& s, bool column, bool divide)
        : my_work(m.new_known_realize_workspace()), my_scale(s), my_column(column), my_divide(divide) {}
    const EigenMatrix_& realize(EigenMatrix_& buffer) {
        my_work->realize_copy(buffer);
        for (Index c = 0; c < buffer.cols(); ++c)
            for (Index r = 0; r < buffer.rows(); ++r) {
                const double s = my_column ? my_scale[c] : my_scale[r];
                buffer(r, c) = my_divide ? buffer(r, c) / s : buffer(r, c) * s;
            }
        return buffer;
    }

private:
    decltype(std::declval<const Inner_&>().new_known_realize_workspace()) my_work;
    const Scale_& my_scale;
    bool my_column, my_divide;
};

}  // namespace detail

template <class EigenVector_ = VectorXd, class EigenMatrix_ = MatrixXd,
          class MatrixPointer_ = const Matrix<EigenVector_, EigenMatrix_>*, class ScalePointer_ = const EigenVector_*>
class ScaledMatrix final : public Matrix<EigenVector_, EigenMatrix_>, public DeviceResident {
public:
    ScaledMatrix(MatrixPointer_ matrix, ScalePointer_ scale, bool column, bool divide)
        : my_matrix(std::move(matrix)), my_scale(std::move(scale)), my_column(column), my_divide(divide) {
        const auto* dev = dynamic_cast<const DeviceResident*>(&(*my_matrix));
        const irlba_hip_matrix* inner = dev ? dev->device_handle() : nullptr;
        if (inner) {
            irlba_hip_matrix* h = nullptr;
            hip_check(irlba_hip_scaled_create(inner, (*my_scale).data(), column ? 1 : 0, divide ? 1 : 0, 0, &h));
            my_handle = detail::own(h);
        }
    }

    Index rows() const { return my_matrix->rows(); }
    Index cols() const { return my_matrix->cols(); }
    const irlba_hip_matrix* device_handle() const { return my_handle.get(); }

    std::unique_ptr<Workspace<EigenVector_> > new_workspace() const {
        if (my_handle) return std::make_unique<detail::HandleWorkspace<EigenVector_> >(my_handle.get());
        return std::make_unique<detail::ScaledWorkspaceT<EigenVector_, Inner, Scale, false> >(*my_matrix, *my_scale,
                                                                                             my_column, my_divide);
    }
    std::unique_ptr<AdjointWorkspace<EigenVector_> > new_adjoint_workspace() const {
        if (my_handle) return std::make_unique<detail::HandleAdjointWorkspace<EigenVector_> >(my_handle.get());
        return std::make_unique<detail::ScaledWorkspaceT<EigenVector_, Inner, Scale, true> >(*my_matrix, *my_scale,
                                                                                            my_column, my_divide);
    }
    std::unique_ptr<RealizeWorkspace<EigenMatrix_> > new_realize_workspace() const {
        if (my_handle)
            return std::make_unique<detail::HandleRealizeWorkspace<EigenMatrix_> >(my_handle.get(), rows(), cols());
        return std::make_unique<detail::ScaledRealizeWorkspace<EigenMatrix_, Inner, Scale> >(*my_matrix, *my_scale,
                                                                                           my_column, my_divide);
    }

private:
    typedef typename std::decay<decltype(*std::declval<MatrixPointer_>())>::type Inner;
    typedef typename std::decay<decltype(*std::declval<ScalePointer_>())>::type Scale;
    MatrixPointer_ my_matrix;
    ScalePointer_ my_scale;
    bool my_column, my_divide;
    detail::SharedHandle my_handle;
};

}  // namespace irlba

#endif
