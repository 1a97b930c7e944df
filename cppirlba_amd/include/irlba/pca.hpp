#ifndef IRLBA_HIP_PCA_HPP
#define IRLBA_HIP_PCA_HPP

// irlba::pca() / PcaResults (include/irlba/pca.hpp:55-204 of the reference).
// Column statistics, the Scaled(Centered(A)) chain, the solve and the score
// scaling all run on the GPU.  The input is a dense host matrix (wrapped as an
// FP64 HipDenseMatrix, like the reference's SimpleMatrix) or any undecorated
// device-resident matrix (HipSparseMatrix, HipDenseMatrix).

#include <type_traits>
#include <utility>

#include "Matrix/hip.hpp"
#include "Options.hpp"
#include "compute.hpp"

namespace irlba {

template <class EigenMatrix_, class EigenVector_>
struct PcaResults {
    EigenMatrix_ scores;    // rows x number
    EigenMatrix_ rotation;  // cols x number
    EigenVector_ variances;
    int iterations;
    bool converged;
};

namespace detail {

template <class OutputEigenMatrix_, class EigenVector_>
std::pair<bool, int> pca_device(const irlba_hip_matrix* h, Index nr, Index nc, bool center, bool scale, Index number,
                                OutputEigenMatrix_& scores, OutputEigenMatrix_& rotation, EigenVector_& variances,
                                const Options<EigenVector_>& options) {
    COptions<EigenVector_> co(options);
    const Index alloc = std::max<Index>(0, std::min<Index>(number, std::min(nr, nc)));
    const Index n = options.cap_number ? alloc : std::max<Index>(number, 0);
    scores.resize(nr, n);
    rotation.resize(nc, n);
    variances.resize(std::max<Index>(n, 1));
    std::int64_t nout = 0;
    std::int32_t conv = 0, iters = 0;
    hip_check(irlba_hip_pca(h, center ? 1 : 0, scale ? 1 : 0, number, &co.o, scores.data(), rotation.data(),
                            variances.data(), &nout, &conv, &iters, 0));
    scores.resize(nr, nout);
    rotation.resize(nc, nout);
    variances.resize(nout);
    return std::make_pair(conv != 0, static_cast<int>(iters));
}

}  // namespace detail

// pca(matrix, center, scale, number, scores, rotation, variances, options) (pca.hpp:55-139).
template <class OutputEigenMatrix_, class EigenVector_>
std::pair<bool, int> pca(const MatrixXd& matrix, bool center, bool scale, Index number, OutputEigenMatrix_& scores,
                         OutputEigenMatrix_& rotation, EigenVector_& variances, const Options<EigenVector_>& options) {
    HipDenseMatrix<EigenVector_, OutputEigenMatrix_> wrapped(matrix);
    return detail::pca_device(wrapped.device_handle(), matrix.rows(), matrix.cols(), center, scale, number, scores,
                              rotation, variances, options);
}

template <class Device_, class OutputEigenMatrix_, class EigenVector_,
          class = typename std::enable_if<std::is_base_of<DeviceResident, Device_>::value>::type>
std::pair<bool, int> pca(const Device_& matrix, bool center, bool scale, Index number, OutputEigenMatrix_& scores,
                         OutputEigenMatrix_& rotation, EigenVector_& variances, const Options<EigenVector_>& options) {
    return detail::pca_device(matrix.device_handle(), matrix.rows(), matrix.cols(), center, scale, number, scores,
                              rotation, variances, options);
}

// pca.hpp:197-204
template <class OutputEigenMatrix_ = MatrixXd, class EigenVector_ = VectorXd, class Input_>
PcaResults<OutputEigenMatrix_, EigenVector_> pca(const Input_& matrix, bool center, bool scale, Index number,
                                                 const Options<EigenVector_>& options) {
    PcaResults<OutputEigenMatrix_, EigenVector_> out;
    const auto st = pca(matrix, center, scale, number, out.scores, out.rotation, out.variances, options);
    out.converged = st.first;
    out.iterations = st.second;
    return out;
}

}  // namespace irlba

#endif
