// Differentiable top-k SVD for libtorch programs (no reference counterpart:
// the reference computes the forward SVD only, include/irlba/compute.hpp:141-308).
//
// The forward is irlba::compute() on the GPU engine (libirlba_hip.so); the
// backward is a torch::autograd::Function whose gradient of
//   L = sum_i w_i s_i + <G_U, U> + <G_V, V>
// with respect to the matrix is the low-rank X Y' of irlba_hip_svd_backward,
// evaluated on the stored entries only (an SDDMM) for sparse input, as a GEMM
// for dense input, plus the gradient of an optional column-centre vector
// (A = B - 1 c').  Link libirlba_torch.so.
#pragma once

#include <ATen/ATen.h>

#include <cstdint>
#include <optional>
#include <tuple>

namespace irlba {
namespace autograd {

struct SvdOptions {
    int64_t extra_work = 7;            // Options.hpp:47
    double convergence_tolerance = 1e-5;  // Options.hpp:34
    int64_t max_iterations = 1000;     // Options.hpp:53
    int64_t seed = 5489;               // Options.hpp:76
    int64_t storage = 0;               // device value type: 0 FP64, 1 FP32 (mixed), 2 BF16 (dense only)
    int64_t complement = 2;            // backward complement: 0 none, 1 rank-k closed form, 2 exact (CG)
};

// U (nrow x k), S (k), V (ncol x k), FP64, on the input's device (host input:
// host outputs).  `values` FP32/FP64, `indices` int32, `pointers` int64.
std::tuple<at::Tensor, at::Tensor, at::Tensor> svd_sparse(const at::Tensor& values, const at::Tensor& indices,
                                                           const at::Tensor& pointers, int64_t nrow, int64_t ncol,
                                                           bool column_major, const std::optional<at::Tensor>& center,
                                                           int64_t k, const SvdOptions& options = SvdOptions());

// Dense input (FP64, FP32 or BF16, any layout).
std::tuple<at::Tensor, at::Tensor, at::Tensor> svd_dense(const at::Tensor& matrix,
                                                          const std::optional<at::Tensor>& center, int64_t k,
                                                          const SvdOptions& options = SvdOptions());

}  // namespace autograd
}  // namespace irlba
