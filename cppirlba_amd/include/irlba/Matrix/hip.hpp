#ifndef IRLBA_HIP_MATRIX_HIP_HPP
#define IRLBA_HIP_MATRIX_HIP_HPP

// GPU-resident matrices behind the reference's Matrix interface:
//   HipSparseMatrix  <- ParallelSparseMatrix (include/irlba/Matrix/sparse.hpp:438-589)
//   HipDenseMatrix   <- SimpleMatrix over a dense matrix (include/irlba/Matrix/simple.hpp:124-178)
// The arrays are copied to HBM once; products, Gram-Schmidt and restart
// rotations then run on the GPU (libirlba_hip.so, include/irlba_hip.h).

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../../../include/irlba_hip.h"
#include "interface.hpp"

namespace irlba {

inline void hip_check(int rc) {
    if (rc != IRLBA_HIP_OK) throw std::runtime_error(irlba_hip_last_error());
}

// One GPU and one stream (no reference counterpart: the reference forks CPU
// threads in parallelize(), include/irlba/parallel.hpp:32-40).
class HipContext {
public:
    explicit HipContext(int device = 0) { hip_check(irlba_hip_context_create(device, &my_ctx)); }
    ~HipContext() { irlba_hip_context_destroy(my_ctx); }
    HipContext(const HipContext&) = delete;
    HipContext& operator=(const HipContext&) = delete;
    irlba_hip_context* get() const { return my_ctx; }
    irlba_hip_stats stats() const {
        irlba_hip_stats s;
        hip_check(irlba_hip_get_stats(my_ctx, &s));
        return s;
    }

private:
    irlba_hip_context* my_ctx = nullptr;
};

inline HipContext& default_hip_context() {
    static HipContext ctx(0);
    return ctx;
}

namespace detail {

struct HandleDeleter {
    void operator()(irlba_hip_matrix* m) const { irlba_hip_matrix_destroy(m); }
};
typedef std::shared_ptr<irlba_hip_matrix> SharedHandle;

inline SharedHandle own(irlba_hip_matrix* m) { return SharedHandle(m, HandleDeleter()); }

template <class EigenVector_>
class HandleWorkspace final : public Workspace<EigenVector_> {
public:
    explicit HandleWorkspace(const irlba_hip_matrix* m) : my_m(m) {}
    void multiply(const EigenVector_& right, EigenVector_& output) {
        hip_check(irlba_hip_multiply(my_m, 0, right.data(), output.data(), 0));
    }

private:
    const irlba_hip_matrix* my_m;
};

template <class EigenVector_>
class HandleAdjointWorkspace final : public AdjointWorkspace<EigenVector_> {
public:
    explicit HandleAdjointWorkspace(const irlba_hip_matrix* m) : my_m(m) {}
    void multiply(const EigenVector_& right, EigenVector_& output) {
        hip_check(irlba_hip_multiply(my_m, 1, right.data(), output.data(), 0));
    }

private:
    const irlba_hip_matrix* my_m;
};

template <class EigenMatrix_>
class HandleRealizeWorkspace final : public RealizeWorkspace<EigenMatrix_> {
public:
    HandleRealizeWorkspace(const irlba_hip_matrix* m, Index nr, Index nc) : my_m(m), my_nr(nr), my_nc(nc) {}
    const EigenMatrix_& realize(EigenMatrix_& buffer) {
        buffer.resize(my_nr, my_nc);
        hip_check(irlba_hip_realize(my_m, buffer.data(), 0));
        return buffer;
    }

private:
    const irlba_hip_matrix* my_m;
    Index my_nr, my_nc;
};

}  // namespace detail

// Common part of every device-resident Matrix: dimensions and workspaces
// that run through the handle (a host vector in, a host vector out).
template <class EigenVector_, class EigenMatrix_>
class HipMatrixBase : public Matrix<EigenVector_, EigenMatrix_>, public DeviceResident {
public:
    Index rows() const { return my_nr; }
    Index cols() const { return my_nc; }
    std::unique_ptr<Workspace<EigenVector_> > new_workspace() const {
        return std::make_unique<detail::HandleWorkspace<EigenVector_> >(my_handle.get());
    }
    std::unique_ptr<AdjointWorkspace<EigenVector_> > new_adjoint_workspace() const {
        return std::make_unique<detail::HandleAdjointWorkspace<EigenVector_> >(my_handle.get());
    }
    std::unique_ptr<RealizeWorkspace<EigenMatrix_> > new_realize_workspace() const {
        return std::make_unique<detail::HandleRealizeWorkspace<EigenMatrix_> >(my_handle.get(), my_nr, my_nc);
    }
    const irlba_hip_matrix* device_handle() const { return my_handle.get(); }
    const detail::SharedHandle& shared_handle() const { return my_handle; }

protected:
    void adopt(detail::SharedHandle h) {
        my_handle = std::move(h);
        std::int64_t r = 0, c = 0;
        hip_check(irlba_hip_matrix_dims(my_handle.get(), &r, &c));
        my_nr = r;
        my_nc = c;
    }

private:
    detail::SharedHandle my_handle;
    Index my_nr = 0, my_nc = 0;
};

// Storage choices of the device copy.
enum class HipPrecision {
    FP64,        // FP64 values, FP64 products
    MIXED,       // FP32 values, FP32 products (Lanczos vectors stay FP64)
    F32_ACC64,   // FP32 values, FP64 products (what a float ValueArray_ does in the reference)
    BF16         // dense only: BF16 values, FP32 products
};

// Drop-in for ParallelSparseMatrix(nrow, ncol, x, i, p, column_major, num_threads):
// `x` holds double or float values, `i` int32 indices, `p` int64 (or size_t) pointers.
template <class EigenVector_ = VectorXd, class EigenMatrix_ = MatrixXd>
class HipSparseMatrix final : public HipMatrixBase<EigenVector_, EigenMatrix_> {
public:
    template <class ValueArray_, class IndexArray_, class PointerArray_>
    HipSparseMatrix(Index nrow, Index ncol, const ValueArray_& x, const IndexArray_& i, const PointerArray_& p,
                    bool column_major, HipPrecision precision = HipPrecision::FP64,
                    HipContext& ctx = default_hip_context()) {
        typedef typename std::decay<decltype(x[0])>::type V;
        static_assert(std::is_same<V, double>::value || std::is_same<V, float>::value, "values must be float or double");
        static_assert(sizeof(i[0]) == 4, "indices must be 32-bit");
        static_assert(sizeof(p[0]) == 8, "pointers must be 64-bit");
        const int in_t = std::is_same<V, double>::value ? IRLBA_HIP_F64 : IRLBA_HIP_F32;
        int st = IRLBA_HIP_F64, acc64 = 1;
        if (precision == HipPrecision::MIXED) {
            st = IRLBA_HIP_F32;
            acc64 = 0;
        } else if (precision == HipPrecision::F32_ACC64) {
            st = IRLBA_HIP_F32;
        } else if (precision == HipPrecision::BF16) {
            throw std::runtime_error("BF16 storage is available for dense matrices only");
        }
        irlba_hip_matrix* h = nullptr;
        hip_check(irlba_hip_sparse_create(ctx.get(), nrow, ncol, in_t, &x[0], reinterpret_cast<const int32_t*>(&i[0]),
                                          reinterpret_cast<const int64_t*>(&p[0]), column_major ? 1 : 0, st, acc64, 0,
                                          &h));
        this->adopt(detail::own(h));
    }
};

// Drop-in for SimpleMatrix over a column-major dense matrix.
template <class EigenVector_ = VectorXd, class EigenMatrix_ = MatrixXd>
class HipDenseMatrix final : public HipMatrixBase<EigenVector_, EigenMatrix_> {
public:
    HipDenseMatrix(const MatrixXd& a, HipPrecision precision = HipPrecision::FP64,
                   HipContext& ctx = default_hip_context()) {
        make(a.rows(), a.cols(), a.data(), precision, ctx);
    }
    HipDenseMatrix(Index nrow, Index ncol, const double* column_major_values, HipPrecision precision = HipPrecision::FP64,
                   HipContext& ctx = default_hip_context()) {
        make(nrow, ncol, column_major_values, precision, ctx);
    }

private:
    void make(Index nr, Index nc, const double* v, HipPrecision precision, HipContext& ctx) {
        const int st = precision == HipPrecision::FP64 ? IRLBA_HIP_F64
                     : precision == HipPrecision::BF16 ? IRLBA_HIP_BF16
                                                        : IRLBA_HIP_F32;
        irlba_hip_matrix* h = nullptr;
        hip_check(irlba_hip_dense_create(ctx.get(), nr, nc, IRLBA_HIP_F64, v, 1, st, 0, &h));
        this->adopt(detail::own(h));
    }
};

}  // namespace irlba

#endif
