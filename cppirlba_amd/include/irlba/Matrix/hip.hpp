#ifndef IRLBA_HIP_MATRIX_HIP_HPP
#define IRLBA_HIP_MATRIX_HIP_HPP

// GPU-resident matrices behind the reference's Matrix interface:
//   HipSparseMatrix  <- ParallelSparseMatrix (include/irlba/Matrix/sparse.hpp:438-589)
//   HipDenseMatrix   <- SimpleMatrix over a dense matrix (include/irlba/Matrix/simple.hpp:124-178)
// The arrays are copied to HBM once; products, Gram-Schmidt and restart
// rotations then run on the GPU (libirlba_hip.so, include/irlba_hip.h).

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

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

template <class A, class = void>
struct has_data : std::false_type {};
template <class A>
struct has_data<A, std::void_t<decltype(std::declval<const A&>().data())> > : std::true_type {};

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

namespace detail {

// A contiguous array of exactly type T holding a[0, n): the caller's own
// storage when it already is one (anything with data() of element type T),
// otherwise a converted copy in `tmp`.
template <typename T, class Array_>
const T* as_contiguous(const Array_& a, std::size_t n, std::vector<T>& tmp) {
    typedef typename std::decay<decltype(a[0])>::type E;
    if constexpr (std::is_same<E, T>::value && has_data<Array_>::value) {
        return a.data();
    } else {
        tmp.resize(n);
        for (std::size_t j = 0; j < n; ++j) tmp[j] = static_cast<T>(a[j]);
        return tmp.data();
    }
}

template <class PointerType_>
struct SparsePartition {
    std::vector<Index> primary, secondary;
    std::vector<std::vector<PointerType_> > nonzero;
};

// The thread partition of ParallelSparseMatrix (sparse.hpp:52-114), for the
// getters of HipSparseMatrix: thread t gets an equal share of the non-zeros
// (the first total % T threads one more), as a range of primary elements and
// as a range of secondary elements, plus, per primary element, the sub-range
// of its non-zeros whose secondary index falls in each thread's secondary
// range.  Empty when num_threads <= 1, where the reference leaves it unset.
template <class PointerType_, class PointerArray_, class IndexArray_>
SparsePartition<PointerType_> sparse_partition(Index n_primary, Index n_secondary, const PointerArray_& ptrs,
                                               const IndexArray_& indices, int num_threads) {
    SparsePartition<PointerType_> part;
    const int T = num_threads;
    if (T <= 1) return part;
    const PointerType_ total = ptrs[n_primary];
    const PointerType_ floor_ = total / T;
    const PointerType_ extra = total % T;
    part.primary.assign(static_cast<std::size_t>(T) + 1, 0);
    Index pc = 0;
    PointerType_ sofar = 0;
    for (int t = 0; t < T; ++t) {
        sofar += floor_ + (static_cast<PointerType_>(t) < extra ? 1 : 0);
        while (pc < n_primary && ptrs[pc + 1] <= sofar) ++pc;
        part.primary[t + 1] = pc;
    }
    std::vector<PointerType_> per(static_cast<std::size_t>(n_secondary), 0);
    for (PointerType_ e = 0; e < total; ++e) ++per[static_cast<std::size_t>(indices[e])];
    part.secondary.assign(static_cast<std::size_t>(T) + 1, 0);
    Index sc = 0;
    PointerType_ cum = 0;
    sofar = 0;
    for (int t = 0; t < T; ++t) {
        sofar += floor_ + (static_cast<PointerType_>(t) < extra ? 1 : 0);
        while (sc < n_secondary && cum <= sofar) cum += per[static_cast<std::size_t>(sc++)];
        part.secondary[t + 1] = sc;
    }
    part.nonzero.assign(static_cast<std::size_t>(T) + 1, std::vector<PointerType_>(static_cast<std::size_t>(n_primary)));
    for (Index c = 0; c < n_primary; ++c) {
        const PointerType_ end = ptrs[c + 1];
        PointerType_ e = ptrs[c];
        part.nonzero[0][c] = e;
        for (int t = 0; t < T; ++t) {
            const Index lim = part.secondary[t + 1];
            while (e < end && static_cast<Index>(indices[e]) < lim) ++e;
            part.nonzero[t + 1][c] = e;
        }
    }
    return part;
}

}  // namespace detail

// Drop-in for ParallelSparseMatrix<EigenVector_, EigenMatrix_, ValueArray_,
// IndexArray_, PointerArray_> (include/irlba/Matrix/sparse.hpp:438-589): the
// same template parameters, the same constructor
//     (nrow, ncol, x, i, p, column_major, num_threads)
// (sparse.hpp:469-471), which keeps x, i and p by value as the reference does,
// and the same getters (sparse.hpp:480-552).  Two trailing arguments, with
// defaults, choose the device precision and context.
//
// The products run on the GPU from device copies of both orientations, built
// at construction: `num_threads` does not change them.  The reference's
// per-thread partition of the primary and secondary dimensions (sparse.hpp:
// 52-114) is still available from the getters for num_threads > 1; it is
// computed on the host on first request (these getters are documented by the
// reference as uninitialised when num_threads <= 1, and return empty vectors
// then).
//
// Values of any arithmetic type are accepted: float and double go to the
// device as they are, other types (e.g. integer counts) as double.  Indices go
// as int32 and pointers as int64; other integer types are converted on the host.
template <class EigenVector_ = VectorXd, class EigenMatrix_ = MatrixXd, class ValueArray_ = std::vector<double>,
          class IndexArray_ = std::vector<int>, class PointerArray_ = std::vector<std::size_t> >
class HipSparseMatrix final : public HipMatrixBase<EigenVector_, EigenMatrix_> {
public:
    typedef typename std::decay<decltype(std::declval<PointerArray_>()[0])>::type PointerType;

    // sparse.hpp:448-451: a default-constructed matrix cannot be used for any operation.
    HipSparseMatrix() {}

    HipSparseMatrix(Index nrow, Index ncol, ValueArray_ x, IndexArray_ i, PointerArray_ p, bool column_major,
                    int num_threads = 1, HipPrecision precision = HipPrecision::FP64,
                    HipContext& ctx = default_hip_context())
        : my_values(std::move(x)), my_indices(std::move(i)), my_ptrs(std::move(p)), my_column_major(column_major),
          my_num_threads(num_threads), my_primary(column_major ? ncol : nrow),
          my_secondary(column_major ? nrow : ncol) {
        typedef typename std::decay<decltype(my_values[0])>::type V;
        static_assert(std::is_arithmetic<V>::value, "values must be of an arithmetic type");
        int st = IRLBA_HIP_F64, acc64 = 1;
        if (precision == HipPrecision::MIXED) {
            st = IRLBA_HIP_F32;
            acc64 = 0;
        } else if (precision == HipPrecision::F32_ACC64) {
            st = IRLBA_HIP_F32;
        } else if (precision == HipPrecision::BF16) {
            throw std::runtime_error("BF16 storage is available for dense matrices only");
        }
        if (nrow < 0 || ncol < 0) throw std::runtime_error("matrix dimensions must be non-negative");
        std::vector<std::int64_t> ptmp;
        const std::int64_t* pp =
            detail::as_contiguous<std::int64_t>(my_ptrs, static_cast<std::size_t>(my_primary) + 1, ptmp);
        const std::int64_t nnz = pp[my_primary];
        if (nnz < 0) throw std::runtime_error("sparse pointers must end at a non-negative count");
        std::vector<std::int32_t> itmp;
        const std::int32_t* ip = nullptr;
        typedef typename std::decay<decltype(my_indices[0])>::type IT;
        if constexpr (sizeof(IT) > 4 || (sizeof(IT) == 4 && std::is_unsigned<IT>::value)) {
            // Narrowed on the host: out-of-range indices must fail, not wrap into range.
            itmp.resize(static_cast<std::size_t>(nnz));
            for (std::int64_t e = 0; e < nnz; ++e) {
                const IT v = my_indices[e];
                bool negative = false;
                if constexpr (std::is_signed<IT>::value) negative = v < 0;
                if (negative || static_cast<std::uint64_t>(v) >= static_cast<std::uint64_t>(my_secondary))
                    throw std::runtime_error("sparse indices must lie within the secondary dimension");
                itmp[e] = static_cast<std::int32_t>(v);
            }
            ip = itmp.data();
        } else {
            ip = detail::as_contiguous<std::int32_t>(my_indices, static_cast<std::size_t>(nnz), itmp);
        }
        irlba_hip_matrix* h = nullptr;
        if constexpr (std::is_same<V, float>::value) {
            std::vector<float> xtmp;
            const float* xp = detail::as_contiguous<float>(my_values, static_cast<std::size_t>(nnz), xtmp);
            hip_check(irlba_hip_sparse_create(ctx.get(), nrow, ncol, IRLBA_HIP_F32, xp, ip, pp, column_major ? 1 : 0,
                                              st, acc64, 0, &h));
        } else {
            std::vector<double> xtmp;
            const double* xp = detail::as_contiguous<double>(my_values, static_cast<std::size_t>(nnz), xtmp);
            hip_check(irlba_hip_sparse_create(ctx.get(), nrow, ncol, IRLBA_HIP_F64, xp, ip, pp, column_major ? 1 : 0,
                                              st, acc64, 0, &h));
        }
        this->adopt(detail::own(h));
    }

    // sparse.hpp:495-517
    const ValueArray_& get_values() const { return my_values; }
    const IndexArray_& get_indices() const { return my_indices; }
    const PointerArray_& get_pointers() const { return my_ptrs; }

    // sparse.hpp:527-552: the reference's nnz-balanced thread partition.
    const std::vector<Index>& get_primary_boundaries() const {
        partition();
        return my_part->primary;
    }
    const std::vector<Index>& get_secondary_boundaries() const {
        partition();
        return my_part->secondary;
    }
    const std::vector<std::vector<PointerType> >& get_secondary_nonzero_boundaries() const {
        partition();
        return my_part->nonzero;
    }

    // No reference counterpart (a storage choice behind the value array): the
    // products stream one code byte per stored value, looked up in a table of
    // the distinct values, instead of the value itself.  Lossless; needs at
    // most 256 distinct values (count data) and throws otherwise.  Applies to
    // every decorator built on this matrix.  Returns the table size (0 when
    // switched off).
    int use_value_dictionary(bool enable = true) {
        std::int32_t n = 0;
        hip_check(irlba_hip_sparse_value_dictionary(const_cast<irlba_hip_matrix*>(this->device_handle()),
                                                     enable ? 1 : 0, &n));
        return n;
    }

private:
    ValueArray_ my_values;
    IndexArray_ my_indices;
    PointerArray_ my_ptrs;
    bool my_column_major = true;
    int my_num_threads = 1;
    Index my_primary = 0, my_secondary = 0;

    typedef detail::SparsePartition<PointerType> Partition;
    mutable std::unique_ptr<Partition> my_part;
    mutable std::once_flag my_part_once;

    void partition() const {
        std::call_once(my_part_once, [this] {
            my_part = std::make_unique<Partition>(
                detail::sparse_partition<PointerType>(my_primary, my_secondary, my_ptrs, my_indices, my_num_threads));
        });
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
