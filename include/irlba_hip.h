/*
 * irlba_hip.h — C ABI of the MI355X (gfx950) IRLBA engine.
 *
 * Plain C types only (pointers, sizes, int status).  Every call returns
 * IRLBA_HIP_OK (0) on success or IRLBA_HIP_ERROR (1), after which
 * irlba_hip_last_error() holds the message (thread-local).  Messages of the
 * errors the reference raises keep the reference's wording
 * (include/irlba/compute.hpp:156,159,174,184; lanczos.hpp:83; pca.hpp:72,80).
 *
 * Arrays are column-major.  `device_pointers` = 0 means host memory (copied
 * in/out by the call); 1 means device memory on the context's GPU.  The
 * constructors (sparse, dense, centered, scaled) synchronise the device before
 * they read a caller's device arrays; the other calls run on the context's
 * stream, which does not wait for other streams, so their device inputs must
 * be complete when the call is made (e.g. after torch.cuda.synchronize()).
 * One context = one device + one stream; calls on a context are stream
 * ordered, and calls from several threads on one context (or on handles of
 * one context) are serialised by a per-context lock, so a matrix handle can be
 * shared by concurrent compute()/pca() calls as a reference Matrix can
 * (include/irlba/Matrix/interface.hpp:171-185).  Threads that need to run at
 * the same time use one context each.  A product callback runs with its
 * context's lock held: it may call back into the same context on its own
 * thread, but must not wait on another thread that uses that context.
 * A matrix handle must not outlive its context; decorator handles share the
 * storage of the handle they wrap (include/irlba/Matrix/centered.hpp:139-159).
 */
#ifndef IRLBA_HIP_H
#define IRLBA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IRLBA_HIP_OK 0
#define IRLBA_HIP_ERROR 1

/* Element types of matrix values (input arrays and device storage). */
#define IRLBA_HIP_F64 0
#define IRLBA_HIP_F32 1
#define IRLBA_HIP_BF16 2

typedef struct irlba_hip_context irlba_hip_context;
typedef struct irlba_hip_matrix irlba_hip_matrix;

/* irlba::Options (include/irlba/Options.hpp:19-83), same fields and defaults. */
typedef struct irlba_hip_options {
    double invariant_subspace_tolerance;   /* Options.hpp:27, default -1 (eps^0.8) */
    double convergence_tolerance;          /* Options.hpp:34, default 1e-5 */
    double singular_value_ratio_tolerance; /* Options.hpp:41, default -1 (= convergence_tolerance) */
    int32_t extra_work;                    /* Options.hpp:47, default 7 */
    int32_t max_iterations;                /* Options.hpp:53, default 1000 */
    int32_t exact_for_small_matrix;        /* Options.hpp:59, default 1 */
    int32_t exact_for_large_number;        /* Options.hpp:65, default 1 */
    int32_t cap_number;                    /* Options.hpp:71, default 0 */
    int32_t reserved;
    uint64_t seed;                         /* Options.hpp:76, default 5489 (mt19937_64::default_seed) */
    const double* initial;                 /* Options.hpp:82, host array of length cols, or NULL */
    int64_t initial_length;
} irlba_hip_options;

/* Counters of the last compute()/pca() on a context (no reference counterpart). */
typedef struct irlba_hip_stats {
    int64_t forward_products;     /* A.v calls */
    int64_t adjoint_products;     /* A'.u calls */
    double forward_kernel_ms;     /* summed main-kernel time of the A.v products (HIP events) */
    double adjoint_kernel_ms;
    double forward_bytes;         /* algorithmic bytes of one A.v product */
    double adjoint_bytes;
    double solve_ms;              /* wall time of the last compute() */
    double host_svd_ms;           /* time in the host FP64 SVD of B */
    int32_t restarts;
    int32_t refills;              /* invariant-subspace refills (lanczos.hpp:101-106,124-130) */
    int64_t forward_timed;        /* A.v products whose main kernel was timed (see set_profiling) */
    int64_t adjoint_timed;
} irlba_hip_stats;

/* Options defaults of Options.hpp:27-82. */
void irlba_hip_options_default(irlba_hip_options* opt);

/* Message of the last failed call on this thread. */
const char* irlba_hip_last_error(void);

/* Device context: one GPU, one stream.  No reference counterpart (the
 * reference's parallelism is parallelize(), include/irlba/parallel.hpp:32-40). */
int irlba_hip_context_create(int device, irlba_hip_context** out);
int irlba_hip_context_destroy(irlba_hip_context* ctx);
int irlba_hip_context_synchronize(irlba_hip_context* ctx);
int irlba_hip_get_stats(const irlba_hip_context* ctx, irlba_hip_stats* out);
/* Time the main kernel of one product in `period` of each direction (A.v and
 * A'.u counted apart) with HIP events (1 = every product, the default; 0 =
 * none).  Each timed product adds two event records to the stream, a few
 * microseconds of idle GPU around the kernel, so long solves are timed with a
 * sampling period such as 8. */
int irlba_hip_context_set_profiling(irlba_hip_context* ctx, int period);
/* The stream the context launches on (a hipStream_t), for callers timing it. */
void* irlba_hip_context_stream(irlba_hip_context* ctx);
/* Stream ordering without a host wait (no reference counterpart): after
 * irlba_hip_context_wait_stream, work the context queues waits for the work
 * queued so far on `stream` (a hipStream_t on the context's device, e.g.
 * torch's current stream, where device inputs are produced); after
 * irlba_hip_stream_wait_context, work queued on `stream` waits for the work
 * queued so far on the context (the outputs of the _async entry points). */
int irlba_hip_context_wait_stream(irlba_hip_context* ctx, void* stream);
int irlba_hip_stream_wait_context(irlba_hip_context* ctx, void* stream);

/* ParallelSparseMatrix(nrow, ncol, x, i, p, column_major, num_threads)
 * (include/irlba/Matrix/sparse.hpp:438-472).  `values` has type `input_type`
 * (F64/F32), `indices` int32, `pointers` int64 (length primary+1).  The values
 * are stored on the device as `storage_type` (F64/F32) and products accumulate
 * in FP64 when `accumulate_f64` != 0, else in FP32.  The arrays are copied:
 * the handle owns device copies of both orientations. */
int irlba_hip_sparse_create(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, int input_type, const void* values,
                            const int32_t* indices, const int64_t* pointers, int column_major, int storage_type,
                            int accumulate_f64, int device_pointers, irlba_hip_matrix** out);

/* SimpleMatrix over a dense matrix (include/irlba/Matrix/simple.hpp:124-178).
 * Storage F64 (FP64 accumulation), F32 or BF16 (FP32 accumulation). */
int irlba_hip_dense_create(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, int input_type, const void* values,
                           int column_major, int storage_type, int device_pointers, irlba_hip_matrix** out);

/* CenteredMatrix(matrix, center) (include/irlba/Matrix/centered.hpp:139-159):
 * column centring, center has length cols.  Returns a new handle. */
int irlba_hip_centered_create(const irlba_hip_matrix* inner, const double* center, int device_pointers,
                              irlba_hip_matrix** out);

/* ScaledMatrix(matrix, scale, column, divide) (include/irlba/Matrix/scaled.hpp:195-223). */
int irlba_hip_scaled_create(const irlba_hip_matrix* inner, const double* scale, int column, int divide,
                            int device_pointers, irlba_hip_matrix** out);

int irlba_hip_matrix_destroy(irlba_hip_matrix* m);

/* Matrix::rows() / cols() (include/irlba/Matrix/interface.hpp:161,166). */
int irlba_hip_matrix_dims(const irlba_hip_matrix* m, int64_t* rows, int64_t* cols);

/* Device layout of one orientation of a sparse handle (no reference
 * counterpart: introspection for tests and measurements).  adjoint = 0 is the
 * row-major layout that A.v streams, 1 the column-major one of A'.u. */
typedef struct irlba_hip_layout {
    int64_t nnz;            /* true entries */
    int64_t slots;          /* stored slots: entries, zero padding, gap-bridging slots */
    int64_t index_bytes;    /* bytes of the index stream (2 per slot, or 1 with 8-bit deltas) */
    int32_t tile_len;       /* minor-dimension tile staged in LDS */
    int32_t n_tiles;
    int32_t vec;            /* slots per lane per step */
    int32_t lanes_per_line;
    int32_t delta;          /* 1: 8-bit delta indices */
    int32_t blocks;         /* workgroups of one product */
    int32_t dictionary;     /* entries of the value dictionary the products read, 0: the value stream */
    int32_t dictionary_run; /* slots per code word group of that stream (4 or 16) */
} irlba_hip_layout;
int irlba_hip_sparse_layout(const irlba_hip_matrix* m, int adjoint, irlba_hip_layout* out);

/* Value dictionary for a sparse handle (no reference counterpart: a storage
 * option of ParallelSparseMatrix's value array, include/irlba/Matrix/sparse.hpp:469,
 * whose products read values through ValueArray_'s operator[]).  enable = 1
 * makes the products of both orientations stream one code byte per stored slot
 * instead of the FP32/FP64 value, looked up in a table of the distinct stored
 * values; lossless, every product term is the same number.  Needs the 8-bit
 * delta layout in both orientations and at most 256 distinct value bit
 * patterns (count data); otherwise returns an error and the products keep the
 * value stream.  The codes are built on the first call (2 more bytes of HBM per
 * slot per orientation; the value stream stays for realize, column statistics
 * and block products) and re-used after; enable = 0 switches back.  It acts on
 * the stored matrix, so on every handle that shares it (decorators).
 * *n_codes receives the table size (0 when switched off). */
int irlba_hip_sparse_value_dictionary(irlba_hip_matrix* m, int enable, int32_t* n_codes);

/* Workspace::multiply (adjoint = 0) and AdjointWorkspace::multiply (adjoint = 1)
 * (include/irlba/Matrix/interface.hpp:49,87): output is fully overwritten. */
int irlba_hip_multiply(const irlba_hip_matrix* m, int adjoint, const double* right, double* output,
                       int device_pointers);

/* Block product: output (k columns) = M right (adjoint 0) or M' right (adjoint 1)
 * for k column-major vectors at once, 1 <= k <= 128.  No reference counterpart
 * (the reference multiplies one vector per call, interface.hpp:49,87); sparse
 * handles stream the matrix once for all k, which block methods and the SVD
 * backward's CG use.  Same precision as irlba_hip_multiply except that sparse
 * products accumulate in FP64. */
int irlba_hip_multiply_block(const irlba_hip_matrix* m, int adjoint, int64_t k, const double* right, double* output,
                             int device_pointers);

/* RealizeWorkspace::realize_copy (include/irlba/Matrix/interface.hpp:119-130):
 * rows x cols, column-major FP64. */
int irlba_hip_realize(const irlba_hip_matrix* m, double* buffer, int device_pointers);

/* compute(matrix, number, outU, outV, outD, options) (include/irlba/compute.hpp:141-308).
 * U is rows x number, V cols x number, D number; *number_out receives the
 * count actually written (smaller than `number` under cap_number). */
int irlba_hip_compute(const irlba_hip_matrix* m, int64_t number, const irlba_hip_options* opt, double* U, double* V,
                      double* D, int64_t* number_out, int32_t* converged, int32_t* iterations, int device_pointers);

/* compute() over a caller-defined Matrix (interface.hpp:141-223): the products
 * run in the caller through `matvec` on host vectors, everything else on the GPU. */
typedef int (*irlba_hip_matvec_fn)(void* user, int adjoint, const double* right, double* output);
int irlba_hip_compute_callback(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, irlba_hip_matvec_fn matvec,
                               void* user, int64_t number, const irlba_hip_options* opt, double* U, double* V,
                               double* D, int64_t* number_out, int32_t* converged, int32_t* iterations);

/* run_lanczos_bidiagonalization(work, W, V, B, eng, start, options)
 * (include/irlba/lanczos.hpp:54-143): one sweep from column `start` on the
 * caller's W (rows x work), V (cols x work) and B (work x work), all host,
 * column-major, updated in place; F (cols) receives the residual vector.  The
 * random engine is std::mt19937_64(seed). */
int irlba_hip_lanczos(const irlba_hip_matrix* m, int64_t work, double* W, double* V, double* B, uint64_t seed,
                      int64_t start, const irlba_hip_options* opt, double* F);

/* orthogonalize_vector(mat, vec, ncols, tmp) (include/irlba/lanczos.hpp:17-21):
 * vec -= M[:, :ncols] (M[:, :ncols]' vec) for a host rows x total_cols matrix M;
 * *norm_out (if not NULL) receives the norm of the result. */
int irlba_hip_orthogonalize(irlba_hip_context* ctx, const double* M, int64_t rows, int64_t total_cols, double* vec,
                            int64_t ncols, double* norm_out);

/* The restart rotation out = M[:, :work] Y[:, :k] (include/irlba/compute.hpp:268,279)
 * on the matrix cores; host arrays, M rows x work, Y work x k, out rows x k. */
int irlba_hip_rotate(irlba_hip_context* ctx, const double* M, int64_t rows, int64_t work, const double* Y, int64_t k,
                     double* out);

/* Column means / standard deviations of pca() (include/irlba/pca.hpp:66-109). */
int irlba_hip_column_stats(const irlba_hip_matrix* m, int center, int scale, double* center_out, double* scale_out,
                           int device_pointers);

/* pca(matrix, center, scale, number, scores, rotation, variances, options)
 * (include/irlba/pca.hpp:55-139): scores rows x number, rotation cols x number. */
int irlba_hip_pca(const irlba_hip_matrix* m, int center, int scale, int64_t number, const irlba_hip_options* opt,
                  double* scores, double* rotation, double* variances, int64_t* number_out, int32_t* converged,
                  int32_t* iterations, int device_pointers);

/* Synthetic CSC matrix generated on the device (no reference counterpart):
 * positions Bernoulli(density), values log1p(NB(mu=2, phi=0.5)); bit-identical
 * to the CPU generator for the same seed. */
int irlba_hip_generate_sparse(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, double density, uint64_t seed,
                              int storage_type, int accumulate_f64, irlba_hip_matrix** out);

/* Raw device CSC of the same generator (values as `value_type`, int32 rows,
 * int64 column pointers) into caller-provided device arrays; call with
 * values == NULL to get *nnz_out first. */
int irlba_hip_generate_csc(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, double density, uint64_t seed,
                           int value_type, void* values, int32_t* indices, int64_t* pointers, int64_t* nnz_out);

/* Backward of a top-k SVD A ~ U diag(D) V' (no reference counterpart: the
 * reference computes the forward only).  Given the k triplets of m (U rows x k,
 * D k, V cols x k, e.g. from irlba_hip_compute) and the gradients of a loss
 * with respect to them (grad_D k, grad_U rows x k, grad_V cols x k; NULL =
 * zero), writes dL/dA = X Y' in factored form: X rows x 2k, Y cols x 2k,
 * X = [U | U M + Z'], Y = [Y' | V] (Z', Y' the complement terms below).
 * `complement` selects the part of dU, dV outside span(U), span(V):
 *   0 = none (in-subspace terms only),
 *   1 = (I - UU') grad_U D^-1 and (I - VV') grad_V D^-1, exact when rank(A) = k,
 *   2 = exact for any A: one CG system per triplet and side on the projected
 *       normal operators (I - UU') A A' (I - UU') and (I - VV') A'A (I - VV'),
 *       all 2k run in lock step through m's own products, to a residual of
 *       cg_tolerance relative (needs D[k-1] > the (k+1)-th singular value).
 * grad_center (cols, nullable) receives -Y (X' 1), the gradient with respect to
 * a column-centre vector c of A = B - 1 c'.  *cg_iterations (nullable) gets the
 * number of lock-step CG iterations.  k <= 64. */
int irlba_hip_svd_backward(const irlba_hip_matrix* m, int64_t k, const double* U, const double* D, const double* V,
                           const double* grad_D, const double* grad_U, const double* grad_V, int complement,
                           int32_t max_cg_iterations, double cg_tolerance, double* X, double* Y, double* grad_center,
                           int32_t* cg_iterations, int device_pointers);

/* irlba_hip_svd_backward that returns with its work queued on the context's
 * stream instead of waiting for it.  No reference counterpart: it is for
 * callers that order their own stream after it (irlba_hip_stream_wait_context),
 * such as an autograd backward, so that the GPU does not idle for host round
 * trips.  Device arrays only, all three gradients given (none NULL).  The
 * rank-k and projected paths do not read back the check that D is positive and
 * non-increasing (the caller checks D; other values give inf or NaN, not an
 * error); the exact path checks it before its CG, which synchronises anyway. */
int irlba_hip_svd_backward_async(const irlba_hip_matrix* m, int64_t k, const double* U, const double* D,
                                 const double* V, const double* grad_D, const double* grad_U, const double* grad_V,
                                 int complement, int32_t max_cg_iterations, double cg_tolerance, double* X, double* Y,
                                 double* grad_center, int32_t* cg_iterations);

/* out[e] = X[row_e, :] . Y[col_e, :] for every stored entry e of a compressed
 * pattern (indices int32, pointers int64; column_major = CSC), with X nrow x rank
 * and Y ncol x rank column-major FP64: the gradient of the stored values of a
 * sparse matrix from the factored dL/dA of irlba_hip_svd_backward.  `out_type`
 * F64 or F32 (the values' type): F32 output rounds the factors to FP32 and
 * accumulates each dot product in FP64.  rank <= 128.  The pattern is checked
 * as the kernel reads it (monotone pointers from 0, indices in range); a bad
 * pattern fails the call after the kernel, its offending entries unwritten or 0. */
int irlba_hip_sddmm(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, const int32_t* indices,
                    const int64_t* pointers, int column_major, int64_t rank, const double* X, const double* Y,
                    void* out, int out_type, int device_pointers);

/* irlba_hip_sddmm that returns with its work queued on the context's stream
 * (no reference counterpart; for the same callers as
 * irlba_hip_svd_backward_async).  Device arrays only, none NULL.  The kernel's
 * pattern checks are not read back, so a bad pattern is not reported (its
 * entries are left unwritten or 0): for a pattern already checked, such as that
 * of a matrix created from the same arrays. */
int irlba_hip_sddmm_async(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, const int32_t* indices,
                          const int64_t* pointers, int column_major, int64_t rank, const double* X, const double* Y,
                          void* out, int out_type);

/* dL/dA = X Y' of irlba_hip_svd_backward written out as a dense nrow x ncol
 * matrix: the dense counterpart of irlba_hip_sddmm (no reference counterpart).
 * X nrow x rank and Y ncol x rank column-major FP64; `out_type` F64, F32 or
 * BF16 (the matrix's own type: every entry is summed in FP64 over the rank in
 * order, then rounded, BF16 through FP32); out column-major (column_major = 1)
 * or row-major.  rank <= 128. */
int irlba_hip_dense_gradient(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, int64_t rank, const double* X,
                             const double* Y, void* out, int out_type, int column_major, int device_pointers);
/* The same, returning with the work queued on the context's stream (device
 * arrays only; for the callers of irlba_hip_svd_backward_async). */
int irlba_hip_dense_gradient_async(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, int64_t rank,
                                   const double* X, const double* Y, void* out, int out_type, int column_major);

/* irlba_hip_sddmm without any device allocation.  No reference counterpart:
 * it exists for callers that cannot allocate device memory (code under stream
 * capture, tight training or serving loops).  Same arguments and results as
 * irlba_hip_sddmm with device_pointers = 1 (every array on the context's
 * device, none NULL); its temporaries (the error flag and both factors packed
 * row-major) live in a fixed per-device scratch array of 32 MiB inside the
 * library, held by every GPU the library runs on until the process ends.  A
 * call whose packed factors need more, 256 + (nrow + ncol) x padded rank x
 * value size bytes (padded rank: rank rounded up to a power-of-two multiple of
 * 2 for F64 and of 4 for F32 output), fails with an error and writes nothing.
 * All calls on one GPU share the array: calls that overlap in time on the same
 * device (through two contexts, or two threads) race on it. */
int irlba_hip_sddmm_noalloc(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, const int32_t* indices,
                            const int64_t* pointers, int column_major, int64_t rank, const double* X, const double* Y,
                            void* out, int out_type);

#ifdef __cplusplus
}
#endif

#endif /* IRLBA_HIP_H */
