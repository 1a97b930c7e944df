// Host side of the MI355X IRLBA engine: device matrices, the decorator
// algebra, the device-resident Lanczos bidiagonalisation, the restart loop
// with an FP64 host SVD of the small B, pca(), and the C ABI of
// include/irlba_hip.h.
//
// Algorithm: include/irlba/compute.hpp:141-308 (restart loop),
// include/irlba/lanczos.hpp:54-143 (sweep), include/irlba/pca.hpp:55-139.
// Every Lanczos vector (V, W, F) lives in HBM in FP64.  A sweep is queued on
// the context's stream without host round trips: the norms S and R_F are
// produced and consumed on the device, and the invariant-subspace tests
// (lanczos.hpp:99-109,123-132) run on the host after the sweep.  If one
// fires, the affected half-step is recomputed, refilled with host normals,
// and the rest of the sweep is queued again, which reproduces the
// reference's step-by-step control flow exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "../../include/irlba_hip.h"
#include <mutex>

#include "common.hpp"
#include "host_svd.hpp"

using namespace irlba_hip;

namespace {

thread_local std::string g_last_error;

template <class Fn>
int guarded(Fn&& fn) {
    try {
        fn();
        return IRLBA_HIP_OK;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return IRLBA_HIP_ERROR;
    } catch (...) {
        g_last_error = "unknown error";
        return IRLBA_HIP_ERROR;
    }
}

void require(bool ok, const char* msg) {
    if (!ok) throw std::runtime_error(msg);
}

// ------------------------------------------------------------------ device memory
template <typename T>
std::shared_ptr<T> dalloc(i64 n) {
    T* p = nullptr;
    const size_t bytes = sizeof(T) * static_cast<size_t>(std::max<i64>(n, 1));
    IRLBA_HIP_CHECK(hipMalloc(&p, bytes));
    return std::shared_ptr<T>(p, [](T* q) { (void)hipFree(q); });
}

std::shared_ptr<void> dalloc_bytes(size_t bytes) {
    void* p = nullptr;
    IRLBA_HIP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 1)));
    return std::shared_ptr<void>(p, [](void* q) { (void)hipFree(q); });
}

// Pinned host memory (hipHostMalloc), freed with its owner.
struct PinnedBuffer {
    double* p = nullptr;
    explicit PinnedBuffer(size_t bytes) {
        IRLBA_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(bytes, 8)));
    }
    ~PinnedBuffer() {
        if (p) (void)hipHostFree(p);
    }
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    double* get() const { return p; }
};

size_t storage_bytes(StorageType st) { return st == ST_F64 ? 8 : (st == ST_F32 ? 4 : 2); }

StorageType to_storage(int t) {
    if (t == IRLBA_HIP_F64) return ST_F64;
    if (t == IRLBA_HIP_F32) return ST_F32;
    if (t == IRLBA_HIP_BF16) return ST_BF16;
    throw std::runtime_error("unknown value type");
}

}  // namespace

// ------------------------------------------------------------------ context
struct irlba_hip_context {
    // Calls on one context are serialised: its stream, workspaces, events and
    // counters are shared, and the reference lets several threads run
    // compute() on one Matrix at once (include/irlba/Matrix/interface.hpp:171-185).
    // Recursive, since some entry points call others.
    mutable std::recursive_mutex mu;
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 256;
    size_t lds_per_block = 65536;
    size_t lds_per_cu = 65536;
    int profiling = 1;     // time one product in `profiling` per direction (0: none)
    i64 profile_seq[2] = {0, 0};
    irlba_hip_stats stats{};
    struct Pending {
        hipEvent_t a, b;
        bool adjoint;
    };
    std::vector<hipEvent_t> pool;
    std::vector<Pending> pending;

    hipEvent_t get_event() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        IRLBA_HIP_CHECK(hipEventCreate(&e));
        return e;
    }
    // Fold the timings of completed products into the stats: all of them
    // (after a stream sync) or the first `upto` (after waiting on an event
    // recorded behind them).
    void resolve_events(size_t upto = static_cast<size_t>(-1)) {
        upto = std::min(upto, pending.size());
        for (size_t i = 0; i < upto; ++i) {
            auto& p = pending[i];
            float ms = 0;
            // A product with an empty dimension launches nothing and records no events.
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                if (p.adjoint) stats.adjoint_kernel_ms += ms; else stats.forward_kernel_ms += ms;
            } else {
                (void)hipGetLastError();  // clear it so the next launch check does not see it
            }
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.erase(pending.begin(), pending.begin() + static_cast<std::ptrdiff_t>(upto));
    }
    void sync() {
        IRLBA_HIP_CHECK(hipStreamSynchronize(stream));
        resolve_events();
    }
    // Grow-only device scratch for the temporaries of one call.  Calls on a
    // context are serialised on its stream and each ends with a sync, so one
    // buffer serves them all.
    std::shared_ptr<void> ws;
    size_t ws_bytes = 0;
    // A second stream (default priority) for host-to-device copies that should
    // overlap the compute stream, and an event to order them; created on first
    // use (a stream costs milliseconds to create), destroyed with the context.
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_done = nullptr;
    void ensure_copy_stream() {
        if (!copy_stream) IRLBA_HIP_CHECK(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
        if (!copy_done) IRLBA_HIP_CHECK(hipEventCreateWithFlags(&copy_done, hipEventDisableTiming));
    }
    // An event that orders this stream against a caller's stream (a wait takes
    // the event's latest record, so one event serves both directions).
    hipEvent_t order_ev = nullptr;
    void order(hipStream_t from, hipStream_t to) {
        if (!order_ev) IRLBA_HIP_CHECK(hipEventCreateWithFlags(&order_ev, hipEventDisableTiming));
        IRLBA_HIP_CHECK(hipEventRecord(order_ev, from));
        IRLBA_HIP_CHECK(hipStreamWaitEvent(to, order_ev, 0));
    }
    // Device address of the fixed scratch of the allocation-free entry point on
    // this context's device (resolved when the context is created).
    void* noalloc_scratch = nullptr;
    // A device int cleared in stream order, for the error bits of checking kernels.
    std::shared_ptr<int> flag_buf;
    int* flag() {
        if (!flag_buf) flag_buf = dalloc<int>(1);
        IRLBA_HIP_CHECK(hipMemsetAsync(flag_buf.get(), 0, sizeof(int), stream));
        return flag_buf.get();
    }
    char* workspace(size_t bytes) { return grow(ws, ws_bytes, bytes); }
    // A second one for the multi-vector products, which run inside calls that
    // hold the first.
    std::shared_ptr<void> ws_mv;
    size_t ws_mv_bytes = 0;
    char* mv_workspace(size_t bytes) { return grow(ws_mv, ws_mv_bytes, bytes); }

    char* grow(std::shared_ptr<void>& buf, size_t& have, size_t bytes) {
        if (bytes > have) {
            IRLBA_HIP_CHECK(hipStreamSynchronize(stream));
            buf.reset();
            buf = dalloc_bytes(bytes);
            have = bytes;
        }
        return static_cast<char*>(buf.get());
    }
};

namespace {

// Carves 256-byte aligned arrays out of one buffer; size() on a null base
// measures what the carving needs.
struct Carve {
    char* base = nullptr;
    size_t off = 0;
    template <typename T>
    T* take(i64 n) {
        off = (off + 255) / 256 * 256;
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += sizeof(T) * static_cast<size_t>(std::max<i64>(n, 1));
        return p;
    }
};

}  // namespace

namespace {

struct DeviceGuard {
    explicit DeviceGuard(int dev) { IRLBA_HIP_CHECK(hipSetDevice(dev)); }
};

// Holds the context's lock for the duration of an entry point, then makes its
// device current.
struct CtxGuard {
    std::lock_guard<std::recursive_mutex> lk;
    DeviceGuard dg;
    explicit CtxGuard(irlba_hip_context* c) : lk(c->mu), dg(c->device) {}
};

// ------------------------------------------------------------------ storage
struct Base {
    irlba_hip_context* ctx = nullptr;
    i64 nr = 0, nc = 0;
    virtual ~Base() {}
    virtual void product(bool adjoint, const Prologue& pro, const Epilogue& epi, hipEvent_t e0, hipEvent_t e1) = 0;
    virtual void realize(double* dense_dev) = 0;      // rows x cols column-major
    virtual void colstats(double* c, double* s) = 0;  // device outputs, nullable
    virtual double product_bytes(bool adjoint) const = 0;
};

struct SparseBase : Base {
    TiledLines by_row, by_col;  // major = rows (A.v), major = cols (A'.u)
    std::vector<std::shared_ptr<void> > keep;
    bool acc64 = true;
    bool column_major = true;
    i64 nnz = 0;
    std::shared_ptr<void> partial_row, partial_col;
    std::shared_ptr<double> gamma_row, gamma_col;

    void product(bool adjoint, const Prologue& pro, const Epilogue& epi, hipEvent_t e0, hipEvent_t e1) override {
        const TiledLines& t = adjoint ? by_col : by_row;
        launch_spmv_tiled(t, pro, epi, acc64, adjoint ? partial_col.get() : partial_row.get(),
                          adjoint ? gamma_col.get() : gamma_row.get(), ctx->stream, e0, e1);
    }
    void realize(double* dense) override {
        IRLBA_HIP_CHECK(hipMemsetAsync(dense, 0, sizeof(double) * nr * nc, ctx->stream));
        launch_sparse_realize(indexed(true), true, nr, dense, ctx->stream);
    }
    // The orientation with 16-bit indices, expanded once from the delta layout
    // for the kernels that read indices out of order (realize, block products).
    const TiledLines& indexed(bool adjoint) {
        TiledLines& t = adjoint ? by_col : by_row;
        if (t.delta && !t.idx && t.slots > 0) {
            auto idx = dalloc<std::uint16_t>(t.slots);
            launch_delta_expand(t, idx.get(), ctx->stream);
            t.idx = idx.get();
            keep.push_back(idx);
        }
        return t;
    }
    void colstats(double* c, double* s) override { launch_sparse_colstats(by_col, nr, c, s, ctx->stream); }
    double product_bytes(bool adjoint) const override {
        // nnz (value + int32 index) + pointers of the input orientation + input and output vectors (FP64).
        const double sval = static_cast<double>(storage_bytes(by_row.st));
        const double nptr = static_cast<double>(column_major ? nc : nr) + 1;
        const double nin = static_cast<double>(adjoint ? nr : nc), nout = static_cast<double>(adjoint ? nc : nr);
        return static_cast<double>(nnz) * (sval + 4.0) + nptr * 8.0 + (nin + nout) * 8.0;
    }
};

struct DenseBase : Base {
    DenseDesc d;
    std::shared_ptr<void> data, xs, partial;
    std::shared_ptr<double> gamma;
    i64 partial_cap = 0;

    void product(bool adjoint, const Prologue& pro, const Epilogue& epi, hipEvent_t e0, hipEvent_t e1) override {
        launch_dense_matvec(d, adjoint, pro, epi, xs.get(), gamma.get(), partial.get(), partial_cap, ctx->stream, e0,
                            e1);
    }
    void realize(double* dense) override { launch_dense_realize(d, dense, ctx->stream); }
    void colstats(double* c, double* s) override { launch_dense_colstats(d, c, s, ctx->stream); }
    double product_bytes(bool adjoint) const override {
        const double nin = static_cast<double>(adjoint ? nr : nc), nout = static_cast<double>(adjoint ? nc : nr);
        return static_cast<double>(nr) * nc * storage_bytes(d.st) + (nin + nout) * 8.0;
    }
};

// The decorators of a handle in closed form:  M = D_r A D_c - sum_k a_k b_k'.
struct Affine {
    std::shared_ptr<double> dr, dc;  // nullable
    int n_terms = 0;
    std::shared_ptr<double> a[kMaxTerms], b[kMaxTerms];  // a nullable (= ones), b never null
};

}  // namespace

struct irlba_hip_matrix {
    irlba_hip_context* ctx = nullptr;
    std::shared_ptr<Base> base;
    Affine aff;
    bool decorated = false;
};

namespace {

// A copy on the context's stream, waited for.  Work already queued on that
// stream (zero fills of a new basis, products) is ordered before it, which a
// null-stream hipMemcpy does not give: the context's stream does not wait for
// the null stream, nor the null stream for it.
void stream_copy(irlba_hip_context* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    if (bytes == 0) return;
    IRLBA_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
    IRLBA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

std::shared_ptr<double> upload_vec(irlba_hip_context* ctx, const double* src, i64 n, bool device_ptrs) {
    auto d = dalloc<double>(n);
    // The caller's device vector may still be written by work on another stream.
    if (device_ptrs) IRLBA_HIP_CHECK(hipDeviceSynchronize());
    IRLBA_HIP_CHECK(hipMemcpyAsync(d.get(), src, sizeof(double) * n,
                                   device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    IRLBA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return d;
}

// Apply one product of a handle: y = M x (or M' x), then y -= (*coef) src.
void apply_handle(const irlba_hip_matrix& m, bool adjoint, const double* x, double* y, const double* axpy_src,
                  const double* axpy_coef) {
    Prologue pro;
    Epilogue epi;
    const Affine& a = m.aff;
    pro.x = x;
    pro.n_terms = a.n_terms;
    epi.out = y;
    epi.n_terms = a.n_terms;
    epi.axpy_src = axpy_src;
    epi.axpy_coef = axpy_coef;
    if (!adjoint) {
        pro.xscale = a.dc.get();
        epi.oscale = a.dr.get();
        for (int k = 0; k < a.n_terms; ++k) {
            pro.tdot[k] = a.b[k].get();
            epi.tvec[k] = a.a[k].get();
        }
    } else {
        pro.xscale = a.dr.get();
        epi.oscale = a.dc.get();
        for (int k = 0; k < a.n_terms; ++k) {
            pro.tdot[k] = a.a[k].get();
            epi.tvec[k] = a.b[k].get();
        }
    }
    irlba_hip_context* ctx = m.ctx;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->profiling > 0 && ctx->profile_seq[adjoint ? 1 : 0]++ % ctx->profiling == 0) {
        e0 = ctx->get_event();
        e1 = ctx->get_event();
        ctx->pending.push_back({e0, e1, adjoint});
        if (adjoint) ++ctx->stats.adjoint_timed; else ++ctx->stats.forward_timed;
    }
    if (adjoint) ++ctx->stats.adjoint_products; else ++ctx->stats.forward_products;
    m.base->product(adjoint, pro, epi, e0, e1);
}

// One input block of a multi-vector product: k columns of X (leading dimension
// ldx) into k columns of Y (ldy).
struct BlockIO {
    const double* X;
    i64 ldx;
    double* Y;
    i64 ldy;
};

// Whether nio blocks of k vectors each fit one pass of the block kernels
// (sparse: up to 64 lanes of 16 bytes per entry; dense: 32 packed columns).
bool blocks_fit_one_pass(const irlba_hip_matrix& m, bool adjoint, int k, int nio) {
    auto* sb = dynamic_cast<SparseBase*>(m.base.get());
    auto* db = dynamic_cast<DenseBase*>(m.base.get());
    const int K = k * nio;
    if (K <= 1 || (!sb && !db)) return false;
    if (db) return spmm_padded_k(K, db->d.st != ST_F64) <= 32;
    const bool x_f32 = (adjoint ? sb->by_col : sb->by_row).st == ST_F32 && !sb->acc64;
    return K * (x_f32 ? 4 : 8) <= kWave * 16;
}

// Y_b[:, :k] = M X_b[:, :k] (or M' X_b) for nio blocks of k column vectors at
// once.  Sparse and dense handles stream the matrix once for all nio k vectors
// (kernels_spmm.hip), the blocks packed side by side; blocks that do not fit
// one pass run one at a time, and a single vector as a single product.
void apply_handle_blocks(const irlba_hip_matrix& m, bool adjoint, const BlockIO* io, int nio, int k) {
    if (nio > kMaxPackBlocks) {
        apply_handle_blocks(m, adjoint, io, kMaxPackBlocks, k);
        apply_handle_blocks(m, adjoint, io + kMaxPackBlocks, nio - kMaxPackBlocks, k);
        return;
    }
    if (nio > 1 && !blocks_fit_one_pass(m, adjoint, k, nio)) {
        for (int b = 0; b < nio; ++b) apply_handle_blocks(m, adjoint, io + b, 1, k);
        return;
    }
    if (nio == 1 && !blocks_fit_one_pass(m, adjoint, k, 1)) {
        for (int v = 0; v < k; ++v)
            apply_handle(m, adjoint, io[0].X + v * io[0].ldx, io[0].Y + v * io[0].ldy, nullptr, nullptr);
        return;
    }
    auto* sb = dynamic_cast<SparseBase*>(m.base.get());
    auto* db = dynamic_cast<DenseBase*>(m.base.get());
    irlba_hip_context* ctx = m.ctx;
    hipStream_t s = ctx->stream;
    const Affine& a = m.aff;
    const int K = k * nio;
    const bool x_f32 = sb ? ((adjoint ? sb->by_col : sb->by_row).st == ST_F32 && !sb->acc64) : db->d.st != ST_F64;
    const int kp = spmm_padded_k(K, x_f32);
    const i64 L_in = adjoint ? m.base->nr : m.base->nc, L_out = adjoint ? m.base->nc : m.base->nr;
    const i64 dscratch = db ? dense_block_scratch(db->d, kp) : 0;
    const double* xscale = adjoint ? a.dr.get() : a.dc.get();
    const double* oscale = adjoint ? a.dc.get() : a.dr.get();
    const double* tdot[kMaxTerms] = {nullptr, nullptr};
    const double* tvec[kMaxTerms] = {nullptr, nullptr};
    bool need_ones = false;
    for (int j = 0; j < a.n_terms; ++j) {
        tdot[j] = adjoint ? a.a[j].get() : a.b[j].get();
        tvec[j] = adjoint ? a.b[j].get() : a.a[j].get();
        need_ones |= tdot[j] == nullptr;
    }
    auto plan = [&](Carve& c, std::vector<void*>& p) {
        p.clear();
        p.push_back(c.take<char>(static_cast<i64>(x_f32 ? 4 : 8) * kp * L_in));       // packed X
        p.push_back(c.take<double>(static_cast<i64>(kp) * L_out));                     // packed Y
        p.push_back(c.take<double>(2 * static_cast<i64>(K)));                          // gammas
        p.push_back(c.take<double>(static_cast<i64>(cgs_dot_blocks(L_in)) * k));      // dot partials
        p.push_back(need_ones ? c.take<double>(L_in) : nullptr);                       // ones
        p.push_back(db ? c.take<double>(dscratch) : nullptr);                          // dense partials
    };
    std::vector<void*> wp;
    Carve sizing;
    plan(sizing, wp);
    Carve carve;
    carve.base = ctx->mv_workspace(sizing.off);
    plan(carve, wp);
    double* gam = static_cast<double*>(wp[2]);
    double* partials = static_cast<double*>(wp[3]);
    double* ones = static_cast<double*>(wp[4]);
    if (need_ones) launch_fill(ones, 1.0, L_in, s);
    StageBlocks in{};
    FinishBlocks out{};
    in.n = out.n = nio;
    for (int b = 0; b < nio; ++b) {
        for (int j = 0; j < a.n_terms; ++j) {
            // gamma_j[v] = tdot_j . X[:, v] on the unscaled input (the single product's prologue).
            launch_cgs_dots(io[b].X, io[b].ldx, k, tdot[j] ? tdot[j] : ones, L_in, partials, s);
            launch_cgs_reduce(partials, L_in, k, gam + (2 * b + j) * k, s);
        }
        in.X[b] = io[b].X;
        in.ld[b] = io[b].ldx;
        out.Y[b] = io[b].Y;
        out.ld[b] = io[b].ldy;
    }
    launch_spmm_stage(in, L_in, k, kp, xscale, x_f32, wp[0], s);
    if (sb)
        launch_spmm_lines(sb->indexed(adjoint), wp[0], K, x_f32, static_cast<double*>(wp[1]), s);
    else
        launch_dense_block(db->d, adjoint, kp, x_f32, wp[0], static_cast<double*>(wp[1]), static_cast<double*>(wp[5]),
                           dscratch, s);
    launch_spmm_finish(static_cast<const double*>(wp[1]), L_out, k, kp, oscale, a.n_terms, tvec[0], tvec[1], gam, out,
                       s);
    if (adjoint) ctx->stats.adjoint_products += K; else ctx->stats.forward_products += K;
}

// Y[:, :k] = M X[:, :k] (or M' X) for k column vectors (column-major, leading
// dimensions ldx, ldy).
void apply_handle_multi(const irlba_hip_matrix& m, bool adjoint, const double* X, i64 ldx, double* Y, i64 ldy,
                        int k) {
    const BlockIO io{X, ldx, Y, ldy};
    apply_handle_blocks(m, adjoint, &io, 1, k);
}

// ------------------------------------------------------------------ sparse construction
TiledLines build_lines(irlba_hip_context* ctx, const std::int32_t* d_idx, const void* d_val, StorageType in_t,
                       const i64* d_ptr, i64 n_primary, bool transpose, i64 n_major_out, i64 n_minor, int tile_len,
                       StorageType st, bool tile_acc64, i64 nnz, std::vector<std::shared_ptr<void> >& keep) {
    hipStream_t s = ctx->stream;
    TiledLines t;
    t.n_major = n_major_out;
    t.n_minor = n_minor;
    // As many tiles as the LDS budget requires, of equal width (rounded up to 64
    // entries, within the budget): a remainder tile of a few entries would hold
    // one short segment per line, whose bookkeeping costs as much as a full one.
    const i64 tmax = std::max<i64>(1, std::min<i64>(tile_len, std::max<i64>(n_minor, 1)));
    const i64 n_tiles = std::max<i64>(1, (n_minor + tmax - 1) / tmax);
    const i64 even = (std::max<i64>(n_minor, 1) + n_tiles - 1) / n_tiles;
    t.tile_len = static_cast<int>(std::min<i64>(tmax, (even + 63) / 64 * 64));
    t.n_tiles = static_cast<int>(std::max<i64>(1, (n_minor + t.tile_len - 1) / t.tile_len));
    require(t.tile_len <= 65536, "input tiles must hold at most 65536 entries (16-bit in-tile indices)");
    t.nnz = nnz;
    t.st = st;
    const i64 nseg = static_cast<i64>(t.n_tiles) * n_major_out;
    require(n_major_out < (static_cast<i64>(1) << 31), "matrix dimension too large for the device layout");

    // Per-lane vector width and lanes per line from the mean segment length.
    const double avg = nseg ? static_cast<double>(nnz) / nseg : 0.0;
    t.vec = avg >= 12 ? 4 : 1;
    const double per = avg / t.vec;
    int lpl = 4;
    while (lpl < 64 && lpl * 2 <= per / 2) lpl *= 2;
    t.lpl = lpl;
    if (const char* e = std::getenv("IRLBA_HIP_LPL")) t.lpl = std::atoi(e);

    auto counts = dalloc<unsigned long long>(nseg);
    IRLBA_HIP_CHECK(hipMemsetAsync(counts.get(), 0, sizeof(unsigned long long) * nseg, s));
    launch_count_entries(d_idx, d_ptr, n_primary, transpose, t.tile_len, n_major_out, counts.get(), s);
    auto plen = dalloc<i64>(nseg);
    // Segment lengths (and so starts) are padded to a multiple of `seg_align` slots.
    // Long segments start on 256-byte boundaries of the FP32 values (64 slots):
    // a lane group's 16-byte loads then never straddle cache lines shared with
    // the neighbouring segment, which the streaming (nontemporal) loads would
    // otherwise fetch twice (measured: 20-27% less HBM traffic, 13-18% faster
    // products on a 9e8-nnz matrix).  The padding is kept under ~6% of the
    // slots by shrinking the alignment for short segments.
    int seg_align = t.vec;
    while (seg_align * 2 <= 64 && seg_align * 2 <= 0.12 * avg) seg_align *= 2;
    if (const char* e = std::getenv("IRLBA_HIP_SEG_ALIGN")) seg_align = std::max(t.vec, std::atoi(e) / t.vec * t.vec);
    launch_pad_counts(counts.get(), plen.get(), nseg, seg_align, s);
    t.seg_align = seg_align;
    auto ptr = dalloc<i64>(nseg + 1);
    exclusive_scan_i64(plen.get(), ptr.get(), nseg, s);
    plen.reset();
    stream_copy(ctx, &t.slots, ptr.get() + nseg, sizeof(i64), hipMemcpyDeviceToHost);

    auto cursor = dalloc<unsigned long long>(nseg);
    IRLBA_HIP_CHECK(hipMemcpyAsync(cursor.get(), ptr.get(), sizeof(i64) * nseg, hipMemcpyDeviceToDevice, s));
    const size_t vb = storage_bytes(st);
    auto sidx = dalloc<std::uint16_t>(t.slots);
    auto sval = dalloc_bytes(vb * t.slots);
    IRLBA_HIP_CHECK(hipMemsetAsync(sidx.get(), 0, sizeof(std::uint16_t) * t.slots, s));
    IRLBA_HIP_CHECK(hipMemsetAsync(sval.get(), 0, vb * t.slots, s));
    launch_scatter_entries(d_idx, d_val, in_t, d_ptr, n_primary, transpose, t.tile_len, n_major_out, cursor.get(),
                           sidx.get(), sval.get(), st, s);
    cursor.reset();

    auto fidx = dalloc<std::uint16_t>(t.slots);
    auto fval = dalloc_bytes(vb * t.slots);
    IRLBA_HIP_CHECK(hipMemsetAsync(fidx.get(), 0, sizeof(std::uint16_t) * t.slots, s));
    IRLBA_HIP_CHECK(hipMemsetAsync(fval.get(), 0, vb * t.slots, s));
    auto dup = dalloc<int>(1);
    IRLBA_HIP_CHECK(hipMemsetAsync(dup.get(), 0, sizeof(int), s));
    launch_sort_segments(ptr.get(), counts.get(), nseg, t.tile_len, sidx.get(), sval.get(), fidx.get(), fval.get(), st,
                         dup.get(), s);
    int dup_h = 0;
    IRLBA_HIP_CHECK(hipMemcpyAsync(&dup_h, dup.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    IRLBA_HIP_CHECK(hipStreamSynchronize(s));
    require(dup_h == 0, "sparse input holds the same row/column pair more than once");
    sidx.reset();
    sval.reset();

    // 8-bit delta indices: one byte per slot instead of two, plus a bridging
    // slot (value 0) for each in-tile gap above 255.  Kept when the bridging
    // slots cost less than the byte they save (3% margin), i.e. for segments
    // of mean gap well under 255.
    std::shared_ptr<std::uint8_t> dlt;
    std::shared_ptr<std::uint16_t> first;
    // IRLBA_HIP_DELTA=0 keeps the 16-bit layout, =2 takes the delta layout whatever it costs.
    const char* de = std::getenv("IRLBA_HIP_DELTA");
    const int delta_mode = de ? std::atoi(de) : 1;
    if (t.vec == 4 && delta_mode != 0) {
        auto lenB = dalloc<unsigned long long>(nseg);
        launch_delta_count(ptr.get(), counts.get(), nseg, fidx.get(), lenB.get(), s);
        auto plenB = dalloc<i64>(nseg);
        launch_pad_counts(lenB.get(), plenB.get(), nseg, seg_align, s);
        lenB.reset();
        auto ptrB = dalloc<i64>(nseg + 1);
        exclusive_scan_i64(plenB.get(), ptrB.get(), nseg, s);
        plenB.reset();
        i64 slotsB = 0;
        stream_copy(ctx, &slotsB, ptrB.get() + nseg, sizeof(i64), hipMemcpyDeviceToHost);
        const double bytes_u16 = static_cast<double>(t.slots) * (vb + 2.0);
        const double bytes_u8 = static_cast<double>(slotsB) * (vb + 1.0);
        if (slotsB > 0 && (bytes_u8 < 0.97 * bytes_u16 || delta_mode == 2)) {
            dlt = dalloc<std::uint8_t>(slotsB);
            first = dalloc<std::uint16_t>(nseg);
            auto vB = dalloc_bytes(vb * slotsB);
            IRLBA_HIP_CHECK(hipMemsetAsync(dlt.get(), 0, slotsB, s));
            IRLBA_HIP_CHECK(hipMemsetAsync(vB.get(), 0, vb * slotsB, s));
            launch_delta_write(ptr.get(), counts.get(), nseg, fidx.get(), fval.get(), st, ptrB.get(), first.get(),
                               dlt.get(), vB.get(), s);
            IRLBA_HIP_CHECK(hipStreamSynchronize(s));
            ptr = ptrB;
            fval = vB;
            fidx.reset();
            t.slots = slotsB;
            t.delta = true;
        }
    }
    counts.reset();

    // nnz-balanced chunks for exactly one resident wave of workgroups: as many
    // blocks per CU as the LDS tile allows, times the CU count.
    const size_t tile_bytes = static_cast<size_t>(t.tile_len) * (st == ST_F64 || tile_acc64 ? 8 : 4) + 1024;
    const int per_cu = static_cast<int>(std::max<size_t>(1, std::min<size_t>(2, ctx->lds_per_cu / tile_bytes)));
    int target = ctx->cu_count * per_cu;
    if (const char* e = std::getenv("IRLBA_HIP_SPMV_BLOCKS")) target = std::max(1, std::atoi(e));
    t.B = static_cast<int>(std::max<i64>(1, std::min<i64>(std::max<i64>(nseg, 1), target)));
    auto chunk = dalloc<i64>(t.B + 1);
    launch_chunk_bounds(ptr.get(), nseg, t.B, chunk.get(), s);
    IRLBA_HIP_CHECK(hipStreamSynchronize(s));

    t.ptr = ptr.get();
    t.idx = fidx.get();
    t.val = fval.get();
    t.chunk = chunk.get();
    t.dlt = dlt.get();
    t.first = first.get();
    keep.push_back(ptr);
    if (fidx) keep.push_back(fidx);
    keep.push_back(fval);
    keep.push_back(chunk);
    if (dlt) {
        keep.push_back(dlt);
        keep.push_back(first);
    }
    return t;
}

// Lanes per line: the best group width depends on the segment lengths, the
// number of tiles, the precision and the stream (values or dictionary codes),
// so each orientation times a plain product at 8, 16, 32 and 64 lanes (4 too
// for codes; HIP events) and keeps the fastest.  The widths are timed in turn
// over a warm-up round and three timed rounds, each keeping its best: timed one
// width after another, the later widths ran at clocks still rising and were
// picked when they were not the fastest (64 lanes for the 30000 x 1e6 A.v,
// 0.777 ms in the solve against 0.739 ms at 32).
void autotune_lanes(irlba_hip_context* ctx, TiledLines& t, bool acc64, void* partial, double* gamma_part) {
    if (std::getenv("IRLBA_HIP_LPL") || t.nnz == 0 || t.n_major == 0) return;
    if (t.nnz < (static_cast<i64>(1) << 20)) return;  // small matrices: launch-bound, keep the heuristic
    hipStream_t s = ctx->stream;
    auto x = dalloc<double>(t.n_minor);
    auto y = dalloc<double>(t.n_major);
    IRLBA_HIP_CHECK(hipMemsetAsync(x.get(), 0, sizeof(double) * t.n_minor, s));
    Prologue pro;
    pro.x = x.get();
    Epilogue epi;
    epi.out = y.get();
    hipEvent_t e0, e1;
    IRLBA_HIP_CHECK(hipEventCreate(&e0));
    IRLBA_HIP_CHECK(hipEventCreate(&e1));
    // The dictionary stream (2 bytes per slot) also tries 4 lanes per line,
    // which suits its many-tile FP64 orientation.
    static const int kPlain[] = {8, 16, 32, 64}, kDict[] = {4, 8, 16, 32, 64};
    const bool dict = t.use_dict && t.cd;
    const std::vector<int> widths = dict ? std::vector<int>(kDict, kDict + 5) : std::vector<int>(kPlain, kPlain + 4);
    std::vector<float> best_ms(widths.size(), 0.0f);
    for (int round = 0; round < 4; ++round) {  // round 0 warms up
        for (size_t w = 0; w < widths.size(); ++w) {
            t.lpl = widths[w];
            IRLBA_HIP_CHECK(hipEventRecord(e0, s));
            launch_spmv_tiled(t, pro, epi, acc64, partial, gamma_part, s);
            IRLBA_HIP_CHECK(hipEventRecord(e1, s));
            IRLBA_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            IRLBA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (round > 0 && (best_ms[w] == 0.0f || ms < best_ms[w])) best_ms[w] = ms;
        }
    }
    size_t pick = 0;
    for (size_t w = 1; w < widths.size(); ++w)
        if (best_ms[w] < best_ms[pick]) pick = w;
    t.lpl = widths[pick];
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

// Largest input tile (entries) of the product kernel: 156 KB of the 160 KB LDS
// of a CU, the rest holding the value-dictionary table (up to 2 KB) and the
// kernel's reduction scratch.  With tiles of equal width this beat 128 KB on
// every measured product (A.v of the 30000 x 1e6 matrix: FP32 0.762 -> 0.747 ms,
// FP64 1.425 -> 1.406 ms).
int tile_elems(irlba_hip_context* ctx, bool acc64) {
    size_t budget = std::min<size_t>(156 * 1024, ctx->lds_per_block > 8192 ? ctx->lds_per_block - 4096 : 4096);
    if (const char* e = std::getenv("IRLBA_HIP_TILE_BYTES")) budget = std::min<size_t>(budget, std::atol(e));
    return static_cast<int>(budget / (acc64 ? 8 : 4));
}

std::shared_ptr<SparseBase> make_sparse(irlba_hip_context* ctx, i64 nrow, i64 ncol, StorageType in_t,
                                        const void* values, const std::int32_t* indices, const i64* pointers,
                                        bool column_major, StorageType st, bool acc64, bool device_ptrs) {
    require(nrow >= 0 && ncol >= 0, "matrix dimensions must be non-negative");
    require(st == ST_F64 || st == ST_F32, "sparse storage must be F64 or F32");
    require(!(st == ST_F64 && !acc64), "FP64 storage accumulates in FP64");
    require(in_t == ST_F64 || in_t == ST_F32, "sparse input values must be F64 or F32");
    hipStream_t s = ctx->stream;
    const i64 primary = column_major ? ncol : nrow, secondary = column_major ? nrow : ncol;
    i64 nnz = 0;
    // The caller's device arrays may still be written by work on another stream.
    if (device_ptrs) IRLBA_HIP_CHECK(hipDeviceSynchronize());
    if (device_ptrs) IRLBA_HIP_CHECK(hipMemcpy(&nnz, pointers + primary, sizeof(i64), hipMemcpyDeviceToHost));
    else nnz = pointers[primary];
    require(nnz >= 0, "sparse pointers must end at a non-negative count");

    std::shared_ptr<void> u_idx, u_val, u_ptr;
    const std::int32_t* d_idx = indices;
    const void* d_val = values;
    const i64* d_ptr = pointers;
    if (!device_ptrs) {
        u_idx = dalloc_bytes(sizeof(std::int32_t) * nnz);
        u_val = dalloc_bytes(storage_bytes(in_t) * nnz);
        u_ptr = dalloc_bytes(sizeof(i64) * (primary + 1));
        stream_copy(ctx, u_idx.get(), indices, sizeof(std::int32_t) * nnz, hipMemcpyHostToDevice);
        stream_copy(ctx, u_val.get(), values, storage_bytes(in_t) * nnz, hipMemcpyHostToDevice);
        stream_copy(ctx, u_ptr.get(), pointers, sizeof(i64) * (primary + 1), hipMemcpyHostToDevice);
        d_idx = static_cast<const std::int32_t*>(u_idx.get());
        d_val = u_val.get();
        d_ptr = static_cast<const i64*>(u_ptr.get());
    }
    {
        auto flag = dalloc<int>(1);
        IRLBA_HIP_CHECK(hipMemsetAsync(flag.get(), 0, sizeof(int), s));
        launch_validate_compressed(d_idx, d_ptr, primary, secondary, nnz, flag.get(), s);
        int f = 0;
        IRLBA_HIP_CHECK(hipMemcpyAsync(&f, flag.get(), sizeof(int), hipMemcpyDeviceToHost, s));
        IRLBA_HIP_CHECK(hipStreamSynchronize(s));
        require(!(f & 1), "sparse pointers must start at 0, be non-decreasing and end at the number of non-zeros");
        require(!(f & 2), "sparse indices must lie within the secondary dimension");
    }

    auto sp = std::make_shared<SparseBase>();
    sp->ctx = ctx;
    sp->nr = nrow;
    sp->nc = ncol;
    sp->nnz = nnz;
    sp->acc64 = acc64;
    sp->column_major = column_major;
    const int tl = tile_elems(ctx, acc64);
    // by_col: lines are columns, the input (minor) dimension is rows.
    sp->by_col = build_lines(ctx, d_idx, d_val, in_t, d_ptr, primary, !column_major, ncol, nrow, tl, st, acc64, nnz,
                             sp->keep);
    sp->by_row = build_lines(ctx, d_idx, d_val, in_t, d_ptr, primary, column_major, nrow, ncol, tl, st, acc64, nnz,
                             sp->keep);
    const size_t ab = acc64 ? 8 : 4;
    sp->partial_row = dalloc_bytes(ab * std::max<i64>(1, sp->by_row.n_tiles > 1 ? sp->by_row.n_tiles * nrow : 1));
    sp->partial_col = dalloc_bytes(ab * std::max<i64>(1, sp->by_col.n_tiles > 1 ? sp->by_col.n_tiles * ncol : 1));
    sp->gamma_row = dalloc<double>(2 * sp->by_row.n_tiles);
    sp->gamma_col = dalloc<double>(2 * sp->by_col.n_tiles);
    u_idx.reset();  // the raw upload is no longer needed
    u_val.reset();
    u_ptr.reset();
    autotune_lanes(ctx, sp->by_row, acc64, sp->partial_row.get(), sp->gamma_row.get());
    autotune_lanes(ctx, sp->by_col, acc64, sp->partial_col.get(), sp->gamma_col.get());
    return sp;
}

// ------------------------------------------------------------------ value dictionary
// Switches the products of both orientations between the value stream and one
// code byte per slot (kernels_sparse.hip, "value dictionary").  The codes are
// built on the first switch-on: the distinct bit patterns of the stored values
// (at most 256, +0.0 included for the padding and bridging slots), numbered in
// increasing order of value so that the table does not depend on the order in
// which blocks met them.  Returns the table size (0 when switched off).
int set_value_dictionary(irlba_hip_context* ctx, SparseBase& sb, bool on) {
    TiledLines* ts[2] = {&sb.by_row, &sb.by_col};
    if (!on) {
        for (TiledLines* t : ts) {
            if (t->use_dict && t->lpl_values) t->lpl = t->lpl_values;
            t->use_dict = false;
        }
        return 0;
    }
    for (TiledLines* t : ts)
        require(t->delta && t->slots > 0,
                "the value dictionary needs the 8-bit delta layout in both orientations (lines of about 12 or more "
                "entries per tile)");
    if (sb.by_row.cd && sb.by_col.cd) {
        for (TiledLines* t : ts) {
            if (!t->use_dict && t->lpl_dict) t->lpl = t->lpl_dict;
            t->use_dict = true;
        }
        return sb.by_row.n_dict;
    }
    hipStream_t s = ctx->stream;
    const StorageType st = sb.by_row.st;
    const bool f64 = st == ST_F64;
    constexpr int kSet = 1 << 10;  // the global hash set of k_dict_collect
    auto gset = dalloc<unsigned long long>(kSet);
    auto cf = dalloc<int>(2);  // distinct count, flags
    IRLBA_HIP_CHECK(hipMemsetAsync(gset.get(), 0xff, sizeof(unsigned long long) * kSet, s));
    IRLBA_HIP_CHECK(hipMemsetAsync(cf.get(), 0, 2 * sizeof(int), s));
    for (TiledLines* t : ts) launch_dict_collect(t->val, t->st, t->slots, gset.get(), cf.get(), cf.get() + 1, s);
    std::vector<unsigned long long> hset(kSet);
    int hcf[2] = {0, 0};
    IRLBA_HIP_CHECK(hipMemcpyAsync(hset.data(), gset.get(), sizeof(unsigned long long) * kSet, hipMemcpyDeviceToHost, s));
    IRLBA_HIP_CHECK(hipMemcpyAsync(hcf, cf.get(), sizeof(hcf), hipMemcpyDeviceToHost, s));
    IRLBA_HIP_CHECK(hipStreamSynchronize(s));
    const char* too_many = "the matrix values take more than 256 distinct values: no value dictionary";
    require(!(hcf[1] & 1), too_many);
    std::vector<unsigned long long> pats;
    for (unsigned long long p : hset)
        if (p != ~0ull) pats.push_back(p);
    if (std::find(pats.begin(), pats.end(), 0ull) == pats.end()) pats.push_back(0ull);
    require(pats.size() <= static_cast<size_t>(kDictMaxCodes), too_many);
    auto as_value = [f64](unsigned long long p) {
        if (f64) {
            double d;
            std::memcpy(&d, &p, sizeof(d));
            return d;
        }
        const std::uint32_t u = static_cast<std::uint32_t>(p);
        float f;
        std::memcpy(&f, &u, sizeof(f));
        return static_cast<double>(f);
    };
    std::sort(pats.begin(), pats.end(), [&](unsigned long long a, unsigned long long b) {
        const double x = as_value(a), y = as_value(b);
        const bool nx = std::isnan(x), ny = std::isnan(y);
        if (nx != ny) return ny;
        if (!nx && x != y) return x < y;
        return a < b;
    });
    const int n = static_cast<int>(pats.size());
    const size_t vb = storage_bytes(st);
    std::vector<unsigned char> table(vb * n);
    for (int c = 0; c < n; ++c) {
        if (f64) {
            std::memcpy(table.data() + 8 * c, &pats[c], 8);
        } else {
            const std::uint32_t u = static_cast<std::uint32_t>(pats[c]);
            std::memcpy(table.data() + 4 * c, &u, 4);
        }
    }
    auto dpats = dalloc<unsigned long long>(n);
    auto dtab = dalloc_bytes(vb * n);
    stream_copy(ctx, dpats.get(), pats.data(), sizeof(unsigned long long) * n, hipMemcpyHostToDevice);
    stream_copy(ctx, dtab.get(), table.data(), vb * n, hipMemcpyHostToDevice);
    std::shared_ptr<void> cds[2];
    for (int o = 0; o < 2; ++o) {
        // Runs of 16 slots (one cross-lane index scan per 16 slots instead of
        // per 4) when every segment holds whole runs.
        ts[o]->dict_run = ts[o]->seg_align % 16 == 0 ? 16 : 4;
        cds[o] = dalloc_bytes(2 * static_cast<size_t>(ts[o]->slots));
        launch_dict_encode(*ts[o], dpats.get(), n, cds[o].get(), cf.get() + 1, s);
    }
    IRLBA_HIP_CHECK(hipMemcpyAsync(hcf, cf.get(), sizeof(hcf), hipMemcpyDeviceToHost, s));
    IRLBA_HIP_CHECK(hipStreamSynchronize(s));
    require(!(hcf[1] & 2), "internal error: a stored value is missing from the value dictionary");
    void* partial[2] = {sb.partial_row.get(), sb.partial_col.get()};
    double* gamma[2] = {sb.gamma_row.get(), sb.gamma_col.get()};
    for (int o = 0; o < 2; ++o) {
        TiledLines* t = ts[o];
        t->cd = cds[o].get();
        t->dict = dtab.get();
        t->n_dict = n;
        t->lpl_values = t->lpl;
        t->use_dict = true;
        autotune_lanes(ctx, *t, sb.acc64, partial[o], gamma[o]);
        t->lpl_dict = t->lpl;
        sb.keep.push_back(cds[o]);
    }
    sb.keep.push_back(dtab);
    return n;
}

// The host FP64 SVD (HostMat, host_svd) lives in host_svd.cpp.

// ------------------------------------------------------------------ random normals
// Box-Muller pairs from two uniforms in (0, 1] of std::mt19937_64, the method
// of aarand::standard_normal used by fill_with_random_normals (utils.hpp:16-37).
double uniform01(std::mt19937_64& eng) {
    const double range = static_cast<double>(std::mt19937_64::max() - std::mt19937_64::min()) + 1.0;
    double v;
    do {
        v = static_cast<double>(eng() - std::mt19937_64::min()) / range;
    } while (v == 0);
    return v;
}

void random_normals(double* v, i64 n, std::mt19937_64& eng) {
    const double two_pi = 6.283185307179586476925286766559;
    const i64 even = n - (n % 2);
    for (i64 i = 0; i < even; i += 2) {
        const double r = std::sqrt(-2.0 * std::log(uniform01(eng)));
        const double th = two_pi * uniform01(eng);
        v[i] = r * std::sin(th);
        v[i + 1] = r * std::cos(th);
    }
    if (n % 2) {
        const double r = std::sqrt(-2.0 * std::log(uniform01(eng)));
        const double th = two_pi * uniform01(eng);
        v[even] = r * std::sin(th);
    }
}

// ------------------------------------------------------------------ options and helpers
struct Opts {
    double invariant_subspace_tolerance = -1;
    double convergence_tolerance = 1e-5;
    double singular_value_ratio_tolerance = -1;
    int extra_work = 7;
    int max_iterations = 1000;
    bool exact_for_small_matrix = true;
    bool exact_for_large_number = true;
    bool cap_number = false;
    std::uint64_t seed = std::mt19937_64::default_seed;
    const double* initial = nullptr;
    i64 initial_len = 0;
};

Opts to_opts(const irlba_hip_options* p) {
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
    o.initial_len = p->initial_length;
    return o;
}

// compute.hpp:49-56
bool requested_ge_half(i64 requested, i64 smaller) {
    const i64 half = smaller / 2;
    if (requested == half) return smaller % 2 == 0;
    return requested > half;
}
// compute.hpp:59-67
i64 choose_work(i64 requested, int extra_work, i64 smaller) {
    if (requested >= smaller) return smaller;
    return requested + std::min<i64>(smaller - requested, extra_work);
}
// compute.hpp:78-111
i64 update_k(i64 k, i64 requested, i64 n_converged, i64 work) {
    if (work <= 3) return 1;
    const i64 limit = work - 3;
    auto below = [&](i64 v) { return v < requested || (v - requested) < n_converged; };
    if (below(k)) {
        if (below(limit)) return limit;
        return std::max<i64>(requested + n_converged, 1);
    }
    return std::max<i64>(std::min(k, limit), 1);
}

// ------------------------------------------------------------------ operators seen by the engine
struct Op {
    irlba_hip_context* ctx;
    i64 nr, nc;
    virtual ~Op() {}
    virtual void apply(bool adjoint, const double* x, double* y, const double* src, const double* coef) = 0;
    virtual void realize_host(HostMat& out) = 0;
};

struct HandleOp : Op {
    const irlba_hip_matrix* m;
    void apply(bool adjoint, const double* x, double* y, const double* src, const double* coef) override {
        apply_handle(*m, adjoint, x, y, src, coef);
    }
    void realize_host(HostMat& out) override;
};

struct CallbackOp : Op {
    irlba_hip_matvec_fn fn = nullptr;
    void* user = nullptr;
    std::vector<double> hin, hout;
    void apply(bool adjoint, const double* x, double* y, const double* src, const double* coef) override {
        const i64 nin = adjoint ? nr : nc, nout = adjoint ? nc : nr;
        hin.resize(nin);
        hout.resize(nout);
        IRLBA_HIP_CHECK(hipMemcpyAsync(hin.data(), x, sizeof(double) * nin, hipMemcpyDeviceToHost, ctx->stream));
        IRLBA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (fn(user, adjoint ? 1 : 0, hin.data(), hout.data()) != 0)
            throw std::runtime_error("matrix product callback failed");
        if (adjoint) ++ctx->stats.adjoint_products; else ++ctx->stats.forward_products;
        IRLBA_HIP_CHECK(hipMemcpyAsync(y, hout.data(), sizeof(double) * nout, hipMemcpyHostToDevice, ctx->stream));
        IRLBA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (src) launch_axpy_dev(y, src, coef, nout, ctx->stream);
    }
    void realize_host(HostMat& out) override {
        out = HostMat(nr, nc);
        std::vector<double> e(nc, 0.0);
        for (i64 c = 0; c < nc; ++c) {
            e[c] = 1;
            if (fn(user, 0, e.data(), out.col(c)) != 0) throw std::runtime_error("matrix product callback failed");
            e[c] = 0;
        }
    }
};

void HandleOp::realize_host(HostMat& out) {
    auto dense = dalloc<double>(nr * nc);
    IRLBA_HIP_CHECK(hipMemsetAsync(dense.get(), 0, sizeof(double) * nr * nc, ctx->stream));
    m->base->realize(dense.get());
    // Decorators on the realized matrix (centered.hpp:118-122, scaled.hpp:158-177) on the host.
    out = HostMat(nr, nc);
    IRLBA_HIP_CHECK(hipMemcpyAsync(out.a.data(), dense.get(), sizeof(double) * nr * nc, hipMemcpyDeviceToHost,
                                   ctx->stream));
    IRLBA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const Affine& a = m->aff;
    auto fetch = [&](const std::shared_ptr<double>& v, i64 n) {
        std::vector<double> h;
        if (v) {
            h.resize(n);
            stream_copy(ctx, h.data(), v.get(), sizeof(double) * n, hipMemcpyDeviceToHost);
        }
        return h;
    };
    const auto dr = fetch(a.dr, nr), dc = fetch(a.dc, nc);
    for (i64 c = 0; c < nc; ++c)
        for (i64 r = 0; r < nr; ++r) {
            double v = out(r, c);
            if (!dr.empty()) v *= dr[r];
            if (!dc.empty()) v *= dc[c];
            out(r, c) = v;
        }
    for (int k = 0; k < a.n_terms; ++k) {
        const auto av = fetch(a.a[k], nr), bv = fetch(a.b[k], nc);
        for (i64 c = 0; c < nc; ++c)
            for (i64 r = 0; r < nr; ++r) out(r, c) -= (av.empty() ? 1.0 : av[r]) * bv[c];
    }
}

// ------------------------------------------------------------------ the engine
struct Engine {
    irlba_hip_context* ctx;
    Op& op;
    const Opts& o;
    hipStream_t s;
    i64 nr, nc, work;
    double eps;
    std::shared_ptr<double> V, W, F, Wn, dnorm, partials, tvec, norm_part, Yt;
    int ldy = 0;
    std::mt19937_64 eng;
    std::vector<double> S, R;  // host copies of the sweep norms
    int refills = 0;

    Engine(irlba_hip_context* c, Op& op_, const Opts& o_, i64 work_)
        : ctx(c), op(op_), o(o_), s(c->stream), nr(op_.nr), nc(op_.nc), work(work_), eng(o_.seed) {
        const double raw = o.invariant_subspace_tolerance;
        eps = raw < 0 ? std::pow(std::numeric_limits<double>::epsilon(), 0.8) : raw;
        V = dalloc<double>(nc * work);
        W = dalloc<double>(nr * work);
        F = dalloc<double>(nc);
        Wn = dalloc<double>(nr);
        dnorm = dalloc<double>(2 * work + 4);
        const int G = std::max(cgs_blocks(nr), cgs_blocks(nc));
        partials = dalloc<double>(static_cast<i64>(G) * work);
        tvec = dalloc<double>(work);
        norm_part = dalloc<double>(G);
        ldy = static_cast<int>((work + 15) / 16 * 16);
        Yt = dalloc<double>(2 * work * ldy);
        IRLBA_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&yt_host), sizeof(double) * 2 * work * ldy));
        IRLBA_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&norms_host), sizeof(double) * (2 * work + 4)));
        IRLBA_HIP_CHECK(hipEventCreateWithFlags(&norms_ready, hipEventDisableTiming));
        ctx->ensure_copy_stream();
        IRLBA_HIP_CHECK(hipMemsetAsync(W.get(), 0, sizeof(double) * nr * work, s));
        IRLBA_HIP_CHECK(hipMemsetAsync(V.get(), 0, sizeof(double) * nc * work, s));
        IRLBA_HIP_CHECK(hipMemsetAsync(dnorm.get(), 0, sizeof(double) * (2 * work + 4), s));
        S.assign(work, 0.0);
        R.assign(work, 0.0);
    }

    double* Vc(i64 j) { return V.get() + j * nc; }
    double* Wc(i64 j) { return W.get() + j * nr; }
    double* dS(i64 j) { return dnorm.get() + j; }
    double* dR(i64 j) { return dnorm.get() + work + 1 + j; }
    double* dTmp() { return dnorm.get() + 2 * work + 2; }

    // f -= M[:, :n] M[:, :n]' f (lanczos.hpp:17-21); leaves the sums of squares of f in norm_part.
    void cgs(double* M, i64 L, i64 n, double* f) {
        if (n == 0) {
            launch_sumsq(f, L, norm_part.get(), s);
            return;
        }
        launch_cgs_dots(M, L, static_cast<int>(n), f, L, partials.get(), s);
        launch_cgs_reduce(partials.get(), L, static_cast<int>(n), tvec.get(), s);
        launch_cgs_update(M, L, static_cast<int>(n), tvec.get(), f, L, norm_part.get(), s);
    }
    void norm_to(double* f, i64 L, double* norm_dev, double* dst) {
        launch_normalize(f, L, norm_part.get(), cgs_blocks(L), norm_dev, dst, s);
    }

    // lanczos.hpp:73-86.  `product_ready`: W_next already holds A V[:, start]
    // (queued by the restart before the host SVD, see run_compute).
    void first(i64 start, bool product_ready = false) {
        if (!product_ready) op.apply(false, Vc(start), Wn.get(), nullptr, nullptr);
        cgs(W.get(), nr, start, Wn.get());
        norm_to(Wn.get(), nr, dS(start), Wc(start));
    }
    // lanczos.hpp:93-113 (half-step 2j)
    void phase_a(i64 j) {
        op.apply(true, Wc(j), F.get(), Vc(j), dS(j));
        cgs(V.get(), nc, j + 1, F.get());
        norm_to(F.get(), nc, dR(j), j + 1 < work ? Vc(j + 1) : nullptr);
    }
    // lanczos.hpp:115-134 (half-step 2j+1)
    void phase_b(i64 j) {
        op.apply(false, Vc(j + 1), Wn.get(), Wc(j), dR(j));
        cgs(W.get(), nr, j + 1, Wn.get());
        norm_to(Wn.get(), nr, dS(j + 1), Wc(j + 1));
    }
    void launch_from(i64 h) {
        const i64 hend = 2 * (work - 1);  // the last half-step is phase A of j = work-1
        for (; h <= hend; ++h) {
            if (h % 2 == 0) phase_a(h / 2); else phase_b(h / 2);
        }
    }
    // The next sweep starts from V[:, k] = F / R_F (compute.hpp:280) whatever k
    // the restart picks, so its first product A V[:, k] can be queued as soon
    // as the sweep's kernels are: F is scaled in place (R_F read on the
    // device) and A F lands in W_next.  After an invariant-subspace refill the
    // sweep recomputes F and W_next, so a queued product is only ever wasted.
    void queue_next_product() {
        launch_scale_copy_dev(F.get(), nc, dR(work - 1), F.get(), s);
        op.apply(false, F.get(), Wn.get(), nullptr, nullptr);
    }
    // Copies the norms to the host and waits for that copy only, so work
    // queued behind it (the speculative first product) keeps the GPU busy
    // while the host runs the invariant checks and the B-SVD.
    void read_norms(bool speculate = false) {
        const i64 n = 2 * work + 4;
        IRLBA_HIP_CHECK(hipMemcpyAsync(norms_host, dnorm.get(), sizeof(double) * n, hipMemcpyDeviceToHost, s));
        IRLBA_HIP_CHECK(hipEventRecord(norms_ready, s));
        const size_t done = ctx->pending.size();
        if (speculate) queue_next_product();
        IRLBA_HIP_CHECK(hipEventSynchronize(norms_ready));
        ctx->resolve_events(done);
        for (i64 j = 0; j < work; ++j) {
            S[j] = norms_host[j];
            R[j] = norms_host[work + 1 + j];
        }
    }
    void set_dev_scalar(double* d, double v) {
        IRLBA_HIP_CHECK(hipMemcpyAsync(d, &v, sizeof(double), hipMemcpyHostToDevice, s));
        IRLBA_HIP_CHECK(hipStreamSynchronize(s));
    }
    void upload_normals(double* dst, i64 n) {
        std::vector<double> h(n);
        random_normals(h.data(), n, eng);
        IRLBA_HIP_CHECK(hipMemcpyAsync(dst, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
        IRLBA_HIP_CHECK(hipStreamSynchronize(s));
    }

    // run_lanczos_bidiagonalization (lanczos.hpp:54-143) from column `start`; fills B.
    // With `speculate`, the next sweep's first product is queued behind every
    // read of the norms (see queue_next_product).
    void sweep(i64 start, HostMat& B, bool product_ready = false, bool speculate = false) {
        first(start, product_ready);
        launch_from(2 * start);
        read_norms(speculate);
        if (S[start] < eps) throw std::runtime_error("starting vector near the null space of the input matrix");
        i64 h = 2 * start;
        const i64 hcheck_end = 2 * (work - 1);  // half-steps with an invariant-subspace test
        while (true) {
            i64 ev = -1;
            for (; h < hcheck_end; ++h) {
                const i64 j = h / 2;
                const double nrm = (h % 2 == 0) ? R[j] : S[j + 1];
                if (nrm < eps) {
                    ev = h;
                    break;
                }
            }
            if (ev < 0) break;
            const i64 j = ev / 2;
            ++refills;
            if (ev % 2 == 0) {
                // lanczos.hpp:101-106: recompute F, refill with normals, re-orthogonalise, R_F = 0.
                op.apply(true, Wc(j), F.get(), Vc(j), dS(j));
                cgs(V.get(), nc, j + 1, F.get());
                upload_normals(F.get(), nc);
                cgs(V.get(), nc, j + 1, F.get());
                norm_to(F.get(), nc, dTmp(), Vc(j + 1));
                set_dev_scalar(dR(j), 0.0);
            } else {
                // lanczos.hpp:124-130: the same for W_next, S = 0.
                op.apply(false, Vc(j + 1), Wn.get(), Wc(j), dR(j));
                cgs(W.get(), nr, j + 1, Wn.get());
                upload_normals(Wn.get(), nr);
                cgs(W.get(), nr, j + 1, Wn.get());
                norm_to(Wn.get(), nr, dTmp(), Wc(j + 1));
                set_dev_scalar(dS(j + 1), 0.0);
            }
            launch_from(ev + 1);
            read_norms(speculate);
            h = ev + 1;
        }
        for (i64 j = start; j < work; ++j) {
            B(j, j) = S[j];
            if (j + 1 < work) B(j, j + 1) = R[j];
        }
    }

    // Rotation coefficients Y[:, :k] go row-major (zero padded to ldy columns)
    // into one of two slots of a pinned staging buffer; upload_y copies slots
    // [0, n) in one transfer, so the V and W rotations of a restart cost one
    // host-to-device copy.  A slot is rewritten only after the next sweep's
    // synchronisation (or an explicit one).
    void stage_y(int slot, const HostMat& Y, i64 k) {
        double* yt = yt_host + static_cast<i64>(slot) * work * ldy;
        std::fill(yt, yt + work * ldy, 0.0);
        for (i64 r = 0; r < work; ++r)
            for (i64 c = 0; c < k; ++c) yt[r * ldy + c] = Y(r, c);
    }
    // The copy runs on a stream of its own, so that it overlaps the product
    // queued before the host SVD instead of waiting behind it on the copy
    // engine; the compute stream waits for it before the rotations.  (Every
    // earlier reader of the device slots precedes the norms read-back the host
    // has waited for.)
    void upload_y(int nslots) {
        IRLBA_HIP_CHECK(hipMemcpyAsync(Yt.get(), yt_host, sizeof(double) * nslots * work * ldy, hipMemcpyHostToDevice,
                                       ctx->copy_stream));
        IRLBA_HIP_CHECK(hipEventRecord(ctx->copy_done, ctx->copy_stream));
        IRLBA_HIP_CHECK(hipStreamWaitEvent(s, ctx->copy_done, 0));
    }
    // out[:, :k] = M[:, :work] Y[:, :k] with the staged slot's Y (out == M allowed).
    // The rotation kernels work in place for up to 128 output columns (every
    // row is read before any of its outputs is written); a wider in-place
    // restart (number + extra_work above ~130) rotates into a scratch basis,
    // as the reference does with Vtmp / Wtmp (compute.hpp:268,279), and
    // copies it back in stream order.
    void rotate_staged(int slot, double* M, i64 L, i64 k, double* out, i64 ld_out) {
        const double* dY = Yt.get() + static_cast<i64>(slot) * work * ldy;
        if (out == M && k > kRotateInPlaceMax) {
            if (rot_tmp_len < L * k) {
                IRLBA_HIP_CHECK(hipStreamSynchronize(s));
                rot_tmp = dalloc<double>(L * k);
                rot_tmp_len = L * k;
            }
            launch_rotate(M, L, static_cast<int>(work), dY, ldy, static_cast<int>(k), rot_tmp.get(), L, L, s);
            IRLBA_HIP_CHECK(hipMemcpy2DAsync(out, sizeof(double) * ld_out, rot_tmp.get(), sizeof(double) * L,
                                             sizeof(double) * L, k, hipMemcpyDeviceToDevice, s));
            return;
        }
        launch_rotate(M, L, static_cast<int>(work), dY, ldy, static_cast<int>(k), out, ld_out, L, s);
    }
    static constexpr i64 kRotateInPlaceMax = 128;
    std::shared_ptr<double> rot_tmp;
    i64 rot_tmp_len = 0;

    double* yt_host = nullptr;
    double* norms_host = nullptr;  // pinned, so the norms copy is asynchronous
    hipEvent_t norms_ready = nullptr;
    ~Engine() {
        (void)hipStreamSynchronize(s);
        (void)hipStreamSynchronize(ctx->copy_stream);
        if (yt_host) (void)hipHostFree(yt_host);
        if (norms_host) (void)hipHostFree(norms_host);
        if (norms_ready) (void)hipEventDestroy(norms_ready);
    }
};

// Exact SVD of the realized matrix (compute.hpp:30-46).
void exact_svd(Op& op, i64 requested, double* U, double* V, double* D, bool device_ptrs, hipStream_t s) {
    HostMat A;
    op.realize_host(A);
    HostMat u, v;
    std::vector<double> sv;
    host_svd(A, u, sv, v);
    const auto kind = device_ptrs ? hipMemcpyHostToDevice : hipMemcpyHostToHost;
    IRLBA_HIP_CHECK(hipMemcpyAsync(D, sv.data(), sizeof(double) * requested, kind, s));
    IRLBA_HIP_CHECK(hipMemcpyAsync(U, u.a.data(), sizeof(double) * u.nr * requested, kind, s));
    IRLBA_HIP_CHECK(hipMemcpyAsync(V, v.a.data(), sizeof(double) * v.nr * requested, kind, s));
    IRLBA_HIP_CHECK(hipStreamSynchronize(s));
}

// compute.hpp:141-308.  Outputs sized for `number`; returns (converged, iterations).
std::pair<bool, int> run_compute(irlba_hip_context* ctx, Op& op, i64 number, const Opts& o, double* U, double* V,
                                 double* D, i64* nout, bool device_ptrs) {
    const auto t0 = std::chrono::steady_clock::now();
    const i64 nr = op.nr, nc = op.nc;
    const i64 smaller = std::min(nr, nc);
    // The reference takes `number` as an Eigen::Index and sizes its outputs
    // from it; a negative count has no meaning there and would size the
    // output copies here from a wrapped value.
    if (number < 0) throw std::runtime_error("requested number of singular values must be non-negative");
    i64 requested = number;
    if (requested > smaller) {
        if (o.cap_number) requested = smaller;
        else throw std::runtime_error("requested number of singular values cannot be greater than the smaller matrix dimension");
    } else if (requested == smaller && !o.exact_for_large_number) {
        throw std::runtime_error("requested number of singular values must be less than the smaller matrix dimension for IRLBA iterations");
    }
    *nout = requested;
    auto& st = ctx->stats;
    st = irlba_hip_stats{};
    if ((o.exact_for_small_matrix && smaller < 6) || (o.exact_for_large_number && requested_ge_half(requested, smaller))) {
        exact_svd(op, requested, U, V, D, device_ptrs, ctx->stream);
        st.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return {true, 0};
    }
    const i64 work = choose_work(requested, o.extra_work, smaller);
    if (work == 0) throw std::runtime_error("number of requested dimensions must be positive");

    hipStream_t s = ctx->stream;
    Engine eg(ctx, op, o, work);
    {
        std::vector<double> v0(nc);
        if (o.initial) {
            if (o.initial_len != nc)
                throw std::runtime_error("initialization vector does not have expected number of rows");
            std::copy(o.initial, o.initial + nc, v0.begin());
        } else {
            random_normals(v0.data(), nc, eg.eng);
        }
        IRLBA_HIP_CHECK(hipMemcpyAsync(eg.Vc(0), v0.data(), sizeof(double) * nc, hipMemcpyHostToDevice, s));
        launch_sumsq(eg.Vc(0), nc, eg.norm_part.get(), s);
        eg.norm_to(eg.Vc(0), nc, eg.dTmp(), eg.Vc(0));
        IRLBA_HIP_CHECK(hipStreamSynchronize(s));
    }

    bool converged = false;
    int iter = 0;
    i64 k = 0;
    HostMat B(work, work), BU, BV;
    std::vector<double> BS, res(work), prevS(work, 0.0);
    const double tol = o.convergence_tolerance;
    const double svtol = o.singular_value_ratio_tolerance >= 0 ? o.singular_value_ratio_tolerance : tol;

    for (; iter < o.max_iterations; ++iter) {
        // The sweep queues the next sweep's first product A (F / R_F) before it
        // waits for the norms, so it runs on the GPU while the host checks the
        // norms and computes the SVD of B; F / R_F is copied into V[:, k] after
        // the V rotation has read the old basis.  (Wasted once, when this
        // restart converges.)
        eg.sweep(k, B, iter > 0, true);
        const double R_F = eg.R[work - 1];
        const auto ts = std::chrono::steady_clock::now();
        host_svd(B, BU, BS, BV);
        st.host_svd_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts).count();
        ++st.restarts;
        if (B(work - 1, work - 1) == 0) {
            converged = true;
            break;
        }
        for (i64 j = 0; j < work; ++j) res[j] = R_F * BU(work - 1, j);
        i64 n_conv = 0;
        if (iter > 0) {
            const double Smax = *std::max_element(BS.begin(), BS.end());
            const double thr = Smax * tol;
            for (i64 j = 0; j < work; ++j) {
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
        eg.stage_y(0, BV, k);
        eg.stage_y(1, BU, k);
        eg.upload_y(2);
        eg.rotate_staged(0, eg.V.get(), nc, k, eg.V.get(), nc);
        IRLBA_HIP_CHECK(hipMemcpyAsync(eg.Vc(k), eg.F.get(), sizeof(double) * nc, hipMemcpyDeviceToDevice, s));
        eg.rotate_staged(1, eg.W.get(), nr, k, eg.W.get(), nr);
        std::fill(B.a.begin(), B.a.end(), 0.0);
        for (i64 l = 0; l < k; ++l) {
            B(l, l) = BS[l];
            B(l, k) = res[l];
        }
    }

    // compute.hpp:298-305 (with the last B-SVD, as the reference does).
    if (BS.empty()) {
        BS.assign(work, 0.0);
        BU = HostMat(work, work);
        BV = HostMat(work, work);
    }
    std::shared_ptr<double> dU, dV;
    double* oU = U;
    double* oV = V;
    if (!device_ptrs) {
        dU = dalloc<double>(nr * requested);
        dV = dalloc<double>(nc * requested);
        oU = dU.get();
        oV = dV.get();
    }
    if (requested > 0) {
        // After max_iterations the last restart's rotations may still be reading
        // the pinned staging slots; finish them before the slots are rewritten.
        IRLBA_HIP_CHECK(hipStreamSynchronize(s));
        eg.stage_y(0, BU, requested);
        eg.stage_y(1, BV, requested);
        eg.upload_y(2);
        eg.rotate_staged(0, eg.W.get(), nr, requested, oU, nr);
        eg.rotate_staged(1, eg.V.get(), nc, requested, oV, nc);
    }
    const auto kind = device_ptrs ? hipMemcpyHostToDevice : hipMemcpyHostToHost;
    IRLBA_HIP_CHECK(hipMemcpyAsync(D, BS.data(), sizeof(double) * requested, kind, s));
    if (!device_ptrs && requested > 0) {
        IRLBA_HIP_CHECK(hipMemcpyAsync(U, oU, sizeof(double) * nr * requested, hipMemcpyDeviceToHost, s));
        IRLBA_HIP_CHECK(hipMemcpyAsync(V, oV, sizeof(double) * nc * requested, hipMemcpyDeviceToHost, s));
    }
    ctx->sync();
    st.refills = eg.refills;
    st.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return {converged, converged ? iter + 1 : iter};
}

// ------------------------------------------------------------------ decorator composition
std::shared_ptr<double> combine(irlba_hip_context* ctx, const double* a, const double* b, i64 n, int op) {
    auto out = dalloc<double>(n);
    launch_vec_combine(out.get(), a, b, n, op, ctx->stream);
    return out;
}

irlba_hip_matrix* decorate_center(const irlba_hip_matrix* in, std::shared_ptr<double> c) {
    auto* m = new irlba_hip_matrix(*in);
    m->decorated = true;
    Affine& a = m->aff;
    for (int k = 0; k < a.n_terms; ++k) {
        if (!a.a[k]) {  // merge with the existing ones-term: b += c
            a.b[k] = combine(m->ctx, a.b[k].get(), c.get(), m->base->nc, 2);
            IRLBA_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
            return m;
        }
    }
    if (a.n_terms >= kMaxTerms) {
        delete m;
        throw std::runtime_error("decorator chain has too many centring terms for the device operator");
    }
    a.a[a.n_terms] = nullptr;
    a.b[a.n_terms] = std::move(c);
    ++a.n_terms;
    return m;
}

irlba_hip_matrix* decorate_scale(const irlba_hip_matrix* in, const std::shared_ptr<double>& sdev, bool column,
                                 bool divide) {
    auto* m = new irlba_hip_matrix(*in);
    m->decorated = true;
    irlba_hip_context* ctx = m->ctx;
    const i64 n = column ? m->base->nc : m->base->nr;
    auto f = divide ? combine(ctx, sdev.get(), nullptr, n, 3) : sdev;  // 1/s or s
    Affine& a = m->aff;
    if (column) {
        a.dc = a.dc ? combine(ctx, a.dc.get(), f.get(), n, 0) : f;
        for (int k = 0; k < a.n_terms; ++k) a.b[k] = combine(ctx, a.b[k].get(), f.get(), n, 0);
    } else {
        a.dr = a.dr ? combine(ctx, a.dr.get(), f.get(), n, 0) : f;
        for (int k = 0; k < a.n_terms; ++k) a.a[k] = a.a[k] ? combine(ctx, a.a[k].get(), f.get(), n, 0) : f;
    }
    IRLBA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return m;
}

// ------------------------------------------------------------------ differentiable top-k SVD
// dL/dA for L = sum_i w_i s_i + <G_U, U> + <G_V, V> as X Y' with
//   X = [U | U M + Z'],  Y = [Y' | V],
//   M = diag(w) + (F o (J - J')) S + S (F o (K - K')),  J = U'G_U, K = V'G_V,
//   F_ij = 1 / (s_j^2 - s_i^2) (i != j).
// Complement columns (the part of dU, dV outside the top-k subspaces):
//   mode 1: Z' = (G_U - U J) S^-1, Y' = (G_V - V K) S^-1 (exact when rank(A) = k);
//   mode 2: Z'_i = s_i z_i + A t_i, Y'_i = A' z_i + s_i t_i with
//           (s_i^2 I - P A A' P) z_i = P g_i on P = I - U U' and
//           (s_i^2 I - Q A'A Q) t_i = Q h_i on Q = I - V V', solved by CG
//           (positive definite when s_k > s_{k+1}).
// Projections are applied as y - U (U'y) (a Gram product and a rotation); no
// dense I - UU' is formed.  The 2k CG systems run in lock step: one iteration
// is 2k products of each kind plus column-batched kernels, with one host
// synchronisation for the convergence test.

// One set of k shifted systems (sig2_i I - P B P) x_i = b_i on the complement
// of `basis` (L x k), B = A A' (left, L = rows) or A' A (right, L = cols).
struct CgSet {
    bool left = true;
    i64 L = 0, Lo = 0;
    const double* basis = nullptr;
    double *b = nullptr, *x = nullptr, *r = nullptr, *p = nullptr, *ap = nullptr, *uc = nullptr, *tmp = nullptr;
    double *proj = nullptr;                                       // k x ldy Gram output (rotation layout)
    double *rr = nullptr, *rr_new = nullptr, *pap = nullptr, *alpha = nullptr, *beta = nullptr;  // k each
    double* sig2 = nullptr;                                       // k
    double* partials = nullptr;                                   // Gram / column-dot partials
};

struct LockstepCg {
    const irlba_hip_matrix* m;
    hipStream_t s;
    int k, ldy;

    // ap holds B in; ap = basis (basis' ap) - ap + sig2 in = sig2 in - P B in.
    void project(CgSet& c, const double* in) {
        launch_gram(c.basis, c.L, c.ap, c.L, k, c.L, c.partials, c.proj, ldy, s);
        launch_rotate(c.basis, c.L, k, c.proj, ldy, k, c.uc, c.L, c.L, s);
        launch_cg_op_finish(c.ap, c.uc, in, c.sig2, c.L, k, s);
    }
    // out = sig2 p - P B p for every column, into set.ap.
    void op(CgSet& c, const double* in) {
        apply_handle_multi(*m, c.left, in, c.L, c.tmp, c.Lo, k);      // A' in (left) or A in (right)
        apply_handle_multi(*m, !c.left, c.tmp, c.Lo, c.ap, c.L, k);   // then A (A' in) or A' (A in)
        project(c, in);
    }
    void start(CgSet& c) {
        // x0 = b / sig2 (the rank-k answer), r0 = b - op(x0), p0 = r0.
        launch_cg_coefs(c.sig2, nullptr, c.alpha, k, 3, s);  // alpha = 1 / sig2
        launch_cols_axpby(c.x, c.L, nullptr, 0.0, c.b, c.L, c.alpha, 1.0, k, c.L, s);
        op(c, c.x);
        launch_cols_axpby(c.r, c.L, nullptr, 0.0, c.b, c.L, nullptr, 1.0, k, c.L, s);
        launch_cols_axpby(c.r, c.L, nullptr, 1.0, c.ap, c.L, nullptr, -1.0, k, c.L, s);
        launch_cols_axpby(c.p, c.L, nullptr, 0.0, c.r, c.L, nullptr, 1.0, k, c.L, s);
        launch_coldots(c.r, c.L, c.r, c.L, k, c.L, c.partials, c.rr, s);
    }
    void step(CgSet& c) {
        apply_handle_multi(*m, c.left, c.p, c.L, c.tmp, c.Lo, k);
        apply_handle_multi(*m, !c.left, c.tmp, c.Lo, c.ap, c.L, k);
        project_update(c);
    }
    // ap holds B p: the projection terms, then the CG update (fused kernels:
    // alpha = rr / p'ap, x += alpha p, r -= alpha ap, beta = rr_new / rr,
    // p = beta p + r; rr_new becomes rr, the two buffers swap).
    void project_update(CgSet& c) {
        launch_gram(c.basis, c.L, c.ap, c.L, k, c.L, c.partials, c.proj, ldy, s);
        launch_rotate(c.basis, c.L, k, c.proj, ldy, k, c.uc, c.L, c.L, s);
        launch_cg_fused(c.ap, c.uc, c.p, c.x, c.r, c.sig2, c.rr, c.rr_new, c.L, k, c.partials, s);
        std::swap(c.rr, c.rr_new);
    }
    void finish(CgSet& c) {
        // Drop the rounding drift of x back out of span(basis).
        launch_gram(c.basis, c.L, c.x, c.L, k, c.L, c.partials, c.proj, ldy, s);
        launch_rotate(c.basis, c.L, k, c.proj, ldy, k, c.uc, c.L, c.L, s);
        launch_cols_axpby(c.x, c.L, nullptr, 1.0, c.uc, c.L, nullptr, -1.0, k, c.L, s);
    }
};

// `queued`: device arrays only; returns with the work queued on the context's
// stream and leaves out the read-back of the singular-value check after the
// rank-k and projected paths (the caller has checked the values).
int run_svd_backward(const irlba_hip_matrix* m, i64 k, const double* U, const double* sv, const double* V,
                     const double* w, const double* GU, const double* GV, int complement, int max_cg, double cg_tol,
                     double* X, double* Y, double* dcenter, bool device_ptrs, bool queued) {
    irlba_hip_context* ctx = m->ctx;
    hipStream_t s = ctx->stream;
    const i64 nr = m->base->nr, nc = m->base->nc, R = 2 * k;
    require(k > 0 && k <= std::min(nr, nc) && k <= 64, "k must lie in [1, min(rows, cols)] and be at most 64");
    const int ki = static_cast<int>(k);
    const int ldy = static_cast<int>((k + 15) / 16 * 16);  // row stride of rotation coefficients
    const bool exact = complement == 2;

    // Workspace: host inputs and outputs are staged in it; device ones are used in place.
    const int gb = gram_blocks(std::max(nr, nc));
    const int cb = std::max(cgs_blocks(nr), cgs_blocks(nc));
    auto plan = [&](Carve& c, std::vector<double*>& p) {
        p.clear();
        p.push_back(device_ptrs ? nullptr : c.take<double>(nr * k));  // 0 U
        p.push_back(device_ptrs ? nullptr : c.take<double>(nc * k));  // 1 V
        p.push_back(device_ptrs ? nullptr : c.take<double>(nr * k));  // 2 G_U
        p.push_back(device_ptrs ? nullptr : c.take<double>(nc * k));  // 3 G_V
        p.push_back(device_ptrs ? nullptr : c.take<double>(nr * R));  // 4 X
        p.push_back(device_ptrs ? nullptr : c.take<double>(nc * R));  // 5 Y
        p.push_back(c.take<double>(static_cast<i64>(std::max(gb, cb)) * std::max<i64>(k * k, R)));  // 6 partials
        p.push_back(c.take<double>(k * ldy));      // 7 J (rotation layout)
        p.push_back(c.take<double>(k * ldy));      // 8 K
        p.push_back(c.take<double>(2 * k * ldy + 4 * k));  // 9 coefficients CX, CY, d, sig2 (+ staged s, w)
        p.push_back(dcenter ? c.take<double>(nr + R + nc) : nullptr);  // 10 ones, column sums, centre gradient
        if (exact) {
            for (int side = 0; side < 2; ++side) {
                const i64 L = side == 0 ? nr : nc, Lo = side == 0 ? nc : nr;
                for (int v = 0; v < 6; ++v) p.push_back(c.take<double>(L * k));  // b x r p ap uc
                p.push_back(c.take<double>(Lo * k));                            // tmp
                p.push_back(c.take<double>(k * ldy));                           // proj
                p.push_back(c.take<double>(6 * k));                             // rr rr_new pap alpha beta (+1)
            }
        }
    };
    std::vector<double*> wp;
    Carve sizing;
    plan(sizing, wp);
    Carve carve;
    carve.base = ctx->workspace(sizing.off);
    plan(carve, wp);

    const double *dU = U, *dV = V, *dGU = GU, *dGV = GV;
    double *Xo = X, *Yo = Y, *dc = dcenter;
    if (!device_ptrs) {
        auto up = [&](double* dst, const double* src, i64 n) {
            IRLBA_HIP_CHECK(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, s));
            return dst;
        };
        dU = up(wp[0], U, nr * k);
        dV = up(wp[1], V, nc * k);
        dGU = up(wp[2], GU, nr * k);
        dGV = up(wp[3], GV, nc * k);
        Xo = wp[4];
        Yo = wp[5];
        dc = dcenter ? wp[10] + nr + R : nullptr;
    }
    double* partials = wp[6];
    double *dJ = wp[7], *dK = wp[8], *coef = wp[9];
    double *dCX = coef, *dCY = coef + k * ldy, *dd = dCY + k * ldy, *dsig2 = dd + k;
    const double *dsv = sv, *dw = w;
    if (!device_ptrs) {  // stage s and w behind the coefficients
        double* st = dsig2 + k;
        IRLBA_HIP_CHECK(hipMemcpyAsync(st, sv, sizeof(double) * k, hipMemcpyHostToDevice, s));
        IRLBA_HIP_CHECK(hipMemcpyAsync(st + k, w, sizeof(double) * k, hipMemcpyHostToDevice, s));
        dsv = st;
        dw = st + k;
    }
    // J = U'G_U, K = V'G_V, laid out as rotation coefficients (J(i, j) at i * ldy + j);
    // the coefficients follow on the device (no host round trip); the check of
    // the singular values is read back with the results.
    int* flag = ctx->flag();
    launch_gram(dU, nr, dGU, nr, ki, nr, partials, dJ, ldy, s);
    launch_gram(dV, nc, dGV, nc, ki, nc, partials, dK, ldy, s);
    launch_backward_coefs(dJ, dK, ldy, dsv, dw, ki, complement, dCX, dCY, dd, dsig2, flag, s);
    auto check_values = [&] {
        int hflag = 0;
        IRLBA_HIP_CHECK(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
        IRLBA_HIP_CHECK(hipStreamSynchronize(s));
        require(!(hflag & 1), "singular values must be positive and non-increasing");
    };
    if (exact) check_values();  // before the CG, which would not converge on them

    // X = [U | U CX (+ G_U S^-1)], Y = [Y' | V].
    double *X0 = Xo, *X1 = Xo + nr * k, *Y0 = Yo, *Y1 = Yo + nc * k;
    IRLBA_HIP_CHECK(hipMemcpyAsync(X0, dU, sizeof(double) * nr * k, hipMemcpyDeviceToDevice, s));
    IRLBA_HIP_CHECK(hipMemcpyAsync(Y1, dV, sizeof(double) * nc * k, hipMemcpyDeviceToDevice, s));
    launch_rotate(dU, nr, ki, dCX, ldy, ki, X1, nr, nr, s);
    int iterations = 0;
    if (complement == 0) {
        launch_fill(Y0, 0.0, nc * k, s);
    } else if (complement == 1) {
        launch_cols_axpby(X1, nr, nullptr, 1.0, dGU, nr, dd, 1.0, ki, nr, s);
        launch_rotate(dV, nc, ki, dCY, ldy, ki, Y0, nc, nc, s);
        launch_cols_axpby(Y0, nc, nullptr, 1.0, dGV, nc, dd, 1.0, ki, nc, s);
    } else {
        LockstepCg cg{m, s, ki, ldy};
        CgSet set[2];
        size_t at = 11;
        for (int side = 0; side < 2; ++side) {
            CgSet& c = set[side];
            c.left = side == 0;
            c.L = c.left ? nr : nc;
            c.Lo = c.left ? nc : nr;
            c.basis = c.left ? dU : dV;
            c.b = wp[at++];
            c.x = wp[at++];
            c.r = wp[at++];
            c.p = wp[at++];
            c.ap = wp[at++];
            c.uc = wp[at++];
            c.tmp = wp[at++];
            c.proj = wp[at++];
            double* sc = wp[at++];
            c.rr = sc;
            c.rr_new = sc + k;
            c.pap = sc + 2 * k;
            c.alpha = sc + 3 * k;
            c.beta = sc + 4 * k;
            c.sig2 = const_cast<double*>(dsig2);
            c.partials = partials;
            // b = P g = g - basis (basis' g); basis' g is J (left) or K (right).
            launch_rotate(c.basis, c.L, ki, c.left ? dJ : dK, ldy, ki, c.b, c.L, c.L, s);
            launch_cols_axpby(c.b, c.L, nullptr, -1.0, c.left ? dGU : dGV, c.L, nullptr, 1.0, ki, c.L, s);
            cg.start(c);
        }
        // Convergence: |r_i| <= tol |b_i| for every system of both sides.  The
        // residual norms of each iteration are copied to pinned host memory and
        // checked while the next iteration already runs on the device, so the
        // GPU does not idle for the check; the step queued past convergence only
        // refines the solutions.
        double* bnorm = wp[at - 1] + 5 * k;  // spare slot of the right set's scalars (k entries)
        launch_coldots(set[0].b, nr, set[0].b, nr, ki, nr, partials, set[0].beta, s);
        launch_coldots(set[1].b, nc, set[1].b, nc, ki, nc, partials, bnorm, s);
        PinnedBuffer host(sizeof(double) * 6 * k);  // |b|^2 (2k), then two slots of |r|^2 (2k each)
        double* bb = host.get();
        IRLBA_HIP_CHECK(hipMemcpyAsync(bb, set[0].beta, sizeof(double) * k, hipMemcpyDeviceToHost, s));
        IRLBA_HIP_CHECK(hipMemcpyAsync(bb + k, bnorm, sizeof(double) * k, hipMemcpyDeviceToHost, s));
        hipEvent_t ready[2] = {ctx->get_event(), ctx->get_event()};
        auto queue_check = [&](int slot) {
            double* rr = bb + 2 * k + slot * 2 * k;
            IRLBA_HIP_CHECK(hipMemcpyAsync(rr, set[0].rr, sizeof(double) * k, hipMemcpyDeviceToHost, s));
            IRLBA_HIP_CHECK(hipMemcpyAsync(rr + k, set[1].rr, sizeof(double) * k, hipMemcpyDeviceToHost, s));
            IRLBA_HIP_CHECK(hipEventRecord(ready[slot], s));
        };
        auto converged = [&](int slot) {
            IRLBA_HIP_CHECK(hipEventSynchronize(ready[slot]));
            const double* rr = bb + 2 * k + slot * 2 * k;
            for (i64 i = 0; i < 2 * k; ++i)
                if (!(rr[i] <= cg_tol * cg_tol * bb[i])) return false;
            return true;
        };
        int slot = 0;
        queue_check(slot);
        // The right set runs half an iteration ahead of the left one, so that
        // each pass over the matrix serves both: A' [p_left | A p_right] gives
        // the left set's first product and the right set's second, then
        // A [A' p_left | p_right] the left set's second and the right set's
        // first of its next iteration.  Two products of 2k vectors per
        // iteration instead of four of k (on the 50000 x 20000 1% matrix a
        // 32-vector product takes ~1.3 times a 16-vector one).  Where 2k
        // vectors do not fit one pass the sets step one after the other.
        const bool paired = blocks_fit_one_pass(*m, true, ki, 2) && blocks_fit_one_pass(*m, false, ki, 2);
        if (paired) apply_handle_multi(*m, false, set[1].p, nc, set[1].tmp, nr, ki);  // A p_right
        while (iterations < max_cg) {
            if (paired) {
                const BlockIO first[2] = {{set[0].p, nr, set[0].tmp, nc}, {set[1].tmp, nr, set[1].ap, nc}};
                apply_handle_blocks(*m, true, first, 2, ki);
                cg.project_update(set[1]);
                const BlockIO second[2] = {{set[0].tmp, nc, set[0].ap, nr}, {set[1].p, nc, set[1].tmp, nr}};
                apply_handle_blocks(*m, false, second, 2, ki);
                cg.project_update(set[0]);
            } else {
                cg.step(set[0]);
                cg.step(set[1]);
            }
            ++iterations;
            queue_check(slot ^ 1);
            if (converged(slot)) break;
            slot ^= 1;
        }
        IRLBA_HIP_CHECK(hipStreamSynchronize(s));  // every check copy has landed before the buffer goes
        ctx->pool.push_back(ready[0]);
        ctx->pool.push_back(ready[1]);
        cg.finish(set[0]);
        cg.finish(set[1]);
        // Z' = Z S + A T into X1;  Y' = A' Z + T S into Y0.  The right set's
        // scratch (rows x k) receives A T.
        apply_handle_multi(*m, false, set[1].x, nc, set[1].tmp, nr, ki);
        apply_handle_multi(*m, true, set[0].x, nr, Y0, nc, ki);
        launch_cols_axpby(X1, nr, nullptr, 1.0, set[1].tmp, nr, nullptr, 1.0, ki, nr, s);
        launch_cg_coefs(dsig2, nullptr, set[0].alpha, ki, 2, s);  // alpha = sqrt(sig2) = s
        launch_cols_axpby(X1, nr, nullptr, 1.0, set[0].x, nr, set[0].alpha, 1.0, ki, nr, s);
        launch_cols_axpby(Y0, nc, nullptr, 1.0, set[1].x, nc, set[0].alpha, 1.0, ki, nc, s);
    }
    if (dcenter) {
        // d/dc of A - 1 c' is -(X Y')' 1 = -Y (X' 1).
        double *ones = wp[10], *xs = wp[10] + nr;
        launch_fill(ones, 1.0, nr, s);
        launch_cgs_dots(Xo, nr, static_cast<int>(R), ones, nr, partials, s);
        launch_cgs_reduce(partials, nr, static_cast<int>(R), xs, s);
        launch_fill(dc, 0.0, nc, s);
        launch_cgs_update(Yo, nc, static_cast<int>(R), xs, dc, nc, partials, s);
    }
    if (!device_ptrs) {
        IRLBA_HIP_CHECK(hipMemcpyAsync(X, Xo, sizeof(double) * nr * R, hipMemcpyDeviceToHost, s));
        IRLBA_HIP_CHECK(hipMemcpyAsync(Y, Yo, sizeof(double) * nc * R, hipMemcpyDeviceToHost, s));
        if (dcenter) IRLBA_HIP_CHECK(hipMemcpyAsync(dcenter, dc, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
    }
    if (!queued) check_values();  // synchronises
    return iterations;
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

void irlba_hip_options_default(irlba_hip_options* opt) {
    if (!opt) return;
    std::memset(opt, 0, sizeof(*opt));
    opt->invariant_subspace_tolerance = -1;
    opt->convergence_tolerance = 1e-5;
    opt->singular_value_ratio_tolerance = -1;
    opt->extra_work = 7;
    opt->max_iterations = 1000;
    opt->exact_for_small_matrix = 1;
    opt->exact_for_large_number = 1;
    opt->cap_number = 0;
    opt->seed = std::mt19937_64::default_seed;
    opt->initial = nullptr;
    opt->initial_length = 0;
}

const char* irlba_hip_last_error(void) { return g_last_error.c_str(); }

int irlba_hip_context_create(int device, irlba_hip_context** out) {
    return guarded([&] {
        require(out != nullptr, "null output pointer");
        int n = 0;
        IRLBA_HIP_CHECK(hipGetDeviceCount(&n));
        require(device >= 0 && device < n, "no such HIP device");
        DeviceGuard g(device);
        auto ctx = std::make_unique<irlba_hip_context>();
        ctx->device = device;
        IRLBA_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        IRLBA_HIP_CHECK(hipGetDeviceProperties(&prop, device));
        ctx->cu_count = prop.multiProcessorCount;
        ctx->lds_per_block = prop.sharedMemPerBlock;
        ctx->lds_per_cu = std::max<size_t>(prop.maxSharedMemoryPerMultiProcessor, prop.sharedMemPerBlock);
        ctx->noalloc_scratch = noalloc_scratch(device);
        *out = ctx.release();
    });
}

int irlba_hip_context_destroy(irlba_hip_context* ctx) {
    return guarded([&] {
        if (!ctx) return;
        DeviceGuard g(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto& p : ctx->pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : ctx->pool) (void)hipEventDestroy(e);
        if (ctx->copy_stream) {
            (void)hipStreamSynchronize(ctx->copy_stream);
            (void)hipStreamDestroy(ctx->copy_stream);
        }
        if (ctx->copy_done) (void)hipEventDestroy(ctx->copy_done);
        if (ctx->order_ev) (void)hipEventDestroy(ctx->order_ev);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int irlba_hip_context_synchronize(irlba_hip_context* ctx) {
    return guarded([&] {
        require(ctx != nullptr, "null context");
        CtxGuard g(ctx);
        ctx->sync();
    });
}

int irlba_hip_get_stats(const irlba_hip_context* ctx, irlba_hip_stats* out) {
    return guarded([&] {
        require(ctx && out, "null argument");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        *out = ctx->stats;
    });
}

int irlba_hip_context_set_profiling(irlba_hip_context* ctx, int period) {
    return guarded([&] {
        require(ctx != nullptr, "null context");
        if (period < 0) throw std::runtime_error("profiling period must be non-negative");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        ctx->profiling = period;
        ctx->profile_seq[0] = ctx->profile_seq[1] = 0;
    });
}

void* irlba_hip_context_stream(irlba_hip_context* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int irlba_hip_context_wait_stream(irlba_hip_context* ctx, void* stream) {
    return guarded([&] {
        require(ctx != nullptr, "null context");
        CtxGuard g(ctx);
        ctx->order(static_cast<hipStream_t>(stream), ctx->stream);
    });
}

int irlba_hip_stream_wait_context(irlba_hip_context* ctx, void* stream) {
    return guarded([&] {
        require(ctx != nullptr, "null context");
        CtxGuard g(ctx);
        ctx->order(ctx->stream, static_cast<hipStream_t>(stream));
    });
}

int irlba_hip_sparse_create(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, int input_type, const void* values,
                            const int32_t* indices, const int64_t* pointers, int column_major, int storage_type,
                            int accumulate_f64, int device_pointers, irlba_hip_matrix** out) {
    return guarded([&] {
        require(ctx && out, "null argument");
        CtxGuard g(ctx);
        auto base = make_sparse(ctx, nrow, ncol, to_storage(input_type), values, indices, pointers, column_major != 0,
                                to_storage(storage_type), accumulate_f64 != 0, device_pointers != 0);
        auto* m = new irlba_hip_matrix;
        m->ctx = ctx;
        m->base = base;
        *out = m;
    });
}

int irlba_hip_dense_create(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, int input_type, const void* values,
                           int column_major, int storage_type, int device_pointers, irlba_hip_matrix** out) {
    return guarded([&] {
        require(ctx && out, "null argument");
        require(nrow >= 0 && ncol >= 0, "matrix dimensions must be non-negative");
        CtxGuard g(ctx);
        const StorageType in_t = to_storage(input_type), st = to_storage(storage_type);
        auto d = std::make_shared<DenseBase>();
        d->ctx = ctx;
        d->nr = nrow;
        d->nc = ncol;
        const i64 n = nrow * ncol;
        d->data = dalloc_bytes(storage_bytes(st) * n);
        // A caller's device buffer may still be written by work queued on
        // another stream (the engine's stream does not wait for them).
        if (device_pointers) IRLBA_HIP_CHECK(hipDeviceSynchronize());
        if (in_t == st) {
            // On the engine's stream and waited for: a device-to-device hipMemcpy
            // returns before the copy ends, and the products run on a stream that
            // does not wait for the null stream (a 5 GB copy was still running
            // when the first product read the matrix).
            IRLBA_HIP_CHECK(hipMemcpyAsync(d->data.get(), values, storage_bytes(st) * n,
                                           device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                           ctx->stream));
            IRLBA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        } else if (device_pointers) {
            launch_convert_values(values, in_t, d->data.get(), st, n, ctx->stream);
            IRLBA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        } else {
            // Host input of another type: copied as given, converted on the device.
            auto raw = dalloc_bytes(storage_bytes(in_t) * n);
            stream_copy(ctx, raw.get(), values, storage_bytes(in_t) * n, hipMemcpyHostToDevice);
            launch_convert_values(raw.get(), in_t, d->data.get(), st, n, ctx->stream);
            IRLBA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        }
        d->d.nr = nrow;
        d->d.nc = ncol;
        d->d.col_major = column_major != 0;
        d->d.st = st;
        d->d.a = d->data.get();
        const size_t ab = st == ST_F64 ? 8 : 4;
        d->xs = dalloc_bytes(ab * std::max<i64>(1, std::max(nrow, ncol)));
        d->partial_cap = dense_partial_elems(d->d);
        d->partial = dalloc_bytes(ab * std::max<i64>(1, d->partial_cap));
        d->gamma = dalloc<double>(2 * 256);
        auto* m = new irlba_hip_matrix;
        m->ctx = ctx;
        m->base = d;
        *out = m;
    });
}

int irlba_hip_centered_create(const irlba_hip_matrix* inner, const double* center, int device_pointers,
                              irlba_hip_matrix** out) {
    return guarded([&] {
        require(inner && out && center, "null argument");
        CtxGuard g(inner->ctx);
        *out = decorate_center(inner, upload_vec(inner->ctx, center, inner->base->nc, device_pointers != 0));
    });
}

int irlba_hip_scaled_create(const irlba_hip_matrix* inner, const double* scale, int column, int divide,
                            int device_pointers, irlba_hip_matrix** out) {
    return guarded([&] {
        require(inner && out && scale, "null argument");
        CtxGuard g(inner->ctx);
        const i64 n = column ? inner->base->nc : inner->base->nr;
        *out = decorate_scale(inner, upload_vec(inner->ctx, scale, n, device_pointers != 0), column != 0, divide != 0);
    });
}

int irlba_hip_matrix_destroy(irlba_hip_matrix* m) {
    return guarded([&] {
        if (!m) return;
        CtxGuard g(m->ctx);
        IRLBA_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
        delete m;
    });
}

int irlba_hip_matrix_dims(const irlba_hip_matrix* m, int64_t* rows, int64_t* cols) {
    return guarded([&] {
        require(m != nullptr, "null matrix");
        if (rows) *rows = m->base->nr;
        if (cols) *cols = m->base->nc;
    });
}

int irlba_hip_sparse_layout(const irlba_hip_matrix* m, int adjoint, irlba_hip_layout* out) {
    return guarded([&] {
        require(m && out, "null argument");
        const auto* sb = dynamic_cast<const SparseBase*>(m->base.get());
        require(sb != nullptr, "not a sparse matrix handle");
        const TiledLines& t = adjoint ? sb->by_col : sb->by_row;
        out->nnz = t.nnz;
        out->slots = t.slots;
        out->index_bytes = t.slots * (t.delta ? 1 : 2);
        out->tile_len = t.tile_len;
        out->n_tiles = t.n_tiles;
        out->vec = t.vec;
        out->lanes_per_line = t.lpl;
        out->delta = t.delta ? 1 : 0;
        out->blocks = t.B;
        out->dictionary = t.use_dict ? t.n_dict : 0;
        out->dictionary_run = t.use_dict ? t.dict_run : 0;
    });
}

int irlba_hip_sparse_value_dictionary(irlba_hip_matrix* m, int enable, int32_t* n_codes) {
    return guarded([&] {
        require(m != nullptr, "null matrix");
        auto* sb = dynamic_cast<SparseBase*>(m->base.get());
        require(sb != nullptr, "not a sparse matrix handle");
        CtxGuard g(m->ctx);
        const int n = set_value_dictionary(m->ctx, *sb, enable != 0);
        if (n_codes) *n_codes = n;
    });
}

int irlba_hip_multiply(const irlba_hip_matrix* m, int adjoint, const double* right, double* output,
                       int device_pointers) {
    return guarded([&] {
        require(m && right && output, "null argument");
        irlba_hip_context* ctx = m->ctx;
        CtxGuard g(ctx);
        const i64 nin = adjoint ? m->base->nr : m->base->nc, nout = adjoint ? m->base->nc : m->base->nr;
        if (device_pointers) {
            apply_handle(*m, adjoint != 0, right, output, nullptr, nullptr);
            ctx->sync();
            return;
        }
        auto din = dalloc<double>(nin);
        auto dout = dalloc<double>(nout);
        IRLBA_HIP_CHECK(hipMemcpyAsync(din.get(), right, sizeof(double) * nin, hipMemcpyHostToDevice, ctx->stream));
        apply_handle(*m, adjoint != 0, din.get(), dout.get(), nullptr, nullptr);
        IRLBA_HIP_CHECK(hipMemcpyAsync(output, dout.get(), sizeof(double) * nout, hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
    });
}

int irlba_hip_multiply_block(const irlba_hip_matrix* m, int adjoint, int64_t k, const double* right, double* output,
                             int device_pointers) {
    return guarded([&] {
        require(m && right && output, "null argument");
        require(k >= 1 && k <= 128, "the block must have between 1 and 128 columns");
        irlba_hip_context* ctx = m->ctx;
        CtxGuard g(ctx);
        const i64 nin = adjoint ? m->base->nr : m->base->nc, nout = adjoint ? m->base->nc : m->base->nr;
        const double* X = right;
        double* Y = output;
        std::shared_ptr<double> din, dout;
        if (!device_pointers) {
            din = dalloc<double>(nin * k);
            dout = dalloc<double>(nout * k);
            IRLBA_HIP_CHECK(
                hipMemcpyAsync(din.get(), right, sizeof(double) * nin * k, hipMemcpyHostToDevice, ctx->stream));
            X = din.get();
            Y = dout.get();
        }
        apply_handle_multi(*m, adjoint != 0, X, nin, Y, nout, static_cast<int>(k));
        if (!device_pointers)
            IRLBA_HIP_CHECK(
                hipMemcpyAsync(output, Y, sizeof(double) * nout * k, hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
    });
}

int irlba_hip_realize(const irlba_hip_matrix* m, double* buffer, int device_pointers) {
    return guarded([&] {
        require(m && buffer, "null argument");
        CtxGuard g(m->ctx);
        HandleOp op;
        op.ctx = m->ctx;
        op.nr = m->base->nr;
        op.nc = m->base->nc;
        op.m = m;
        HostMat A;
        op.realize_host(A);
        stream_copy(m->ctx, buffer, A.a.data(), sizeof(double) * A.a.size(),
                    device_pointers ? hipMemcpyHostToDevice : hipMemcpyHostToHost);
    });
}

int irlba_hip_compute(const irlba_hip_matrix* m, int64_t number, const irlba_hip_options* opt, double* U, double* V,
                      double* D, int64_t* number_out, int32_t* converged, int32_t* iterations, int device_pointers) {
    return guarded([&] {
        require(number >= 0, "requested number of singular values must be non-negative");
        require(m != nullptr, "null matrix");
        CtxGuard g(m->ctx);
        HandleOp op;
        op.ctx = m->ctx;
        op.nr = m->base->nr;
        op.nc = m->base->nc;
        op.m = m;
        const Opts o = to_opts(opt);
        i64 nout = 0;
        auto r = run_compute(m->ctx, op, number, o, U, V, D, &nout, device_pointers != 0);
        m->ctx->stats.forward_bytes = m->base->product_bytes(false);
        m->ctx->stats.adjoint_bytes = m->base->product_bytes(true);
        if (number_out) *number_out = nout;
        if (converged) *converged = r.first;
        if (iterations) *iterations = r.second;
    });
}

int irlba_hip_compute_callback(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, irlba_hip_matvec_fn matvec,
                               void* user, int64_t number, const irlba_hip_options* opt, double* U, double* V,
                               double* D, int64_t* number_out, int32_t* converged, int32_t* iterations) {
    return guarded([&] {
        require(number >= 0, "requested number of singular values must be non-negative");
        require(ctx && matvec, "null argument");
        CtxGuard g(ctx);
        CallbackOp op;
        op.ctx = ctx;
        op.nr = nrow;
        op.nc = ncol;
        op.fn = matvec;
        op.user = user;
        const Opts o = to_opts(opt);
        i64 nout = 0;
        auto r = run_compute(ctx, op, number, o, U, V, D, &nout, false);
        if (number_out) *number_out = nout;
        if (converged) *converged = r.first;
        if (iterations) *iterations = r.second;
    });
}

int irlba_hip_lanczos(const irlba_hip_matrix* m, int64_t work, double* W, double* V, double* B, uint64_t seed,
                      int64_t start, const irlba_hip_options* opt, double* F) {
    return guarded([&] {
        require(m && W && V && B && F, "null argument");
        require(work > 0 && start >= 0 && start < work, "start must lie within the working subspace");
        irlba_hip_context* ctx = m->ctx;
        CtxGuard g(ctx);
        HandleOp op;
        op.ctx = ctx;
        op.nr = m->base->nr;
        op.nc = m->base->nc;
        op.m = m;
        Opts o = to_opts(opt);
        o.seed = seed;
        Engine eg(ctx, op, o, work);
        const i64 nr = op.nr, nc = op.nc;
        // After the engine's zero fills of W and V, which are queued on the same stream.
        stream_copy(ctx, eg.W.get(), W, sizeof(double) * nr * work, hipMemcpyHostToDevice);
        stream_copy(ctx, eg.V.get(), V, sizeof(double) * nc * work, hipMemcpyHostToDevice);
        HostMat hb(work, work);
        std::copy(B, B + work * work, hb.a.begin());
        eg.sweep(start, hb);
        std::copy(hb.a.begin(), hb.a.end(), B);
        stream_copy(ctx, W, eg.W.get(), sizeof(double) * nr * work, hipMemcpyDeviceToHost);
        stream_copy(ctx, V, eg.V.get(), sizeof(double) * nc * work, hipMemcpyDeviceToHost);
        stream_copy(ctx, F, eg.F.get(), sizeof(double) * nc, hipMemcpyDeviceToHost);
    });
}

int irlba_hip_orthogonalize(irlba_hip_context* ctx, const double* M, int64_t rows, int64_t total_cols, double* vec,
                            int64_t ncols, double* norm_out) {
    return guarded([&] {
        require(ctx && M && vec, "null argument");
        require(ncols >= 0 && ncols <= total_cols, "ncols must not exceed the columns of the basis");
        CtxGuard g(ctx);
        hipStream_t s = ctx->stream;
        auto dM = dalloc<double>(rows * std::max<i64>(total_cols, 1));
        auto df = dalloc<double>(rows);
        const int G = cgs_blocks(rows);
        auto part = dalloc<double>(static_cast<i64>(G) * std::max<i64>(ncols, 1));
        auto t = dalloc<double>(std::max<i64>(ncols, 1));
        auto np = dalloc<double>(G);
        auto nrm = dalloc<double>(1);
        stream_copy(ctx, dM.get(), M, sizeof(double) * rows * total_cols, hipMemcpyHostToDevice);
        stream_copy(ctx, df.get(), vec, sizeof(double) * rows, hipMemcpyHostToDevice);
        if (ncols > 0) {
            launch_cgs_dots(dM.get(), rows, static_cast<int>(ncols), df.get(), rows, part.get(), s);
            launch_cgs_reduce(part.get(), rows, static_cast<int>(ncols), t.get(), s);
            launch_cgs_update(dM.get(), rows, static_cast<int>(ncols), t.get(), df.get(), rows, np.get(), s);
        } else {
            launch_sumsq(df.get(), rows, np.get(), s);
        }
        launch_normalize(df.get(), rows, np.get(), G, nrm.get(), nullptr, s);
        IRLBA_HIP_CHECK(hipStreamSynchronize(s));
        stream_copy(ctx, vec, df.get(), sizeof(double) * rows, hipMemcpyDeviceToHost);
        if (norm_out) stream_copy(ctx, norm_out, nrm.get(), sizeof(double), hipMemcpyDeviceToHost);
    });
}

int irlba_hip_rotate(irlba_hip_context* ctx, const double* M, int64_t rows, int64_t work, const double* Y, int64_t k,
                     double* out) {
    return guarded([&] {
        require(ctx && M && Y && out, "null argument");
        require(k >= 0 && k <= work, "k must not exceed work");
        CtxGuard g(ctx);
        hipStream_t s = ctx->stream;
        const int ldy = static_cast<int>((work + 15) / 16 * 16);
        std::vector<double> yt(static_cast<size_t>(work * ldy), 0.0);
        for (i64 r = 0; r < work; ++r)
            for (i64 c = 0; c < k; ++c) yt[r * ldy + c] = Y[r + c * work];
        auto dM = dalloc<double>(rows * work);
        auto dY = dalloc<double>(work * ldy);
        auto dO = dalloc<double>(rows * std::max<i64>(k, 1));
        stream_copy(ctx, dM.get(), M, sizeof(double) * rows * work, hipMemcpyHostToDevice);
        stream_copy(ctx, dY.get(), yt.data(), sizeof(double) * yt.size(), hipMemcpyHostToDevice);
        launch_rotate(dM.get(), rows, static_cast<int>(work), dY.get(), ldy, static_cast<int>(k), dO.get(), rows, rows,
                      s);
        IRLBA_HIP_CHECK(hipStreamSynchronize(s));
        stream_copy(ctx, out, dO.get(), sizeof(double) * rows * k, hipMemcpyDeviceToHost);
    });
}

int irlba_hip_column_stats(const irlba_hip_matrix* m, int center, int scale, double* center_out, double* scale_out,
                           int device_pointers) {
    return guarded([&] {
        require(m != nullptr, "null matrix");
        require(!m->decorated, "column statistics are computed on an undecorated matrix");
        irlba_hip_context* ctx = m->ctx;
        CtxGuard g(ctx);
        const i64 nr = m->base->nr, nc = m->base->nc;
        if (center && nr < 1) throw std::runtime_error("cannot center with no observations");
        if (scale && nr < 2) throw std::runtime_error("cannot scale with fewer than two observations");
        auto c = dalloc<double>(nc), s = dalloc<double>(nc);
        m->base->colstats(center ? c.get() : nullptr, scale ? s.get() : nullptr);
        const auto kind = device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        if (center) IRLBA_HIP_CHECK(hipMemcpyAsync(center_out, c.get(), sizeof(double) * nc, kind, ctx->stream));
        if (scale) IRLBA_HIP_CHECK(hipMemcpyAsync(scale_out, s.get(), sizeof(double) * nc, kind, ctx->stream));
        ctx->sync();
    });
}

int irlba_hip_pca(const irlba_hip_matrix* m, int center, int scale, int64_t number, const irlba_hip_options* opt,
                  double* scores, double* rotation, double* variances, int64_t* number_out, int32_t* converged,
                  int32_t* iterations, int device_pointers) {
    return guarded([&] {
        require(number >= 0, "requested number of singular values must be non-negative");
        require(m != nullptr, "null matrix");
        require(!m->decorated, "pca() takes an undecorated matrix");
        irlba_hip_context* ctx = m->ctx;
        CtxGuard g(ctx);
        const i64 nr = m->base->nr, nc = m->base->nc;
        // pca.hpp:66-109
        if (center && nr < 1) throw std::runtime_error("cannot center with no observations");
        if (scale && nr < 2) throw std::runtime_error("cannot scale with fewer than two observations");
        std::shared_ptr<double> c, s;
        if (center) c = dalloc<double>(nc);
        if (scale) s = dalloc<double>(nc);
        if (center || scale) m->base->colstats(center ? c.get() : nullptr, scale ? s.get() : nullptr);
        IRLBA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        // pca.hpp:113-126: Scaled(Centered(A), column = true, divide = true).
        std::unique_ptr<irlba_hip_matrix> chain(new irlba_hip_matrix(*m));
        if (center) chain.reset(decorate_center(chain.get(), c));
        if (scale) chain.reset(decorate_scale(chain.get(), s, true, true));
        HandleOp op;
        op.ctx = ctx;
        op.nr = nr;
        op.nc = nc;
        op.m = chain.get();
        const Opts o = to_opts(opt);
        // Run on device buffers, then scores = U diag(d), variances = d^2/(nr-1) (pca.hpp:130-136).
        const i64 nalloc = std::max<i64>(number, 0);
        auto dU = dalloc<double>(nr * std::max<i64>(nalloc, 1));
        auto dV = dalloc<double>(nc * std::max<i64>(nalloc, 1));
        auto dD = dalloc<double>(std::max<i64>(nalloc, 1));
        i64 nout = 0;
        auto r = run_compute(ctx, op, number, o, dU.get(), dV.get(), dD.get(), &nout, true);
        ctx->stats.forward_bytes = m->base->product_bytes(false);
        ctx->stats.adjoint_bytes = m->base->product_bytes(true);
        launch_scale_columns(dU.get(), nr, nr, static_cast<int>(nout), dD.get(), ctx->stream);
        std::vector<double> d(nout);
        IRLBA_HIP_CHECK(hipMemcpyAsync(d.data(), dD.get(), sizeof(double) * nout, hipMemcpyDeviceToHost, ctx->stream));
        IRLBA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (nc > 1) {
            for (auto& v : d) v = v * v / static_cast<double>(nr - 1);
        }
        const auto kind = device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        IRLBA_HIP_CHECK(hipMemcpyAsync(scores, dU.get(), sizeof(double) * nr * nout, kind, ctx->stream));
        IRLBA_HIP_CHECK(hipMemcpyAsync(rotation, dV.get(), sizeof(double) * nc * nout, kind, ctx->stream));
        IRLBA_HIP_CHECK(hipMemcpyAsync(variances, d.data(), sizeof(double) * nout,
                                       device_pointers ? hipMemcpyHostToDevice : hipMemcpyHostToHost, ctx->stream));
        ctx->sync();
        if (number_out) *number_out = nout;
        if (converged) *converged = r.first;
        if (iterations) *iterations = r.second;
    });
}

int irlba_hip_generate_csc(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, double density, uint64_t seed,
                           int value_type, void* values, int32_t* indices, int64_t* pointers, int64_t* nnz_out) {
    return guarded([&] {
        require(ctx != nullptr, "null context");
        require(value_type == IRLBA_HIP_F32 || value_type == IRLBA_HIP_F64, "generated values are F32 or F64");
        CtxGuard g(ctx);
        hipStream_t s = ctx->stream;
        double t = std::floor(density * 4294967296.0);
        t = std::min(std::max(t, 0.0), 4294967295.0);
        const std::uint32_t thr = static_cast<std::uint32_t>(t);
        auto counts = dalloc<i64>(ncol);
        launch_gen_nb_count(nrow, ncol, thr, seed, counts.get(), s);
        auto ptr = dalloc<i64>(ncol + 1);
        exclusive_scan_i64(counts.get(), ptr.get(), ncol, s);
        i64 nnz = 0;
        stream_copy(ctx, &nnz, ptr.get() + ncol, sizeof(i64), hipMemcpyDeviceToHost);
        if (nnz_out) *nnz_out = nnz;
        if (!values) return;
        double table[32];
        for (int k = 0; k < 32; ++k) table[k] = std::log1p(static_cast<double>(k));
        auto dtab = dalloc<double>(32);
        stream_copy(ctx, dtab.get(), table, sizeof(table), hipMemcpyHostToDevice);
        launch_gen_nb_fill(nrow, ncol, thr, seed, ptr.get(), dtab.get(), value_type == IRLBA_HIP_F32, values, indices,
                           s);
        IRLBA_HIP_CHECK(hipMemcpyAsync(pointers, ptr.get(), sizeof(i64) * (ncol + 1), hipMemcpyDeviceToDevice, s));
        IRLBA_HIP_CHECK(hipStreamSynchronize(s));
    });
}

int irlba_hip_generate_sparse(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, double density, uint64_t seed,
                              int storage_type, int accumulate_f64, irlba_hip_matrix** out) {
    return guarded([&] {
        require(ctx && out, "null argument");
        CtxGuard g(ctx);
        i64 nnz = 0;
        if (irlba_hip_generate_csc(ctx, nrow, ncol, density, seed, IRLBA_HIP_F32, nullptr, nullptr, nullptr, &nnz) !=
            IRLBA_HIP_OK)
            throw std::runtime_error(g_last_error);
        const int vt = storage_type == IRLBA_HIP_F64 ? IRLBA_HIP_F64 : IRLBA_HIP_F32;
        auto vals = dalloc_bytes((vt == IRLBA_HIP_F64 ? 8 : 4) * static_cast<size_t>(nnz));
        auto idx = dalloc<std::int32_t>(nnz);
        auto ptr = dalloc<i64>(ncol + 1);
        if (irlba_hip_generate_csc(ctx, nrow, ncol, density, seed, vt, vals.get(), idx.get(), ptr.get(), &nnz) !=
            IRLBA_HIP_OK)
            throw std::runtime_error(g_last_error);
        auto base = make_sparse(ctx, nrow, ncol, to_storage(vt), vals.get(), idx.get(), ptr.get(), true,
                                to_storage(storage_type), accumulate_f64 != 0, true);
        auto* m = new irlba_hip_matrix;
        m->ctx = ctx;
        m->base = base;
        *out = m;
    });
}

namespace {

// The backward entry points; `queued` returns with the work queued on the
// context's stream (device arrays and every gradient given, so that nothing
// the queued kernels read is freed on return).
int svd_backward_entry(const irlba_hip_matrix* m, int64_t k, const double* U, const double* D, const double* V,
                       const double* grad_D, const double* grad_U, const double* grad_V, int complement,
                       int32_t max_cg_iterations, double cg_tolerance, double* X, double* Y, double* grad_center,
                       int32_t* cg_iterations, int device_pointers, bool queued) {
    return guarded([&] {
        require(m && U && D && V && X && Y, "null argument");
        require(complement >= 0 && complement <= 2, "complement must be 0, 1 or 2");
        require(max_cg_iterations >= 0 && cg_tolerance >= 0, "invalid CG settings");
        if (queued)
            require(device_pointers && grad_D && grad_U && grad_V,
                    "the queued backward takes device arrays and all three gradients");
        irlba_hip_context* ctx = m->ctx;
        CtxGuard g(ctx);
        const i64 nr = m->base->nr, nc = m->base->nc;
        require(k > 0 && k <= std::min(nr, nc) && k <= 64, "k must lie in [1, min(rows, cols)] and be at most 64");
        // Absent gradients are zero.
        std::vector<double> zk, zr, zc;
        std::shared_ptr<double> dzr, dzc, dzk;
        auto zeros = [&](i64 n, std::vector<double>& h, std::shared_ptr<double>& d) -> const double* {
            if (!device_pointers) {
                h.assign(static_cast<size_t>(n), 0.0);
                return h.data();
            }
            d = dalloc<double>(n);
            IRLBA_HIP_CHECK(hipMemsetAsync(d.get(), 0, sizeof(double) * n, ctx->stream));
            return d.get();
        };
        const double* w = grad_D ? grad_D : zeros(k, zk, dzk);
        const double* gu = grad_U ? grad_U : zeros(nr * k, zr, dzr);
        const double* gv = grad_V ? grad_V : zeros(nc * k, zc, dzc);
        const int it = run_svd_backward(m, k, U, D, V, w, gu, gv, complement, max_cg_iterations, cg_tolerance, X, Y,
                                        grad_center, device_pointers != 0, queued);
        if (!queued) ctx->sync();
        if (cg_iterations) *cg_iterations = it;
    });
}

// The SDDMM entry points; `queued` takes device arrays only and returns with
// the work queued, without reading back the layout checks of the kernel.
int sddmm_entry(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, const int32_t* indices, const int64_t* pointers,
                int column_major, int64_t rank, const double* X, const double* Y, void* out, int out_type,
                int device_pointers, bool queued) {
    return guarded([&] {
        require(ctx && pointers && X && Y, "null argument");
        require(nrow >= 0 && ncol >= 0, "negative dimension");
        require(rank >= 1 && rank <= 128, "rank must lie in [1, 128]");
        require(out_type == IRLBA_HIP_F64 || out_type == IRLBA_HIP_F32, "out_type must be F64 or F32");
        if (queued) require(device_pointers && indices && out, "the queued SDDMM takes device arrays");
        CtxGuard g(ctx);
        hipStream_t s = ctx->stream;
        const bool f32 = out_type == IRLBA_HIP_F32;
        const size_t ob = f32 ? 4 : 8;
        const i64 n_lines = column_major ? ncol : nrow, n_other = column_major ? nrow : ncol;
        // The entry count is needed on the host to stage host arrays, or to
        // accept null index/output pointers for an empty pattern; device arrays
        // are otherwise used as they are, their layout checked by the kernel.
        i64 nnz = -1;
        if (!device_pointers || !indices || !out) {
            IRLBA_HIP_CHECK(hipMemcpyAsync(&nnz, pointers + n_lines, sizeof(i64),
                                           device_pointers ? hipMemcpyDeviceToHost : hipMemcpyHostToHost, s));
            IRLBA_HIP_CHECK(hipStreamSynchronize(s));
            require(nnz >= 0, "pointers are not a valid compressed layout");
            require(nnz == 0 || (indices && out), "null argument");
        }
        const int Rp = sddmm_padded_rank(static_cast<int>(rank), f32);
        // Device arrays are used in place; host ones are staged in the workspace.
        const i64 stage_nnz = device_pointers ? 0 : nnz;
        auto plan = [&](Carve& c, std::vector<void*>& p) {
            p.clear();
            p.push_back(c.take<char>(static_cast<i64>(ob) * Rp * (nrow + ncol)));  // 0 packed factors
            p.push_back(device_pointers ? nullptr : c.take<i64>(n_lines + 1));     // 1 pointers
            p.push_back(device_pointers ? nullptr : c.take<std::int32_t>(stage_nnz));  // 2 indices
            p.push_back(device_pointers ? nullptr : c.take<double>(nrow * rank));  // 3 X
            p.push_back(device_pointers ? nullptr : c.take<double>(ncol * rank));  // 4 Y
            p.push_back(device_pointers ? nullptr : c.take<char>(static_cast<i64>(ob) * stage_nnz));  // 5 out
        };
        std::vector<void*> wp;
        Carve sizing;
        plan(sizing, wp);
        Carve carve;
        carve.base = ctx->workspace(sizing.off);
        plan(carve, wp);
        const i64* dptr = pointers;
        const std::int32_t* didx = indices;
        const double *dX = X, *dY = Y;
        void* dout = out;
        if (!device_pointers) {
            IRLBA_HIP_CHECK(hipMemcpyAsync(wp[1], pointers, sizeof(i64) * (n_lines + 1), hipMemcpyHostToDevice, s));
            if (nnz)
                IRLBA_HIP_CHECK(
                    hipMemcpyAsync(wp[2], indices, sizeof(std::int32_t) * nnz, hipMemcpyHostToDevice, s));
            IRLBA_HIP_CHECK(hipMemcpyAsync(wp[3], X, sizeof(double) * nrow * rank, hipMemcpyHostToDevice, s));
            IRLBA_HIP_CHECK(hipMemcpyAsync(wp[4], Y, sizeof(double) * ncol * rank, hipMemcpyHostToDevice, s));
            dptr = static_cast<const i64*>(wp[1]);
            didx = static_cast<const std::int32_t*>(wp[2]);
            dX = static_cast<const double*>(wp[3]);
            dY = static_cast<const double*>(wp[4]);
            dout = wp[5];
        }
        // Both factors packed row-major (stride Rp) in the output precision.
        char* packed = static_cast<char*>(wp[0]);
        char* pX = packed;
        char* pY = packed + ob * Rp * nrow;
        int* flag = ctx->flag();
        launch_pack_rows(dX, nrow, nrow, static_cast<int>(rank), Rp, f32, pX, s);
        launch_pack_rows(dY, ncol, ncol, static_cast<int>(rank), Rp, f32, pY, s);
        if (nnz != 0)
            launch_sddmm(didx, dptr, n_lines, n_other, column_major ? pY : pX, column_major ? pX : pY,
                         static_cast<int>(rank), f32, dout, flag, s);
        if (queued) return;
        int hflag = 0;
        IRLBA_HIP_CHECK(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
        IRLBA_HIP_CHECK(hipStreamSynchronize(s));
        // The same checks as matrix creation: monotone pointers, indices in range.
        require(!(hflag & 1), "pointers are not a valid compressed layout");
        require(!(hflag & 2), "indices out of range");
        if (!device_pointers && nnz > 0) {
            IRLBA_HIP_CHECK(hipMemcpyAsync(out, dout, ob * nnz, hipMemcpyDeviceToHost, s));
            IRLBA_HIP_CHECK(hipStreamSynchronize(s));
        }
    });
}

// The dense-gradient entry points (X Y' as a dense matrix of the output
// type); `queued` takes device arrays and returns with the work queued.
int dense_gradient_entry(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, int64_t rank, const double* X,
                         const double* Y, void* out, int out_type, int column_major, int device_pointers,
                         bool queued) {
    return guarded([&] {
        require(ctx && X && Y && out, "null argument");
        require(nrow >= 0 && ncol >= 0, "negative dimension");
        require(rank >= 1 && rank <= 128, "rank must lie in [1, 128]");
        const StorageType st = to_storage(out_type);
        if (queued) require(device_pointers, "the queued dense gradient takes device arrays");
        CtxGuard g(ctx);
        hipStream_t s = ctx->stream;
        const i64 n_out = nrow * ncol;
        const size_t ob = storage_bytes(st);
        const double *dX = X, *dY = Y;
        void* dout = out;
        if (!device_pointers) {
            auto plan = [&](Carve& c, std::vector<void*>& p) {
                p.clear();
                p.push_back(c.take<double>(nrow * rank));
                p.push_back(c.take<double>(ncol * rank));
                p.push_back(c.take<char>(static_cast<i64>(ob) * n_out));
            };
            std::vector<void*> wp;
            Carve sizing;
            plan(sizing, wp);
            Carve carve;
            carve.base = ctx->workspace(sizing.off);
            plan(carve, wp);
            IRLBA_HIP_CHECK(hipMemcpyAsync(wp[0], X, sizeof(double) * nrow * rank, hipMemcpyHostToDevice, s));
            IRLBA_HIP_CHECK(hipMemcpyAsync(wp[1], Y, sizeof(double) * ncol * rank, hipMemcpyHostToDevice, s));
            dX = static_cast<const double*>(wp[0]);
            dY = static_cast<const double*>(wp[1]);
            dout = wp[2];
        }
        launch_dense_gradient(dX, nrow, dY, ncol, static_cast<int>(rank), nrow, ncol, st, column_major != 0, dout, s);
        if (!device_pointers && n_out > 0)
            IRLBA_HIP_CHECK(hipMemcpyAsync(out, dout, ob * n_out, hipMemcpyDeviceToHost, s));
        if (!queued) ctx->sync();
    });
}

}  // namespace

int irlba_hip_svd_backward(const irlba_hip_matrix* m, int64_t k, const double* U, const double* D, const double* V,
                           const double* grad_D, const double* grad_U, const double* grad_V, int complement,
                           int32_t max_cg_iterations, double cg_tolerance, double* X, double* Y, double* grad_center,
                           int32_t* cg_iterations, int device_pointers) {
    return svd_backward_entry(m, k, U, D, V, grad_D, grad_U, grad_V, complement, max_cg_iterations, cg_tolerance, X,
                              Y, grad_center, cg_iterations, device_pointers, false);
}

int irlba_hip_svd_backward_async(const irlba_hip_matrix* m, int64_t k, const double* U, const double* D,
                                 const double* V, const double* grad_D, const double* grad_U, const double* grad_V,
                                 int complement, int32_t max_cg_iterations, double cg_tolerance, double* X, double* Y,
                                 double* grad_center, int32_t* cg_iterations) {
    return svd_backward_entry(m, k, U, D, V, grad_D, grad_U, grad_V, complement, max_cg_iterations, cg_tolerance, X,
                              Y, grad_center, cg_iterations, 1, true);
}

int irlba_hip_sddmm(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, const int32_t* indices,
                    const int64_t* pointers, int column_major, int64_t rank, const double* X, const double* Y,
                    void* out, int out_type, int device_pointers) {
    return sddmm_entry(ctx, nrow, ncol, indices, pointers, column_major, rank, X, Y, out, out_type, device_pointers,
                       false);
}

int irlba_hip_dense_gradient(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, int64_t rank, const double* X,
                             const double* Y, void* out, int out_type, int column_major, int device_pointers) {
    return dense_gradient_entry(ctx, nrow, ncol, rank, X, Y, out, out_type, column_major, device_pointers, false);
}

int irlba_hip_dense_gradient_async(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, int64_t rank,
                                   const double* X, const double* Y, void* out, int out_type, int column_major) {
    return dense_gradient_entry(ctx, nrow, ncol, rank, X, Y, out, out_type, column_major, 1, true);
}

int irlba_hip_sddmm_async(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, const int32_t* indices,
                          const int64_t* pointers, int column_major, int64_t rank, const double* X, const double* Y,
                          void* out, int out_type) {
    return sddmm_entry(ctx, nrow, ncol, indices, pointers, column_major, rank, X, Y, out, out_type, 1, true);
}

int irlba_hip_sddmm_noalloc(irlba_hip_context* ctx, int64_t nrow, int64_t ncol, const int32_t* indices,
                            const int64_t* pointers, int column_major, int64_t rank, const double* X, const double* Y,
                            void* out, int out_type) {
    return guarded([&] {
        require(ctx && indices && pointers && X && Y && out, "null argument");
        require(nrow >= 0 && ncol >= 0, "negative dimension");
        require(rank >= 1 && rank <= 128, "rank must lie in [1, 128]");
        require(out_type == IRLBA_HIP_F64 || out_type == IRLBA_HIP_F32, "out_type must be F64 or F32");
        const bool f32 = out_type == IRLBA_HIP_F32;
        const size_t ob = f32 ? 4 : 8;
        const int Rp = sddmm_padded_rank(static_cast<int>(rank), f32);
        // The error flag in the first 256 bytes, then both factors packed
        // row-major (stride Rp) in the output precision, as irlba_hip_sddmm
        // packs them in its workspace.
        const size_t need = 256 + ob * static_cast<size_t>(Rp) * static_cast<size_t>(nrow + ncol);
        if (need > kNoAllocScratchBytes)
            throw std::runtime_error("irlba_hip_sddmm_noalloc: the packed factors need " + std::to_string(need) +
                                     " bytes of scratch, above its fixed " + std::to_string(kNoAllocScratchBytes) +
                                     " bytes ((nrow + ncol) x padded rank x value size + 256); use irlba_hip_sddmm");
        CtxGuard g(ctx);
        require(ctx->noalloc_scratch != nullptr, "the fixed scratch is not available on this device");
        hipStream_t s = ctx->stream;
        char* base = static_cast<char*>(ctx->noalloc_scratch);
        int* flag = reinterpret_cast<int*>(base);
        char* pX = base + 256;
        char* pY = pX + ob * static_cast<size_t>(Rp) * static_cast<size_t>(nrow);
        const i64 n_lines = column_major ? ncol : nrow, n_other = column_major ? nrow : ncol;
        launch_clear_flag(flag, s);
        launch_pack_rows(X, nrow, nrow, static_cast<int>(rank), Rp, f32, pX, s);
        launch_pack_rows(Y, ncol, ncol, static_cast<int>(rank), Rp, f32, pY, s);
        launch_sddmm(indices, pointers, n_lines, n_other, column_major ? pY : pX, column_major ? pX : pY,
                     static_cast<int>(rank), f32, out, flag, s);
        int hflag = 0;
        IRLBA_HIP_CHECK(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
        IRLBA_HIP_CHECK(hipStreamSynchronize(s));
        require(!(hflag & 1), "pointers are not a valid compressed layout");
        require(!(hflag & 2), "indices out of range");
    });
}

}  // extern "C"
