// libtorch binding of the differentiable top-k SVD: a torch::autograd::Function
// over the engine's C ABI (irlba_hip.h), registered as the operators
// torch.ops.irlba_hip.svd_sparse / svd_dense.  Host code only; every
// computation runs in libirlba_hip.so.
#include <ATen/ATen.h>
#include <c10/hip/HIPStream.h>
#include <torch/csrc/autograd/custom_function.h>
#include <torch/cuda.h>
#include <torch/custom_class.h>
#include <torch/library.h>

#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "irlba/torch.hpp"
#include "irlba_hip.h"

namespace irlba {
namespace autograd {
namespace {

using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

void check(int rc) {
    if (rc != IRLBA_HIP_OK) throw std::runtime_error(irlba_hip_last_error());
}

// One engine context per GPU, for the life of the process.
irlba_hip_context* context_for(int device) {
    static std::mutex mu;
    static std::map<int, irlba_hip_context*> ctxs;
    std::lock_guard<std::mutex> lock(mu);
    auto it = ctxs.find(device);
    if (it != ctxs.end()) return it->second;
    irlba_hip_context* c = nullptr;
    check(irlba_hip_context_create(device, &c));
    ctxs[device] = c;
    return c;
}

// The device handles of one forward, kept for its backward.
struct Handles : torch::CustomClassHolder {
    irlba_hip_context* ctx = nullptr;
    irlba_hip_matrix* base = nullptr;
    irlba_hip_matrix* op = nullptr;  // base, or the centred view of it
    ~Handles() override {
        if (op && op != base) irlba_hip_matrix_destroy(op);
        if (base) irlba_hip_matrix_destroy(base);
    }
};

int value_type(const at::Tensor& t) {
    switch (t.scalar_type()) {
        case at::kDouble: return IRLBA_HIP_F64;
        case at::kFloat: return IRLBA_HIP_F32;
        case at::kBFloat16: return IRLBA_HIP_BF16;
        default: throw std::runtime_error("matrix values must be float64, float32 or bfloat16");
    }
}

// Device index of the inputs (-1 = host); all must agree.  Device inputs are
// produced on torch's stream and the engine runs on its own, so wait for torch.
int common_device(std::initializer_list<const at::Tensor*> ts) {
    int dev = -2;
    for (const at::Tensor* t : ts) {
        if (!t || !t->defined()) continue;
        const int d = t->is_cuda() ? t->get_device() : -1;
        if (dev != -2 && d != dev) throw std::runtime_error("all inputs must be on the same device");
        dev = d;
    }
    if (dev >= 0) torch::cuda::synchronize(dev);
    return dev < 0 ? -1 : dev;
}

irlba_hip_options make_options(const SvdOptions& o) {
    irlba_hip_options opt;
    irlba_hip_options_default(&opt);
    opt.extra_work = static_cast<int32_t>(o.extra_work);
    opt.convergence_tolerance = o.convergence_tolerance;
    opt.max_iterations = static_cast<int32_t>(o.max_iterations);
    opt.seed = static_cast<uint64_t>(o.seed);
    return opt;
}

at::TensorOptions f64_on(int dev) {
    auto t = at::TensorOptions().dtype(at::kDouble);
    return dev >= 0 ? t.device(at::Device(at::kCUDA, static_cast<c10::DeviceIndex>(dev))) : t;
}

// torch's current stream on `dev` (a hipStream_t), for the engine's stream ordering.
void* torch_stream(int dev) {
    return static_cast<void*>(c10::hip::getCurrentHIPStream(static_cast<c10::DeviceIndex>(dev)).stream());
}

// Column-major (rows x k) storage of a (rows x k) gradient, FP64 on `dev`.
at::Tensor col_major(const at::Tensor& g, int dev) {
    return g.to(f64_on(dev).device(), at::kDouble).t().contiguous();
}

// Shared forward: decorate, solve, save.  U, V are returned as transposed views
// of column-major (k x rows) buffers.
variable_list solve(AutogradContext* ctx, c10::intrusive_ptr<Handles> h, const std::optional<at::Tensor>& center,
                    int64_t k, const SvdOptions& o, int dev) {
    int64_t nr = 0, nc = 0;
    check(irlba_hip_matrix_dims(h->base, &nr, &nc));
    h->op = h->base;
    at::Tensor c;
    if (center.has_value() && center->defined()) {
        TORCH_CHECK(center->dim() == 1 && center->size(0) == nc, "center must have one entry per column");
        c = center->detach().to(f64_on(dev).device(), at::kDouble).contiguous();
        check(irlba_hip_centered_create(h->base, c.data_ptr<double>(), dev >= 0, &h->op));
    }
    TORCH_CHECK(k >= 1 && k <= std::min(nr, nc), "k must lie in [1, min(rows, cols)]");
    at::Tensor Ucm = at::empty({k, nr}, f64_on(dev)), Vcm = at::empty({k, nc}, f64_on(dev)),
               S = at::empty({k}, f64_on(dev));
    irlba_hip_options opt = make_options(o);
    int64_t nout = 0;
    int32_t conv = 0, iters = 0;
    check(irlba_hip_compute(h->op, k, &opt, Ucm.data_ptr<double>(), Vcm.data_ptr<double>(), S.data_ptr<double>(),
                            &nout, &conv, &iters, dev >= 0));
    TORCH_CHECK(nout == k, "the solver returned fewer triplets than requested");
    // The backward on device arrays does not read back the engine's check of
    // the singular values (positive, non-increasing); they are checked here,
    // where the solve has already waited for the device.
    const at::Tensor Sh = S.to(at::kCPU);
    const double* sp = Sh.data_ptr<double>();
    bool values_ok = true;
    for (int64_t i = 0; i < k; ++i) values_ok = values_ok && sp[i] > 0 && (i == 0 || sp[i] <= sp[i - 1]);
    ctx->saved_data["values_ok"] = values_ok;
    ctx->saved_data["handles"] = h;
    ctx->saved_data["k"] = k;
    ctx->saved_data["dev"] = static_cast<int64_t>(dev);
    ctx->saved_data["complement"] = o.complement;
    ctx->saved_data["center"] = c.defined();
    ctx->save_for_backward({Ucm, S, Vcm});
    return {Ucm.t(), S, Vcm.t()};
}

// Factored dL/dA = X Y' (X: 2k x rows, Y: 2k x cols, both column-major views)
// and the centre gradient, from the saved state and the incoming gradients.
// On device arrays the engine's work is stream ordered against torch's: its
// stream waits for torch's current stream after every conversion and
// allocation the engine reads or writes (so that no block the caching
// allocator hands out is still in use by earlier torch work), and torch's
// stream waits for the engine's before anything reads the results.  The
// tensors the engine reads are held here until then.
struct Factored {
    at::Tensor X, Y, dcenter;
    at::Tensor gU, gS, gV;
    c10::intrusive_ptr<Handles> h;
    int dev = -1;
};

// The torch side: gradient conversions and output allocations.  `center_edge`
// is the centre's position among the tensor inputs (the indices that
// needs_input_grad() takes).
Factored prepare_grad(AutogradContext* ctx, const variable_list& g, size_t center_edge) {
    auto saved = ctx->get_saved_variables();
    const at::Tensor &Ucm = saved[0], &Vcm = saved[2];
    Factored f;
    f.h = ctx->saved_data["handles"].toCustomClass<Handles>();
    const int64_t k = ctx->saved_data["k"].toInt();
    f.dev = static_cast<int>(ctx->saved_data["dev"].toInt());
    const int dev = f.dev;
    const bool want_center = ctx->saved_data["center"].toBool() && ctx->needs_input_grad(center_edge);
    const int64_t nr = Ucm.size(1), nc = Vcm.size(1);
    // On the device every gradient is given (zeros for an unused output), so
    // that the engine allocates nothing that its queued work would still read.
    if (g[0].defined()) f.gU = col_major(g[0], dev);
    else if (dev >= 0) f.gU = at::zeros({k, nr}, f64_on(dev));
    if (g[1].defined()) f.gS = g[1].to(f64_on(dev).device(), at::kDouble).contiguous();
    else if (dev >= 0) f.gS = at::zeros({k}, f64_on(dev));
    if (g[2].defined()) f.gV = col_major(g[2], dev);
    else if (dev >= 0) f.gV = at::zeros({k, nc}, f64_on(dev));
    f.X = at::empty({2 * k, nr}, f64_on(dev));
    f.Y = at::empty({2 * k, nc}, f64_on(dev));
    if (want_center) f.dcenter = at::empty({nc}, f64_on(dev));
    return f;
}

double* ptr_of(const at::Tensor& t) { return t.defined() ? t.data_ptr<double>() : nullptr; }

// The engine side: on the device its stream first waits for torch's (every
// torch-side step of this backward has been queued by now), then the backward
// is queued; host arrays run synchronously.
void run_grad(AutogradContext* ctx, Factored& f) {
    auto saved = ctx->get_saved_variables();
    const at::Tensor &Ucm = saved[0], &S = saved[1], &Vcm = saved[2];
    const int64_t k = ctx->saved_data["k"].toInt();
    const int complement = static_cast<int>(ctx->saved_data["complement"].toInt());
    int32_t cg = 0;
    if (f.dev >= 0) {
        if (!ctx->saved_data["values_ok"].toBool())
            throw std::runtime_error("singular values must be positive and non-increasing");
        check(irlba_hip_context_wait_stream(f.h->ctx, torch_stream(f.dev)));
        check(irlba_hip_svd_backward_async(f.h->op, k, ptr_of(Ucm), ptr_of(S), ptr_of(Vcm), ptr_of(f.gS),
                                           ptr_of(f.gU), ptr_of(f.gV), complement, 1000, 1e-12, ptr_of(f.X),
                                           ptr_of(f.Y), ptr_of(f.dcenter), &cg));
    } else {
        check(irlba_hip_svd_backward(f.h->op, k, ptr_of(Ucm), ptr_of(S), ptr_of(Vcm), ptr_of(f.gS), ptr_of(f.gU),
                                     ptr_of(f.gV), complement, 1000, 1e-12, ptr_of(f.X), ptr_of(f.Y),
                                     ptr_of(f.dcenter), &cg, 0));
    }
}

// torch's current stream waits for the engine's queued work (device arrays),
// once: at now(), or when the scope ends if an error unwinds it first, so that
// no tensor the engine still uses goes back to torch's allocator unordered.
struct TorchAfterEngine {
    const Factored& f;
    bool done = false;
    explicit TorchAfterEngine(const Factored& fa) : f(fa) {}
    void now() {
        done = true;
        if (f.dev >= 0) check(irlba_hip_stream_wait_context(f.h->ctx, torch_stream(f.dev)));
    }
    ~TorchAfterEngine() {
        if (!done && f.dev >= 0) (void)irlba_hip_stream_wait_context(f.h->ctx, torch_stream(f.dev));
    }
};

class SparseSvd : public torch::autograd::Function<SparseSvd> {
public:
    // Arguments: 0 values, 1 indices, 2 pointers, 3 nrow, 4 ncol, 5 column_major,
    // 6 center, 7 k, 8 options; tensor inputs: values, indices, pointers, center.
    static variable_list forward(AutogradContext* ctx, at::Tensor values, at::Tensor indices, at::Tensor pointers,
                                 int64_t nrow, int64_t ncol, bool column_major, std::optional<at::Tensor> center,
                                 int64_t k, SvdOptions o) {
        TORCH_CHECK(indices.scalar_type() == at::kInt, "indices must be int32");
        TORCH_CHECK(pointers.scalar_type() == at::kLong, "pointers must be int64");
        TORCH_CHECK(values.scalar_type() == at::kDouble || values.scalar_type() == at::kFloat,
                    "sparse values must be float64 or float32");
        const at::Tensor* cptr = center.has_value() ? &*center : nullptr;
        const int dev = common_device({&values, &indices, &pointers, cptr});
        at::Tensor x = values.detach().contiguous(), i = indices.contiguous(), pp = pointers.contiguous();
        auto h = c10::make_intrusive<Handles>();
        h->ctx = context_for(dev < 0 ? 0 : dev);
        const int storage = static_cast<int>(o.storage);
        check(irlba_hip_sparse_create(h->ctx, nrow, ncol, value_type(x), x.data_ptr(), i.data_ptr<int32_t>(),
                                      pp.data_ptr<int64_t>(), column_major, storage, storage == IRLBA_HIP_F64,
                                      dev >= 0, &h->base));
        ctx->saved_data["nrow"] = nrow;
        ctx->saved_data["ncol"] = ncol;
        ctx->saved_data["column_major"] = column_major;
        ctx->saved_data["indices"] = i;
        ctx->saved_data["pointers"] = pp;
        ctx->saved_data["values_dtype"] = static_cast<int64_t>(values.scalar_type());
        return solve(ctx, h, center, k, o, dev);
    }

    static variable_list backward(AutogradContext* ctx, variable_list g) {
        Factored f = prepare_grad(ctx, g, 3);
        // The gradient of the values comes out in their own type (FP32 values:
        // FP32 factors with FP64 dot products).
        const auto vt = static_cast<at::ScalarType>(ctx->saved_data["values_dtype"].toInt());
        const bool f32 = vt == at::kFloat;
        at::Tensor i = ctx->saved_data["indices"].toTensor(), pp = ctx->saved_data["pointers"].toTensor();
        at::Tensor dvalues;
        if (ctx->needs_input_grad(0))
            dvalues = at::empty({i.numel()}, f64_on(f.dev).dtype(f32 ? at::kFloat : at::kDouble));
        TorchAfterEngine order(f);
        run_grad(ctx, f);
        if (dvalues.defined()) {
            const int64_t nrow = ctx->saved_data["nrow"].toInt(), ncol = ctx->saved_data["ncol"].toInt();
            const int cm = ctx->saved_data["column_major"].toBool();
            const int ot = f32 ? IRLBA_HIP_F32 : IRLBA_HIP_F64;
            // On the device the pattern is that of the matrix the forward
            // created from these arrays, which checked it.
            if (f.dev >= 0)
                check(irlba_hip_sddmm_async(f.h->ctx, nrow, ncol, i.data_ptr<int32_t>(), pp.data_ptr<int64_t>(), cm,
                                            f.X.size(0), ptr_of(f.X), ptr_of(f.Y), dvalues.data_ptr(), ot));
            else
                check(irlba_hip_sddmm(f.h->ctx, nrow, ncol, i.data_ptr<int32_t>(), pp.data_ptr<int64_t>(), cm,
                                      f.X.size(0), ptr_of(f.X), ptr_of(f.Y), dvalues.data_ptr(), ot, 0));
        }
        order.now();
        if (dvalues.defined() && dvalues.scalar_type() != vt) dvalues = dvalues.to(vt);
        return {dvalues, at::Tensor(), at::Tensor(), at::Tensor(), at::Tensor(), at::Tensor(), f.dcenter,
                at::Tensor(), at::Tensor()};
    }
};

class DenseSvd : public torch::autograd::Function<DenseSvd> {
public:
    // Arguments: 0 matrix, 1 center, 2 k, 3 options; tensor inputs: matrix, center.
    static variable_list forward(AutogradContext* ctx, at::Tensor matrix, std::optional<at::Tensor> center, int64_t k,
                                 SvdOptions o) {
        TORCH_CHECK(matrix.dim() == 2, "matrix must be two-dimensional");
        const at::Tensor* cptr = center.has_value() ? &*center : nullptr;
        const int dev = common_device({&matrix, cptr});
        at::Tensor a = matrix.detach();
        int column_major = 0;
        if (!a.is_contiguous()) {
            if (a.t().is_contiguous()) column_major = 1;
            else a = a.contiguous();
        }
        auto h = c10::make_intrusive<Handles>();
        h->ctx = context_for(dev < 0 ? 0 : dev);
        check(irlba_hip_dense_create(h->ctx, a.size(0), a.size(1), value_type(a), a.data_ptr(), column_major,
                                     static_cast<int>(o.storage), dev >= 0, &h->base));
        ctx->saved_data["matrix_dtype"] = static_cast<int64_t>(matrix.scalar_type());
        return solve(ctx, h, center, k, o, dev);
    }

    static variable_list backward(AutogradContext* ctx, variable_list g) {
        Factored f = prepare_grad(ctx, g, 1);
        const auto mt = static_cast<at::ScalarType>(ctx->saved_data["matrix_dtype"].toInt());
        const int64_t nr = f.X.size(1), nc = f.Y.size(1);
        // On the device the engine writes X Y' in the matrix's type, queued
        // behind the backward (row-major, the layout of a contiguous matrix).
        at::Tensor dA;
        if (ctx->needs_input_grad(0) && f.dev >= 0) dA = at::empty({nr, nc}, f64_on(f.dev).dtype(mt));
        TorchAfterEngine order(f);
        run_grad(ctx, f);
        if (dA.defined())
            check(irlba_hip_dense_gradient_async(f.h->ctx, nr, nc, f.X.size(0), ptr_of(f.X), ptr_of(f.Y),
                                                 dA.data_ptr(), value_type(dA), 0));
        order.now();
        if (ctx->needs_input_grad(0) && !dA.defined()) dA = at::mm(f.X.t(), f.Y).to(mt);
        return {dA, f.dcenter, at::Tensor(), at::Tensor()};
    }
};

SvdOptions options_of(int64_t extra_work, double tol, int64_t max_iterations, int64_t seed, int64_t storage,
                      int64_t complement) {
    TORCH_CHECK(complement >= 0 && complement <= 2, "complement must be 0, 1 or 2");
    SvdOptions o;
    o.extra_work = extra_work;
    o.convergence_tolerance = tol;
    o.max_iterations = max_iterations;
    o.seed = seed;
    o.storage = storage;
    o.complement = complement;
    return o;
}

std::tuple<at::Tensor, at::Tensor, at::Tensor> op_sparse(const at::Tensor& values, const at::Tensor& indices,
                                                          const at::Tensor& pointers, int64_t nrow, int64_t ncol,
                                                          bool column_major, const std::optional<at::Tensor>& center,
                                                          int64_t k, int64_t extra_work, double tol,
                                                          int64_t max_iterations, int64_t seed, int64_t storage,
                                                          int64_t complement) {
    return svd_sparse(values, indices, pointers, nrow, ncol, column_major, center, k,
                      options_of(extra_work, tol, max_iterations, seed, storage, complement));
}

std::tuple<at::Tensor, at::Tensor, at::Tensor> op_dense(const at::Tensor& matrix,
                                                         const std::optional<at::Tensor>& center, int64_t k,
                                                         int64_t extra_work, double tol, int64_t max_iterations,
                                                         int64_t seed, int64_t storage, int64_t complement) {
    return svd_dense(matrix, center, k, options_of(extra_work, tol, max_iterations, seed, storage, complement));
}

}  // namespace

std::tuple<at::Tensor, at::Tensor, at::Tensor> svd_sparse(const at::Tensor& values, const at::Tensor& indices,
                                                           const at::Tensor& pointers, int64_t nrow, int64_t ncol,
                                                           bool column_major, const std::optional<at::Tensor>& center,
                                                           int64_t k, const SvdOptions& options) {
    auto out = SparseSvd::apply(values, indices, pointers, nrow, ncol, column_major, center, k, options);
    return {out[0], out[1], out[2]};
}

std::tuple<at::Tensor, at::Tensor, at::Tensor> svd_dense(const at::Tensor& matrix,
                                                          const std::optional<at::Tensor>& center, int64_t k,
                                                          const SvdOptions& options) {
    auto out = DenseSvd::apply(matrix, center, k, options);
    return {out[0], out[1], out[2]};
}

}  // namespace autograd
}  // namespace irlba

TORCH_LIBRARY(irlba_hip, m) {
    m.class_<irlba::autograd::Handles>("Handles");
    m.def(
        "svd_sparse(Tensor values, Tensor indices, Tensor pointers, int nrow, int ncol, bool column_major, "
        "Tensor? center, int k, int extra_work, float tol, int max_iterations, int seed, int storage, "
        "int complement) -> (Tensor, Tensor, Tensor)");
    m.def(
        "svd_dense(Tensor matrix, Tensor? center, int k, int extra_work, float tol, int max_iterations, int seed, "
        "int storage, int complement) -> (Tensor, Tensor, Tensor)");
}

TORCH_LIBRARY_IMPL(irlba_hip, CompositeImplicitAutograd, m) {
    m.impl("svd_sparse", &irlba::autograd::op_sparse);
    m.impl("svd_dense", &irlba::autograd::op_dense);
}
