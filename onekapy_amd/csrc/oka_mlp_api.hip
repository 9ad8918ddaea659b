// oka_mlp_api.hip -- extern "C" entry points of the surrogate network
// (include/oka.h, "learned control variate").  The network, its Adam state and
// its per-batch workspace live in device memory owned by the oka_mlp object;
// every matrix product goes through the BF16 MFMA kernels of oka_mlp.hip.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "oka_ctx.h"
#include "oka_mlp.h"

using namespace oka;

struct oka_mlp {
    oka_ctx *ctx = nullptr;
    int nin = 0, kin = 64, nh = 0, nout = 0, max_batch = 0;
    int64_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0, np = 0;
    float *p = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *w1h = nullptr, *w1t = nullptr, *w2h = nullptr, *w2t = nullptr, *w3h = nullptr,
             *w3t = nullptr;
    uint16_t *x0 = nullptr, *x0t = nullptr, *h1 = nullptr, *h1t = nullptr, *h2 = nullptr,
             *h2t = nullptr, *g3 = nullptr, *g3t = nullptr, *g2 = nullptr, *g2t = nullptr,
             *g1 = nullptr, *g1t = nullptr;
    float *z1 = nullptr, *z2 = nullptr;
    uint8_t *ym = nullptr;       // the batch's target rows [max_batch x n_out]
    float *acc = nullptr;        // split-K partial slices, acc_cap floats
    size_t acc_cap = 0;
    // training: split-K slices of dW and db per layer (kMaxSplitSlices each),
    // summed by Adam
    float *ws[3] = {nullptr, nullptr, nullptr};
    float *bs[3] = {nullptr, nullptr, nullptr};
    int64_t *dstate = nullptr;   // graph replay: [0] step within the call, [1] steps begun
    int64_t step = 0;
    std::vector<void *> allocs;
};

namespace {

int alloc(oka_mlp *M, size_t bytes, void **out)
{
    void *p = nullptr;
    HIPCHK(M->ctx, hipMalloc(&p, bytes ? bytes : 8));
    M->allocs.push_back(p);
    *out = p;
    return OKA_OK;
}

template <typename T>
int alloc_t(oka_mlp *M, size_t count, T **out)
{
    void *p;
    TRY(alloc(M, count * sizeof(T), &p));
    *out = (T *)p;
    return OKA_OK;
}

GemmArgs gemm(const uint16_t *A, int lda, const uint16_t *B, int ldb, int Mr, int N, int K)
{
    GemmArgs g{};
    g.A = A;
    g.lda = lda;
    g.B = B;
    g.ldb = ldb;
    g.M = Mr;
    g.N = N;
    g.K = K;
    g.mvalid = Mr;
    g.splits = 1;
    return g;
}

// Split-K plan of one layer's backward launch: dW [n_out_l, n_in_l] over a K
// of B (the batch) beside dX [B, n_in_l] over a K of n_out_l.  The two
// products' K per workgroup are made equal first (the launch ends with its
// slowest workgroup), then both are split further while the launch has fewer
// than two workgroups per compute unit.  dW is split only when Adam takes its
// slices (split_dw); dX split more than once leaves FP32 slices in M->acc for
// launch_gelu_bwd, else the product's own epilogue applies gelu'.
struct SplitPlan {
    int dw = 1, dx = 1;
};

SplitPlan backward_plan(oka_mlp *M, int B, int nout_l, int nin_l, bool split_dw)
{
    SplitPlan p;
    // K per LDS stage is 64; a slice keeps at least 128 of K
    auto can = [](int K, int s, size_t slice, size_t cap) {
        return 2 * s <= kMaxSplitSlices && K % (2 * s * 64) == 0 && K / (2 * s) >= 128 &&
               slice * 2 * s <= cap;
    };
    const size_t wslice = (size_t)nout_l * nin_l, xslice = (size_t)B * nin_l;
    const size_t wcap = (size_t)kMaxSplitSlices * wslice;
    auto dw_ok = [&] { return split_dw && can(B, p.dw, wslice, wcap); };
    auto dx_ok = [&] { return can(nout_l, p.dx, xslice, M->acc_cap); };
    while (B / p.dw > nout_l / p.dx && dw_ok())
        p.dw *= 2;
    while (nout_l / p.dx > B / p.dw && dx_ok())
        p.dx *= 2;
    auto workgroups = [&] {
        GemmArgs w{}, x{};
        w.M = nout_l;
        w.N = nin_l;
        w.splits = p.dw;
        x.M = B;
        x.N = nin_l;
        x.splits = p.dx;
        return gemm2_workgroups(w, &x);
    };
    while (workgroups() < 2 * M->ctx->num_cu && dw_ok() && dx_ok()) {
        p.dw *= 2;
        p.dx *= 2;
    }
    return p;
}

// Forward through the two hidden layers for rows of x (rowidx/row0) into the
// workspace; Bp rows (a multiple of 64), the first `valid` of them real.
int hidden_forward(oka_mlp *M, int Bp, int valid, const float *x, const int32_t *rowidx,
                   int64_t row0, bool train, int64_t *dstate = nullptr,
                   const uint8_t *marks = nullptr)
{
    oka_ctx *c = M->ctx;
    hipStream_t s = c->stream;
    HIPCHK(c, launch_prep_input(x, M->nin, M->kin, rowidx, row0, Bp, valid, M->x0, M->x0t, marks,
                                M->nout, M->ym, s, dstate));
    // inference keeps neither the FP32 pre-activations nor the transposed copies
    GemmArgs g = gemm(M->x0, M->kin, M->w1h, M->kin, Bp, M->nh, M->kin);
    g.bias = M->p + M->ob1;
    g.Cf = train ? M->z1 : nullptr;
    g.ldc = M->nh;
    g.Ch = M->h1;
    g.Cht = train ? M->h1t : nullptr;
    HIPCHK(c, launch_gemm(kEpiGeluFwd, g, s));
    g = gemm(M->h1, M->nh, M->w2h, M->nh, Bp, M->nh, M->nh);
    g.bias = M->p + M->ob2;
    g.Cf = train ? M->z2 : nullptr;
    g.ldc = M->nh;
    g.Ch = M->h2;
    g.Cht = train ? M->h2t : nullptr;
    HIPCHK(c, launch_gemm(kEpiGeluFwd, g, s));
    return OKA_OK;
}

// dW1 (the first layer's weight gradient: a few output tiles over a K of B)
// in split-K slices of M->ws[0] / M->bs[0] when the result is above 1, for
// Adam to sum.
int w1_split_plan(int B)
{
    int s = 1;
    while (s < kMaxSplitSlices && B % (2 * s * 64) == 0 && B / (2 * s) >= 256)
        s *= 2;
    return s;
}

// The split-K plans of one training step's backward launches (layer 3, 2
// and 1's dW).
struct StepPlan {
    SplitPlan layer[2];
    int w1 = 1;
};

StepPlan step_plan(oka_mlp *M, int B, bool split_dw)
{
    StepPlan p;
    p.layer[0] = backward_plan(M, B, M->nout, M->nh, split_dw);
    p.layer[1] = backward_plan(M, B, M->nh, M->nh, split_dw);
    p.w1 = split_dw ? w1_split_plan(B) : 1;
    return p;
}

// Loss (when `loss` is set) and full gradient for one batch of B rows, in 7
// or more launches, each dW product sharing its launch with the layer's dX
// product and forming its bias gradient.  Gradients go to M->grad, except the
// dW and db that `plan` splits over K: those stay as slices in M->ws / M->bs
// for Adam.
int forward_backward(oka_mlp *M, int B, const float *x, const uint8_t *marks, float tscale,
                     const int32_t *rowidx, int64_t row0, double *loss, const StepPlan &plan,
                     int64_t *dstate = nullptr)
{
    oka_ctx *c = M->ctx;
    hipStream_t s = c->stream;
    TRY(hidden_forward(M, B, B, x, rowidx, row0, true, dstate, marks));
    GemmArgs g = gemm(M->h2, M->nh, M->w3h, M->nh, B, M->nout, M->nh);
    g.bias = M->p + M->ob3;
    g.marks = M->ym;
    g.dstate = dstate;
    g.tscale = tscale;
    g.gscale = 1.0f / ((float)B * (float)M->nout);
    g.Ch = M->g3;
    g.Cht = M->g3t;
    g.loss = loss;
    HIPCHK(c, launch_gemm(kEpiLossGrad, g, s));
    struct Layer {
        const uint16_t *gt, *xt;     // G^T [n_out_l, B], X^T [n_in_l, B]
        int nout_l, nin_l;
        int64_t oW, ob;
        const uint16_t *g, *wt;      // G [B, n_out_l], W^T [n_in_l, n_out_l] (dX product)
        const float *z;              // pre-activation of the layer input
        uint16_t *gin, *gint;        // dX * gelu'(z): G of the layer below
        float *ws, *bs;              // dW and db slices
    };
    const Layer layers[3] = {
        {M->g3t, M->h2t, M->nout, M->nh, M->oW3, M->ob3, M->g3, M->w3t, M->z2, M->g2, M->g2t,
         M->ws[2], M->bs[2]},
        {M->g2t, M->h1t, M->nh, M->nh, M->oW2, M->ob2, M->g2, M->w2t, M->z1, M->g1, M->g1t,
         M->ws[1], M->bs[1]},
        {M->g1t, M->x0t, M->nh, M->kin, M->oW1, M->ob1, nullptr, nullptr, nullptr, nullptr,
         nullptr, M->ws[0], M->bs[0]}};
    for (int l = 0; l < 3; ++l) {
        const Layer &L = layers[l];
        const int dw_splits = l < 2 ? plan.layer[l].dw : plan.w1;
        // dW = G^T X and db = row sums of G^T
        GemmArgs dw = gemm(L.gt, B, L.xt, B, L.nout_l, L.nin_l, B);
        dw.Cf = M->grad + L.oW;
        dw.ldc = L.nin_l;
        dw.dbias = M->grad + L.ob;
        int ew = kEpiStoreF32;
        if (dw_splits > 1) {
            dw.Cf = L.ws;
            dw.dbias = L.bs;
            dw.splits = dw_splits;
            ew = kEpiPartialF32;
        }
        if (!L.g) {
            HIPCHK(c, launch_gemm(ew, dw, s));
            break;
        }
        const int dx_splits = plan.layer[l].dx;
        GemmArgs dx = gemm(L.g, L.nout_l, L.wt, L.nout_l, B, L.nin_l, L.nout_l);
        int ex = kEpiGeluBwd;
        if (dx_splits == 1) {
            dx.Zsaved = L.z;
            dx.Ch = L.gin;
            dx.Cht = L.gint;
        } else {
            dx.Cf = M->acc;
            dx.ldc = L.nin_l;
            dx.splits = dx_splits;
            ex = kEpiPartialF32;
        }
        HIPCHK(c, launch_gemm2(ew, dw, ex, dx, s));
        if (dx_splits > 1)
            HIPCHK(c, launch_gelu_bwd(M->acc, dx_splits, L.z, B, L.nin_l, L.gin, L.gint, s));
    }
    return OKA_OK;
}

// One captured training step's inputs (device pointers).
struct TrainStep {
    int batch;
    const float *x;
    const uint8_t *marks;
    float tscale;
    const int32_t *order;
    double *loss;
    const AdamArgs *adam;
    StepPlan plan;   // the dW / db split-K slices that Adam sums
};

constexpr int kStepsPerGraph = 16;

// Capture `steps` consecutive training steps on the context stream as one
// executable graph.
int capture_steps(oka_mlp *M, const TrainStep &t, int steps, hipGraphExec_t *out)
{
    oka_ctx *c = M->ctx;
    *out = nullptr;
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = OKA_OK;
    for (int k = 0; k < steps && rc == OKA_OK; ++k) {
        rc = forward_backward(M, t.batch, t.x, t.marks, t.tscale, t.order, 0, t.loss, t.plan,
                              M->dstate);
        if (rc == OKA_OK && launch_adam(*t.adam, c->stream) != hipSuccess)
            rc = fail(c, OKA_ERR_DEVICE, "adam launch failed");
    }
    const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    if (rc != OKA_OK) {
        if (graph)
            (void)hipGraphDestroy(graph);
        return rc;
    }
    HIPCHK(c, ec);
    const hipError_t ei = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(c, ei);
    return OKA_OK;
}

int refresh_bf16(oka_mlp *M)
{
    hipStream_t s = M->ctx->stream;
    HIPCHK(M->ctx, launch_refresh_bf16(M->p + M->oW1, M->nh, M->kin, M->w1h, M->w1t, s));
    HIPCHK(M->ctx, launch_refresh_bf16(M->p + M->oW2, M->nh, M->nh, M->w2h, M->w2t, s));
    HIPCHK(M->ctx, launch_refresh_bf16(M->p + M->oW3, M->nout, M->nh, M->w3h, M->w3t, s));
    return OKA_OK;
}

// Copy between the ABI layout (W1 with n_in columns) and the padded one.
int copy_params(oka_mlp *M, float *dst, bool dst_padded, const float *src, hipMemcpyKind kind)
{
    oka_ctx *c = M->ctx;
    const int64_t abi_w1 = (int64_t)M->nh * M->nin;
    const size_t dpitch = (dst_padded ? M->kin : M->nin) * sizeof(float);
    const size_t spitch = (dst_padded ? M->nin : M->kin) * sizeof(float);
    HIPCHK(c, hipMemcpy2DAsync(dst, dpitch, src, spitch, M->nin * sizeof(float), M->nh, kind,
                               c->stream));
    const int64_t rest = M->np - M->ob1;
    float *d2 = dst + (dst_padded ? M->ob1 : abi_w1);
    const float *s2 = src + (dst_padded ? abi_w1 : M->ob1);
    HIPCHK(c, hipMemcpyAsync(d2, s2, rest * sizeof(float), kind, c->stream));
    return OKA_OK;
}

hipMemcpyKind kind_to_dev(uint32_t flags)
{
    return (flags & OKA_PTR_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
}

hipMemcpyKind kind_from_dev(uint32_t flags)
{
    return (flags & OKA_PTR_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
}

int sync_unless_async(oka_ctx *c, uint32_t flags)
{
    if (!(flags & OKA_ASYNC) || !(flags & OKA_PTR_DEVICE))
        HIPCHK(c, hipStreamSynchronize(c->stream));
    return OKA_OK;
}

}  // namespace

extern "C" {

int oka_mlp_create(oka_ctx *c, int32_t n_in, int32_t n_hidden, int32_t n_out, int32_t max_batch,
                   oka_mlp **out)
{
    if (!c || !out)
        return OKA_ERR_ARG;
    *out = nullptr;
    if (n_in < 1 || n_in > 64 || n_hidden < 64 || n_hidden % 64 || n_out < 64 || n_out % 64 ||
        max_batch < 64 || max_batch % 64)
        return fail(c, OKA_ERR_ARG, "mlp sizes: n_in 1..64, n_hidden/n_out/max_batch multiples of 64");
    HIPCHK(c, hipSetDevice(c->device));
    oka_mlp *M = new oka_mlp();
    M->ctx = c;
    M->nin = n_in;
    M->nh = n_hidden;
    M->nout = n_out;
    M->max_batch = max_batch;
    M->oW1 = 0;
    M->ob1 = M->oW1 + (int64_t)M->nh * M->kin;
    M->oW2 = M->ob1 + M->nh;
    M->ob2 = M->oW2 + (int64_t)M->nh * M->nh;
    M->oW3 = M->ob2 + M->nh;
    M->ob3 = M->oW3 + (int64_t)M->nout * M->nh;
    M->np = M->ob3 + M->nout;
    const size_t B = (size_t)max_batch;
    int rc = OKA_OK;
    rc = rc ? rc : alloc_t(M, M->np, &M->p);
    rc = rc ? rc : alloc_t(M, M->np, &M->grad);
    rc = rc ? rc : alloc_t(M, M->np, &M->m);
    rc = rc ? rc : alloc_t(M, M->np, &M->v);
    rc = rc ? rc : alloc_t(M, (size_t)M->nh * M->kin, &M->w1h);
    rc = rc ? rc : alloc_t(M, (size_t)M->nh * M->kin, &M->w1t);
    rc = rc ? rc : alloc_t(M, (size_t)M->nh * M->nh, &M->w2h);
    rc = rc ? rc : alloc_t(M, (size_t)M->nh * M->nh, &M->w2t);
    rc = rc ? rc : alloc_t(M, (size_t)M->nout * M->nh, &M->w3h);
    rc = rc ? rc : alloc_t(M, (size_t)M->nout * M->nh, &M->w3t);
    rc = rc ? rc : alloc_t(M, B * M->kin, &M->x0);
    rc = rc ? rc : alloc_t(M, B * M->kin, &M->x0t);
    rc = rc ? rc : alloc_t(M, B * M->nh, &M->h1);
    rc = rc ? rc : alloc_t(M, B * M->nh, &M->h1t);
    rc = rc ? rc : alloc_t(M, B * M->nh, &M->h2);
    rc = rc ? rc : alloc_t(M, B * M->nh, &M->h2t);
    rc = rc ? rc : alloc_t(M, B * M->nout, &M->g3);
    rc = rc ? rc : alloc_t(M, B * M->nout, &M->g3t);
    rc = rc ? rc : alloc_t(M, B * M->nh, &M->g2);
    rc = rc ? rc : alloc_t(M, B * M->nh, &M->g2t);
    rc = rc ? rc : alloc_t(M, B * M->nh, &M->g1);
    rc = rc ? rc : alloc_t(M, B * M->nh, &M->g1t);
    rc = rc ? rc : alloc_t(M, B * M->nh, &M->z1);
    rc = rc ? rc : alloc_t(M, B * M->nh, &M->z2);
    rc = rc ? rc : alloc_t(M, B * M->nout, &M->ym);
    M->acc_cap = std::max(B * M->nh, (size_t)4 * c->num_cu * 64 * 64);
    rc = rc ? rc : alloc_t(M, M->acc_cap, &M->acc);
    {
        const int rows[3] = {M->nh, M->nh, M->nout}, cols[3] = {M->kin, M->nh, M->nh};
        for (int l = 0; l < 3; ++l) {
            rc = rc ? rc : alloc_t(M, (size_t)kMaxSplitSlices * rows[l] * cols[l], &M->ws[l]);
            rc = rc ? rc : alloc_t(M, (size_t)kMaxSplitSlices * rows[l], &M->bs[l]);
        }
    }
    rc = rc ? rc : alloc_t(M, 2, &M->dstate);
    if (rc == OKA_OK) {
        // zero state (padding columns of W1 must stay exactly zero)
        for (float *q : {M->p, M->grad, M->m, M->v})
            rc = rc ? rc : fill_u32(c, q, M->np);
    }
    if (rc == OKA_OK)
        rc = refresh_bf16(M);
    if (rc == OKA_OK && hipStreamSynchronize(c->stream) != hipSuccess)
        rc = fail(c, OKA_ERR_DEVICE, "mlp init failed");
    if (rc != OKA_OK) {
        oka_mlp_destroy(M);
        return rc;
    }
    *out = M;
    return OKA_OK;
}

void oka_mlp_destroy(oka_mlp *M)
{
    if (!M)
        return;
    (void)hipSetDevice(M->ctx->device);
    (void)hipStreamSynchronize(M->ctx->stream);
    for (void *p : M->allocs)
        (void)hipFree(p);
    delete M;
}

int64_t oka_mlp_param_count(const oka_mlp *M)
{
    return M ? M->np - (int64_t)M->nh * (M->kin - M->nin) : 0;
}

int oka_mlp_set_params(oka_mlp *M, const float *params, uint32_t flags)
{
    if (!M || !params)
        return OKA_ERR_ARG;
    oka_ctx *c = M->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    TRY(fill_u32(c, M->p, M->np));
    TRY(copy_params(M, M->p, true, params, kind_to_dev(flags)));
    TRY(fill_u32(c, M->m, M->np));
    TRY(fill_u32(c, M->v, M->np));
    M->step = 0;
    TRY(refresh_bf16(M));
    return sync_unless_async(c, flags);
}

int oka_mlp_get_params(oka_mlp *M, float *params, uint32_t flags)
{
    if (!M || !params)
        return OKA_ERR_ARG;
    HIPCHK(M->ctx, hipSetDevice(M->ctx->device));
    TRY(copy_params(M, params, false, M->p, kind_from_dev(flags)));
    return sync_unless_async(M->ctx, flags);
}

int oka_mlp_grad(oka_mlp *M, int32_t batch, const float *x, const uint8_t *marks,
                 float target_scale, float *grad, float *loss, uint32_t flags)
{
    if (!M || !x || !marks || !grad)
        return OKA_ERR_ARG;
    oka_ctx *c = M->ctx;
    if (batch < 64 || batch % 64 || batch > M->max_batch)
        return fail(c, OKA_ERR_ARG, "batch must be a multiple of 64 up to max_batch");
    HIPCHK(c, hipSetDevice(c->device));
    const void *dx, *dm;
    TRY(stage_in(c, kSlotMlpX, x, sizeof(float) * batch * M->nin, flags, &dx));
    TRY(stage_in(c, kSlotMlpMarks, marks, (size_t)batch * M->nout, flags, &dm));
    void *dl = nullptr;
    if (loss) {
        TRY(ensure(c, kSlotMlpLoss, sizeof(double), &dl));
        HIPCHK(c, launch_zero_f64((double *)dl, 1, c->stream));
    }
    // every gradient into M->grad: no dW split
    TRY(forward_backward(M, batch, (const float *)dx, (const uint8_t *)dm, target_scale, nullptr, 0,
                         (double *)dl, step_plan(M, batch, false)));
    TRY(copy_params(M, grad, false, M->grad, kind_from_dev(flags)));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (loss) {
        double h;
        HIPCHK(c, hipMemcpy(&h, dl, sizeof(double), hipMemcpyDeviceToHost));
        *loss = (float)(h / ((double)batch * M->nout));
    }
    return OKA_OK;
}

int oka_mlp_train(oka_mlp *M, int64_t nsamples, const float *x, const uint8_t *marks,
                  float target_scale, const int32_t *order, int32_t batch, int32_t nsteps,
                  float lr, float beta1, float beta2, float eps, float *losses, uint32_t flags)
{
    if (!M || !x || !marks || !order || nsteps < 0 || nsamples < 1)
        return OKA_ERR_ARG;
    oka_ctx *c = M->ctx;
    if (batch < 64 || batch % 64 || batch > M->max_batch)
        return fail(c, OKA_ERR_ARG, "batch must be a multiple of 64 up to max_batch");
    HIPCHK(c, hipSetDevice(c->device));
    const void *dx, *dm, *dord;
    TRY(stage_in(c, kSlotMlpX, x, sizeof(float) * nsamples * M->nin, flags, &dx));
    TRY(stage_in(c, kSlotMlpMarks, marks, (size_t)nsamples * M->nout, flags, &dm));
    TRY(stage_in(c, kSlotMlpOrder, order, sizeof(int32_t) * (size_t)nsteps * batch, flags, &dord));
    void *dl = nullptr;
    if (losses && nsteps > 0) {
        TRY(ensure(c, kSlotMlpLoss, sizeof(double) * nsteps, &dl));
        HIPCHK(c, launch_zero_f64((double *)dl, nsteps, c->stream));
    }
    // Training steps (forward, backward, Adam with the BF16 refresh: 10
    // launches each) are captured kStepsPerGraph at a time as a HIP graph and
    // replayed; the step-dependent values (batch rows, loss slot, Adam step)
    // come from the device counters M->dstate, which each step's first and
    // last kernels advance.
    const int64_t st0[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(M->dstate, st0, sizeof(st0), hipMemcpyHostToDevice, c->stream));
    if (nsteps > 0) {
        AdamArgs a{};
        a.p = M->p;
        a.grad = M->grad;
        a.m = M->m;
        a.v = M->v;
        a.lr = lr;
        a.b1 = beta1;
        a.b2 = beta2;
        a.eps = eps;
        a.tbase = M->step;
        a.dstate = M->dstate;
        const StepPlan plan = step_plan(M, batch, true);
        const int split[3] = {plan.w1, plan.layer[1].dw, plan.layer[0].dw};
        const int64_t oW[3] = {M->oW1, M->oW2, M->oW3}, ob[3] = {M->ob1, M->ob2, M->ob3};
        const int rows[3] = {M->nh, M->nh, M->nout}, cols[3] = {M->kin, M->nh, M->nh};
        uint16_t *wh[3] = {M->w1h, M->w2h, M->w3h}, *wt[3] = {M->w1t, M->w2t, M->w3t};
        for (int l = 0; l < 3; ++l) {
            const bool sl = split[l] > 1;
            a.seg[2 * l] = AdamSeg{oW[l], rows[l], cols[l], wh[l], wt[l], sl ? M->ws[l] : nullptr,
                                   sl ? split[l] : 0};
            a.seg[2 * l + 1] = AdamSeg{ob[l], rows[l] / 32, 32, nullptr, nullptr,
                                       sl ? M->bs[l] : nullptr, sl ? split[l] : 0};
        }
        a.nseg = 6;
        const TrainStep ts{batch, (const float *)dx, (const uint8_t *)dm, target_scale,
                           (const int32_t *)dord, (double *)dl, &a, plan};
        const int unroll = std::min<int32_t>(nsteps, kStepsPerGraph);
        const int full = nsteps / unroll, rest = nsteps % unroll;
        hipGraphExec_t ex_full = nullptr, ex_rest = nullptr;
        TRY(capture_steps(M, ts, unroll, &ex_full));
        if (rest) {
            const int rc = capture_steps(M, ts, rest, &ex_rest);
            if (rc != OKA_OK) {
                (void)hipGraphExecDestroy(ex_full);
                return rc;
            }
        }
        hipError_t el = hipSuccess;
        for (int i = 0; i < full && el == hipSuccess; ++i)
            el = hipGraphLaunch(ex_full, c->stream);
        if (rest && el == hipSuccess)
            el = hipGraphLaunch(ex_rest, c->stream);
        (void)hipGraphExecDestroy(ex_full);
        if (ex_rest)
            (void)hipGraphExecDestroy(ex_rest);
        HIPCHK(c, el);
        M->step += nsteps;
    }
    if (losses && nsteps > 0) {
        std::vector<double> h(nsteps);
        HIPCHK(c, hipMemcpyAsync(h.data(), dl, sizeof(double) * nsteps, hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int32_t st = 0; st < nsteps; ++st)
            losses[st] = (float)(h[st] / ((double)batch * M->nout));
        return OKA_OK;
    }
    return sync_unless_async(c, flags);
}

static int infer(oka_mlp *M, int64_t n, const float *x, const uint8_t *marks, float tscale,
                 float *probs, double *acc[4], uint32_t flags)
{
    oka_ctx *c = M->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    const void *dx, *dm = nullptr;
    TRY(stage_in(c, kSlotMlpX, x, sizeof(float) * n * M->nin, flags, &dx));
    if (marks)
        TRY(stage_in(c, kSlotMlpMarks, marks, (size_t)n * M->nout, flags, &dm));
    void *dp = nullptr;
    void *da[4] = {nullptr, nullptr, nullptr, nullptr};
    if (probs)
        TRY(stage_out(c, kSlotMlpOut, probs, sizeof(float) * n * M->nout, flags, false, &dp));
    if (acc) {
        // the four FP64 accumulators share one staging buffer
        void *base = nullptr;
        if (!(flags & OKA_PTR_DEVICE))
            TRY(ensure(c, kSlotMlpOut2, sizeof(double) * 4 * M->nout, &base));
        for (int k = 0; k < 4; ++k) {
            if (!acc[k])
                continue;
            if (flags & OKA_PTR_DEVICE) {
                da[k] = acc[k];
            } else {
                da[k] = (double *)base + (size_t)k * M->nout;
                HIPCHK(c, hipMemcpyAsync(da[k], acc[k], sizeof(double) * M->nout,
                                         hipMemcpyHostToDevice, c->stream));
            }
        }
    }
    const int chunk = M->max_batch;
    for (int64_t r0 = 0; r0 < n; r0 += chunk) {
        const int valid = (int)std::min<int64_t>(chunk, n - r0);
        const int Bp = (valid + 63) / 64 * 64;
        TRY(hidden_forward(M, Bp, valid, (const float *)dx, nullptr, r0, false));
        GemmArgs g = gemm(M->h2, M->nh, M->w3h, M->nh, Bp, M->nout, M->nh);
        g.bias = M->p + M->ob3;
        g.mvalid = valid;
        if (probs) {
            g.Cf = (float *)dp + r0 * M->nout;
            g.ldc = M->nout;
            HIPCHK(c, launch_gemm(kEpiProbs, g, c->stream));
        } else {
            g.sums = (double *)da[0];
            g.sums_sq = (double *)da[1];
            g.diffs = (double *)da[2];
            g.diffs_sq = (double *)da[3];
            g.marks = marks ? (const uint8_t *)dm : nullptr;
            g.row0 = r0;
            g.tscale = tscale;
            HIPCHK(c, launch_gemm(kEpiAccum, g, c->stream));
        }
    }
    if (probs)
        TRY(copy_back(c, probs, dp, sizeof(float) * n * M->nout, flags));
    if (acc)
        for (int k = 0; k < 4; ++k)
            if (acc[k])
                TRY(copy_back(c, acc[k], da[k], sizeof(double) * M->nout, flags));
    return sync_unless_async(c, flags);
}

int oka_mlp_predict(oka_mlp *M, int64_t n, const float *x, float *probs, uint32_t flags)
{
    if (!M || !x || !probs || n < 0)
        return OKA_ERR_ARG;
    if (n == 0)
        return OKA_OK;
    return infer(M, n, x, nullptr, 0.f, probs, nullptr, flags);
}

int oka_mlp_accumulate(oka_mlp *M, int64_t n, const float *x, const uint8_t *marks,
                       float target_scale, double *sums, double *sums_sq, double *diffs,
                       double *diffs_sq, uint32_t flags)
{
    if (!M || !x || !sums || n < 0 || (!marks && (diffs || diffs_sq)))
        return OKA_ERR_ARG;
    if (n == 0)
        return OKA_OK;
    double *acc[4] = {sums, sums_sq, diffs, diffs_sq};
    return infer(M, n, x, marks, target_scale, nullptr, acc, flags);
}

}  // extern "C"

// ---- realization draws and the device-side surrogate term -----------------

namespace {

int check_dists(oka_ctx *c, int32_t nwells, const oka_dist *dists)
{
    if (nwells < 0 || !dists)
        return fail(c, OKA_ERR_ARG, "bad distributions");
    for (int32_t i = 0; i < nwells + 3; ++i) {
        const oka_dist &d = dists[i];
        if (d.kind < OKA_DIST_CONST || d.kind > OKA_DIST_LOGNORMAL)
            return fail(c, OKA_ERR_ARG, "unknown distribution kind %d", d.kind);
        if (d.kind == OKA_DIST_TRIANGULAR && !(d.a <= d.b && d.b <= d.c && d.a < d.c))
            return fail(c, OKA_ERR_ARG, "triangular needs lo <= mode <= hi, lo < hi");
        if (!(d.p_negative >= 0.0 && d.p_negative <= 1.0))
            return fail(c, OKA_ERR_ARG, "p_negative must be in [0, 1]");
    }
    return OKA_OK;
}

}  // namespace

extern "C" {

// OKA_QMC preconditions: not with OKA_ANTITHETIC, point indices below 2^32,
// at most kSobolMaxDims dimensions.
static int check_qmc(oka_ctx *c, int32_t nwells, const oka_dist *dists, int64_t n, int64_t first,
                     uint32_t flags)
{
    if (!(flags & OKA_QMC))
        return OKA_OK;
    if (flags & OKA_ANTITHETIC)
        return fail(c, OKA_ERR_ARG, "OKA_QMC and OKA_ANTITHETIC exclude each other");
    if (first + n > (int64_t(1) << 32))
        return fail(c, OKA_ERR_ARG, "QMC point indices must stay below 2^32");
    if (qmc_dims(nwells, dists) > kSobolMaxDims)
        return fail(c, OKA_ERR_ARG, "QMC draws need %d dimensions, at most %d",
                    qmc_dims(nwells, dists), kSobolMaxDims);
    return OKA_OK;
}

// The direction numbers a QMC draw of these distributions uses and the
// scramble keys of nrep replicates (seeds), on the device.  The copies are
// synchronous, after the context's stream has drained, so no pending work
// still reads the table (the direction numbers live for the whole process).
static int stage_qmc(oka_ctx *c, int32_t nwells, const oka_dist *dists, const uint64_t *seeds,
                     int32_t nrep, DrawArgs *a)
{
    const int dims = qmc_dims(nwells, dists);
    const size_t bytes_v = sizeof(uint32_t) * 32 * (size_t)dims;
    std::vector<uint32_t> keys((size_t)nrep * dims);
    for (int32_t k = 0; k < nrep; ++k)
        for (int d = 0; d < dims; ++d)
            keys[(size_t)k * dims + d] = qmc_key(seeds[k], d);
    void *p;
    TRY(ensure(c, kSlotQmc, bytes_v + sizeof(uint32_t) * keys.size(), &p));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(p, sobol_directions(), bytes_v, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy((char *)p + bytes_v, keys.data(), sizeof(uint32_t) * keys.size(),
                        hipMemcpyHostToDevice));
    a->sobol = (const uint32_t *)p;
    a->qmc_keys = (const uint32_t *)((char *)p + bytes_v);
    a->qmc_dims = dims;
    return OKA_OK;
}

int oka_draw_realizations(oka_ctx *c, int32_t nwells, const oka_dist *dists, int64_t n,
                          uint64_t seed, int64_t first, double *q, double *K, double *por,
                          double *thick, double *z, uint32_t flags)
{
    if (!c)
        return OKA_ERR_ARG;
    TRY(check_dists(c, nwells, dists));
    if (n < 0 || first < 0)
        return fail(c, OKA_ERR_ARG, "negative count or offset");
    if ((flags & OKA_ANTITHETIC) && ((n | first) & 1))
        return fail(c, OKA_ERR_ARG, "antithetic draws need an even count and offset");
    TRY(check_qmc(c, nwells, dists, n, first, flags));
    if (n == 0)
        return OKA_OK;
    if ((nwells > 0 && !q) || !K || !por || !thick || !z)
        return fail(c, OKA_ERR_ARG, "missing outputs");
    HIPCHK(c, hipSetDevice(c->device));
    DrawArgs a{};
    a.nwells = nwells;
    a.n = n;
    a.first = first;
    a.seed = seed;
    a.antithetic = (flags & OKA_ANTITHETIC) != 0;
    if (flags & OKA_QMC)
        TRY(stage_qmc(c, nwells, dists, &seed, 1, &a));
    const void *dd;
    TRY(stage_in(c, kSlotCvDists, dists, sizeof(oka_dist) * (nwells + 3), 0, &dd));
    a.dists = (const oka_dist *)dd;
    void *o;
    TRY(stage_out(c, kSlotQ, q, sizeof(double) * n * nwells, flags, false, &o));
    a.q = (double *)o;
    TRY(stage_out(c, kSlotK, K, sizeof(double) * n, flags, false, &o));
    a.K = (double *)o;
    TRY(stage_out(c, kSlotN, por, sizeof(double) * n, flags, false, &o));
    a.por = (double *)o;
    TRY(stage_out(c, kSlotB, thick, sizeof(double) * n, flags, false, &o));
    a.thick = (double *)o;
    TRY(stage_out(c, kSlotZ, z, sizeof(double) * n * 6, flags, false, &o));
    a.z = (double *)o;
    HIPCHK(c, launch_draw(a, c->stream));
    TRY(copy_back(c, q, a.q, sizeof(double) * n * nwells, flags));
    TRY(copy_back(c, K, a.K, sizeof(double) * n, flags));
    TRY(copy_back(c, por, a.por, sizeof(double) * n, flags));
    TRY(copy_back(c, thick, a.thick, sizeof(double) * n, flags));
    TRY(copy_back(c, z, a.z, sizeof(double) * n * 6, flags));
    return sync_unless_async(c, flags);
}

int oka_mark_moments(oka_ctx *c, int64_t n, int32_t ncell, const uint8_t *marks, uint64_t *sums,
                     uint64_t *sums_sq, uint32_t flags)
{
    if (!c || !marks || !sums || n < 0 || ncell < 1)
        return OKA_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const void *dm;
    TRY(stage_in(c, kSlotMlpMarks, marks, (size_t)n * ncell, flags, &dm));
    uint64_t *d_sums = sums, *d_sq = sums_sq;
    if (!(flags & OKA_PTR_DEVICE)) {
        void *base;
        TRY(ensure(c, kSlotMlpOut2, sizeof(uint64_t) * 2 * ncell, &base));
        d_sums = (uint64_t *)base;
        d_sq = sums_sq ? d_sums + ncell : nullptr;
        HIPCHK(c, hipMemcpyAsync(d_sums, sums, sizeof(uint64_t) * ncell, hipMemcpyHostToDevice, s));
        if (sums_sq)
            HIPCHK(c, hipMemcpyAsync(d_sq, sums_sq, sizeof(uint64_t) * ncell,
                                     hipMemcpyHostToDevice, s));
    }
    HIPCHK(c, launch_mark_moments((const uint8_t *)dm, n, ncell, d_sums, d_sq, s));
    if (!(flags & OKA_PTR_DEVICE)) {
        HIPCHK(c, hipMemcpyAsync(sums, d_sums, sizeof(uint64_t) * ncell, hipMemcpyDeviceToHost, s));
        if (sums_sq)
            HIPCHK(c, hipMemcpyAsync(sums_sq, d_sq, sizeof(uint64_t) * ncell,
                                     hipMemcpyDeviceToHost, s));
    }
    return sync_unless_async(c, flags);
}

// The surrogate sums of nrep runs over realizations first .. first + n - 1,
// run k with seeds[k] into row k of sums / sums_sq (nrep x n_out): the site
// is staged once and every run is enqueued on the context's stream.
static int cv_sums(oka_ctx *c, oka_mlp *M, const oka_site *site, const oka_dist *dists,
                   int64_t n, const uint64_t *seeds, int32_t nrep, int64_t first,
                   const double *feat_mean, const double *feat_std, double *sums,
                   double *sums_sq, uint32_t flags)
{
    if (!c || !M || !site || !feat_mean || !feat_std || !sums || !seeds)
        return OKA_ERR_ARG;
    if (nrep < 0)
        return fail(c, OKA_ERR_ARG, "negative replicate count");
    TRY(check_dists(c, site->nwells, dists));
    const int nw = site->nwells, nobs = site->nobs, nin = nw + 9;
    if (M->nin != nin)
        return fail(c, OKA_ERR_ARG, "network has %d inputs, the site needs %d", M->nin, nin);
    if (nobs < 6 || !site->obs || (nw > 0 && !site->wells_xy))
        return fail(c, OKA_ERR_ARG, "site needs wells and at least 6 observations");
    if (n < 0 || first < 0)
        return fail(c, OKA_ERR_ARG, "negative count or offset");
    const bool pairs = (flags & OKA_ANTITHETIC) != 0;
    // pairs (2k, 2k + 1) never straddle a chunk or a network pass: both are
    // multiples of max_batch (a multiple of 64) from an even `first`
    if (pairs && ((n | first) & 1))
        return fail(c, OKA_ERR_ARG, "antithetic draws need an even count and offset");
    TRY(check_qmc(c, nw, dists, n, first, flags));
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    DrawArgs qmc{};   // the QMC tables (unused without OKA_QMC)
    if (flags & OKA_QMC)
        TRY(stage_qmc(c, nw, dists, seeds, nrep, &qmc));
    // QMC replicates whose point count is a multiple of 1024 run as one
    // sequence of nrep * n realizations (replicate = index / n), so every
    // launch has full chunks; the network passes (multiples of 1024 rows)
    // then never straddle two replicates' sums
    const bool batched = (flags & OKA_QMC) && nrep > 1 && n % 1024 == 0 && first == 0 &&
                         M->max_batch % 1024 == 0;
    const int64_t total = batched ? n * nrep : n;
    const int32_t runs = batched ? 1 : nrep;
    // site, distributions, standardization (host memory)
    const size_t site_doubles = (size_t)2 * nw + 4 * nobs + (size_t)nobs * nw + 2 * nin;
    void *sb;
    TRY(ensure(c, kSlotCvSite, sizeof(double) * site_doubles, &sb));
    double *d_wells = (double *)sb, *d_obs = d_wells + 2 * nw, *d_logr2 = d_obs + 4 * nobs;
    double *d_mean = d_logr2 + (size_t)nobs * nw, *d_std = d_mean + nin;
    HIPCHK(c, hipMemcpyAsync(d_wells, site->wells_xy, sizeof(double) * 2 * nw,
                             hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(d_obs, site->obs, sizeof(double) * 4 * nobs, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(d_mean, feat_mean, sizeof(double) * nin, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(d_std, feat_std, sizeof(double) * nin, hipMemcpyHostToDevice, s));
    const void *dd;
    TRY(stage_in(c, kSlotCvDists, dists, sizeof(oka_dist) * (nw + 3), 0, &dd));
    // per-chunk scratch: draws, fits and features go 64 Ki realizations per
    // launch (enough lanes to fill the GPU), the network max_batch rows per pass
    const int64_t chunk = std::max<int64_t>(M->max_batch, 65536 / M->max_batch * M->max_batch);
    const size_t per = (size_t)nw + 3 + 6 + 6;   // q, K, por, thick, z, coef (doubles)
    void *wb;
    TRY(ensure(c, kSlotCvDraw, (sizeof(double) * per + sizeof(float) * nin + 4) * chunk + 256,
               &wb));
    double *q = (double *)wb, *K = q + chunk * nw, *por = K + chunk, *thick = por + chunk;
    double *z = thick + chunk, *coef = z + chunk * 6;
    float *x = (float *)(coef + chunk * 6);
    int32_t *status = (int32_t *)(x + chunk * nin);
    // accumulators
    void *acc_base = nullptr;
    double *d_sums = sums, *d_sq = sums_sq;
    const size_t acc_bytes = sizeof(double) * (size_t)nrep * M->nout;
    if (!(flags & OKA_PTR_DEVICE)) {
        TRY(ensure(c, kSlotMlpOut2, 2 * acc_bytes, &acc_base));
        d_sums = (double *)acc_base;
        d_sq = sums_sq ? d_sums + (size_t)nrep * M->nout : nullptr;
        HIPCHK(c, hipMemcpyAsync(d_sums, sums, acc_bytes, hipMemcpyHostToDevice, s));
        if (sums_sq)
            HIPCHK(c, hipMemcpyAsync(d_sq, sums_sq, acc_bytes, hipMemcpyHostToDevice, s));
    }
    FitArgs fa{};
    fa.nwells = nw;
    fa.nobs = nobs;
    fa.wells_xy = d_wells;
    fa.obs = d_obs;
    fa.xo = site->xo;
    fa.yo = site->yo;
    fa.base = site->base;
    fa.logr2 = d_logr2;
    fa.K = K;
    fa.b = thick;
    fa.q = q;
    fa.z = z;
    fa.coef = coef;
    fa.status = status;
    fa.q_well_stride = chunk;
    HIPCHK(c, launch_fit_logs(fa, d_logr2, s));
    for (int32_t k = 0; k < runs; ++k) {
        for (int64_t r0 = 0; r0 < total; r0 += chunk) {
            const uint64_t seed = seeds[k];
            const int valid = (int)std::min<int64_t>(chunk, total - r0);
            DrawArgs da{};
            da.nwells = nw;
            da.dists = (const oka_dist *)dd;
            da.n = valid;
            da.first = first + r0;
            da.seed = seed;
            da.q = q;
            da.K = K;
            da.por = por;
            da.thick = thick;
            da.z = z;
            da.antithetic = pairs;
            da.sobol = qmc.sobol;
            da.qmc_dims = qmc.qmc_dims;
            da.qmc_keys = qmc.qmc_keys ? qmc.qmc_keys + (batched ? 0 : (size_t)k * qmc.qmc_dims)
                                       : nullptr;
            da.qmc_npts = batched ? n : 0;
            da.q_well_stride = chunk;   // well-major: the fit's loads are contiguous
            HIPCHK(c, launch_draw(da, s));
            fa.nreal = valid;
            HIPCHK(c, launch_fit_draw(fa, s));
            FeatureArgs fe{};
            fe.nwells = nw;
            fe.n = valid;
            fe.q = q;
            fe.K = K;
            fe.por = por;
            fe.thick = thick;
            fe.coef = coef;
            fe.mean = d_mean;
            fe.std = d_std;
            fe.x = x;
            fe.q_well_stride = chunk;
            HIPCHK(c, launch_features(fe, s));
            for (int r1 = 0; r1 < valid; r1 += M->max_batch) {
                const int nv = std::min(M->max_batch, valid - r1);
                const int Bp = (nv + 63) / 64 * 64;
                TRY(hidden_forward(M, Bp, nv, x, nullptr, r1, false));
                GemmArgs g = gemm(M->h2, M->nh, M->w3h, M->nh, Bp, M->nout, M->nh);
                g.bias = M->p + M->ob3;
                g.mvalid = nv;
                g.pairs = pairs;
                g.sums = d_sums + (size_t)k * M->nout;
                g.sums_sq = d_sq ? d_sq + (size_t)k * M->nout : nullptr;
                if (batched) {   // rows of replicate (r0 + r1 + m) / n
                    g.sum_row0 = r0 + r1;
                    g.sum_rows = n;
                }
                HIPCHK(c, launch_gemm(kEpiAccum, g, s));
            }
        }
    }
    if (!(flags & OKA_PTR_DEVICE)) {
        HIPCHK(c, hipMemcpyAsync(sums, d_sums, acc_bytes, hipMemcpyDeviceToHost, s));
        if (sums_sq)
            HIPCHK(c, hipMemcpyAsync(sums_sq, d_sq, acc_bytes, hipMemcpyDeviceToHost, s));
    }
    return sync_unless_async(c, flags);
}

}  // extern "C"

// ---- node-resolution correction sums ---------------------------------------

namespace {

// The node-resolution sums take each window node's controls from the 3 x 3
// cells around its own (kNbr of them, neighbour m = cell (r + m / 3 - 1,
// c + m % 3 - 1); a neighbour outside the window is absent and counts as 0),
// and the products of two controls from pairs of cells kPairs half-offsets
// apart (the offsets (dr, dc) of a 5 x 5 block with dr > 0, or dr = 0 and
// dc >= 0; the other half by symmetry).
constexpr int kNbr = 9;
constexpr int kPairs = 13;

__device__ __forceinline__ void pair_offset(int h, int &dr, int &dc)
{
    if (h < 3) {
        dr = 0;
        dc = h;
    } else if (h < 8) {
        dr = 1;
        dc = h - 5;
    } else {
        dr = 2;
        dc = h - 10;
    }
}

// The window's cells over rows [lo, hi) of one chunk (probs [rows x nout],
// nout = nr nc cells), in gridDim.y row ranges: thread (cell a, split sp)
// writes part[(sp nout + a) (1 + kPairs) + f]: f = 0 the sum of s_a, 1 + h
// the sum of s_a s_b for b = a + offset h (0 when b is outside the window).
// FP64 accumulators in row order.
__global__ void __launch_bounds__(256) k_node_cells(const float *probs, int nr, int nc, int rows,
                                                    double *part)
{
    const int nout = nr * nc;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nout)
        return;
    const int sp = blockIdx.y, splits = gridDim.y;
    const int lo = (int)((int64_t)rows * sp / splits);
    const int hi = (int)((int64_t)rows * (sp + 1) / splits);
    const int ar = a / nc, ac = a % nc;
    int bi[kPairs];
#pragma unroll
    for (int h = 0; h < kPairs; ++h) {
        int dr, dc;
        pair_offset(h, dr, dc);
        const int br = ar + dr, bc = ac + dc;
        bi[h] = (br < nr && bc >= 0 && bc < nc) ? br * nc + bc : -1;
    }
    double s1 = 0.0, q[kPairs];
#pragma unroll
    for (int h = 0; h < kPairs; ++h)
        q[h] = 0.0;
    for (int r = lo; r < hi; ++r) {
        const float *row = probs + (int64_t)r * nout;
        const double sa = (double)row[a];
        s1 += sa;
#pragma unroll
        for (int h = 0; h < kPairs; ++h)
            if (bi[h] >= 0)
                q[h] = fma(sa, (double)row[bi[h]], q[h]);
    }
    double *o = part + ((int64_t)sp * nout + a) * (1 + kPairs);
    o[0] = s1;
#pragma unroll
    for (int h = 0; h < kPairs; ++h)
        o[1 + h] = q[h];
}

// The window's nodes over the same rows: thread (cell a, node k, split sp)
// writes part[((sp nout + a) nk + k) (1 + kNbr) + f]: f = 0 the rows with
// node k of cell a visited (bit k of the cell's mask), 1 + m the sum over
// them of neighbour m's s.
__global__ void __launch_bounds__(256) k_node_nodes(const float *probs, const uint16_t *masks,
                                                    int nr, int nc, int nk, int rows, double *part)
{
    const int nout = nr * nc;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nout * nk)
        return;
    const int a = (int)(t / nk), k = (int)(t % nk);
    const int sp = blockIdx.y, splits = gridDim.y;
    const int lo = (int)((int64_t)rows * sp / splits);
    const int hi = (int)((int64_t)rows * (sp + 1) / splits);
    const int ar = a / nc, ac = a % nc;
    int ni[kNbr];
#pragma unroll
    for (int m = 0; m < kNbr; ++m) {
        const int br = ar + m / 3 - 1, bc = ac + m % 3 - 1;
        ni[m] = (br >= 0 && br < nr && bc >= 0 && bc < nc) ? br * nc + bc : -1;
    }
    double y = 0.0, ys[kNbr];
#pragma unroll
    for (int m = 0; m < kNbr; ++m)
        ys[m] = 0.0;
    for (int r = lo; r < hi; ++r) {
        if (!((masks[(int64_t)r * nout + a] >> k) & 1u))
            continue;
        const float *row = probs + (int64_t)r * nout;
        y += 1.0;
#pragma unroll
        for (int m = 0; m < kNbr; ++m)
            if (ni[m] >= 0)
                ys[m] += (double)row[ni[m]];
    }
    double *o = part + (((int64_t)sp * nout + a) * nk + k) * (1 + kNbr);
    o[0] = y;
#pragma unroll
    for (int m = 0; m < kNbr; ++m)
        o[1 + m] = ys[m];
}

// acc[i] += sum over splits (in order) of part[sp * n + i].
__global__ void k_node_reduce(const double *part, int splits, int64_t n, double *acc)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    double t = acc[i];
    for (int sp = 0; sp < splits; ++sp)
        t += part[(int64_t)sp * n + i];
    acc[i] = t;
}

constexpr int kNodeSplits = 32;

// Device copies of a correction set's trace inputs (or the caller's device
// arrays as they are), with the site and options pointing at them.
struct NodeInputs {
    oka_site site;
    oka_trace_opts opts;
    const double *K, *n, *b, *q, *coef;
};

int stage_node_inputs(oka_ctx *c, const oka_site *site, const oka_trace_opts *opts, int32_t nreal,
                      const double *K, const double *n, const double *b, const double *q,
                      const double *coef, uint32_t flags, NodeInputs *out)
{
    NodeInputs &in = *out;
    in.site = *site;
    in.opts = *opts;
    in.K = K;
    in.n = n;
    in.b = b;
    in.q = q;
    in.coef = coef;
    in.site.obs = nullptr;
    in.site.nobs = 0;
    if (flags & OKA_PTR_DEVICE)
        return OKA_OK;
    const int nw = site->nwells, P = opts->npaths;
    const size_t parts[7] = {2 * (size_t)nw, 2 * (size_t)P, (size_t)nreal, (size_t)nreal,
                             (size_t)nreal, (size_t)nreal * nw, 6 * (size_t)nreal};
    const double *src[7] = {site->wells_xy, opts->start_xy, K, n, b, q, coef};
    size_t total = 0;
    for (size_t v : parts)
        total += v;
    void *buf;
    TRY(ensure(c, kSlotNodeIn, sizeof(double) * std::max<size_t>(total, 1), &buf));
    const double *dst[7];
    double *d = (double *)buf;
    for (int i = 0; i < 7; ++i) {
        if (parts[i] && !src[i])
            return fail(c, OKA_ERR_ARG, "missing realization inputs");
        if (parts[i])
            HIPCHK(c, hipMemcpyAsync(d, src[i], sizeof(double) * parts[i], hipMemcpyHostToDevice,
                                     c->stream));
        dst[i] = d;
        d += parts[i];
    }
    in.site.wells_xy = dst[0];
    in.opts.start_xy = dst[1];
    in.K = dst[2];
    in.n = dst[3];
    in.b = dst[4];
    in.q = dst[5];
    in.coef = dst[6];
    return OKA_OK;
}

int check_node_args(oka_ctx *c, int64_t nout, const oka_site *site, const oka_trace_opts *opts,
                    const oka_grid *grid, const oka_window *win, int32_t nreal)
{
    if (!site || !opts || !grid || !win || nreal < 0)
        return fail(c, OKA_ERR_ARG, "bad node-trace arguments");
    if (win->cell < 1 || win->cell > 4 || (nout > 0 && (int64_t)win->nrows * win->ncols != nout))
        return fail(c, OKA_ERR_ARG, "the window must have the network's %lld cells of at most "
                    "4 x 4 nodes", (long long)nout);
    if (opts->npaths > 0 && !opts->start_xy)
        return fail(c, OKA_ERR_ARG, "start_xy is NULL");
    return OKA_OK;
}

void add_stats(oka_trace_stats &tot, const oka_trace_stats &st)
{
    tot.attempted += st.attempted;
    tot.accepted += st.accepted;
    for (int j = 0; j < 4; ++j)
        tot.status_count[j] += st.status_count[j];
    tot.escaped += st.escaped;
    tot.visited += st.visited;
    tot.candidates += st.candidates;
    tot.kernel_ms += st.kernel_ms;
}

// Traces rows [r0, r0 + rows) of the staged set: node masks into `masks`
// (rows x cells, device), cell marks into `marks` (scratch), counts added.
int trace_node_rows(oka_ctx *c, const NodeInputs &in, int64_t r0, int rows, const oka_grid *grid,
                    const oka_window *win, uint16_t *masks, uint8_t *marks, int32_t *dcounts,
                    oka_trace_stats &tot)
{
    const int nw = in.site.nwells;
    oka_trace_stats st{};
    TRY(oka_trace_marks(c, &in.site, rows, in.K + r0, in.n + r0, in.b + r0, in.q + r0 * nw,
                        in.coef + 6 * r0, &in.opts, grid, win, marks, masks, dcounts, &st,
                        OKA_PTR_DEVICE));
    add_stats(tot, st);
    return OKA_OK;
}

// The network on rows [r0, r0 + rows) of dx and the node sums of those rows
// (masks: their rows x cells node masks) added to one part's accumulators.
int reduce_node_rows(oka_mlp *M, const float *dx, int64_t r0, int rows, const uint16_t *masks,
                     const oka_window *win, float *probs, double *part, double *accp)
{
    oka_ctx *c = M->ctx;
    const hipStream_t s = c->stream;
    const int nk = win->cell * win->cell;
    const int64_t nout = M->nout;
    const int64_t ncf = nout * (1 + kPairs), nnf = nout * nk * (1 + kNbr);
    const int Bp = (rows + 63) / 64 * 64;
    TRY(hidden_forward(M, Bp, rows, dx, nullptr, r0, false));
    GemmArgs g = gemm(M->h2, M->nh, M->w3h, M->nh, Bp, M->nout, M->nh);
    g.bias = M->p + M->ob3;
    g.mvalid = rows;
    g.Cf = probs;
    g.ldc = M->nout;
    HIPCHK(c, launch_gemm(kEpiProbs, g, s));
    double *pc = part, *pn = part + kNodeSplits * ncf;
    hipLaunchKernelGGL(k_node_cells, dim3((unsigned)((nout + 255) / 256), kNodeSplits), dim3(256),
                       0, s, (const float *)probs, win->nrows, win->ncols, rows, pc);
    hipLaunchKernelGGL(k_node_nodes, dim3((unsigned)((nout * nk + 255) / 256), kNodeSplits),
                       dim3(256), 0, s, (const float *)probs, masks, win->nrows, win->ncols, nk,
                       rows, pn);
    hipLaunchKernelGGL(k_node_reduce, dim3((unsigned)((ncf + 255) / 256)), dim3(256), 0, s,
                       (const double *)pc, kNodeSplits, ncf, accp);
    hipLaunchKernelGGL(k_node_reduce, dim3((unsigned)((nnf + 255) / 256)), dim3(256), 0, s,
                       (const double *)pn, kNodeSplits, nnf, accp + ncf);
    HIPCHK(c, hipGetLastError());
    return OKA_OK;
}

// The part accumulators [part]([cell][1 + kPairs], [cell][node][1 + kNbr])
// unpacked into the four host outputs.
void unpack_node_sums(const std::vector<double> &h, int nparts, int64_t nout, int nk,
                      double *sum_s, double *sum_ss, double *sum_y, double *sum_ys)
{
    const int fc = 1 + kPairs, fn = 1 + kNbr;
    const int64_t ncf = nout * fc, nnf = nout * nk * fn;
    for (int p = 0; p < nparts; ++p) {
        const double *hc = h.data() + (size_t)p * (ncf + nnf), *hn = hc + ncf;
        for (int64_t a = 0; a < nout; ++a) {
            sum_s[p * nout + a] = hc[a * fc];
            for (int q = 0; q < kPairs; ++q)
                sum_ss[(p * nout + a) * kPairs + q] = hc[a * fc + 1 + q];
            for (int k = 0; k < nk; ++k) {
                const double *v = hn + (a * nk + k) * fn;
                sum_y[(p * nout + a) * nk + k] = v[0];
                for (int m = 0; m < kNbr; ++m)
                    sum_ys[((p * nout + a) * nk + k) * kNbr + m] = v[1 + m];
            }
        }
    }
}

// Work buffers of the node reduction: probabilities and masks of one chunk,
// the split partial sums and the part accumulators (zeroed).
int node_buffers(oka_mlp *M, int nk, int nparts, float **probs, double **part, double **acc)
{
    oka_ctx *c = M->ctx;
    const int64_t nout = M->nout;
    const int64_t ncf = nout * (1 + kPairs), nnf = nout * nk * (1 + kNbr);
    void *p1, *p2, *p3;
    TRY(ensure(c, kSlotNodeProbs, sizeof(float) * M->max_batch * nout, &p1));
    TRY(ensure(c, kSlotNodePart, sizeof(double) * kNodeSplits * (ncf + nnf), &p2));
    TRY(ensure(c, kSlotNodeAcc, sizeof(double) * nparts * (ncf + nnf), &p3));
    TRY(fill_u32(c, p3, 2 * (int64_t)nparts * (ncf + nnf)));
    *probs = (float *)p1;
    *part = (double *)p2;
    *acc = (double *)p3;
    return OKA_OK;
}

int stage_features(oka_mlp *M, const float *x, int64_t nreal, uint32_t flags, const float **dx)
{
    oka_ctx *c = M->ctx;
    *dx = x;
    if (flags & OKA_PTR_DEVICE)
        return OKA_OK;
    void *xv;
    TRY(ensure(c, kSlotNodeX, sizeof(float) * std::max<int64_t>(nreal * M->nin, 1), &xv));
    HIPCHK(c, hipMemcpyAsync(xv, x, sizeof(float) * (size_t)nreal * M->nin, hipMemcpyHostToDevice,
                             c->stream));
    *dx = (const float *)xv;
    return OKA_OK;
}

}  // namespace

struct oka_node_set {
    oka_ctx *c = nullptr;
    int64_t nreal = 0;
    oka_window win{};
    uint16_t *masks = nullptr;      // [nreal][cells] node masks on the device
};

extern "C" {

int oka_cv_node_sums(oka_ctx *c, oka_mlp *M, const oka_site *site, int32_t nreal,
                     const double *K, const double *n, const double *b, const double *q,
                     const double *coef, const float *x, const oka_trace_opts *opts,
                     const oka_grid *grid, const oka_window *win, int32_t nparts, int32_t *counts,
                     double *sum_s, double *sum_ss, double *sum_y, double *sum_ys,
                     oka_trace_stats *stats, uint32_t flags)
{
    if (!c || !M || M->ctx != c)
        return c ? fail(c, OKA_ERR_ARG, "the network belongs to another context") : OKA_ERR_ARG;
    TRY(check_node_args(c, M->nout, site, opts, grid, win, nreal));
    if (!x || !sum_s || !sum_ss || !sum_y || !sum_ys || nparts < 1 || nparts > nreal)
        return fail(c, OKA_ERR_ARG, "bad node-sum arguments");
    HIPCHK(c, hipSetDevice(c->device));
    const int nk = win->cell * win->cell;
    const int64_t nout = M->nout;
    NodeInputs in;
    TRY(stage_node_inputs(c, site, opts, nreal, K, n, b, q, coef, flags, &in));
    const float *dx;
    TRY(stage_features(M, x, nreal, flags, &dx));
    void *dcounts = nullptr;
    const int64_t nnodes = (int64_t)grid->nrows * grid->ncols;
    if (counts)
        TRY(stage_out(c, kSlotNodeCounts, counts, sizeof(int32_t) * nnodes, flags, true, &dcounts));
    const int chunk = M->max_batch;
    void *masks, *marks;
    TRY(ensure(c, kSlotNodeMasks, sizeof(uint16_t) * chunk * nout, &masks));
    TRY(ensure(c, kSlotNodeMarks, (size_t)chunk * nout, &marks));
    float *probs;
    double *part, *acc;
    TRY(node_buffers(M, nk, nparts, &probs, &part, &acc));
    const int64_t per_part = nout * (1 + kPairs) + nout * nk * (1 + kNbr);
    oka_trace_stats tot{};
    // each part's rows in chunks of the network's batch: trace (marks and node
    // masks stay on the device), forward, then the part's node sums
    for (int p = 0; p < nparts; ++p) {
        const int64_t lo = (int64_t)nreal * p / nparts, hi = (int64_t)nreal * (p + 1) / nparts;
        for (int64_t r0 = lo; r0 < hi; r0 += chunk) {
            const int rows = (int)std::min<int64_t>(chunk, hi - r0);
            TRY(trace_node_rows(c, in, r0, rows, grid, win, (uint16_t *)masks, (uint8_t *)marks,
                                (int32_t *)dcounts, tot));
            TRY(reduce_node_rows(M, dx, r0, rows, (const uint16_t *)masks, win, probs, part,
                                 acc + p * per_part));
        }
    }
    std::vector<double> h((size_t)nparts * per_part);
    HIPCHK(c, hipMemcpyAsync(h.data(), acc, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                             c->stream));
    if (counts)
        TRY(copy_back(c, counts, dcounts, sizeof(int32_t) * nnodes, flags));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unpack_node_sums(h, nparts, nout, nk, sum_s, sum_ss, sum_y, sum_ys);
    if (stats)
        *stats = tot;
    return OKA_OK;
}

int oka_cv_node_trace(oka_ctx *c, const oka_site *site, int32_t nreal, const double *K,
                      const double *n, const double *b, const double *q, const double *coef,
                      const oka_trace_opts *opts, const oka_grid *grid, const oka_window *win,
                      int32_t *counts, oka_node_set **out, oka_trace_stats *stats, uint32_t flags)
{
    if (!c || !out)
        return c ? fail(c, OKA_ERR_ARG, "out is NULL") : OKA_ERR_ARG;
    *out = nullptr;
    TRY(check_node_args(c, 0, site, opts, grid, win, nreal));
    if ((int64_t)win->nrows * win->ncols > 16384)
        return fail(c, OKA_ERR_ARG, "at most 16384 window cells");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t ncell = (int64_t)win->nrows * win->ncols;
    NodeInputs in;
    TRY(stage_node_inputs(c, site, opts, nreal, K, n, b, q, coef, flags, &in));
    void *dcounts = nullptr;
    const int64_t nnodes = (int64_t)grid->nrows * grid->ncols;
    if (counts)
        TRY(stage_out(c, kSlotNodeCounts, counts, sizeof(int32_t) * nnodes, flags, true, &dcounts));
    oka_node_set *S = new oka_node_set();
    S->c = c;
    S->nreal = nreal;
    S->win = *win;
    if (hipMalloc((void **)&S->masks, sizeof(uint16_t) * std::max<int64_t>(nreal * ncell, 1)) !=
        hipSuccess) {
        delete S;
        return fail(c, OKA_ERR_NOMEM, "no device memory for %d x %lld node masks", nreal,
                    (long long)ncell);
    }
    // chunks of realizations: the cell marks go to a chunk-sized scratch
    const int chunk = 16384;
    void *marks;
    int rc = ensure(c, kSlotNodeMarks, (size_t)std::min<int64_t>(chunk, std::max(nreal, 1)) * ncell,
                    &marks);
    oka_trace_stats tot{};
    for (int64_t r0 = 0; rc == OKA_OK && r0 < nreal; r0 += chunk) {
        const int rows = (int)std::min<int64_t>(chunk, nreal - r0);
        rc = trace_node_rows(c, in, r0, rows, grid, win, S->masks + r0 * ncell, (uint8_t *)marks,
                             (int32_t *)dcounts, tot);
    }
    if (rc == OKA_OK && counts)
        rc = copy_back(c, counts, dcounts, sizeof(int32_t) * nnodes, flags);
    if (rc == OKA_OK && hipStreamSynchronize(c->stream) != hipSuccess)
        rc = fail(c, OKA_ERR_DEVICE, "node trace failed");
    if (rc != OKA_OK) {
        oka_node_set_destroy(S);
        return rc;
    }
    if (stats)
        *stats = tot;
    *out = S;
    return OKA_OK;
}

void oka_node_set_destroy(oka_node_set *S)
{
    if (!S)
        return;
    if (S->masks) {
        (void)hipSetDevice(S->c->device);
        (void)hipFree(S->masks);
    }
    delete S;
}

int oka_cv_node_reduce(oka_mlp *M, const oka_node_set *S, const float *x, int32_t nparts,
                       double *sum_s, double *sum_ss, double *sum_y, double *sum_ys,
                       uint32_t flags)
{
    if (!M)
        return OKA_ERR_ARG;
    oka_ctx *c = M->ctx;
    if (!S || !x || !sum_s || !sum_ss || !sum_y || !sum_ys || nparts < 1 || nparts > S->nreal)
        return fail(c, OKA_ERR_ARG, "bad node-reduce arguments");
    if (S->c->device != c->device)
        return fail(c, OKA_ERR_ARG, "the node set lives on another device");
    if ((int64_t)S->win.nrows * S->win.ncols != M->nout)
        return fail(c, OKA_ERR_ARG, "the node set's window has %d cells, the network %d outputs",
                    S->win.nrows * S->win.ncols, M->nout);
    HIPCHK(c, hipSetDevice(c->device));
    const int nk = S->win.cell * S->win.cell;
    const int64_t nout = M->nout, nreal = S->nreal;
    const float *dx;
    TRY(stage_features(M, x, nreal, flags, &dx));
    float *probs;
    double *part, *acc;
    TRY(node_buffers(M, nk, nparts, &probs, &part, &acc));
    const int64_t per_part = nout * (1 + kPairs) + nout * nk * (1 + kNbr);
    const int chunk = M->max_batch;
    for (int p = 0; p < nparts; ++p) {
        const int64_t lo = nreal * p / nparts, hi = nreal * (p + 1) / nparts;
        for (int64_t r0 = lo; r0 < hi; r0 += chunk) {
            const int rows = (int)std::min<int64_t>(chunk, hi - r0);
            TRY(reduce_node_rows(M, dx, r0, rows, S->masks + r0 * nout, &S->win, probs, part,
                                 acc + p * per_part));
        }
    }
    std::vector<double> h((size_t)nparts * per_part);
    HIPCHK(c, hipMemcpyAsync(h.data(), acc, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unpack_node_sums(h, nparts, nout, nk, sum_s, sum_ss, sum_y, sum_ys);
    return OKA_OK;
}

int oka_cv_surrogate_sums(oka_ctx *c, oka_mlp *M, const oka_site *site, const oka_dist *dists,
                          int64_t n, uint64_t seed, int64_t first, const double *feat_mean,
                          const double *feat_std, double *sums, double *sums_sq, uint32_t flags)
{
    return cv_sums(c, M, site, dists, n, &seed, 1, first, feat_mean, feat_std, sums, sums_sq,
                   flags);
}

int oka_cv_surrogate_sums_qmc(oka_ctx *c, oka_mlp *M, const oka_site *site,
                              const oka_dist *dists, int64_t npts, int32_t nrep,
                              const uint64_t *seeds, const double *feat_mean,
                              const double *feat_std, double *sums, uint32_t flags)
{
    if (c && (flags & OKA_ANTITHETIC))
        return fail(c, OKA_ERR_ARG, "OKA_QMC and OKA_ANTITHETIC exclude each other");
    return cv_sums(c, M, site, dists, npts, seeds, nrep, 0, feat_mean, feat_std, sums, nullptr,
                   flags | OKA_QMC);
}

}  // extern "C"
