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
    return g;
}

// Forward through the two hidden layers for rows of x (rowidx/row0) into the
// workspace; Bp rows (a multiple of 64), the first `valid` of them real.
int hidden_forward(oka_mlp *M, int Bp, int valid, const float *x, const int32_t *rowidx,
                   int64_t row0)
{
    oka_ctx *c = M->ctx;
    hipStream_t s = c->stream;
    HIPCHK(c, launch_prep_input(x, M->nin, M->kin, rowidx, row0, Bp, valid, M->x0, M->x0t, s));
    GemmArgs g = gemm(M->x0, M->kin, M->w1h, M->kin, Bp, M->nh, M->kin);
    g.bias = M->p + M->ob1;
    g.Cf = M->z1;
    g.ldc = M->nh;
    g.Ch = M->h1;
    g.Cht = M->h1t;
    HIPCHK(c, launch_gemm(kEpiGeluFwd, g, s));
    g = gemm(M->h1, M->nh, M->w2h, M->nh, Bp, M->nh, M->nh);
    g.bias = M->p + M->ob2;
    g.Cf = M->z2;
    g.ldc = M->nh;
    g.Ch = M->h2;
    g.Cht = M->h2t;
    HIPCHK(c, launch_gemm(kEpiGeluFwd, g, s));
    return OKA_OK;
}

// Loss and full gradient (into M->grad) for one batch of B rows.
int forward_backward(oka_mlp *M, int B, const float *x, const uint8_t *marks, float tscale,
                     const int32_t *rowidx, int64_t row0, double *loss)
{
    oka_ctx *c = M->ctx;
    hipStream_t s = c->stream;
    TRY(hidden_forward(M, B, B, x, rowidx, row0));
    GemmArgs g = gemm(M->h2, M->nh, M->w3h, M->nh, B, M->nout, M->nh);
    g.bias = M->p + M->ob3;
    g.marks = marks;
    g.rowidx = rowidx;
    g.row0 = row0;
    g.tscale = tscale;
    g.gscale = 1.0f / ((float)B * (float)M->nout);
    g.Ch = M->g3;
    g.Cht = M->g3t;
    g.loss = loss;
    HIPCHK(c, launch_gemm(kEpiLossGrad, g, s));
    // layer 3: dW3 = G3^T H2, db3, G2 = (G3 W3) * gelu'(Z2)
    g = gemm(M->g3t, B, M->h2t, B, M->nout, M->nh, B);
    g.Cf = M->grad + M->oW3;
    g.ldc = M->nh;
    HIPCHK(c, launch_gemm(kEpiStoreF32, g, s));
    HIPCHK(c, launch_bias_grad(M->g3t, M->nout, B, M->grad + M->ob3, s));
    g = gemm(M->g3, M->nout, M->w3t, M->nout, B, M->nh, M->nout);
    g.Zsaved = M->z2;
    g.Ch = M->g2;
    g.Cht = M->g2t;
    HIPCHK(c, launch_gemm(kEpiGeluBwd, g, s));
    // layer 2
    g = gemm(M->g2t, B, M->h1t, B, M->nh, M->nh, B);
    g.Cf = M->grad + M->oW2;
    g.ldc = M->nh;
    HIPCHK(c, launch_gemm(kEpiStoreF32, g, s));
    HIPCHK(c, launch_bias_grad(M->g2t, M->nh, B, M->grad + M->ob2, s));
    g = gemm(M->g2, M->nh, M->w2t, M->nh, B, M->nh, M->nh);
    g.Zsaved = M->z1;
    g.Ch = M->g1;
    g.Cht = M->g1t;
    HIPCHK(c, launch_gemm(kEpiGeluBwd, g, s));
    // layer 1
    g = gemm(M->g1t, B, M->x0t, B, M->nh, M->kin, B);
    g.Cf = M->grad + M->oW1;
    g.ldc = M->kin;
    HIPCHK(c, launch_gemm(kEpiStoreF32, g, s));
    HIPCHK(c, launch_bias_grad(M->g1t, M->nh, B, M->grad + M->ob1, s));
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
    void *dl;
    TRY(ensure(c, kSlotMlpLoss, sizeof(double), &dl));
    HIPCHK(c, launch_zero_f64((double *)dl, 1, c->stream));
    TRY(forward_backward(M, batch, (const float *)dx, (const uint8_t *)dm, target_scale, nullptr, 0,
                         (double *)dl));
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
    void *dl;
    TRY(ensure(c, kSlotMlpLoss, sizeof(double) * std::max(nsteps, 1), &dl));
    HIPCHK(c, launch_zero_f64((double *)dl, nsteps, c->stream));
    for (int32_t st = 0; st < nsteps; ++st) {
        TRY(forward_backward(M, batch, (const float *)dx, (const uint8_t *)dm, target_scale,
                             (const int32_t *)dord, (int64_t)st * batch, (double *)dl + st));
        M->step += 1;
        const float c1 = (float)(1.0 / (1.0 - pow((double)beta1, (double)M->step)));
        const float c2 = (float)(1.0 / (1.0 - pow((double)beta2, (double)M->step)));
        HIPCHK(c, launch_adam(M->p, M->grad, M->m, M->v, M->np, lr, beta1, beta2, eps, c1, c2,
                              c->stream));
        TRY(refresh_bf16(M));
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
        TRY(hidden_forward(M, Bp, valid, (const float *)dx, nullptr, r0));
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
