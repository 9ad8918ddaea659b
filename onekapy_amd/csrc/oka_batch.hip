// oka_batch.hip -- many sites in one persistent, graph-captured run
// (include/oka.h, "batches").  Every site is staged once in buffers the
// batch owns; one run (per-site fit + MVN draw, then the pool tracer over all
// sites) is captured as a HIP graph and replayed.
#include <string.h>

#include <algorithm>
#include <vector>

#include "oka_ctx.h"

using namespace oka;

struct oka_batch {
    oka_ctx *c = nullptr;
    int nsites = 0;
    std::vector<TraceArgs> args;
    std::vector<FitArgs> fits;
    std::vector<int64_t> nnodes;
    unsigned long long *stats = nullptr;   // [nsites][kStatCount]
    std::vector<void *> allocs;
    oka_trace_opts tune{};
    int precision = OKA_FP32;   // every site's opts.precision
    PoolShape shape;
    void *d_sites = nullptr, *d_first = nullptr, *d_queue = nullptr;
    void *d_fits = nullptr, *d_fit_first = nullptr;   // fit table, its two prefix arrays
    int64_t fit_logs = 0, fit_reals = 0;
    void *d_hbm = nullptr, *d_rec = nullptr, *d_beta = nullptr;
    size_t hbm_cap = 0, rec_cap = 0, first_cap = 0, beta_cap = 0;
    hipGraphExec_t exec = nullptr;
};

namespace {

int balloc(oka_batch *B, size_t bytes, void **out)
{
    void *p = nullptr;
    HIPCHK(B->c, hipMalloc(&p, bytes ? bytes : 8));
    B->allocs.push_back(p);
    *out = p;
    return OKA_OK;
}

// A device copy of host data (NULL stays NULL).
template <typename T>
int bcopy(oka_batch *B, const T *src, size_t count, const T **out)
{
    *out = nullptr;
    if (!src)
        return OKA_OK;
    void *p;
    TRY(balloc(B, sizeof(T) * count, &p));
    HIPCHK(B->c, hipMemcpy(p, src, sizeof(T) * count, hipMemcpyHostToDevice));
    *out = (const T *)p;
    return OKA_OK;
}

int stage_site(oka_batch *B, int s, const oka_batch_site &bs)
{
    oka_ctx *c = B->c;
    const oka_site *site = bs.site;
    const oka_trace_opts &o = bs.opts;
    const oka_grid &g = bs.grid;
    if (!site || bs.nreal < 0 || site->nwells < 0 || site->nobs < 0 || o.npaths < 0)
        return fail(c, OKA_ERR_ARG, "site %d: bad sizes", s);
    if (site->nobs < 6 && bs.nreal > 0)
        return fail(c, OKA_ERR_LINALG, "site %d: fewer than 6 observations", s);
    if (bs.nreal > 0 && (!bs.K || !bs.n || !bs.b || (site->nwells > 0 && (!bs.q || !site->wells_xy))
                         || !site->obs))
        return fail(c, OKA_ERR_ARG, "site %d: missing realization inputs", s);
    if (o.npaths > 0 && !o.start_xy)
        return fail(c, OKA_ERR_ARG, "site %d: start_xy is NULL", s);
    if (o.far_order > kFarTerms)
        return fail(c, OKA_ERR_ARG, "site %d: far_order %d above %d", s, o.far_order,
                    kFarTerms);
    if (o.precision != B->precision)
        return fail(c, OKA_ERR_ARG, "site %d: every site of a batch traces in one precision", s);
    if (g.nrows <= 0 || g.ncols <= 0 || !(g.dx > 0.0) || !(g.dy > 0.0))
        return fail(c, OKA_ERR_ARG, "site %d: bad grid", s);
    const int nw = site->nwells, nr = bs.nreal;
    FitArgs &f = B->fits[s];
    TraceArgs &a = B->args[s];
    f.nwells = a.nwells = nw;
    f.nobs = site->nobs;
    f.nreal = a.nreal = nr;
    f.xo = a.xo = site->xo;
    f.yo = a.yo = site->yo;
    f.base = site->base;
    TRY(bcopy(B, site->wells_xy, (size_t)2 * nw, &f.wells_xy));
    a.wells_xy = f.wells_xy;
    TRY(bcopy(B, site->obs, (size_t)4 * site->nobs, &f.obs));
    TRY(bcopy(B, bs.K, (size_t)nr, &f.K));
    a.K = f.K;
    TRY(bcopy(B, bs.n, (size_t)nr, &a.n));
    TRY(bcopy(B, bs.b, (size_t)nr, &f.b));
    a.b = f.b;
    TRY(bcopy(B, bs.q, (size_t)nr * nw, &f.q));
    a.q = f.q;
    TRY(bcopy(B, bs.z, (size_t)6 * nr, &f.z));
    TRY(bcopy(B, o.start_xy, (size_t)2 * o.npaths, &a.start_xy));
    void *p;
    TRY(balloc(B, sizeof(double) * site->nobs * std::max(nw, 1), &p));
    f.logr2 = (const double *)p;
    TRY(balloc(B, sizeof(double) * 6 * std::max(nr, 1), &p));
    f.coef = (double *)p;
    a.coef = f.coef;
    TRY(balloc(B, sizeof(int32_t) * std::max(nr, 1), &p));
    f.status = (int32_t *)p;
    a.npaths = o.npaths;
    a.confined = o.confined ? 1 : 0;
    a.duration = o.duration;
    a.tol = o.tol;
    a.maxstep = o.maxstep;
    a.umbra = o.umbra;
    a.max_steps = o.max_steps > 0 ? o.max_steps : 1000000;
    a.gxmin = g.xmin;
    a.gymin = g.ymin;
    a.gxmax = g.xmax;
    a.gymax = g.ymax;
    a.gdx = g.dx;
    a.gdy = g.dy;
    a.gnrows = g.nrows;
    a.gncols = g.ncols;
    a.rasterize = 1;
    a.tiles_x = (g.ncols + 31) / 32;
    a.ntiles = a.tiles_x * ((g.nrows + 31) / 32);
    const oka_epoch whole{0, g.xmin, g.ymin, g.nrows, g.ncols, 0, 0};
    TRY(bcopy(B, &whole, 1, &a.epochs));
    a.nepochs = 1;
    B->nnodes[s] = (int64_t)g.nrows * g.ncols;
    TRY(balloc(B, sizeof(int32_t) * B->nnodes[s], &p));
    a.counts = (int32_t *)p;
    a.stats = B->stats + (size_t)s * kStatCount;
    FarPlan P;
    far_plan(a, site->wells_xy, o.precision, o.far_order, &P);
    if (P.order > 0) {
        const int32_t *d[2];
        TRY(bcopy(B, P.near.data(), P.near.size(), &d[0]));
        TRY(bcopy(B, P.wbox.data(), P.wbox.size(), &d[1]));
        far_args(P, d[0], d[1], a);
    }
    return OKA_OK;
}

// Plans the pool launch for the batch's knobs, (re)allocates its buffers and
// captures the run.
int capture(oka_batch *B)
{
    oka_ctx *c = B->c;
    if (B->exec) {
        (void)hipGraphExecDestroy(B->exec);
        B->exec = nullptr;
    }
    // Queue order: the sites with the longest backtraces first, so that the
    // launch does not end on their long traces.  The order changes no result:
    // each site's outputs are addressed through its own arguments.
    std::vector<TraceArgs> order(B->args);
    std::stable_sort(order.begin(), order.end(), [](const TraceArgs &x, const TraceArgs &y) {
        return fabs(x.duration) > fabs(y.duration);
    });
    TRY(pool_shape(c, order.data(), B->nsites, &B->tune, B->precision, &B->shape));
    HIPCHK(c, hipMemcpy(B->d_sites, order.data(), sizeof(TraceArgs) * B->nsites,
                        hipMemcpyHostToDevice));
    const PoolShape &sh = B->shape;
    const size_t first_bytes = sizeof(int64_t) * sh.first.size();
    if (!B->d_first || B->first_cap < first_bytes) {
        TRY(balloc(B, first_bytes, &B->d_first));
        B->first_cap = first_bytes;
    }
    HIPCHK(c, hipMemcpy(B->d_first, sh.first.data(), first_bytes, hipMemcpyHostToDevice));
    if (!B->d_hbm || B->hbm_cap < (size_t)sh.hbm_words * 4) {
        TRY(balloc(B, (size_t)sh.hbm_words * 4, &B->d_hbm));
        B->hbm_cap = (size_t)sh.hbm_words * 4;
    }
    if (sh.rec_words > 0 && (!B->d_rec || B->rec_cap < (size_t)sh.rec_words * 4)) {
        TRY(balloc(B, (size_t)sh.rec_words * 4, &B->d_rec));
        B->rec_cap = (size_t)sh.rec_words * 4;
    }
    const size_t beta_bytes = sizeof(float2) * pool_beta_count(sh) * sh.far_stride;
    if (beta_bytes > 0 && (!B->d_beta || B->beta_cap < beta_bytes)) {
        TRY(balloc(B, beta_bytes, &B->d_beta));
        B->beta_cap = beta_bytes;
    }
    const PoolArgs p =
        pool_args(sh, B->d_sites, B->d_first, B->d_queue, B->d_hbm, B->d_rec, B->d_beta);
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = fill_u32(c, B->stats, 2 * (int64_t)kStatCount * B->nsites);
    for (int s = 0; s < B->nsites && rc == OKA_OK; ++s)
        rc = fill_u32(c, B->args[s].counts, B->nnodes[s]);
    // every site's fits in two launches
    const int64_t *ff = (const int64_t *)B->d_fit_first;
    if (rc == OKA_OK &&
        (launch_fit_logs_multi((const FitArgs *)B->d_fits, B->nsites, ff, B->fit_logs,
                               c->stream) != hipSuccess ||
         launch_fit_draw_multi((const FitArgs *)B->d_fits, B->nsites, ff + B->nsites + 1,
                               B->fit_reals, c->stream) != hipSuccess))
        rc = fail(c, OKA_ERR_DEVICE, "fit launch failed");
    if (rc == OKA_OK)
        rc = pool_launch(c, sh, p);
    const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    if (rc != OKA_OK) {
        if (graph)
            (void)hipGraphDestroy(graph);
        return rc;
    }
    HIPCHK(c, ec);
    const hipError_t ei = hipGraphInstantiate(&B->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(c, ei);
    return OKA_OK;
}

}  // namespace

extern "C" {

int oka_batch_create(oka_ctx *c, int32_t nsites, const oka_batch_site *sites,
                     const oka_trace_opts *tune, oka_batch **out)
{
    if (!c || !out || nsites < 1 || !sites)
        return OKA_ERR_ARG;
    *out = nullptr;
    HIPCHK(c, hipSetDevice(c->device));
    oka_batch *B = new oka_batch();
    B->c = c;
    B->nsites = nsites;
    B->args.assign(nsites, TraceArgs{});
    B->fits.assign(nsites, FitArgs{});
    B->nnodes.assign(nsites, 0);
    if (tune)
        B->tune = *tune;
    B->precision = sites[0].opts.precision;
    void *p;
    int rc = balloc(B, sizeof(unsigned long long) * kStatCount * nsites, &p);
    B->stats = (unsigned long long *)p;
    for (int s = 0; s < nsites && rc == OKA_OK; ++s)
        rc = stage_site(B, s, sites[s]);
    rc = rc ? rc : balloc(B, sizeof(TraceArgs) * nsites, &B->d_sites);
    if (rc == OKA_OK && hipMemcpy(B->d_sites, B->args.data(), sizeof(TraceArgs) * nsites,
                                  hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(c, OKA_ERR_DEVICE, "batch staging failed");
    rc = rc ? rc : balloc(B, sizeof(unsigned long long), &B->d_queue);
    if (rc == OKA_OK) {
        // fit table and the prefix sums of its (observation, well) pairs and realizations
        std::vector<int64_t> ff(2 * ((size_t)nsites + 1), 0);
        for (int s = 0; s < nsites; ++s) {
            const FitArgs &f = B->fits[s];
            const int64_t logs = f.nreal > 0 ? (int64_t)f.nobs * f.nwells : 0;
            ff[s + 1] = ff[s] + logs;
            ff[nsites + 1 + s + 1] = ff[nsites + 1 + s] + f.nreal;
        }
        B->fit_logs = ff[nsites];
        B->fit_reals = ff[2 * (size_t)nsites + 1];
        rc = balloc(B, sizeof(FitArgs) * nsites, &B->d_fits);
        rc = rc ? rc : balloc(B, sizeof(int64_t) * ff.size(), &B->d_fit_first);
        if (rc == OKA_OK &&
            (hipMemcpy(B->d_fits, B->fits.data(), sizeof(FitArgs) * nsites,
                       hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(B->d_fit_first, ff.data(), sizeof(int64_t) * ff.size(),
                       hipMemcpyHostToDevice) != hipSuccess))
            rc = fail(c, OKA_ERR_DEVICE, "batch staging failed");
    }
    rc = rc ? rc : capture(B);
    if (rc != OKA_OK) {
        oka_batch_destroy(B);
        return rc;
    }
    *out = B;
    return OKA_OK;
}

int oka_batch_tune(oka_batch *B, const oka_trace_opts *tune)
{
    if (!B || !tune)
        return OKA_ERR_ARG;
    HIPCHK(B->c, hipSetDevice(B->c->device));
    HIPCHK(B->c, hipStreamSynchronize(B->c->stream));
    B->tune = *tune;
    return capture(B);
}

int oka_batch_run(oka_batch *B, float *ms)
{
    if (!B || !B->exec)
        return OKA_ERR_ARG;
    oka_ctx *c = B->c;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, hipGraphLaunch(B->exec, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (ms)
        HIPCHK(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
    return OKA_OK;
}

int oka_batch_results(oka_batch *B, int32_t s, int32_t *counts, double *coef,
                      oka_trace_stats *stats)
{
    if (!B || s < 0 || s >= B->nsites)
        return OKA_ERR_ARG;
    oka_ctx *c = B->c;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const TraceArgs &a = B->args[s];
    if (counts)
        HIPCHK(c, hipMemcpy(counts, a.counts, sizeof(int32_t) * B->nnodes[s],
                            hipMemcpyDeviceToHost));
    if (coef && a.nreal > 0)
        HIPCHK(c, hipMemcpy(coef, a.coef, sizeof(double) * 6 * a.nreal, hipMemcpyDeviceToHost));
    if (stats) {
        unsigned long long h[kStatCount];
        HIPCHK(c, hipMemcpy(h, a.stats, sizeof(h), hipMemcpyDeviceToHost));
        stats->attempted = (int64_t)h[kStatAttempted];
        stats->accepted = (int64_t)h[kStatAccepted];
        for (int j = 0; j < 4; ++j)
            stats->status_count[j] = (int64_t)h[kStatStatus0 + j];
        stats->escaped = (int64_t)h[kStatEscaped];
        stats->visited = (int64_t)h[kStatVisited];
        stats->candidates = (int64_t)h[kStatCandidates];
        stats->kernel_ms = 0.f;
    }
    if (a.nreal > 0) {
        std::vector<int32_t> st(a.nreal);
        HIPCHK(c, hipMemcpy(st.data(), B->fits[s].status, sizeof(int32_t) * a.nreal,
                            hipMemcpyDeviceToHost));
        for (int32_t v : st)
            if (v != 0)
                return fail(c, v, "site %d: a realization's fit failed", s);
    }
    return OKA_OK;
}

void oka_batch_destroy(oka_batch *B)
{
    if (!B)
        return;
    (void)hipSetDevice(B->c->device);
    (void)hipStreamSynchronize(B->c->stream);
    if (B->exec)
        (void)hipGraphExecDestroy(B->exec);
    for (void *p : B->allocs)
        (void)hipFree(p);
    delete B;
}

}  // extern "C"
