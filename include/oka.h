/*
 * oka.h -- C ABI of liboka.so, the MI355X (gfx950) implementation of OnekaPy's
 * Monte Carlo capture-zone hot path.
 *
 * The reference has no FFI layer: its boundary is the Python API in
 * oneka/model.py, oneka/capturezone.py, oneka/probabilityfield.py and
 * oneka/stochastic.py.  Each entry point below replaces the reference function
 * cited beside it; the Python package onekapy_amd binds them with ctypes and
 * keeps the reference's Python signatures.
 *
 * Conventions
 *   - Every function returns OKA_OK (0) or a negative OKA_ERR_* code; the text
 *     of the last failure on a context is oka_last_error(ctx).
 *   - Array arguments are host pointers (C-contiguous, row-major) unless the
 *     call's flags include OKA_PTR_DEVICE, in which case every array argument
 *     of that call is a device pointer on the context's device.
 *   - Calls are synchronous unless flags include OKA_ASYNC (only honoured with
 *     OKA_PTR_DEVICE): then the work is enqueued on the context's stream.
 *   - A context is used by one host thread at a time.
 */
#ifndef OKA_H
#define OKA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OKA_ABI_VERSION 5

enum {
    OKA_OK = 0,
    OKA_ERR_ARG = -1,         /* invalid argument                                  */
    OKA_ERR_RANGE = -2,       /* RangeError (model.py:51-53, probabilityfield.py:47-49) */
    OKA_ERR_AQUIFER = -3,     /* AquiferError (model.py:56-58)                     */
    OKA_ERR_LINALG = -4,      /* numpy.linalg.LinAlgError (model.py:598-604)        */
    OKA_ERR_DEVICE = -5,      /* HIP runtime failure                               */
    OKA_ERR_NOMEM = -6,       /* allocation failure                                */
    OKA_ERR_CAPACITY = -7     /* an output buffer is too small (size reported)     */
};

/* Per-trace status.  The reference either finishes a trace, or truncates it
 * when feval raises (capturezone.py:249-253).  STALLED and STEP_CAP are
 * deliberate deviations: the reference loops forever when dt becomes NaN or 0
 * (capturezone.py:221,247), and has no step cap. */
enum {
    OKA_TRACE_OK = 0,
    OKA_TRACE_TRUNCATED = 1,
    OKA_TRACE_STALLED = 2,
    OKA_TRACE_STEP_CAP = 3
};

/* Arithmetic of the particle tracer. */
enum {
    OKA_FP64_EXACT = 0,  /* reference operation order, no contraction, correctly rounded pow/log */
    OKA_FP32 = 1         /* FP32 tracing in target-relative coordinates, compensated positions */
};

/* What oka_model_eval computes at each point. */
enum {
    OKA_EVAL_POTENTIAL = 0,          /* model.py:207-237        -> 1 double  */
    OKA_EVAL_POTENTIAL_WELLS = 1,    /* model.py:240-266        -> 1 double  */
    OKA_EVAL_DISCHARGE = 2,          /* model.py:269-315        -> 2 doubles */
    OKA_EVAL_HEAD = 3,               /* model.py:318-350        -> 1 double  */
    OKA_EVAL_VELOCITY = 4,           /* model.py:353-389        -> 2 doubles */
    OKA_EVAL_VELOCITY_CONFINED = 5   /* model.py:392-427        -> 2 doubles */
};

#define OKA_PTR_DEVICE 0x1u
#define OKA_ASYNC 0x2u
#define OKA_RASTER_FP32 0x4u   /* oka_rasterize: the FP32 tracers' rasterizer */
#define OKA_ANTITHETIC 0x8u    /* oka_draw_realizations, oka_cv_surrogate_sums: antithetic pairs */
#define OKA_QMC 0x10u          /* oka_draw_realizations, oka_cv_surrogate_sums: scrambled Sobol' points */

typedef struct oka_ctx oka_ctx;

/* ---- context ------------------------------------------------------------ */
int oka_abi_version(void);

/* What the context's device offers the launch planners. */
typedef struct {
    int32_t num_cu;              /* compute units                                   */
    int32_t warp_size;           /* lanes per wave                                  */
    int64_t lds_per_workgroup;   /* largest LDS allocation of one workgroup (bytes)  */
    int64_t lds_per_cu;          /* LDS shared by the workgroups of one CU (bytes)   */
} oka_device_info;
int oka_ctx_device_info(oka_ctx *ctx, oka_device_info *out);
int oka_ctx_create(oka_ctx **out, int device);
void oka_ctx_destroy(oka_ctx *ctx);
const char *oka_last_error(const oka_ctx *ctx);
/* Use the caller's HIP stream (a hipStream_t), or NULL for the library's own. */
int oka_ctx_set_stream(oka_ctx *ctx, void *hip_stream);
void *oka_ctx_stream(oka_ctx *ctx);
int oka_ctx_synchronize(oka_ctx *ctx);

/* ---- one model: Model.compute_* (model.py:207-427) ----------------------- */
typedef struct {
    double base, conductivity, porosity, thickness;
    double xo, yo;
    double coef[6];
    int32_t nwells;
    int32_t reserved;
    const double *wells;     /* nwells x 4: xw, yw, rw, qw (always host memory) */
} oka_model;

/* Evaluates `kind` at npts points xy (npts x 2).  out is npts x 1 or npts x 2.
 * status (npts int32, may be NULL) receives 0 or OKA_ERR_AQUIFER per point;
 * out holds NaN where status is nonzero.  FP64, reference operation order. */
int oka_model_eval(oka_ctx *ctx, const oka_model *m, int32_t kind, int64_t npts,
                   const double *xy, double *out, int32_t *status, uint32_t flags);

/* ---- a site: geometry shared by every realization ------------------------ */
typedef struct {
    int32_t nwells;          /* wells, including the target well                */
    int32_t nobs;            /* observations                                     */
    const double *wells_xy;  /* nwells x 2 (host, or device with OKA_PTR_DEVICE) */
    const double *obs;       /* nobs x 4: x, y, z_ev, z_std                      */
    double xo, yo;           /* origin of the regional flow = the target well    */
    double base;             /* aquifer base elevation                           */
} oka_site;

/* ---- fit + draw: Model.fit_regional_flow (model.py:430-606) and the MVN
 *      draw of stochastic.py:239-241, for nreal realizations at once ----------
 * For realization r: ev[r] (6), cov[r] (6x6) = inv(WA^T WA) from the weighted
 * least squares of construct_fit with conductivity K[r], thickness b[r] and
 * discharges q[r, :] (nreal x nwells); then coef[r] = ev[r] + L z[r] with L the
 * lower Cholesky factor of cov[r] (z: nreal x 6 standard normals; NULL gives
 * coef = ev, the deterministic path of deterministic.py:197-199).  ev, cov may
 * be NULL.  status[r] (may be NULL) is 0, OKA_ERR_RANGE (an observation at or
 * below the base) or OKA_ERR_LINALG; the call returns the first nonzero one.
 * FP64: column-equilibrated normal equations, Cholesky, one refinement step. */
int oka_fit_draw(oka_ctx *ctx, const oka_site *site, int32_t nreal, const double *K,
                 const double *b, const double *q, const double *z, double *ev, double *cov,
                 double *coef, int32_t *status, uint32_t flags);

/* ---- trace options: compute_capturezone / compute_backtrace arguments ---- */
typedef struct {
    int32_t npaths;          /* particles per realization (capturezone.py:113)      */
    int32_t precision;       /* OKA_FP64_EXACT or OKA_FP32                           */
    int32_t confined;        /* 1: compute_velocity_confined, 0: compute_velocity    */
    int32_t schedule;        /* OKA_SCHED_*                                          */
    double duration, tol, maxstep, umbra;
    int64_t max_steps;       /* attempted-step cap per trace; <= 0: 1,000,000        */
    const double *start_xy;  /* npaths x 2 start points (capturezone.py:114-115)     */
    int32_t bitmap_mode;     /* visited-node bitmap of a realization: OKA_BITMAP_*   */
    int32_t tile_slots;      /* OKA_BITMAP_TILED: LDS tile slots (<= 0: as many as fit) */
    int32_t pool_slots;      /* OKA_SCHED_POOL: realizations in flight per wave (<= 0: 2) */
    int32_t pool_waves_per_cu; /* OKA_SCHED_POOL: resident waves per CU (<= 0: the most
                                  of 16, 12, 8, 6, 4 whose LDS share fits the slots) */
    int32_t pool_chunk;      /* OKA_SCHED_POOL: particles per work item (<= 0: 64)   */
    int32_t far_order;       /* OKA_SCHED_POOL, FP32, confined: wells beyond the 3 x 3
                                boxes around a point summed as a local expansion of
                                this order (1..16); 0: order 16 from 64 wells on,
                                where the plan pays; < 0: every well summed directly */
} oka_trace_opts;

/* How particles are scheduled on the GPU.  WORKGROUP: one 256-lane workgroup
 * per realization, its lanes refilled with the realization's particles.
 * POOL (no window marks): a persistent grid of single waves
 * working through a queue shared by all waves of chunks of pool_chunk
 * particles; each wave keeps pool_slots chunks in flight and refills a lane
 * with a particle of any of them, so uneven particle lifetimes and uneven
 * realizations (even of different sites) balance across the GPU.  The chunks
 * of one realization meet in an HBM record flushed by the last of them.
 * AUTO is POOL where the far-well expansions apply (FP32, confined, a plan
 * taken for far_order, no window marks, bitmap_mode and tile_slots left
 * automatic, and a pool that fits: its LDS holds the site's wells and 8
 * tiles per slot at pool_waves_per_cu, or 4 when that is 0), else WORKGROUP.
 * Both schedules give the same counts when every well is summed directly. */
enum { OKA_SCHED_AUTO = 0, OKA_SCHED_WORKGROUP = 1, OKA_SCHED_POOL = 2 };

/* Where a realization's visited-node bitmap lives.  AUTO picks LDS_FLAT when
 * the whole bitmap fits in LDS, else TILED, else HBM_FLAT.  TILED keeps
 * 32 x 32-node tiles in LDS slots claimed on first touch, with the tiles that
 * find no free slot in HBM; the flush visits only the claimed tiles.  Every
 * mode gives the same counts. */
enum {
    OKA_BITMAP_AUTO = 0,
    OKA_BITMAP_LDS_FLAT = 1,
    OKA_BITMAP_TILED = 2,
    OKA_BITMAP_HBM_FLAT = 3
};

/* A node lattice: node (i, j) is at (xmin + j*dx, ymin + i*dy); row i is y.
 * xmax/ymax are the field's own attributes (probabilityfield.py:140-143,
 * 212-213, 229-245), used for the "strictly inside" test. */
typedef struct {
    double xmin, ymin, xmax, ymax, dx, dy;
    int32_t nrows, ncols;
} oka_grid;

/* The field's state at one moment of the expansion history: traces with global
 * index >= first_trace (index r*npaths + p) and below the next epoch's
 * first_trace were rasterized into a grid with this origin and size
 * (probabilityfield.py:334-335).  Trace-time node (i, j) is final node
 * (i + row_off, j + col_off). */
typedef struct {
    int64_t first_trace;
    double xmin, ymin;
    int32_t nrows, ncols;
    int32_t row_off, col_off;
} oka_epoch;

typedef struct {
    int64_t attempted;        /* attempted Dormand-Prince steps (all traces)         */
    int64_t accepted;         /* accepted steps = segments                           */
    int64_t status_count[4];  /* traces per OKA_TRACE_* status                       */
    int64_t escaped;          /* traces whose vertex bbox is not strictly inside     */
    int64_t visited;          /* (realization, node) marks added to counts           */
    int64_t candidates;       /* nodes in the clipped candidate windows of segments  */
    double kernel_ms;         /* device time of the tracing kernel(s)                */
} oka_trace_stats;

/* ---- trace + chronicle: compute_capturezone (capturezone.py:51-123) plus
 *      rasterize/insert/register (probabilityfield.py:264-359) for nreal
 *      realizations of one site ---------------------------------------------
 * Realization r has conductivity K[r], porosity n[r], thickness b[r],
 * discharges q[r, :] and regional coefficients coef[r, :] (A..F).  Each of its
 * npaths traces is rasterized with the vector-to-raster range umbra; every node
 * marked by any trace of r adds 1 to counts (nrows x ncols int32 of `grid`,
 * accumulated).  With nepochs == 0 the grid does not grow and the clip window
 * is the whole grid (a pre-expanded ProbabilityField); traces that would have
 * grown it are counted in stats->escaped.  With nepochs > 0, epochs (sorted by
 * first_trace, from oka_grid_history) reproduce the reference's expansion
 * history exactly.  bbox (ntraces x 4: xmin, xmax, ymin, ymax), trace_status
 * and trace_steps (attempted steps) are optional per-trace outputs, ntraces =
 * nreal * npaths.  With rasterize == 0 only the traces and the optional
 * outputs are computed. */
int oka_trace_chronicle(oka_ctx *ctx, const oka_site *site, int32_t nreal, const double *K,
                        const double *n, const double *b, const double *q,
                        const double *coef, const oka_trace_opts *opts, const oka_grid *grid,
                        int32_t nepochs, const oka_epoch *epochs, int32_t rasterize,
                        int32_t *counts, double *bbox, int32_t *trace_status,
                        int64_t *trace_steps, oka_trace_stats *stats, uint32_t flags);

/* ---- trace + chronicle without allocating ---------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc -- this is oka_trace_chronicle with the same arguments and
 * results, whose scratch memory (statistics accumulators, the whole-grid
 * epoch, and the HBM bitmaps of grids too large for LDS) is a device array of
 * OKA_NOALLOC_SCRATCH_BYTES inside the library, which each GPU holds from the
 * library's first context on it until the process ends.  It calls no
 * allocator.  Restrictions:
 *   - flags must include OKA_PTR_DEVICE: every array argument (site->wells_xy,
 *     opts->start_xy and epochs included) is device memory;
 *   - the schedule is the workgroup one (OKA_SCHED_POOL is OKA_ERR_ARG);
 *   - HBM bitmaps (tiled grids' overflow tiles: ntiles x 128 bytes, flat HBM
 *     bitmaps: nrows x ncols / 8 bytes, per workgroup in flight) share the
 *     scratch, so large grids run fewer workgroups at once, and a grid whose
 *     bitmap exceeds the scratch alone is OKA_ERR_CAPACITY.
 * Every call on one GPU uses that GPU's one scratch array, which keeps the
 * last call's contents (each call clears what must start at zero): calls that
 * overlap on one GPU (several contexts, streams or host threads) race and
 * must be serialized by the caller.  With stats (host memory) the call is
 * synchronous; without it, OKA_ASYNC leaves it enqueued on the context's
 * stream. */
#define OKA_NOALLOC_SCRATCH_BYTES (64u << 20)
int oka_trace_chronicle_noalloc(oka_ctx *ctx, const oka_site *site, int32_t nreal,
                                const double *K, const double *n, const double *b,
                                const double *q, const double *coef,
                                const oka_trace_opts *opts, const oka_grid *grid,
                                int32_t nepochs, const oka_epoch *epochs, int32_t rasterize,
                                int32_t *counts, double *bbox, int32_t *trace_status,
                                int64_t *trace_steps, oka_trace_stats *stats, uint32_t flags);

/* ---- per-realization marks pooled over a window: the training targets and
 *      the correction term of the learned control variate --------------------
 * No reference counterpart: the reference registers each realization into
 * pgrid (probabilityfield.py:342-359) and keeps nothing per realization.
 * The window is nrows x ncols coarse cells of cell x cell fine nodes, whose
 * fine node (0, 0) is fine node (row0, col0) of `grid` (cells may lie partly
 * outside the grid).  For realization r, marks[r * nrows * ncols + c] is the
 * number of fine nodes of cell c that the realization visited (capped at 255).
 * node_masks (may be NULL; cells of at most 4 x 4 nodes) receives which of
 * them: bit i * cell + j of node_masks[r * nrows * ncols + c] is fine node
 * (i, j) of cell c.  Traces run on the fixed grid as in oka_trace_chronicle
 * with nepochs == 0; counts (may be NULL) accumulates as there. */
typedef struct {
    int32_t row0, col0;      /* fine-grid index of the window's first node     */
    int32_t cell;            /* fine nodes per cell side (1..15)               */
    int32_t nrows, ncols;    /* cells                                          */
} oka_window;

int oka_trace_marks(oka_ctx *ctx, const oka_site *site, int32_t nreal, const double *K,
                    const double *n, const double *b, const double *q, const double *coef,
                    const oka_trace_opts *opts, const oka_grid *grid, const oka_window *win,
                    uint8_t *marks, uint16_t *node_masks, int32_t *counts,
                    oka_trace_stats *stats, uint32_t flags);

/* ---- batches: many sites in one persistent, graph-captured run ------------
 * (the realization loop of stochastic.py:187-207 for many independent sites
 * at once; no reference counterpart for the batching itself)
 * oka_batch_create stages every site on the device and captures one run as a
 * HIP graph: per site the fit and MVN draw of every realization (as
 * oka_fit_draw), then every particle of every site in one launch of the pool
 * tracer (OKA_SCHED_POOL: FP32, confined velocity), each site rasterized into
 * its own fixed lattice.  All site arrays are host memory, read at create. */
typedef struct {
    const oka_site *site;    /* wells, origin (the target), observations, base    */
    int32_t nreal;
    const double *K, *n, *b; /* nreal                                              */
    const double *q;         /* nreal x nwells                                     */
    const double *z;         /* nreal x 6 standard normals, or NULL: coef = ev     */
    oka_trace_opts opts;     /* npaths, duration, tol, maxstep, umbra, start_xy    */
    oka_grid grid;           /* the site's fixed lattice                          */
} oka_batch_site;

typedef struct oka_batch oka_batch;

/* tune: the pool knobs (pool_slots, pool_waves_per_cu, pool_chunk) or NULL. */
int oka_batch_create(oka_ctx *ctx, int32_t nsites, const oka_batch_site *sites,
                     const oka_trace_opts *tune, oka_batch **out);
/* Re-captures the run with other pool knobs; the staged sites are kept. */
int oka_batch_tune(oka_batch *batch, const oka_trace_opts *tune);
/* Replays the captured run (counts start from zero); *ms (may be NULL)
 * receives its device time. */
int oka_batch_run(oka_batch *batch, float *ms);
/* Site s after the last run: counts (nrows x ncols), coef (nreal x 6) and
 * stats, each may be NULL.  Returns OKA_ERR_RANGE or OKA_ERR_LINALG (after
 * copying) when a realization's fit failed, as oka_fit_draw would. */
int oka_batch_results(oka_batch *batch, int32_t s, int32_t *counts, double *coef,
                      oka_trace_stats *stats);
void oka_batch_destroy(oka_batch *batch);

/* ---- expansion history (probabilityfield.py:175-261 applied per trace in
 *      order, capturezone.py:117-120) -- host computation ---------------------
 * Given the field before the batch (init; init->nrows == 0 means an empty
 * field) and each trace's bbox in order, computes the field after the batch
 * (final) and the epochs.  epochs may be NULL to query *nepochs only; returns
 * OKA_ERR_CAPACITY if max_epochs is too small. */
int oka_grid_history(const oka_grid *init, int64_t ntraces, const double *bbox,
                     oka_grid *final_grid, oka_epoch *epochs, int64_t max_epochs,
                     int64_t *nepochs);

/* ---- vertices of individual traces: compute_backtrace (capturezone.py:127-253)
 * For each realization r and particle p, the accepted vertices (start point
 * first) go to verts[(r*npaths + p) * cap * 2 ...]; nverts receives the full
 * count (which may exceed cap: then only the first cap are stored). */
int oka_trace_vertices(oka_ctx *ctx, const oka_site *site, int32_t nreal, const double *K,
                       const double *n, const double *b, const double *q, const double *coef,
                       const oka_trace_opts *opts, int64_t cap, double *verts,
                       int64_t *nverts, int32_t *trace_status, int64_t *trace_steps,
                       uint32_t flags);

/* ---- one polyline into a registration grid: ProbabilityField.rasterize's
 *      insert loop (probabilityfield.py:264-339) on a fixed grid -------------
 * Marks rgrid (nrows x ncols uint8, nonzero = visited) for every node strictly
 * within umbra of any segment of xy (nverts x 2).  No expansion is done.
 * With OKA_RASTER_FP32 in flags the segments go through the FP32 tracers'
 * rasterizer (float coordinates relative to the grid's centre node), which
 * may differ from the FP64-exact one for nodes within rounding of the umbra
 * circle; otherwise the FP64-exact rasterizer is used. */
int oka_rasterize(oka_ctx *ctx, const oka_grid *grid, int64_t nverts, const double *xy,
                  double umbra, uint8_t *rgrid, uint32_t flags);

/* ---- learned control variate: the surrogate MLP --------------------------
 * No reference counterpart (the reference traces every realization).  A
 * 3-layer perceptron n_in -> n_hidden -> n_hidden -> n_out with GELU after the
 * hidden layers maps a realization's standardized parameters to per-cell
 * capture logits.  Weights are FP32 masters with BF16 copies; every matrix
 * product runs in hand-written BF16 MFMA kernels with FP32 accumulation.  The
 * network lives in device memory owned by the library.
 *
 * Parameter vector (FP32, this order): W1 [n_hidden x n_in], b1 [n_hidden],
 * W2 [n_hidden x n_hidden], b2, W3 [n_out x n_hidden], b3 [n_out].
 * Inputs x are FP32 [n x n_in] row-major; targets are window marks (uint8,
 * oka_trace_marks) times target_scale, i.e. fractions in [0, 1].
 * Requirements: n_in <= 64, n_hidden % 64 == 0, n_out % 64 == 0.  All array
 * arguments follow the call's flags (OKA_PTR_DEVICE or host). */
typedef struct oka_mlp oka_mlp;

int oka_mlp_create(oka_ctx *ctx, int32_t n_in, int32_t n_hidden, int32_t n_out,
                   int32_t max_batch, oka_mlp **out);
void oka_mlp_destroy(oka_mlp *mlp);
int64_t oka_mlp_param_count(const oka_mlp *mlp);
int oka_mlp_set_params(oka_mlp *mlp, const float *params, uint32_t flags);
int oka_mlp_get_params(oka_mlp *mlp, float *params, uint32_t flags);
/* Gradient of the mean binary cross-entropy (logits vs targets, averaged over
 * batch x n_out) with respect to every parameter, for one batch. */
int oka_mlp_grad(oka_mlp *mlp, int32_t batch, const float *x, const uint8_t *marks,
                 float target_scale, float *grad, float *loss, uint32_t flags);
/* nsteps Adam steps; step s uses samples order[s*batch .. (s+1)*batch).
 * losses (nsteps, may be NULL) receives each step's mean loss. */
int oka_mlp_train(oka_mlp *mlp, int64_t nsamples, const float *x, const uint8_t *marks,
                  float target_scale, const int32_t *order, int32_t batch, int32_t nsteps,
                  float lr, float beta1, float beta2, float eps, float *losses, uint32_t flags);
/* probs [n x n_out] = sigmoid(logits) (FP32). */
int oka_mlp_predict(oka_mlp *mlp, int64_t n, const float *x, float *probs, uint32_t flags);
/* Per output c, over the n samples, with s = sigmoid(logit_c) and, when marks
 * is not NULL, d = marks * target_scale - s (FP64 accumulators, added to):
 *   sums[c] += sum s,  sums_sq[c] += sum s^2,  diffs[c] += sum d,
 *   diffs_sq[c] += sum d^2
 * -- the terms of the control-variate estimator and of its variance.  Only
 * sums is required; the others may be NULL. */
int oka_mlp_accumulate(oka_mlp *mlp, int64_t n, const float *x, const uint8_t *marks,
                       float target_scale, double *sums, double *sums_sq, double *diffs,
                       double *diffs_sq, uint32_t flags);

/* ---- realization draws on the device --------------------------------------
 * The distribution vocabulary of generate_random_variate (stochastic.py:274-306:
 * constant, uniform (lo, hi), triangular (lo, mode, hi)) plus lognormal
 * (mu, sigma of the underlying normal), each optionally sign-flipped with
 * probability p_negative (injection wells).  Draws come from Philox4x32-10
 * (Salmon et al. 2011) keyed by seed, with counter (realization, slot), so any
 * realization's draws are reproducible on their own. */
enum { OKA_DIST_CONST = 0, OKA_DIST_UNIFORM = 1, OKA_DIST_TRIANGULAR = 2, OKA_DIST_LOGNORMAL = 3 };

typedef struct {
    int32_t kind;
    int32_t reserved;
    double a, b, c;          /* const: a; uniform: a, b; triangular: a, b (mode), c; lognormal: a (mu), b (sigma) */
    double p_negative;
} oka_dist;

/* Realizations first .. first + n - 1 of the stream `seed`: q (n x nwells),
 * K, por (porosity), thick (n) and z (n x 6 standard normals).  dists holds
 * nwells discharge distributions then conductivity, porosity, thickness.
 * With OKA_ANTITHETIC in flags (n and first even), realization 2k + 1 is the
 * antithetic partner of 2k: it takes 2k's Philox uniforms u as 1 - u (the
 * inverse-CDF and sign variates) and 2k's standard normals (z and the
 * lognormal's) negated; realization 2k is drawn as without the flag.
 * With OKA_QMC instead (first + n <= 2^32), realization j takes its uniforms
 * from point j of a Sobol' sequence (natural order; direction numbers from
 * the primitive polynomials over GF(2) by degree) under a nested (Owen-type)
 * hash scramble keyed by seed: each realization is distributed as without
 * the flag, and runs with different seeds are independent.  Dimensions,
 * most influential first: K, porosity, thickness, the 6 normals, the
 * discharges, then the lognormals' second uniforms and the sign uniforms (up
 * to 1024; constants take none). */
int oka_draw_realizations(oka_ctx *ctx, int32_t nwells, const oka_dist *dists, int64_t n,
                          uint64_t seed, int64_t first, double *q, double *K, double *por,
                          double *thick, double *z, uint32_t flags);

/* The correction term of the control variate at node resolution.  Traces the
 * nreal correction realizations (as oka_trace_marks, window `win` of the
 * network's n_out cells of cell x cell nodes, cell <= 4; counts, may be NULL,
 * accumulates every node of the grid as in oka_trace_chronicle), evaluates
 * the network on their standardized inputs x [nreal x n_in], and reduces, for
 * nparts contiguous parts of the realizations (part p: rows
 * [p nreal / nparts, (p + 1) nreal / nparts)), with s_a the network's output
 * for cell a (the window's cells row-major, nrows x ncols) and y the visit of
 * node k (= i * cell + j, fine node (i, j)) of cell a:
 *   sum_s[p][a] = sum s_a,
 *   sum_ss[p][a][h] = sum s_a s_b, b = a + offset h (0 if b is outside the
 *     window), for the 13 offsets (dr, dc) = (0, 0..2), (1, -2..2), (2, -2..2),
 *   sum_y[p][a][k] = sum y,
 *   sum_ys[p][a][k][m] = sum y s_b, b = cell (r + m / 3 - 1, c + m % 3 - 1) of
 *     cell a = (r, c) (0 if b is outside the window), m = 0..8
 * (FP64, in a fixed order: the same inputs give the same sums) -- the terms
 * of a per-node linear control on the 3 x 3 cells around the node's cell.
 * The marks and node masks never leave the device.  Arrays follow flags;
 * the four sums are host memory and are overwritten. */
int oka_cv_node_sums(oka_ctx *ctx, oka_mlp *mlp, const oka_site *site, int32_t nreal,
                     const double *K, const double *n, const double *b, const double *q,
                     const double *coef, const float *x, const oka_trace_opts *opts,
                     const oka_grid *grid, const oka_window *win, int32_t nparts, int32_t *counts,
                     double *sum_s, double *sum_ss, double *sum_y, double *sum_ys,
                     oka_trace_stats *stats, uint32_t flags);

/* The same sums in two calls, so that the correction set can be traced
 * while the network trains (on another context): oka_cv_node_trace traces
 * nreal realizations as oka_cv_node_sums does and keeps their node masks in
 * device memory owned by *out (nreal x cells x 2 bytes; cells of at most
 * 4 x 4 nodes); oka_cv_node_reduce runs a network on the set's standardized
 * inputs x [nreal x n_in] and returns the sums of oka_cv_node_sums over
 * nparts parts (any context of the set's device; the set must not be in use
 * by another call).  oka_node_set_destroy frees the set. */
typedef struct oka_node_set oka_node_set;
int oka_cv_node_trace(oka_ctx *ctx, const oka_site *site, int32_t nreal, const double *K,
                      const double *n, const double *b, const double *q, const double *coef,
                      const oka_trace_opts *opts, const oka_grid *grid, const oka_window *win,
                      int32_t *counts, oka_node_set **out, oka_trace_stats *stats,
                      uint32_t flags);
int oka_cv_node_reduce(oka_mlp *mlp, const oka_node_set *set, const float *x, int32_t nparts,
                       double *sum_s, double *sum_ss, double *sum_y, double *sum_ys,
                       uint32_t flags);
void oka_node_set_destroy(oka_node_set *set);

/* Exact moments of window marks [n x ncell] (uint8): per cell c,
 * sums[c] += sum over rows of marks[r, c] and, when sums_sq is set,
 * sums_sq[c] += sum of marks[r, c]^2 -- the plain Monte Carlo estimate of the
 * cells' capture probability and its variance.  marks, sums and sums_sq
 * follow flags. */
int oka_mark_moments(oka_ctx *ctx, int64_t n, int32_t ncell, const uint8_t *marks, uint64_t *sums,
                     uint64_t *sums_sq, uint32_t flags);

/* Histogram of a count grid (probabilityfield.py pgrid with weight 1:
 * realizations that marked each node), for the impact curve and deciles of
 * visualize.py:382-385 and oneka.py:279-288: hist[k] += number of the n
 * counts equal to k, 0 <= k <= max_count.  Counts outside [0, max_count] are
 * an error (OKA_ERR_RANGE).  counts and hist follow flags. */
int oka_count_histogram(oka_ctx *ctx, int64_t n, const int32_t *counts, int32_t max_count,
                        int64_t *hist, uint32_t flags);

/* ---- the probability field after the run -----------------------------------
 * Pr = pgrid / total_weight on an nrows x ncols node lattice
 * (visualize.py:230), and when radius > 0 its Gaussian smoothing as
 * scipy.ndimage.gaussian_filter(Pr, sigma, mode='constant', cval=0.0) computes
 * it for the probability map (visualize.py:232-233): a correlation along axis
 * 0 then axis 1 with the 2 radius + 1 symmetric weights (scipy's normalized
 * kernel, radius = int(4 sigma + 0.5), computed by the caller), zeros outside
 * the lattice, in scipy's summation order, FP64.  out is nrows x ncols.
 * pgrid and out follow flags; weights are host memory. */
int oka_field_smooth(oka_ctx *ctx, int32_t nrows, int32_t ncols, const double *pgrid,
                     double total_weight, int32_t radius, const double *weights, double *out,
                     uint32_t flags);

/* The impact curve's probabilities (visualize.py:382): pr = the n values of
 * pgrid / total_weight sorted in descending order (a radix sort on the device;
 * for fields of whole counts oka_count_histogram gives the same array).
 * pgrid and pr follow flags. */
int oka_field_sorted_probabilities(oka_ctx *ctx, int64_t n, const double *pgrid,
                                   double total_weight, double *pr, uint32_t flags);

/* The surrogate term of the control variate for n fresh realizations of the
 * stream `seed` (starting at realization `first`), computed entirely on the
 * device in chunks: draws, fit + MVN draw (as oka_fit_draw), standardized
 * features (q, K, porosity, thickness, A..F) -> (f - mean) / std, network, and
 * per output c: sums[c] += sum of s_c, sums_sq[c] += sum of s_c^2.  site and
 * dists are host memory; feat_mean / feat_std (n_in doubles) too; sums and
 * sums_sq follow flags.  With OKA_ANTITHETIC (n and first even) the
 * realizations are drawn in antithetic pairs as in oka_draw_realizations,
 * and sums_sq[c] += sum over the n / 2 pairs of ((s_c,2k + s_c,2k+1) / 2)^2,
 * so that the variance of the sums' mean is that of the pair means / (n / 2).
 * With OKA_QMC the realizations are Sobol' points as in oka_draw_realizations;
 * the variance of the mean then comes from independent scrambles (seeds). */
int oka_cv_surrogate_sums(oka_ctx *ctx, oka_mlp *mlp, const oka_site *site, const oka_dist *dists,
                          int64_t n, uint64_t seed, int64_t first, const double *feat_mean,
                          const double *feat_std, double *sums, double *sums_sq, uint32_t flags);

/* nrep independent randomized-QMC replicates of the surrogate term in one
 * call: replicate k is oka_cv_surrogate_sums with OKA_QMC over Sobol' points
 * 0 .. npts - 1 under the scramble of seeds[k] (host memory), and adds its
 * column sums to row k of sums (nrep x n_out, follows flags).  The site is
 * staged once and the passes are enqueued back to back on the context's
 * stream with one synchronization; the replicate means and their sample
 * variance give the surrogate mean and its variance.  OKA_ANTITHETIC is an
 * error here. */
int oka_cv_surrogate_sums_qmc(oka_ctx *ctx, oka_mlp *mlp, const oka_site *site,
                              const oka_dist *dists, int64_t npts, int32_t nrep,
                              const uint64_t *seeds, const double *feat_mean,
                              const double *feat_std, double *sums, uint32_t flags);

/* ---- exact scalar helpers shared by host and device code (for tests) ----- */
double oka_pow_fifth(double x);   /* correctly rounded pow(x, 0.2) (capturezone.py:247) */
double oka_log(double x);         /* correctly rounded log(x) (model.py:264)            */

#ifdef __cplusplus
}
#endif
#endif
