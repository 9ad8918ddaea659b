/*
 * oka_oracle.h -- CPU FP64 restatement of OnekaPy's Monte Carlo capture-zone
 * hot path.  TEST INFRASTRUCTURE ONLY: this library is the checker that the
 * HIP product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.
 *
 * Every function follows the reference operation by operation (same order of
 * floating-point operations, no FMA contraction: built with -ffp-contract=off).
 * Citations are path:line in the OnekaPy reference tree.
 *
 * Math modes (the only place where the oracle has a choice):
 *   OR_MATH_GLIBC : x**0.2 and log() are glibc's pow/log, which is what the
 *                   reference's numpy scalar arithmetic calls on this image.
 *   OR_MATH_CR    : x**0.2 and log() are correctly rounded (computed in
 *                   binary128 with libquadmath, then rounded once).  The GPU
 *                   product implements correctly-rounded versions of both, so
 *                   this mode is the bit-exact parity target for the GPU.
 * glibc log is correctly rounded on every sample tested; glibc pow(x, 0.2) is
 * misrounded by one ulp on about 0.1 % of inputs.
 */
#ifndef OKA_ORACLE_H
#define OKA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_MATH_GLIBC = 0, OR_MATH_CR = 1 };

/* Trace status codes (shared meaning with the product). */
enum {
    OR_TRACE_OK = 0,        /* reached |t| == |duration|                         */
    OR_TRACE_TRUNCATED = 1, /* an AquiferError-equivalent ended the trace         */
    OR_TRACE_STALLED = 2,   /* dt became NaN or 0: the reference loops forever    */
    OR_TRACE_STEP_CAP = 3   /* attempted-step cap reached: reference would go on  */
};

/* model.py:131-197 (Model).  wells is nwells x 4 = (xw, yw, rw, qw). */
typedef struct {
    double base, conductivity, porosity, thickness;
    double xo, yo;
    double coef[6];
    int32_t nwells;
    int32_t pad_;
    const double *wells;
} or_model;

/* model.py:240-266 */
double or_potential_wells(const or_model *m, double x, double y, int32_t mathmode);
/* model.py:207-237 */
double or_potential(const or_model *m, double x, double y, int32_t mathmode);
/* model.py:269-315; out = (Qx, Qy) */
void or_discharge(const or_model *m, double x, double y, double *out);
/* model.py:318-350; returns 0 or -1 (AquiferError) */
int32_t or_head(const or_model *m, double x, double y, int32_t mathmode, double *head);
/* model.py:353-389; returns 0 or -1 (AquiferError) */
int32_t or_velocity(const or_model *m, double x, double y, int32_t mathmode, double *out);
/* model.py:392-427 */
void or_velocity_confined(const or_model *m, double x, double y, double *out);

/* capturezone.py:127-253 (compute_backtrace) */
typedef struct {
    double duration, tol, maxstep;
    int32_t confined, mathmode;
    int64_t max_steps;      /* attempted-step cap; <= 0 means no cap */
} or_trace_opts;

typedef struct {
    int64_t attempted, accepted, nverts;
    double t_end;
    int32_t status, pad_;
} or_trace_result;

/* Writes up to cap vertices (x, y pairs) into verts; nverts reports the full
 * count even when it exceeds cap.  Returns the trace status. */
int32_t or_backtrace(const or_model *m, double xs, double ys, const or_trace_opts *o,
                     double *verts, int64_t cap, or_trace_result *res);

/* probabilityfield.py:53-427 (ProbabilityField) */
typedef struct {
    double deltax, deltay;
    double xmin, xmax, ymin, ymax;
    int64_t nrows, ncols;
    double *pgrid;      /* nrows x ncols, row = y index */
    uint8_t *rgrid;     /* nrows x ncols */
    double total_weight;
} or_field;

int32_t or_field_init(or_field *f, double dx, double dy, double xo, double yo);
void or_field_free(or_field *f);
int32_t or_field_expand(or_field *f, double xmin, double xmax, double ymin, double ymax);
void or_field_insert(or_field *f, double ax, double ay, double bx, double by, double umbra);
int32_t or_field_rasterize(or_field *f, const double *xy, int64_t n, double umbra);
void or_field_register(or_field *f, double weight);
void or_field_reset(or_field *f);
double or_distancesquared(double ax, double ay, double bx, double by, double cx, double cy);

/* capturezone.py:51-123 (compute_capturezone) with start points supplied by
 * the caller (numpy computes them, capturezone.py:113-115).  Accumulates the
 * per-trace step counts into stats[0..1] (attempted, accepted) and counts of
 * each status into stats[2..5]. */
int32_t or_capturezone(const or_model *m, int32_t npaths, const double *start_xy,
                       or_field *f, double umbra, double weight,
                       const or_trace_opts *o, int64_t *stats);

/* model.py:430-606 (fit_regional_flow) restated as column-equilibrated normal
 * equations + one refinement step; cov = inv(WA^T WA).  obs is nobs x 4.
 * Returns 0, -1 (RangeError: observation below base) or -2 (not positive
 * definite: the reference's LinAlgError). */
int32_t or_fit(const or_model *m, int32_t nobs, const double *obs, double xo, double yo,
               int32_t mathmode, double *ev, double *cov);

/* stochastic.py:240-241: coef = ev + L z with L the lower Cholesky factor of
 * cov.  Returns 0 or -2 (cov not positive definite). */
int32_t or_mvn(const double *ev, const double *cov, const double *z, double *coef);

/* Fixed-grid batch: realizations r in [0, nreal) are traced into a grid that
 * does not grow.  Per realization: wells (xw, yw) shared, q[r*nw..], K, n, b,
 * coef[r*6..].  counts (nrows x ncols int32) accumulates one per realization
 * per visited node.  A trace whose vertex bounding box is not strictly inside
 * the grid (the reference would expand) is counted in stats[6].  stats:
 * [0] attempted, [1] accepted, [2..5] status counts, [6] escaped traces,
 * [7] visited (realization, node) pairs.  OpenMP over realizations. */
typedef struct {
    double xmin, ymin, xmax, ymax, dx, dy;  /* xmax/ymax: the field's attributes */
    int64_t nrows, ncols;
} or_grid;

int32_t or_chronicle_fixed(int32_t nwells, const double *wells_xy, double xo, double yo,
                           double base, int32_t nreal, const double *K, const double *n,
                           const double *b, const double *q, const double *coef,
                           int32_t npaths, const double *start_xy, double umbra,
                           const or_trace_opts *o, const or_grid *g, int32_t *counts,
                           int64_t *stats, int32_t nthreads);

/* Number of threads OpenMP would use (for reporting). */
int32_t or_max_threads(void);

/* Building blocks exposed for tests. */
double or_pow02(double x, int32_t mathmode);
double or_log(double x, int32_t mathmode);

#ifdef __cplusplus
}
#endif
#endif
