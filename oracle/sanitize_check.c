/*
 * sanitize_check.c -- drives every entry point of the CPU oracle once on a
 * small synthetic site, for a build with AddressSanitizer and
 * UndefinedBehaviorSanitizer (make -C oracle sanitize).  TEST INFRASTRUCTURE
 * ONLY, like the oracle itself: tests/test_oracle_sanitize.py runs it.
 *
 * The site: a target well at the origin and three neighbours (one injecting),
 * a quadratic regional flow, eight observations whose heads come from the
 * oracle's own potential; paths through the fit, the MVN draw, single
 * backtraces (confined, unconfined, step-capped, a start on a well), a
 * growing field (expand, rasterize, register, reset), the fixed-grid batch in
 * both math modes and two threads, and the math building blocks at their
 * edge inputs.  Exit status 0 and "ok" on success; a sanitizer report aborts.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oka_oracle.h"

static int fails = 0;

static void check(int cond, const char *what)
{
    if (!cond) {
        fprintf(stderr, "check failed: %s\n", what);
        ++fails;
    }
}

int main(void)
{
    const double wells[4 * 4] = {0.0,    0.0,   0.25, 500.0,  /* target */
                                 350.0,  -120.0, 0.25, 300.0,
                                 -260.0, 410.0,  0.25, 450.0,
                                 500.0,  380.0,  0.25, -200.0}; /* injection */
    or_model m;
    memset(&m, 0, sizeof(m));
    m.base = 0.0;
    m.conductivity = 10.0;
    m.porosity = 0.25;
    m.thickness = 25.0;
    m.xo = 0.0;
    m.yo = 0.0;
    const double truth[6] = {-1e-5, -2e-5, 0.0, 1e-3, -5e-4, 1000.0};
    memcpy(m.coef, truth, sizeof(truth));
    m.nwells = 4;
    m.wells = wells;

    /* observations from the model's own potential, as heads */
    double obs[8 * 4];
    const double ox[8] = {-900.0, 640.0, -310.0, 820.0, 150.0, -720.0, 430.0, -40.0};
    const double oy[8] = {700.0, -560.0, -880.0, 390.0, 910.0, -150.0, 60.0, -420.0};
    for (int i = 0; i < 8; ++i) {
        const double x = ox[i], y = oy[i];
        double h = 0.0;
        check(or_head(&m, x, y, OR_MATH_CR, &h) == 0, "head of an observation");
        obs[4 * i] = x;
        obs[4 * i + 1] = y;
        obs[4 * i + 2] = h;
        obs[4 * i + 3] = 0.1;
    }
    double ev[6], cov[36], coef[6];
    for (int mode = 0; mode < 2; ++mode) {
        const int rc = or_fit(&m, 8, obs, 0.0, 0.0, mode, ev, cov);
        check(rc == 0, "fit");
        check(isfinite(ev[5]) && cov[35] > 0.0, "fit is finite");
    }
    const double z[6] = {0.3, -1.1, 0.7, 0.0, 2.0, -0.4};
    check(or_mvn(ev, cov, z, coef) == 0, "mvn");
    /* an observation below the base is a RangeError */
    double low[4] = {10.0, 10.0, -5.0, 0.1};
    double obs2[9 * 4];
    memcpy(obs2, obs, sizeof(obs));
    memcpy(obs2 + 32, low, sizeof(low));
    check(or_fit(&m, 9, obs2, 0.0, 0.0, OR_MATH_CR, ev, cov) == -1, "fit range error");

    /* single traces */
    or_trace_opts o = {3650.0, 1.0, 10.0, 1, OR_MATH_CR, 0};
    double verts[2 * 256];
    or_trace_result res;
    check(or_backtrace(&m, 20.0, 5.0, &o, verts, 256, &res) == OR_TRACE_OK, "confined trace");
    check(res.nverts > 2, "confined trace has vertices");
    o.confined = 0;
    o.mathmode = OR_MATH_GLIBC;
    (void)or_backtrace(&m, 20.0, 5.0, &o, verts, 256, &res);
    o.confined = 1;
    o.max_steps = 17;
    check(or_backtrace(&m, -30.0, 12.0, &o, verts, 8, &res) == OR_TRACE_STEP_CAP, "step cap");
    o.max_steps = 0;
    (void)or_backtrace(&m, 350.0, -120.0, &o, verts, 256, &res);   /* on a well: 0 / 0 */
    check(res.status == OR_TRACE_STALLED || res.status == OR_TRACE_OK ||
              res.status == OR_TRACE_TRUNCATED, "trace from a well ends");
    /* a dry aquifer: the potential below zero everywhere */
    or_model dry = m;
    dry.coef[5] = -1e9;
    double hv = 0.0, vel[2];
    check(or_head(&dry, 0.0, 0.0, OR_MATH_CR, &hv) != 0, "dry head raises");
    check(or_velocity(&dry, 10.0, 10.0, OR_MATH_CR, vel) != 0, "dry velocity raises");
    o.confined = 0;
    check(or_backtrace(&dry, 20.0, 5.0, &o, verts, 256, &res) == OR_TRACE_TRUNCATED,
          "dry trace is truncated");
    o.confined = 1;

    /* a growing field */
    or_field f;
    check(or_field_init(&f, 20.0, 20.0, 0.0, 0.0) == 0, "field init");
    double starts[2 * 12];
    for (int p = 0; p < 12; ++p) {
        starts[2 * p] = 0.25 * cos(2.0 * M_PI * p / 12.0);
        starts[2 * p + 1] = 0.25 * sin(2.0 * M_PI * p / 12.0);
    }
    int64_t stats[8] = {0};
    check(or_capturezone(&m, 12, starts, &f, 40.0, 1.0, &o, stats) == 0, "capture zone");
    check(f.nrows > 3 && f.ncols > 3 && f.total_weight == 1.0, "field grew and registered");
    check(or_field_expand(&f, -2000.0, 2000.0, -1500.0, 1500.0) == 0, "expand");
    const double poly[8] = {-100.0, -100.0, 0.0, 0.0, 0.0, 0.0, 400.0, 250.0};
    check(or_field_rasterize(&f, poly, 4, 30.0) == 0, "rasterize");
    or_field_register(&f, 0.5);
    or_field_reset(&f);
    (void)or_distancesquared(0.0, 0.0, 0.0, 0.0, 3.0, 4.0);   /* zero-length segment */
    or_field_free(&f);

    /* the fixed-grid batch: 3 realizations x 12 particles, both math modes */
    const double wxy[8] = {0.0, 0.0, 350.0, -120.0, -260.0, 410.0, 500.0, 380.0};
    const double K[3] = {8.0, 10.0, 14.0}, n[3] = {0.22, 0.25, 0.28}, b[3] = {21.0, 25.0, 29.0};
    double q[3 * 4], cf[3 * 6];
    for (int r = 0; r < 3; ++r) {
        for (int w = 0; w < 4; ++w)
            q[4 * r + w] = wells[4 * w + 3] * (0.9 + 0.1 * r);
        memcpy(cf + 6 * r, coef, sizeof(coef));
    }
    or_grid g = {-2000.0, -2000.0, 2000.0, 2000.0, 40.0, 40.0, 101, 101};
    int32_t *counts = calloc((size_t)g.nrows * g.ncols, sizeof(int32_t));
    for (int mode = 0; mode < 2; ++mode) {
        o.mathmode = mode;
        memset(stats, 0, sizeof(stats));
        check(or_chronicle_fixed(4, wxy, 0.0, 0.0, 0.0, 3, K, n, b, q, cf, 12, starts, 40.0, &o,
                                 &g, counts, stats, 2) == 0,
              "chronicle");
        check(stats[0] > 0 && stats[7] > 0, "chronicle traced and visited");
    }
    free(counts);

    /* building blocks at their edges */
    const double xs[] = {0.0, 5e-324, 2.2250738585072014e-308, 1.0, 0.5, 1e300, INFINITY, NAN};
    for (size_t i = 0; i < sizeof(xs) / sizeof(xs[0]); ++i)
        for (int mode = 0; mode < 2; ++mode) {
            (void)or_pow02(xs[i], mode);
            (void)or_log(xs[i], mode);
        }
    check(or_max_threads() >= 1, "threads");

    if (fails) {
        fprintf(stderr, "%d checks failed\n", fails);
        return 1;
    }
    printf("ok\n");
    return 0;
}
