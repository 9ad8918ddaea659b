// oka_fp64_dev.h -- FP64-exact tracing device code shared by the trace
// kernels (the workgroup tracer in oka_fp64.hip, the wave pool in
// oka_pool64.hip).
//
// Every floating-point expression is evaluated in the reference's operation
// order with contraction disabled, so that for identical inputs the results
// are bit-identical to the reference's numpy scalar arithmetic:
//   velocity           oneka/model.py:299-315, 343-350, 377-389, 423-427
//   feval              oneka/stochastic.py:253-260
//   Dormand-Prince     oneka/capturezone.py:198-253
//   segment raster     oneka/probabilityfield.py:292-310, 407-427
// Translation units that include it are compiled with -ffp-contract=off.
#pragma once

#include <math.h>

#include "oka_internal.h"
#include "oka_math.h"

#pragma clang fp contract(off)

namespace oka {
namespace fp64 {

// capturezone.py:202-209, each entry the correctly rounded quotient.
constexpr double A21 = 1.0 / 5.0;
constexpr double A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
constexpr double A41 = 44.0 / 45.0, A42 = -56.0 / 15.0, A43 = 32.0 / 9.0;
constexpr double A51 = 19372.0 / 6561.0, A52 = -25360.0 / 2187.0, A53 = 64448.0 / 6561.0,
                 A54 = -212.0 / 729.0;
constexpr double A61 = 9017.0 / 3168.0, A62 = -355.0 / 33.0, A63 = 46732.0 / 5247.0,
                 A64 = 49.0 / 176.0, A65 = -5103.0 / 18656.0;
constexpr double B1 = 35.0 / 384.0, B3 = 500.0 / 1113.0, B4 = 125.0 / 192.0,
                 B5 = -2187.0 / 6784.0, B6 = 11.0 / 84.0;
constexpr double E1 = 71.0 / 57600.0, E2 = -1.0 / 40.0, E3 = -71.0 / 16695.0,
                 E4 = 71.0 / 1920.0, E5 = -17253.0 / 339200.0, E6 = 22.0 / 525.0;

constexpr double kC2Pi = 0.15915494309189535;  // model.py:312-313
constexpr double kC4Pi = 0.07957747154594767;  // model.py:264
constexpr double kEps = 2.220446049250313e-16; // capturezone.py:199

// One well of a realization's field: position and -q (16-byte aligned, so
// that x and y load as one 128-bit LDS read and one address serves the well).
struct __align__(16) Well64 {
    double x, y, nq, pad;
};

// One realization's velocity field.  Wells are typically in LDS.
struct Field64 {
    const Well64 *w;
    int nw;
    bool plain;          // every well passes plain_coord / plain_strength
    double xo, yo;
    double A, B, C, D, E, F;
    double A2, B2;       // 2.0*A, 2.0*B (exact)
    double K, bt, bn;    // conductivity, thickness, thickness*porosity
    double por, half_kb2;
};

__device__ __forceinline__ Field64 make_field(const TraceArgs &a, int r, const Well64 *w,
                                              bool plain)
{
    Field64 f;
    f.w = w;
    f.plain = plain;
    f.nw = a.nwells;
    f.xo = a.xo;
    f.yo = a.yo;
    const double *c = a.coef + 6 * (int64_t)r;
    f.A = c[0];
    f.B = c[1];
    f.C = c[2];
    f.D = c[3];
    f.E = c[4];
    f.F = c[5];
    f.A2 = 2.0 * f.A;
    f.B2 = 2.0 * f.B;
    f.K = a.K[r];
    f.bt = a.b[r];
    f.por = a.n[r];
    f.bn = f.bt * f.por;
    f.half_kb2 = (0.5 * f.K) * (f.bt * f.bt);
    return f;
}

// a1 / b and a2 / b, each correctly rounded (IEEE division, as numpy's).
// gfx950 has no division instruction; the compiler expands every FP64 `/`
// into one fixed sequence:
//   d = div_scale(b; b, a)  r = rcp(d)  r += r (1 - d r)  r += r (1 - d r)
//   n, s = div_scale(a; b, a)  m = n r  q = div_fmas(fma(-d, m, n), r, m; s)
//   q = div_fixup(q, b, a)
// Only d and r depend on the numerator, through d's scaling of extreme
// exponents.  When both numerators leave d the same (always, in the normal
// range), this runs that sequence once for d and r and twice for the rest,
// so each result has the bits `a / b` would have; otherwise it divides twice.
__device__ __forceinline__ void div_pair(double a1, double a2, double b, double &q1, double &q2)
{
    bool s1, s2;
    const double d = __builtin_amdgcn_div_scale(a1, b, false, &s1);
    const double d2 = __builtin_amdgcn_div_scale(a2, b, false, &s2);
    if (__double_as_longlong(d) != __double_as_longlong(d2)) {
        q1 = a1 / b;
        q2 = a2 / b;
        return;
    }
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    r = fma(r, fma(-d, r, 1.0), r);
    const double n1 = __builtin_amdgcn_div_scale(a1, b, true, &s1);
    const double n2 = __builtin_amdgcn_div_scale(a2, b, true, &s2);
    const double m1 = n1 * r, m2 = n2 * r;
    q1 = __builtin_amdgcn_div_fixup(
        __builtin_amdgcn_div_fmas(fma(-d, m1, n1), r, m1, s1), b, a1);
    q2 = __builtin_amdgcn_div_fixup(
        __builtin_amdgcn_div_fmas(fma(-d, m2, n2), r, m2, s2), b, a2);
}

// a1 / b and a2 / b, each correctly rounded, for operands that need none of
// div_scale's scalings (see plain_coord): then div_scale returns its operand
// unchanged and div_fmas is a plain fma, so the sequence above reduces to
// this one with the same bits.  div_fixup stays: it gives a zero numerator
// its signed zero and 0/0 its NaN, exactly as in the full sequence.
__device__ __forceinline__ void div_pair_plain(double a1, double a2, double b, double &q1,
                                               double &q2)
{
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    r = fma(r, fma(-b, r, 1.0), r);
    const double m1 = a1 * r, m2 = a2 * r;
    q1 = __builtin_amdgcn_div_fixup(fma(fma(-b, m1, a1), r, m1), b, a1);
    q2 = __builtin_amdgcn_div_fixup(fma(fma(-b, m2, a2), r, m2), b, a2);
}

// The well terms' divisions (-q dx) / r2 and (-q dy) / r2 need no scaling
// when every coordinate involved (evaluation point and wells) is 0 or has a
// magnitude in [2^-200, 2^60), and every |q| is 0 or in [2^-100, 2^100]:
// a nonzero difference of two such coordinates is a multiple of 2^-252 below
// 2^61, so a nonzero r2 lies in [2^-504, 2^123), a nonzero numerator in
// [2^-353, 2^161], their exponents differ by less than 768 and no quotient
// is subnormal -- none of div_scale's conditions (tiny or huge denominator,
// exponent gap >= 768, subnormal quotient, numerator exponent <= 53) holds.
// r2 = 0 implies zero numerators: 0 / 0 gives NaN in both sequences.
__device__ __forceinline__ bool plain_coord(double v)
{
    const double a = fabs(v);
    return a == 0.0 || (a >= 0x1p-200 && a < 0x1p60);
}

__device__ __forceinline__ bool plain_strength(double nq)
{
    const double a = fabs(nq);
    return a == 0.0 || (a >= 0x1p-100 && a <= 0x1p100);
}

// The wells' part of the discharge at (x, y) (model.py:307-313) and, for the
// unconfined field, of the potential (model.py:258-264).
template <bool kConfined, bool kPlain>
__device__ __forceinline__ void well_sums(const Field64 &f, double x, double y, double &Qx,
                                          double &Qy, double &pw)
{
    // two wells per iteration in the confined field: one LDS address and
    // loop step serve both (2 % on the 25-well site; by 4 gains nothing, and
    // the unconfined loop, with its logarithms, is slower unrolled)
    constexpr int kUnroll = kConfined ? 2 : 1;
#pragma unroll kUnroll
    for (int w = 0; w < f.nw; ++w) {
        const Well64 wl = f.w[w];
        const double ddx = x - wl.x;
        const double ddy = y - wl.y;
        const double r2 = ddx * ddx + ddy * ddy;
        double sx, sy;
        if (kPlain)
            div_pair_plain(wl.nq * ddx, wl.nq * ddy, r2, sx, sy);
        else
            div_pair(wl.nq * ddx, wl.nq * ddy, r2, sx, sy);
        Qx = Qx + sx * kC2Pi;
        Qy = Qy + sy * kC2Pi;
        if (!kConfined)
            pw = pw + ((-wl.nq) * log_rn(r2)) * kC4Pi;
    }
}

// feval(xy) = -V(xy) (stochastic.py:253-260).  Returns false where the
// reference raises AquiferError (unconfined only).
template <bool kConfined>
__device__ __forceinline__ bool feval(const Field64 &f, double x, double y, double &kx, double &ky)
{
    const double dx = x - f.xo;
    const double dy = y - f.yo;
    double Qx = -((f.A2 * dx + f.C * dy) + f.D);
    double Qy = -((f.B2 * dy + f.C * dx) + f.E);
    double pw = 0.0;
    if (f.plain && plain_coord(x) && plain_coord(y))
        well_sums<kConfined, true>(f, x, y, Qx, Qy, pw);
    else
        well_sums<kConfined, false>(f, x, y, Qx, Qy, pw);
    if (kConfined) {
        double vx, vy;
        div_pair(Qx, Qy, f.bn, vx, vy);
        kx = -vx;
        ky = -vy;
        return true;
    }
    double pot = f.A * (dx * dx) + f.B * (dy * dy);
    pot = pot + (f.C * dx) * dy;
    pot = pot + f.D * dx;
    pot = pot + f.E * dy;
    pot = pot + f.F;
    pot = pot + pw;
    if (pot <= 0.0)
        return false;
    const double head = pot < f.half_kb2 ? sqrt((2.0 * pot) / f.K) : (pot + f.half_kb2) / (f.K * f.bt);
    if (head <= 0.0)
        return false;
    const double den = head > f.bt ? f.bn : head * f.por;
    kx = -(Qx / den);
    ky = -(Qy / den);
    return true;
}

struct Integ {
    double dur, absdur, tol, maxstep, sg;
    int64_t cap;
};

struct Trace64 {
    double x, y, t, dt, k1x, k1y;
    double bx0, bx1, by0, by1;
    int64_t att, acc;
    int status;
};

// Starts a trace at (xs, ys) (capturezone.py:211-219).  Returns true if the
// integration loop has steps to take.
template <bool kConfined>
__device__ __forceinline__ bool dp5_start(const Field64 &f, const Integ &I, Trace64 &s, double xs,
                                          double ys)
{
    s.x = xs;
    s.y = ys;
    s.t = 0.0;
    s.dt = 0.1 * I.sg;
    s.att = 0;
    s.acc = 0;
    s.bx0 = s.bx1 = xs;
    s.by0 = s.by1 = ys;
    s.status = OKA_TRACE_OK;
    if (!feval<kConfined>(f, xs, ys, s.k1x, s.k1y)) {
        s.status = OKA_TRACE_TRUNCATED;
        return false;
    }
    return 0.0 < I.absdur;
}

// One attempted step of capturezone.py:221-247.  Returns bit 0 = accepted
// (segment (ox, oy) -> (s.x, s.y)), bit 1 = the trace has ended.
template <bool kConfined>
__device__ __forceinline__ int dp5_attempt(const Field64 &f, const Integ &I, Trace64 &s,
                                           double &ox, double &oy)
{
    double dt = s.dt;
    if (fabs(s.t + dt) > I.absdur)
        dt = I.dur - s.t;
    s.att += 1;
    const double x = s.x, y = s.y, k1x = s.k1x, k1y = s.k1y;
    double k2x, k2y, k3x, k3y, k4x, k4y, k5x, k5y, k6x, k6y, k7x, k7y;
    if (!feval<kConfined>(f, x + dt * (A21 * k1x), y + dt * (A21 * k1y), k2x, k2y))
        goto trunc;
    if (!feval<kConfined>(f, x + dt * (A31 * k1x + A32 * k2x), y + dt * (A31 * k1y + A32 * k2y),
                          k3x, k3y))
        goto trunc;
    if (!feval<kConfined>(f, x + dt * ((A41 * k1x + A42 * k2x) + A43 * k3x),
                          y + dt * ((A41 * k1y + A42 * k2y) + A43 * k3y), k4x, k4y))
        goto trunc;
    if (!feval<kConfined>(f, x + dt * (((A51 * k1x + A52 * k2x) + A53 * k3x) + A54 * k4x),
                          y + dt * (((A51 * k1y + A52 * k2y) + A53 * k3y) + A54 * k4y), k5x, k5y))
        goto trunc;
    if (!feval<kConfined>(
            f, x + dt * ((((A61 * k1x + A62 * k2x) + A63 * k3x) + A64 * k4x) + A65 * k5x),
            y + dt * ((((A61 * k1y + A62 * k2y) + A63 * k3y) + A64 * k4y) + A65 * k5y), k6x, k6y))
        goto trunc;
    {
        const double xn = x + dt * ((((B1 * k1x + B3 * k3x) + B4 * k4x) + B5 * k5x) + B6 * k6x);
        const double yn = y + dt * ((((B1 * k1y + B3 * k3y) + B4 * k4y) + B5 * k5y) + B6 * k6y);
        if (!feval<kConfined>(f, xn, yn, k7x, k7y))
            goto trunc;
        const double ex =
            dt * (((((E1 * k1x + E2 * k7x) + E3 * k3x) + E4 * k4x) + E5 * k5x) + E6 * k6x);
        const double ey =
            dt * (((((E1 * k1y + E2 * k7y) + E3 * k3y) + E4 * k4y) + E5 * k5y) + E6 * k6y);
        const double ax = fabs(ex), ay = fabs(ey);
        const double est = (isnan(ax) || isnan(ay)) ? (double)NAN : (ay > ax ? ay : ax);
        const double d0 = xn - x, d1 = yn - y;
        const double ds = sqrt(fma_(d1, d1, d0 * d0));  // numpy norm: sqrt of a contracted ddot
        int res = 0;
        if (est < I.tol && ds < I.maxstep) {
            s.t = s.t + dt;
            s.k1x = k7x;
            s.k1y = k7y;
            ox = x;
            oy = y;
            s.x = xn;
            s.y = yn;
            s.acc += 1;
            s.bx0 = xn < s.bx0 ? xn : s.bx0;
            s.bx1 = xn > s.bx1 ? xn : s.bx1;
            s.by0 = yn < s.by0 ? yn : s.by0;
            s.by1 = yn > s.by1 ? yn : s.by1;
            res = 1;
        }
        double c = pow_fifth(I.tol / (est + kEps));
        const double c2 = I.maxstep / (ds + kEps);
        if (c2 < c)
            c = c2;
        if (10.0 < c)
            c = 10.0;
        dt = (0.9 * c) * dt;
        s.dt = dt;
        if (fabs(s.t) < I.absdur) {
            if (isnan(dt) || dt == 0.0) {
                s.status = OKA_TRACE_STALLED;
                res |= 2;
            } else if (I.cap > 0 && s.att >= I.cap) {
                s.status = OKA_TRACE_STEP_CAP;
                res |= 2;
            }
        } else {
            res |= 2;
        }
        return res;
    }
trunc:
    s.status = OKA_TRACE_TRUNCATED;
    return 2;
}

struct EpochReg {
    double xmin, ymin;
    int nrows, ncols, roff, coff;
};

__device__ __forceinline__ EpochReg find_epoch(const TraceArgs &a, int64_t k)
{
    int lo = 0, hi = a.nepochs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.epochs[mid].first_trace <= k)
            lo = mid;
        else
            hi = mid - 1;
    }
    const oka_epoch &e = a.epochs[lo];
    return {e.xmin, e.ymin, e.nrows, e.ncols, e.row_off, e.col_off};
}

// probabilityfield.py:292-310 for one segment, marking nodes of the final grid
// in `bits` (FlatBits or TileBits).  Already-marked nodes are skipped, as the
// reference does.
template <typename Bits>
__device__ __forceinline__ int raster_segment(Bits &bits, const EpochReg &e, double dx, double dy,
                                               double umbra, double u2, double ax, double ay,
                                               double bx, double by)
{
    const double lox = bx < ax ? bx : ax, hix = bx > ax ? bx : ax;
    const double loy = by < ay ? by : ay, hiy = by > ay ? by : ay;
    double qxl, qxr, qyl, qyr;
    div_pair((lox - umbra) - e.xmin, (hix + umbra) - e.xmin, dx, qxl, qxr);
    div_pair((loy - umbra) - e.ymin, (hiy + umbra) - e.ymin, dy, qyl, qyr);
    double jl = floor(qxl);
    double jr = floor(qxr) + 1.0;
    double il = floor(qyl);
    double ir = floor(qyr) + 1.0;
    if (jl < 0.0) jl = 0.0;
    if (il < 0.0) il = 0.0;
    if (jr > (double)e.ncols) jr = (double)e.ncols;
    if (ir > (double)e.nrows) ir = (double)e.nrows;
    if (!(jl < jr) || !(il < ir))
        return 0;
    const int j0 = (int)jl, j1 = (int)jr, i0 = (int)il, i1 = (int)ir;
    const double bax = bx - ax, bay = by - ay;
    const double l2 = bax * bax + bay * bay;
    for (int j = j0; j < j1; ++j) {
        const double cx = e.xmin + (double)j * dx;
        const double cax = cx - ax;
        const int col = j + e.coff;
        for (int i = i0; i < i1; ++i) {
            const int row = i + e.roff;
            if (bits.test(row, col))
                continue;
            const double cy = e.ymin + (double)i * dy;
            const double cay = cy - ay;
            const double perp = bax * cay - bay * cax;
            const double dot = bax * cax + bay * cay;
            double al, be;
            div_pair(perp * perp, dot * dot, l2, al, be);
            double d2;
            if (dot < 0.0)
                d2 = al + be;
            else if (be > l2)
                d2 = ((al + be) - 2.0 * dot) + l2;
            else
                d2 = al;
            if (d2 < u2)
                bits.set(row, col);
        }
    }
    return (j1 - j0) * (i1 - i0);
}

__device__ __forceinline__ Integ make_integ(const TraceArgs &a)
{
    Integ I;
    I.dur = a.duration;
    I.absdur = fabs(a.duration);
    I.tol = a.tol;
    I.maxstep = a.maxstep;
    I.sg = a.duration > 0.0 ? 1.0 : (a.duration < 0.0 ? -1.0 : (a.duration == 0.0 ? 0.0 : (double)NAN));
    I.cap = a.max_steps;
    return I;
}

}  // namespace fp64
}  // namespace oka
