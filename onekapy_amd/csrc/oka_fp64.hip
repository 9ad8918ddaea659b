// oka_fp64.hip -- FP64-exact kernels for gfx950.
//
// Every floating-point expression below is evaluated in the reference's
// operation order with contraction disabled, so that for identical inputs the
// results are bit-identical to the reference's numpy scalar arithmetic:
//   velocity           oneka/model.py:299-315, 343-350, 377-389, 423-427
//   feval              oneka/stochastic.py:253-260
//   Dormand-Prince     oneka/capturezone.py:198-253
//   segment raster     oneka/probabilityfield.py:292-310, 407-427
//
// Work layout: one workgroup per realization (grid-stride over realizations),
// one particle per lane with lanes refilled from an LDS counter as their traces
// end, the realization's wells staged in LDS, and the realization's visited-node
// bitmap in LDS (or HBM for grids too large for it).  At the end of the
// realization the bitmap is flushed into the int32 count grid with one atomic
// add per set bit, so a node counts once per realization
// (probabilityfield.py:357-359).
#include <math.h>

#include "oka_internal.h"
#include "oka_math.h"

#pragma clang fp contract(off)

namespace oka {
namespace {

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

// One realization's velocity field.  Wells are (xw, yw, -qw), typically in LDS.
struct Field64 {
    const double *xw, *yw, *nq;
    int nw;
    double xo, yo;
    double A, B, C, D, E, F;
    double A2, B2;       // 2.0*A, 2.0*B (exact)
    double K, bt, bn;    // conductivity, thickness, thickness*porosity
    double por, half_kb2;
};

__device__ __forceinline__ Field64 make_field(const TraceArgs &a, int r, const double *xw,
                                              const double *yw, const double *nq)
{
    Field64 f;
    f.xw = xw;
    f.yw = yw;
    f.nq = nq;
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
    for (int w = 0; w < f.nw; ++w) {
        const double ddx = x - f.xw[w];
        const double ddy = y - f.yw[w];
        const double r2 = ddx * ddx + ddy * ddy;
        const double nq = f.nq[w];
        double sx, sy;
        div_pair(nq * ddx, nq * ddy, r2, sx, sy);
        Qx = Qx + sx * kC2Pi;
        Qy = Qy + sy * kC2Pi;
        if (!kConfined)
            pw = pw + ((-nq) * cr_log(r2)) * kC4Pi;
    }
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


// Trace kernel modes.
enum { kModeLdsBitmap = 0, kModeHbmBitmap = 1, kModeVertices = 2, kModeTiledBitmap = 3 };

// The chronicle step of one kernel mode: marks a segment, flushes a realization.
template <int kMode>
struct Chron;

template <>
struct Chron<kModeLdsBitmap> {
    FlatBits bits;
    __device__ void begin(const TraceArgs &a)
    {
        for (int w = threadIdx.x; w < a.nwords; w += blockDim.x)
            bits.bm[w] = 0u;
    }
    __device__ unsigned long long flush(const TraceArgs &a, unsigned int *s_win)
    {
        return flush_bitmap(a, bits.bm, s_win);
    }
};

template <>
struct Chron<kModeHbmBitmap> : Chron<kModeLdsBitmap> {};

template <>
struct Chron<kModeTiledBitmap> {
    TileBits bits;
    __device__ void begin(const TraceArgs &) {}
    __device__ unsigned long long flush(const TraceArgs &a, unsigned int *s_win)
    {
        return flush_tiles(a, bits, s_win);
    }
};

template <>
struct Chron<kModeVertices> {
    FlatBits bits;   // unused
    __device__ void begin(const TraceArgs &) {}
    __device__ unsigned long long flush(const TraceArgs &, unsigned int *) { return 0; }
};

// One workgroup per realization (grid-stride), one trace per lane, lanes
// refilled from the LDS counter s_next when their trace ends.
// The confined tracer is held to 128 VGPRs (4 waves per SIMD; one 64-bit
// value spills, stored before the realization loop and read after it), which
// is 3 % faster than the 150 the compiler picks; the unconfined one would
// spill 43 registers and keeps the compiler's choice.
template <bool kConfined, int kMode>
__global__ void __launch_bounds__(kTraceBlock)
__attribute__((amdgpu_waves_per_eu(kConfined ? 4 : 1))) k_trace_fp64(TraceArgs a)
{
    constexpr bool kVerts = kMode == kModeVertices;
    extern __shared__ double smem[];
    double *s_xw = smem;
    double *s_yw = smem + a.nwells;
    double *s_nq = smem + 2 * a.nwells;
    __shared__ int s_next;
    __shared__ unsigned long long s_stat[kStatCount];

    const int tid = threadIdx.x;
    const int P = a.npaths;
    const Integ I = make_integ(a);
    const double umbra = a.umbra;
    const double u2 = umbra * umbra;
    const bool raster = !kVerts && a.rasterize;
    uint32_t *lds_tail = reinterpret_cast<uint32_t *>(smem + 3 * a.nwells);
    Chron<kMode> ch;
    if constexpr (kMode == kModeLdsBitmap) {
        ch.bits = FlatBits{lds_tail, a.gncols};
        lds_tail += a.rasterize ? a.nwords : 0;
    } else if constexpr (kMode == kModeHbmBitmap) {
        ch.bits = FlatBits{a.gbitmap + (int64_t)blockIdx.x * a.nwords, a.gncols};
    } else if constexpr (kMode == kModeTiledBitmap) {
        ch.bits = tile_bits(a, lds_tail);
        tile_init(ch.bits);
        lds_tail += tile_lds_words(a.ntiles, a.tile_slots);
    }
    unsigned int *s_win = (raster && a.wmarks) ? lds_tail : nullptr;
    if (s_win)
        for (int c = threadIdx.x; c < a.w_nr * a.w_nc; c += blockDim.x)
            s_win[c] = 0u;

    for (int r = blockIdx.x; r < a.nreal; r += gridDim.x) {
        if (tid < kStatCount)
            s_stat[tid] = 0;
        for (int i = tid; i < a.nwells; i += blockDim.x) {
            s_xw[i] = a.wells_xy[2 * i];
            s_yw[i] = a.wells_xy[2 * i + 1];
            s_nq[i] = -a.q[(int64_t)r * a.nwells + i];
        }
        if (raster)
            ch.begin(a);
        if (tid == 0)
            s_next = blockDim.x;
        __syncthreads();

        const Field64 f = make_field(a, r, s_xw, s_yw, s_nq);
        unsigned long long n_att = 0, n_acc = 0, n_esc = 0, n_cand = 0;
        unsigned long long n_st[4] = {0, 0, 0, 0};

        Trace64 s;
        EpochReg ep;
        int p = tid;
        bool have = p < P;
        bool live = false;
        for (;;) {
            bool ended = false;
            int64_t k = (int64_t)r * P + p;
            if (!live) {
                if (!have)
                    break;
                if (raster)
                    ep = find_epoch(a, k);
                const double xs = a.start_xy[2 * p], ys = a.start_xy[2 * p + 1];
                if (kVerts && a.vcap > 0) {
                    a.verts[2 * (k * a.vcap)] = xs;
                    a.verts[2 * (k * a.vcap) + 1] = ys;
                }
                live = dp5_start<kConfined>(f, I, s, xs, ys);
                ended = !live;
            } else {
                double ox, oy;
                const int res = dp5_attempt<kConfined>(f, I, s, ox, oy);
                if (res & 1) {
                    if (raster)
                        n_cand += (unsigned long long)raster_segment(ch.bits, ep, a.gdx, a.gdy,
                                                                     umbra, u2, ox, oy, s.x, s.y);
                    if (kVerts && s.acc < a.vcap) {
                        a.verts[2 * (k * a.vcap + s.acc)] = s.x;
                        a.verts[2 * (k * a.vcap + s.acc) + 1] = s.y;
                    }
                }
                if (res & 2) {
                    live = false;
                    ended = true;
                }
            }
            if (ended) {
                n_att += (unsigned long long)s.att;
                n_acc += (unsigned long long)s.acc;
                n_st[s.status & 3] += 1;
                if (!(s.bx0 > a.gxmin && s.bx1 < a.gxmax && s.by0 > a.gymin && s.by1 < a.gymax))
                    n_esc += 1;
                if (a.bbox) {
                    a.bbox[4 * k] = s.bx0;
                    a.bbox[4 * k + 1] = s.bx1;
                    a.bbox[4 * k + 2] = s.by0;
                    a.bbox[4 * k + 3] = s.by1;
                }
                if (a.trace_status)
                    a.trace_status[k] = s.status;
                if (a.trace_steps)
                    a.trace_steps[k] = s.att;
                if (kVerts)
                    a.nverts[k] = s.acc + 1;
                p = atomicAdd(&s_next, 1);
                have = p < P;
            }
        }
        atomicAdd(&s_stat[kStatAttempted], n_att);
        atomicAdd(&s_stat[kStatAccepted], n_acc);
        atomicAdd(&s_stat[kStatEscaped], n_esc);
        atomicAdd(&s_stat[kStatCandidates], n_cand);
        for (int j = 0; j < 4; ++j)
            if (n_st[j])
                atomicAdd(&s_stat[kStatStatus0 + j], n_st[j]);
        __syncthreads();

        if (raster)
            atomicAdd(&s_stat[kStatVisited], ch.flush(a, s_win));
        __syncthreads();
        if (s_win) {
            store_window(a, r, s_win);
            __syncthreads();
        }
        if (tid < kStatCount && a.stats)
            atomicAdd(&a.stats[tid], s_stat[tid]);
        __syncthreads();
    }
}

// Model.compute_* at many points (model.py:207-427), one point per lane.
__global__ void __launch_bounds__(256) k_eval_fp64(EvalArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.npts)
        return;
    const double x = a.xy[2 * i], y = a.xy[2 * i + 1];
    const double dx = x - a.xo, dy = y - a.yo;
    const double *c = a.coef;
    double Qx = -(((2.0 * c[0]) * dx + c[2] * dy) + c[3]);
    double Qy = -(((2.0 * c[1]) * dy + c[2] * dx) + c[4]);
    double pw = 0.0;
    const bool need_q = a.kind == OKA_EVAL_DISCHARGE || a.kind == OKA_EVAL_VELOCITY ||
                        a.kind == OKA_EVAL_VELOCITY_CONFINED;
    const bool need_p = !need_q || a.kind == OKA_EVAL_VELOCITY;
    for (int w = 0; w < a.nwells; ++w) {
        const double *wl = a.wells + 4 * w;
        const double ddx = x - wl[0], ddy = y - wl[1];
        const double r2 = ddx * ddx + ddy * ddy;
        if (need_q) {
            const double nq = -wl[3];
            double sx, sy;
            div_pair(nq * ddx, nq * ddy, r2, sx, sy);
            Qx = Qx + sx * kC2Pi;
            Qy = Qy + sy * kC2Pi;
        }
        if (need_p)
            pw = pw + (wl[3] * cr_log(r2)) * kC4Pi;
    }
    int st = 0;
    double o0 = 0.0, o1 = 0.0;
    double pot = c[0] * (dx * dx) + c[1] * (dy * dy);
    pot = pot + (c[2] * dx) * dy;
    pot = pot + c[3] * dx;
    pot = pot + c[4] * dy;
    pot = pot + c[5];
    pot = pot + pw;
    const double half_kb2 = (0.5 * a.K) * (a.b * a.b);
    double head = 0.0;
    if (a.kind == OKA_EVAL_HEAD || a.kind == OKA_EVAL_VELOCITY) {
        if (pot <= 0.0)
            st = OKA_ERR_AQUIFER;
        else if (pot < half_kb2)
            head = sqrt((2.0 * pot) / a.K);
        else
            head = (pot + half_kb2) / (a.K * a.b);
    }
    switch (a.kind) {
    case OKA_EVAL_POTENTIAL: o0 = pot; break;
    case OKA_EVAL_POTENTIAL_WELLS: o0 = pw; break;
    case OKA_EVAL_DISCHARGE: o0 = Qx; o1 = Qy; break;
    case OKA_EVAL_HEAD: o0 = head; break;
    case OKA_EVAL_VELOCITY_CONFINED:
        o0 = Qx / (a.b * a.n);
        o1 = Qy / (a.b * a.n);
        break;
    case OKA_EVAL_VELOCITY:
        if (st == 0 && head <= 0.0)
            st = OKA_ERR_AQUIFER;
        if (st == 0) {
            const double den = head > a.b ? a.b * a.n : head * a.n;
            o0 = Qx / den;
            o1 = Qy / den;
        }
        break;
    default: st = OKA_ERR_ARG; break;
    }
    if (st) {
        o0 = (double)NAN;
        o1 = (double)NAN;
    }
    const int width = (a.kind == OKA_EVAL_DISCHARGE || a.kind == OKA_EVAL_VELOCITY ||
                       a.kind == OKA_EVAL_VELOCITY_CONFINED) ? 2 : 1;
    a.out[width * i] = o0;
    if (width == 2)
        a.out[2 * i + 1] = o1;
    if (a.status)
        a.status[i] = st;
}

// ProbabilityField.insert over the segments of one polyline (one per lane).
__global__ void __launch_bounds__(256) k_raster_fp64(RasterArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nseg)
        return;
    EpochReg e{a.xmin, a.ymin, a.nrows, a.ncols, 0, 0};
    FlatBits bits{a.bitmap, a.ncols};
    raster_segment(bits, e, a.dx, a.dy, a.umbra, a.umbra * a.umbra, a.xy[2 * i], a.xy[2 * i + 1],
                   a.xy[2 * i + 2], a.xy[2 * i + 3]);
}

template <int kMode>
hipError_t launch_mode(const TraceArgs &a, int grid_blocks, size_t lds, hipStream_t s)
{
    dim3 grid(grid_blocks), block(kTraceBlock);
    if (a.confined)
        hipLaunchKernelGGL((k_trace_fp64<true, kMode>), grid, block, lds, s, a);
    else
        hipLaunchKernelGGL((k_trace_fp64<false, kMode>), grid, block, lds, s, a);
    return hipGetLastError();
}

}  // namespace

size_t trace_fp64_wells_lds(int nwells) { return (size_t)3 * nwells * sizeof(double); }

hipError_t launch_trace_fp64(const TraceArgs &a, int grid_blocks, int kind, hipStream_t s)
{
    size_t lds = trace_fp64_wells_lds(a.nwells) + (a.wmarks ? (size_t)a.w_nr * a.w_nc * 4 : 0);
    if (!a.rasterize || kind == kBmLdsFlat)
        return launch_mode<kModeLdsBitmap>(a, grid_blocks,
                                           lds + (a.rasterize ? (size_t)a.nwords * 4 : 0), s);
    if (kind == kBmTiled)
        return launch_mode<kModeTiledBitmap>(
            a, grid_blocks, lds + (size_t)tile_lds_words(a.ntiles, a.tile_slots) * 4, s);
    return launch_mode<kModeHbmBitmap>(a, grid_blocks, lds, s);
}

hipError_t launch_vertices_fp64(const TraceArgs &a, hipStream_t s)
{
    const int blocks = a.nreal < 1 ? 1 : a.nreal;
    return launch_mode<kModeVertices>(a, blocks, trace_fp64_wells_lds(a.nwells), s);
}

hipError_t launch_eval_fp64(const EvalArgs &a, hipStream_t s)
{
    const int64_t blocks = (a.npts + 255) / 256;
    hipLaunchKernelGGL(k_eval_fp64, dim3((unsigned)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_raster_fp64(const RasterArgs &a, hipStream_t s)
{
    const int64_t blocks = (a.nseg + 255) / 256;
    hipLaunchKernelGGL(k_raster_fp64, dim3((unsigned)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

double host_pow_fifth(double x) { return pow_fifth(x); }
double host_log(double x) { return cr_log(x); }

}  // namespace oka
