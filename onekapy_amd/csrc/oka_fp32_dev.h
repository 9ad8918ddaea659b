// oka_fp32_dev.h -- FP32 tracing device code shared by the trace kernels:
// the Dormand-Prince 5(4) step with the reference's step control
// (oneka/capturezone.py:198-253), the confined velocity field (model.py:
// 299-315, 423-427), the unconfined one (model.py:207-266, 318-389) and the
// segment rasterization (probabilityfield.py:292-310, 407-427), computed in
// FP32:
//   - coordinates are relative to the regional origin (the target well), so a
//     float resolves millimetres anywhere on a 2000 x 2000 node grid;
//   - particle positions are Kahan-compensated sums of their increments;
//   - per realization, -1/(b n) and 1/(2 pi) are folded into the regional
//     coefficients and the well strengths: k = -V = sum_w c_w d/r^2 + k_reg;
//   - time t and step dt stay in FP64 (a handful of operations per step) so the
//     termination rule |t| < |duration| behaves as in the reference.
#pragma once

#include <math.h>

#include "oka_internal.h"

namespace oka {
namespace fp32 {

constexpr float A21 = 1.0f / 5.0f;
constexpr float A31 = 3.0f / 40.0f, A32 = 9.0f / 40.0f;
constexpr float A41 = 44.0f / 45.0f, A42 = -56.0f / 15.0f, A43 = 32.0f / 9.0f;
constexpr float A51 = 19372.0f / 6561.0f, A52 = -25360.0f / 2187.0f, A53 = 64448.0f / 6561.0f,
                A54 = -212.0f / 729.0f;
constexpr float A61 = 9017.0f / 3168.0f, A62 = -355.0f / 33.0f, A63 = 46732.0f / 5247.0f,
                A64 = 49.0f / 176.0f, A65 = -5103.0f / 18656.0f;
constexpr float B1 = 35.0f / 384.0f, B3 = 500.0f / 1113.0f, B4 = 125.0f / 192.0f,
                B5 = -2187.0f / 6784.0f, B6 = 11.0f / 84.0f;
constexpr float E1 = 71.0f / 57600.0f, E2 = -1.0f / 40.0f, E3 = -71.0f / 16695.0f,
                E4 = 71.0f / 1920.0f, E5 = -17253.0f / 339200.0f, E6 = 22.0f / 525.0f;
constexpr double kInv2Pi = 0.15915494309189535;
constexpr double kInv4Pi = 0.07957747154594767;
constexpr double kLn2 = 0.6931471805599453;

struct Field32 {
    const float4 *w;   // (xw - xo, yw - yo, q / (2 pi b n), q ln 2 / (4 pi h)), h = K b^2 / 2
    int nw;
    float a2, b2, c, d, e;  // regional: k_reg = ((a2 dx + c dy) + d, (b2 dy + c dx) + e)
    float pa, pb, pc, pd, pe, pf;   // regional potential over h (unconfined only)
};

__device__ __forceinline__ void well_term(const float4 &w, float x, float y, float &vx, float &vy)
{
    const float ddx = x - w.x;
    const float ddy = y - w.y;
    const float r2 = fmaf(ddx, ddx, ddy * ddy);
    const float s = w.z * __builtin_amdgcn_rcpf(r2);
    vx = fmaf(s, ddx, vx);
    vy = fmaf(s, ddy, vy);
}

// The confined well term plus the well's potential over h, q ln(r^2) /
// (4 pi h), as w.w log2(r^2) with the hardware base-2 logarithm.
__device__ __forceinline__ void well_term_pot(const float4 &w, float x, float y, float &vx,
                                              float &vy, float &ph)
{
    const float ddx = x - w.x;
    const float ddy = y - w.y;
    const float r2 = fmaf(ddx, ddx, ddy * ddy);
    const float s = w.z * __builtin_amdgcn_rcpf(r2);
    vx = fmaf(s, ddx, vx);
    vy = fmaf(s, ddy, vy);
    ph = fmaf(w.w, __builtin_amdgcn_logf(r2), ph);
}

// k = -V at (x, y).  Wells are summed in index order; their LDS reads are
// issued 8 at a time so that one trace alone (a long trace after its
// realization's others ended) does not wait a full LDS round trip per well.
//
// Unconfined (kConfined false), with phi = potential / h (model.py:207-266):
//   - phi <= 0: the reference raises AquiferError (model.py:342-343); `dry`
//     is set and k is NaN, which ends the attempted step unaccepted;
//   - phi < 1: head = sqrt(2 potential / K) = b sqrt(phi) < b, so V = Q / (head
//     n) is the confined velocity times 1 / sqrt(phi) (model.py:344-345, 384-386);
//   - otherwise head = (potential + h) / (K b) >= b and V = Q / (b n), the
//     confined velocity (model.py:346-348, 381-383).
template <bool kConfined = true>
__device__ __forceinline__ void feval32(const Field32 &f, float x, float y, float &kx, float &ky,
                                        bool &dry)
{
    float vx = fmaf(f.a2, x, fmaf(f.c, y, f.d));
    float vy = fmaf(f.b2, y, fmaf(f.c, x, f.e));
    float ph = 0.f;
    int i = 0;
    for (; i + 8 <= f.nw; i += 8) {
        float4 w[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            w[k] = f.w[i + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if constexpr (kConfined)
                well_term(w[k], x, y, vx, vy);
            else
                well_term_pot(w[k], x, y, vx, vy, ph);
        }
    }
    for (; i < f.nw; ++i) {
        if constexpr (kConfined)
            well_term(f.w[i], x, y, vx, vy);
        else
            well_term_pot(f.w[i], x, y, vx, vy, ph);
    }
    if constexpr (!kConfined) {
        ph += fmaf(fmaf(f.pa, x, fmaf(f.pc, y, f.pd)), x, fmaf(fmaf(f.pb, y, f.pe), y, f.pf));
        if (ph <= 0.f) {
            dry = true;
            vx = vy = __builtin_nanf("");
        } else if (ph < 1.f) {
            const float g = __builtin_amdgcn_rsqf(ph);
            vx *= g;
            vy *= g;
        }
    }
    kx = vx;
    ky = vy;
}

template <bool kConfined = true>
__device__ __forceinline__ void feval32(const Field32 &f, float x, float y, float &kx, float &ky)
{
    bool dry = false;
    feval32<kConfined>(f, x, y, kx, ky, dry);
}

// Far wells as local expansions (the FP32 pool on sites of many wells).  The
// plane is cut into square boxes of side L (target-relative coordinates,
// box (i, j) has its lower-left corner at (x0 + i L, y0 + j L)); a well is
// near a box when it lies in one of the 3 x 3 boxes around it, else far.
// With F(z) = sum_w s_w / (z - z_w) (z = x + i y, s_w the well's folded
// strength), a well term of k = -V is s_w (dx, dy) / r^2 = conj(s_w / (z - z_w)),
// so k = conj(F) plus the regional part.  For z in box b with centre c, u =
// (z - c) / L (|u| <= 1/sqrt(2)) and v_w = (z_w - c) / L (|v_w| >= 3/2 for a
// far well):
//   s_w / (z - z_w) = -(s_w / (L v_w)) sum_k (u / v_w)^k,
// so the far wells sum to B_b(u) = sum_{k < order} beta_k u^k with
//   beta_k = -(1/L) sum_{far w} s_w / v_w^(k+1),
// truncated where |u / v_w|^order <= (sqrt(2) / 3)^order (6e-6 at order 16).
// Every box keeps kFarTerms coefficients (zero from the order up), so the
// evaluation is one fixed-length Horner sum.  The near wells are summed
// directly, as in feval32.
// Loads through global-memory pointers (address space 1): global loads,
// counted apart from LDS reads, of clang vector types.
typedef float VecF4 __attribute__((ext_vector_type(4)));
typedef int VecI4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) VecF4 *GlobalF4;
typedef const __attribute__((address_space(1))) VecI4 *GlobalI4;

struct FarField {
    int nbx, nby, order, near_stride;
    float x0, y0, len, ilen;        // box grid origin and side, 1 / side
    const int *near;                // [nbx nby][near_stride] (a multiple of 4):
                                    // per box the count of near wells, their
                                    // indices, then -1
    float2 *beta;                   // [nbx nby][kFarTerms] of this slot
};

// beta of boxes lane, lane + stride, ... from the wells w[0, nw): each lane
// walks the wells once per kFarChunk coefficients of one box, keeping only
// those in registers; the power s_w / v_w^(k0+1) that starts a chunk is
// s_w / v_w times (1 / v_w^4)^(k0 / 4).  A confined slot's well records carry
// their box in the fourth component (far_box), so the near test reads LDS
// only.  Coefficients from the order up to kFarTerms are zero.
constexpr int kFarChunk = 4;
static_assert(kFarTerms % kFarChunk == 0 && kFarTerms % 2 == 0, "whole chunks and float4 pairs");

// A well's box packed into a float's bits (column low, row high 16 bits).
__device__ __forceinline__ float far_box(int bx, int by)
{
    return __int_as_float((int)((uint32_t)bx | ((uint32_t)by << 16)));
}

__device__ __forceinline__ void far_coefficients(const FarField &F, const float4 *w, int nw,
                                                 float2 *beta, int lane, int stride)
{
    const int nb = F.nbx * F.nby;
    for (int b = lane; b < nb; b += stride) {
        const int bx = b % F.nbx, by = b / F.nbx;
        const float cx = F.x0 + ((float)bx + 0.5f) * F.len;
        const float cy = F.y0 + ((float)by + 0.5f) * F.len;
        for (int k0 = 0; k0 < kFarTerms; k0 += kFarChunk) {
            float ar[kFarChunk], ai[kFarChunk];
#pragma unroll
            for (int k = 0; k < kFarChunk; ++k)
                ar[k] = ai[k] = 0.0f;
            for (int i = 0; k0 < F.order && i < nw; ++i) {
                const float4 wl = w[i];
                const uint32_t wb = (uint32_t)__float_as_int(wl.w);
                if (abs((int)(wb & 0xffffu) - bx) <= 1 && abs((int)(wb >> 16) - by) <= 1)
                    continue;   // near: summed directly
                const float vx = (wl.x - cx) * F.ilen, vy = (wl.y - cy) * F.ilen;
                const float iv = 1.0f / fmaf(vx, vx, vy * vy);
                const float qx = vx * iv, qy = -vy * iv;   // 1 / v_w
                float tx = wl.z * qx, ty = wl.z * qy;      // s_w / v_w
                if (k0 > 0) {
                    const float q2x = fmaf(qx, qx, -qy * qy), q2y = 2.0f * qx * qy;
                    const float q4x = fmaf(q2x, q2x, -q2y * q2y), q4y = 2.0f * q2x * q2y;
                    for (int k = 0; k < k0; k += kFarChunk) {
                        const float nx = fmaf(tx, q4x, -ty * q4y);
                        ty = fmaf(tx, q4y, ty * q4x);
                        tx = nx;
                    }
                }
#pragma unroll
                for (int k = 0; k < kFarChunk; ++k) {
                    ar[k] -= tx;
                    ai[k] -= ty;
                    const float nx = fmaf(tx, qx, -ty * qy);
                    ty = fmaf(tx, qy, ty * qx);
                    tx = nx;
                }
            }
#pragma unroll
            for (int k = 0; k < kFarChunk; ++k) {
                const bool on = k0 + k < F.order;
                beta[(size_t)b * kFarTerms + k0 + k] =
                    make_float2(on ? ar[k] * F.ilen : 0.0f, on ? ai[k] * F.ilen : 0.0f);
            }
        }
    }
}

// The well terms of four near wells (index -1: none).
__device__ __forceinline__ void near_terms(const float4 *w, VecI4 n, float x, float y, float &vx,
                                           float &vy)
{
    const float4 w0 = w[max(n.x, 0)], w1 = w[max(n.y, 0)];
    const float4 w2 = w[max(n.z, 0)], w3 = w[max(n.w, 0)];
    if (n.x >= 0)
        well_term(w0, x, y, vx, vy);
    if (n.y >= 0)
        well_term(w1, x, y, vx, vy);
    if (n.z >= 0)
        well_term(w2, x, y, vx, vy);
    if (n.w >= 0)
        well_term(w3, x, y, vx, vy);
}

// k = -V at (x, y) with the far wells of its box taken from the expansion;
// false (k untouched) when (x, y) lies outside the boxes.  The box's
// coefficients and the first four words of its near row (the count and three
// indices) are loaded together, and each next four indices while the current
// four are summed.
__device__ __forceinline__ bool feval32_far(const Field32 &f, const FarField &F, float x, float y,
                                            float &kx, float &ky)
{
    const float fx = (x - F.x0) * F.ilen, fy = (y - F.y0) * F.ilen;
    const float bxf = floorf(fx), byf = floorf(fy);
    if (!(bxf >= 0.0f && bxf < (float)F.nbx && byf >= 0.0f && byf < (float)F.nby))
        return false;
    const int b = (int)byf * F.nbx + (int)bxf;
    const GlobalI4 row = (GlobalI4)(F.near + (size_t)b * F.near_stride);
    VecI4 cur = row[0];
    const GlobalF4 a = (GlobalF4)(F.beta + (size_t)b * kFarTerms);
    VecF4 c[kFarTerms / 2];   // coefficients 2i (x, y) and 2i + 1 (z, w)
#pragma unroll
    for (int i = 0; i < kFarTerms / 2; ++i)
        c[i] = a[i];
    const float ux = fx - bxf - 0.5f, uy = fy - byf - 0.5f;
    // Horner in u from the highest coefficient
    float gr = c[kFarTerms / 2 - 1].z, gi = c[kFarTerms / 2 - 1].w;
#pragma unroll
    for (int k = kFarTerms - 2; k >= 0; --k) {
        const VecF4 ck = c[k >> 1];
        const float ar = (k & 1) ? ck.z : ck.x, ai = (k & 1) ? ck.w : ck.y;
        const float nr = fmaf(gr, ux, fmaf(-gi, uy, ar));
        gi = fmaf(gr, uy, fmaf(gi, ux, ai));
        gr = nr;
    }
    float vx = fmaf(f.a2, x, fmaf(f.c, y, f.d)) + gr;
    float vy = fmaf(f.b2, y, fmaf(f.c, x, f.e)) - gi;
    const VecI4 none = {-1, -1, -1, -1};
    const int groups = (cur.x + 4) >> 2;   // groups of 4 words holding the count and indices
    cur.x = -1;
    for (int g = 0; g < groups; ++g) {
        const VecI4 next = g + 1 < groups ? row[g + 1] : none;
        near_terms(f.w, cur, x, y, vx, vy);
        cur = next;
    }
    kx = vx;
    ky = vy;
    return true;
}

// The confined field through the far-well expansions where the point lies
// in a box, directly elsewhere, as finish32's evaluator.
struct LaneEvalFar {
    const Field32 &f;
    const FarField &F;
    __device__ __forceinline__ void operator()(float x, float y, float &kx, float &ky) const
    {
        if (!feval32_far(f, F, x, y, kx, ky))
            feval32<true>(f, x, y, kx, ky);
    }
};

// The lane's own field evaluation, as finish32's evaluator; `dry` collects the
// unconfined field's AquiferError.
template <bool kConfined = true>
struct LaneEval {
    const Field32 &f;
    bool &dry;
    __device__ __forceinline__ void operator()(float x, float y, float &kx, float &ky) const
    {
        feval32<kConfined>(f, x, y, kx, ky, dry);
    }
};

struct Trace32 {
    float x, y, cx, cy;   // Kahan sums (x, y) and their compensations (cx, cy)
    float k1x, k1y;
    float bx0, bx1, by0, by1;
    double t, dt;
    int att, acc;
    int status;
};

__device__ __forceinline__ void kahan_add(float &s, float &c, float v)
{
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

struct Integ32 {
    double dur, absdur, sg;
    float tol, maxstep;
    int64_t cap;
};

// Opens a trace at (xs, ys).  Its k1 = f(start) is not evaluated here: the
// trace is `fresh`, and the caller's next field evaluation, the one other
// lanes spend on their first Dormand-Prince stage, is at the start point
// (first_point) and gives k1 (take_k1).  A lane that opens a trace therefore
// costs its wave no evaluation of its own.
__device__ __forceinline__ void open32(Trace32 &s, const Integ32 &I, float xs, float ys)
{
    s.x = xs;
    s.y = ys;
    s.cx = 0.f;
    s.cy = 0.f;
    s.t = 0.0;
    s.dt = 0.1 * I.sg;
    s.att = 0;
    s.acc = 0;
    s.bx0 = s.bx1 = xs;
    s.by0 = s.by1 = ys;
    s.status = OKA_TRACE_OK;
}

// The point of this iteration's first field evaluation and the step it
// starts (capturezone.py:221-231): the start of a fresh trace, else the first
// stage with dt clamped to the remaining duration.
__device__ __forceinline__ void first_point(const Integ32 &I, const Trace32 &s, bool fresh,
                                            double &dtd, float &dt, float &px, float &py)
{
    dtd = s.dt;
    if (fabs(s.t + dtd) > I.absdur)
        dtd = I.dur - s.t;
    dt = (float)dtd;
    px = fresh ? s.x : s.x + dt * (A21 * s.k1x);
    py = fresh ? s.y : s.y + dt * (A21 * s.k1y);
}

// A fresh trace's k1; false when its duration is zero (the trace is only its
// start vertex, capturezone.py:221).
__device__ __forceinline__ bool take_k1(const Integ32 &I, Trace32 &s, float kx, float ky)
{
    s.k1x = kx;
    s.k1y = ky;
    return 0.0 < I.absdur;
}

// The rest of one attempted step (capturezone.py:221-247) given k2 = f(first
// stage): bit 0 of the result = accepted (the segment is (ox, oy) -> (s.x,
// s.y)), bit 1 = the trace ended.  ev(x, y, kx, ky) evaluates the field (the
// lane's own evaluation, or one that the whole wave shares).
template <typename Eval>
__device__ __forceinline__ int finish32(const Eval &ev, const Integ32 &I, Trace32 &s,
                                        double dtd, float dt, float k2x, float k2y, float &ox,
                                        float &oy)
{
    s.att += 1;
    const float x = s.x, y = s.y, k1x = s.k1x, k1y = s.k1y;
    float k3x, k3y, k4x, k4y, k5x, k5y, k6x, k6y, k7x, k7y;
    ev(x + dt * (A31 * k1x + A32 * k2x), y + dt * (A31 * k1y + A32 * k2y), k3x, k3y);
    ev(x + dt * (A41 * k1x + A42 * k2x + A43 * k3x),
       y + dt * (A41 * k1y + A42 * k2y + A43 * k3y), k4x, k4y);
    ev(x + dt * (A51 * k1x + A52 * k2x + A53 * k3x + A54 * k4x),
       y + dt * (A51 * k1y + A52 * k2y + A53 * k3y + A54 * k4y), k5x, k5y);
    ev(x + dt * (A61 * k1x + A62 * k2x + A63 * k3x + A64 * k4x + A65 * k5x),
       y + dt * (A61 * k1y + A62 * k2y + A63 * k3y + A64 * k4y + A65 * k5y), k6x, k6y);
    const float dxs = dt * (B1 * k1x + B3 * k3x + B4 * k4x + B5 * k5x + B6 * k6x);
    const float dys = dt * (B1 * k1y + B3 * k3y + B4 * k4y + B5 * k5y + B6 * k6y);
    const float xn = x + dxs, yn = y + dys;
    ev(xn, yn, k7x, k7y);
    const float ex = dt * (E1 * k1x + E2 * k7x + E3 * k3x + E4 * k4x + E5 * k5x + E6 * k6x);
    const float ey = dt * (E1 * k1y + E2 * k7y + E3 * k3y + E4 * k4y + E5 * k5y + E6 * k6y);
    const float est = fmaxf(fabsf(ex), fabsf(ey));
    const float ds = __builtin_amdgcn_sqrtf(fmaf(dxs, dxs, dys * dys));
    int res = 0;
    const bool finite = !(isnan(ex) || isnan(ey));
    if (finite && est < I.tol && ds < I.maxstep) {
        s.t = s.t + dtd;
        s.k1x = k7x;
        s.k1y = k7y;
        ox = x;
        oy = y;
        kahan_add(s.x, s.cx, dxs);
        kahan_add(s.y, s.cy, dys);
        s.acc += 1;
        s.bx0 = fminf(s.bx0, s.x);
        s.bx1 = fmaxf(s.bx1, s.x);
        s.by0 = fminf(s.by0, s.y);
        s.by1 = fmaxf(s.by1, s.y);
        res = 1;
    }
    // step-size factors with hardware reciprocal and square root (about 1 ulp)
    float cf = __expf(0.2f * __logf(I.tol * __builtin_amdgcn_rcpf(est + 2.220446049250313e-16f)));
    const float c2 = I.maxstep * __builtin_amdgcn_rcpf(ds + 2.220446049250313e-16f);
    if (c2 < cf)
        cf = c2;
    if (10.0f < cf)
        cf = 10.0f;
    const double ndt = (double)(0.9f * cf) * dtd;
    s.dt = ndt;
    if (fabs(s.t) < I.absdur) {
        if (!finite || isnan(ndt) || ndt == 0.0) {
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

struct Epoch32 {
    float xmin, ymin;   // relative to the regional origin
    int nrows, ncols, roff, coff;
};

__device__ __forceinline__ Epoch32 find_epoch32(const TraceArgs &a, int64_t k)
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
    return {(float)(e.xmin - a.xo), (float)(e.ymin - a.yo), e.nrows, e.ncols, e.row_off, e.col_off};
}

// Per-node form of the segment test (probabilityfield.py:296-310, 407-427),
// used for segments within 1e-20 of an axis direction, where the row-span
// form below would divide by a vanishing component.
template <typename Bits>
__device__ __forceinline__ void raster32_nodes(Bits &bits, const Epoch32 &e, float dx, float dy,
                                               float u2, float ax, float ay, float bax, float bay,
                                               float l2, int j0, int j1, int i0, int i1)
{
    const float il2 = 1.0f / l2;
    for (int j = j0; j < j1; ++j) {
        const float cax = fmaf((float)j, dx, e.xmin) - ax;
        const int col = j + e.coff;
        for (int i = i0; i < i1; ++i) {
            const int row = i + e.roff;
            if (bits.test(row, col))
                continue;
            const float cay = fmaf((float)i, dy, e.ymin) - ay;
            const float perp = bax * cay - bay * cax;
            const float dot = fmaf(bax, cax, bay * cay);
            const float al = perp * perp * il2;
            const float be = dot * dot * il2;
            float d2;
            if (dot < 0.0f)
                d2 = al + be;
            else if (be > l2)
                d2 = (al + be) - 2.0f * dot + l2;
            else
                d2 = al;
            if (d2 < u2)
                bits.set(row, col);
        }
    }
}

// Marks the nodes strictly within umbra of segment A-B among the reference's
// candidate window (probabilityfield.py:296-301) and returns the window's
// size.  The reference's distance (probabilityfield.py:407-427) is the
// distance to the segment, so the marked nodes of a row are those inside the
// row's chord of the capsule {P : dist(P, AB) < umbra}: an open interval, the
// hull of the chords of the disks around A and B and of the band
// {0 <= (P - A).(B - A) <= |B - A|^2, |(B - A) x (P - A)| < umbra |B - A|}.
// Each row costs its chord and a word-wise mark instead of a test per node;
// the two forms differ only for nodes within rounding of the umbra circle.
// The band with the disk around B is convex too, so leaving out a disk around
// A that is marked already still leaves one interval per row.
// disk_a = false skips the disk around A: the caller's previous segment of
// the same trace ended at A and marked that disk already (its `covered` came
// back true).  covered is set when the disk around B is marked (or lies off
// the grid), i.e. unless the segment has zero length.
template <typename Bits>
__device__ __forceinline__ int raster32(Bits &bits, const Epoch32 &e, float dx, float dy,
                                         float idx_, float idy, float umbra, float u2, float ax,
                                         float ay, float bx, float by, bool disk_a, bool &covered)
{
    covered = true;
    const float lox = fminf(ax, bx), hix = fmaxf(ax, bx);
    const float loy = fminf(ay, by), hiy = fmaxf(ay, by);
    float jl = floorf((lox - umbra - e.xmin) * idx_);
    float jr = floorf((hix + umbra - e.xmin) * idx_) + 1.0f;
    float il = floorf((loy - umbra - e.ymin) * idy);
    float ir = floorf((hiy + umbra - e.ymin) * idy) + 1.0f;
    jl = fmaxf(jl, 0.0f);
    il = fmaxf(il, 0.0f);
    jr = fminf(jr, (float)e.ncols);
    ir = fminf(ir, (float)e.nrows);
    if (!(jl < jr) || !(il < ir))
        return 0;
    const int j0 = (int)jl, j1 = (int)jr, i0 = (int)il, i1 = (int)ir;
    const float bax = bx - ax, bay = by - ay;
    const float l2 = fmaf(bax, bax, bay * bay);
    if (!(l2 > 0.0f)) {
        covered = false;
        return 0;
    }
    if (!(fabsf(bax) > 1e-20f && fabsf(bay) > 1e-20f)) {
        raster32_nodes(bits, e, dx, dy, u2, ax, ay, bax, bay, l2, j0, j1, i0, i1);
        return (j1 - j0) * (i1 - i0);
    }
    // In row coordinates s = x - ax, c = y_row - ay the band's chord is
    // kb c + [min(0, l2 / bax), max(0, l2 / bax)] cut by ka c -+ wr.
    const float ibax = __builtin_amdgcn_rcpf(bax), ibay = __builtin_amdgcn_rcpf(bay);
    const float kb = -bay * ibax, l2b = l2 * ibax;
    const float blo = fminf(l2b, 0.0f), bhi = fmaxf(l2b, 0.0f);
    const float ka = bax * ibay, wr = umbra * __builtin_amdgcn_sqrtf(l2) * fabsf(ibay);
    const float ox = (ax - e.xmin) * idx_;   // A's column coordinate
    const float jlast = jr - 1.0f;
    const int stride = bits.stride();
    int base = (i0 + e.roff) * stride;       // flat index of the row's first node
    for (int i = i0; i < i1; ++i, base += stride) {
        const float c = fmaf((float)i, dy, e.ymin) - ay;
        float lo = fmaxf(fmaf(kb, c, blo), fmaf(ka, c, -wr));
        float hi = fminf(fmaf(kb, c, bhi), fmaf(ka, c, wr));
        if (!(lo < hi)) {
            lo = INFINITY;
            hi = -INFINITY;
        }
        if (disk_a) {
            const float ha = fmaf(-c, c, u2);
            if (ha > 0.0f) {
                const float r = __builtin_amdgcn_sqrtf(ha);
                lo = fminf(lo, -r);
                hi = fmaxf(hi, r);
            }
        }
        const float cb = c - bay;
        const float hb = fmaf(-cb, cb, u2);
        if (hb > 0.0f) {
            const float r = __builtin_amdgcn_sqrtf(hb);
            lo = fminf(lo, bax - r);
            hi = fmaxf(hi, bax + r);
        }
        // nodes with lo < x_j - ax < hi, clamped to the window (both bounds
        // are finite and inside it whenever fl <= fh)
        const float fl = fmaxf(floorf(fmaf(lo, idx_, ox + 1.0f)), jl);
        const float fh = fminf(ceilf(fmaf(hi, idx_, ox - 1.0f)), jlast);
        if (fl <= fh)
            bits.mark(i + e.roff, base, (int)fl + e.coff, (int)fh + e.coff);
    }
    return (j1 - j0) * (i1 - i0);
}

}  // namespace fp32
}  // namespace oka
