// oka_pool.hip -- the persistent pool tracer (oka_pool_impl.h) in FP32.
//
// The step, field and rasterizer are those of the FP32 workgroup tracer
// (oka_fp32_dev.h: oneka/capturezone.py:198-253, model.py:207-266, 299-315,
// 318-389, 423-427, probabilityfield.py:292-310, 407-427), with each slot's
// constants folded exactly as that tracer folds them per realization.  The
// dispatch between the FP32 and FP64-exact pools is here too.
#include <math.h>

#include "oka_fp32_dev.h"
#include "oka_internal.h"
#include "oka_pool_impl.h"

namespace oka {
namespace {

using namespace fp32;
using namespace pool;

struct PoolFP32 {
    using Well = float4;   // (xw - xo, yw - yo, q / (2 pi b n), q ln 2 / (4 pi h) or the box)
    struct Consts {
        float a2, b2, c, d, e;              // regional velocity over -(b n)
        float pa, pb, pc, pd, pe, pf;       // regional potential over h (unconfined)
        float tol, maxstep, gdx, gdy, igdx, igdy, umbra, u2;
        float gx0, gx1, gy0, gy1;           // grid bounds relative to the origin
        double dur, absdur, sg, xo, yo;
        long long cap;
        FarField far;                       // far.order == 0: every well summed directly
    };
    struct Lane {
        Trace32 tr;
        Epoch32 ep;
        unsigned long long cand;   // candidate nodes of the trace
        uint32_t last_t, last_v;   // the lane's tile cache (PoolBits::lookup)
        bool fresh;                // k1 not yet evaluated
        bool a_done;               // the disk around the trace's current point is marked
    };
    // At most 128 VGPRs (one 64-bit value spills, in the flush's scan of an
    // overfull map), so that 16 waves fit a CU when the LDS allows them.
    static constexpr int kWavesPerEU = 4;

    __device__ static bool stage_wells(float4 *w, const TraceArgs &A, int64_t r)
    {
        const double bn = A.b[r] * A.n[r];
        // h = K b^2 / 2, the potential at which the head reaches the top of
        // the aquifer (model.py:344)
        const double h = A.confined ? 1.0 : 0.5 * A.K[r] * (A.b[r] * A.b[r]);
        // the fourth component: the unconfined potential's coefficient, or for
        // a confined slot with far-well expansions the well's box (far_box)
        const bool far = A.confined && A.far_order > 0;
        for (int i = threadIdx.x; i < A.nwells; i += kWave) {
            const double qw = A.q[r * A.nwells + i];
            w[i] = make_float4((float)(A.wells_xy[2 * i] - A.xo),
                               (float)(A.wells_xy[2 * i + 1] - A.yo), (float)(qw * kInv2Pi / bn),
                               far ? far_box(A.far_wbox[2 * i], A.far_wbox[2 * i + 1])
                                   : (A.confined ? 0.0f : (float)(qw * (kInv4Pi * kLn2) / h)));
        }
        return true;
    }

    __device__ static void stage_consts(Consts &k, const PoolArgs &P, const TraceArgs &A,
                                        int64_t r, int s, bool)
    {
        const double bn = A.b[r] * A.n[r];
        const double h = A.confined ? 1.0 : 0.5 * A.K[r] * (A.b[r] * A.b[r]);
        const double *c = A.coef + 6 * r;
        k.a2 = (float)(2.0 * c[0] / bn);
        k.b2 = (float)(2.0 * c[1] / bn);
        k.c = (float)(c[2] / bn);
        k.d = (float)(c[3] / bn);
        k.e = (float)(c[4] / bn);
        k.pa = (float)(c[0] / h);
        k.pb = (float)(c[1] / h);
        k.pc = (float)(c[2] / h);
        k.pd = (float)(c[3] / h);
        k.pe = (float)(c[4] / h);
        k.pf = (float)(c[5] / h);
        k.dur = A.duration;
        k.absdur = fabs(A.duration);
        k.sg = A.duration > 0.0 ? 1.0
                                : (A.duration < 0.0 ? -1.0 : (A.duration == 0.0 ? 0.0 : (double)NAN));
        k.cap = A.max_steps;
        k.tol = (float)A.tol;
        k.maxstep = (float)A.maxstep;
        k.gdx = (float)A.gdx;
        k.gdy = (float)A.gdy;
        k.igdx = (float)(1.0 / A.gdx);
        k.igdy = (float)(1.0 / A.gdy);
        k.umbra = (float)A.umbra;
        k.u2 = k.umbra * k.umbra;
        k.gx0 = (float)(A.gxmin - A.xo);
        k.gx1 = (float)(A.gxmax - A.xo);
        k.gy0 = (float)(A.gymin - A.yo);
        k.gy1 = (float)(A.gymax - A.yo);
        k.xo = A.xo;
        k.yo = A.yo;
        FarField &F = k.far;
        F.order = A.confined && A.far_order > 0 ? min(A.far_order, kFarTerms) : 0;
        F.nbx = A.far_nbx;
        F.nby = A.far_nby;
        F.x0 = A.far_x0;
        F.y0 = A.far_y0;
        F.len = A.far_len;
        F.ilen = 1.0f / A.far_len;
        F.near_stride = A.far_near_stride;
        F.near = A.far_near;
        F.beta = P.far_beta + ((int64_t)blockIdx.x * P.slots + s) * P.far_stride;   // 128-byte rows
    }

    // The far-well coefficients of the slot's realization: a table that the
    // chunks of a realization may share.
    __device__ static bool table_shared(const Consts &k) { return k.far.order > 0; }
    __device__ static void table_to(Consts &k, float2 *beta) { k.far.beta = beta; }

    // The far-well coefficients of the slot's realization, one box per lane.
    __device__ static bool stage_slot(const Consts &k, const float4 *w, int nw)
    {
        if (k.far.order <= 0)
            return false;
        far_coefficients(k.far, w, nw, k.far.beta, threadIdx.x, kWave);
        return true;
    }

    __device__ static Integ32 integ(const Consts &k)
    {
        Integ32 I;
        I.dur = k.dur;
        I.absdur = k.absdur;
        I.sg = k.sg;
        I.tol = k.tol;
        I.maxstep = k.maxstep;
        I.cap = k.cap;
        return I;
    }

    __device__ static Field32 field(const Consts &k, const float4 *w, int nw, bool confined)
    {
        Field32 f;
        f.w = w;
        f.nw = nw;
        f.a2 = k.a2;
        f.b2 = k.b2;
        f.c = k.c;
        f.d = k.d;
        f.e = k.e;
        if (!confined) {
            f.pa = k.pa;
            f.pb = k.pb;
            f.pc = k.pc;
            f.pd = k.pd;
            f.pe = k.pe;
            f.pf = k.pf;
        }
        return f;
    }

    __device__ static void open(Lane &ln, const Consts &k, const TraceArgs &A, int64_t r, int p,
                                bool raster)
    {
        if (raster)
            ln.ep = find_epoch32(A, r * A.npaths + p);
        ln.cand = 0;
        ln.last_t = kKeyEmpty;
        ln.last_v = kValNone;
        open32(ln.tr, integ(k), (float)(A.start_xy[2 * p] - A.xo),
               (float)(A.start_xy[2 * p + 1] - A.yo));
        ln.fresh = true;
        ln.a_done = false;
    }

    // One field evaluation per live lane: k1 of a fresh trace (which costs
    // the wave no evaluation of its own), else the first stage of an
    // attempted step and then the rest of the step.
    template <bool kConfined>
    __device__ static int advance(Lane &ln, const Consts &k, const float4 *w, int nw,
                                  PoolBits &bits, bool raster)
    {
        const Field32 f = field(k, w, nw, kConfined);
        const Integ32 I = integ(k);
        double dtd;
        float dt, px, py, kx, ky;
        bool dry = false;   // an evaluation raised AquiferError (unconfined only)
        first_point(I, ln.tr, ln.fresh, dtd, dt, px, py);
        const bool far = kConfined && k.far.order > 0;
        if (far)
            LaneEvalFar{f, k.far}(px, py, kx, ky);
        else
            feval32<kConfined>(f, px, py, kx, ky, dry);
        if (ln.fresh) {
            ln.fresh = false;
            if (dry) {   // k1 raised: the trace is its start vertex (capturezone.py:218, 249)
                ln.tr.status = OKA_TRACE_TRUNCATED;
                return 2;
            }
            return take_k1(I, ln.tr, kx, ky) ? 0 : 2;
        }
        float ox, oy;
        const int res = far ? finish32(LaneEvalFar{f, k.far}, I, ln.tr, dtd, dt, kx, ky, ox, oy)
                            : finish32(LaneEval<kConfined>{f, dry}, I, ln.tr, dtd, dt, kx, ky, ox,
                                       oy);
        // a stage raised: the step is not accepted and the trace ends
        if (!kConfined && dry)
            ln.tr.status = OKA_TRACE_TRUNCATED;
        if ((res & 1) && raster) {
            bits.last_t = ln.last_t;
            bits.last_v = ln.last_v;
            ln.cand += (unsigned long long)raster32(bits, ln.ep, k.gdx, k.gdy, k.igdx, k.igdy,
                                                    k.umbra, k.u2, ox, oy, ln.tr.x, ln.tr.y,
                                                    !ln.a_done, ln.a_done);
            ln.last_t = bits.last_t;
            ln.last_v = bits.last_v;
        }
        return res;
    }

    __device__ static void summary(const Lane &ln, const TraceArgs &A, long long &att,
                                   long long &acc, unsigned long long &cand, int &status,
                                   double bb[4], bool &escaped)
    {
        const Trace32 &t = ln.tr;
        att = t.att;
        acc = t.acc;
        cand = ln.cand;
        status = t.status;
        bb[0] = A.xo + (double)t.bx0;
        bb[1] = A.xo + (double)t.bx1;
        bb[2] = A.yo + (double)t.by0;
        bb[3] = A.yo + (double)t.by1;
        const float gx0 = (float)(A.gxmin - A.xo), gx1 = (float)(A.gxmax - A.xo);
        const float gy0 = (float)(A.gymin - A.yo), gy1 = (float)(A.gymax - A.yo);
        escaped = !(t.bx0 > gx0 && t.bx1 < gx1 && t.by0 > gy0 && t.by1 < gy1);
    }
};

}  // namespace

size_t pool_lds_bytes(int precision, int slots, int max_nwells, int pool_tiles)
{
    if (precision == OKA_FP64_EXACT)
        return pool64_lds_bytes(slots, max_nwells, pool_tiles);
    return pool::lds_bytes<PoolFP32>(slots, max_nwells, pool_tiles);
}

int pool_max_tiles() { return (int)((1u << (32 - pool::kValBits)) - 1u); }
int pool_max_pool_tiles() { return (int)pool::kValHbm; }

hipError_t launch_trace_pool(const PoolArgs &p, int waves, int precision, int conf_mode,
                             hipStream_t s)
{
    if (precision == OKA_FP64_EXACT)
        return launch_trace_pool64(p, waves, conf_mode, s);
    return pool::launch<PoolFP32>(p, waves, conf_mode, s);
}

}  // namespace oka
