// oka_pool64.hip -- the persistent pool tracer (oka_pool_impl.h) in the
// FP64-exact arithmetic of the workgroup tracer (oka_fp64_dev.h: the same
// feval, Dormand-Prince step and per-node rasterizer, in the reference's
// operation order with contraction off), so that the pool's traces, counts
// and statistics are those of oka_fp64.hip bit for bit: each trace's
// arithmetic is lane-local and the schedule only changes which lane runs it.
#include <math.h>

#include "oka_fp64_dev.h"
#include "oka_internal.h"
#include "oka_pool_impl.h"

#pragma clang fp contract(off)

namespace oka {
namespace {

using namespace fp64;
using namespace pool;

struct PoolFP64 {
    using Well = Well64;
    struct Consts {
        double xo, yo, A, B, C, D, E, F, A2, B2, K, bt, bn, por, half_kb2;   // Field64
        double dur, absdur, tol, maxstep, sg;                               // Integ
        double gdx, gdy, umbra, u2;
        long long cap;
        int plain, pad;
    };
    struct Lane {
        Trace64 s;
        EpochReg ep;
        double xs, ys;             // start point of a fresh trace
        unsigned long long cand;   // candidate nodes of the trace
        uint32_t last_t, last_v;   // the lane's tile cache (PoolBits::lookup)
        bool fresh;                // not yet started (k1 not evaluated)
    };
    static constexpr int kWavesPerEU = 2;

    // Realization r's wells (x, y, -q); returns whether they all allow the
    // short division sequence (plain_coord, plain_strength).
    __device__ static bool stage_wells(Well64 *w, const TraceArgs &A, int64_t r)
    {
        int ok = 1;
        for (int i = threadIdx.x; i < A.nwells; i += kWave) {
            Well64 wl;
            wl.x = A.wells_xy[2 * i];
            wl.y = A.wells_xy[2 * i + 1];
            wl.nq = -A.q[r * A.nwells + i];
            wl.pad = 0.0;
            w[i] = wl;
            ok &= plain_coord(wl.x) && plain_coord(wl.y) && plain_strength(wl.nq);
        }
        return __ballot(!ok) == 0ull;
    }

    // make_field and make_integ of the workgroup tracer, per slot.
    __device__ static void stage_consts(Consts &k, const PoolArgs &, const TraceArgs &A, int64_t r,
                                        int, bool plain)
    {
        const double *c = A.coef + 6 * r;
        k.xo = A.xo;
        k.yo = A.yo;
        k.A = c[0];
        k.B = c[1];
        k.C = c[2];
        k.D = c[3];
        k.E = c[4];
        k.F = c[5];
        k.A2 = 2.0 * k.A;
        k.B2 = 2.0 * k.B;
        k.K = A.K[r];
        k.bt = A.b[r];
        k.por = A.n[r];
        k.bn = k.bt * k.por;
        k.half_kb2 = (0.5 * k.K) * (k.bt * k.bt);
        k.dur = A.duration;
        k.absdur = fabs(A.duration);
        k.tol = A.tol;
        k.maxstep = A.maxstep;
        k.sg = A.duration > 0.0 ? 1.0
                                : (A.duration < 0.0 ? -1.0 : (A.duration == 0.0 ? 0.0 : (double)NAN));
        k.cap = A.max_steps;
        k.gdx = A.gdx;
        k.gdy = A.gdy;
        k.umbra = A.umbra;
        k.u2 = A.umbra * A.umbra;
        k.plain = plain ? 1 : 0;
    }

    __device__ static bool stage_slot(const Consts &, const Well64 *, int) { return false; }
    __device__ static bool table_shared(const Consts &) { return false; }
    __device__ static void table_to(Consts &, float2 *) {}

    __device__ static Field64 field(const Consts &k, const Well64 *w, int nw)
    {
        Field64 f;
        f.w = w;
        f.nw = nw;
        f.plain = k.plain != 0;
        f.xo = k.xo;
        f.yo = k.yo;
        f.A = k.A;
        f.B = k.B;
        f.C = k.C;
        f.D = k.D;
        f.E = k.E;
        f.F = k.F;
        f.A2 = k.A2;
        f.B2 = k.B2;
        f.K = k.K;
        f.bt = k.bt;
        f.bn = k.bn;
        f.por = k.por;
        f.half_kb2 = k.half_kb2;
        return f;
    }

    __device__ static Integ integ(const Consts &k)
    {
        Integ I;
        I.dur = k.dur;
        I.absdur = k.absdur;
        I.tol = k.tol;
        I.maxstep = k.maxstep;
        I.sg = k.sg;
        I.cap = k.cap;
        return I;
    }

    __device__ static void open(Lane &ln, const Consts &, const TraceArgs &A, int64_t r, int p,
                                bool raster)
    {
        if (raster)
            ln.ep = find_epoch(A, r * A.npaths + p);
        ln.xs = A.start_xy[2 * p];
        ln.ys = A.start_xy[2 * p + 1];
        ln.cand = 0;
        ln.last_t = kKeyEmpty;
        ln.last_v = kValNone;
        ln.fresh = true;
    }

    // A fresh trace evaluates k1 at its start (capturezone.py:211-219); every
    // other iteration is one attempted step (capturezone.py:221-247).
    template <bool kConfined>
    __device__ static int advance(Lane &ln, const Consts &k, const Well64 *w, int nw,
                                  PoolBits &bits, bool raster)
    {
        const Field64 f = field(k, w, nw);
        const Integ I = integ(k);
        if (ln.fresh) {
            ln.fresh = false;
            return dp5_start<kConfined>(f, I, ln.s, ln.xs, ln.ys) ? 0 : 2;
        }
        double ox, oy;
        const int res = dp5_attempt<kConfined>(f, I, ln.s, ox, oy);
        if ((res & 1) && raster) {
            bits.last_t = ln.last_t;
            bits.last_v = ln.last_v;
            ln.cand += (unsigned long long)raster_segment(bits, ln.ep, k.gdx, k.gdy, k.umbra, k.u2,
                                                          ox, oy, ln.s.x, ln.s.y);
            ln.last_t = bits.last_t;
            ln.last_v = bits.last_v;
        }
        return res;
    }

    __device__ static void summary(const Lane &ln, const TraceArgs &A, long long &att,
                                   long long &acc, unsigned long long &cand, int &status,
                                   double bb[4], bool &escaped)
    {
        const Trace64 &s = ln.s;
        att = s.att;
        acc = s.acc;
        cand = ln.cand;
        status = s.status;
        bb[0] = s.bx0;
        bb[1] = s.bx1;
        bb[2] = s.by0;
        bb[3] = s.by1;
        escaped = !(s.bx0 > A.gxmin && s.bx1 < A.gxmax && s.by0 > A.gymin && s.by1 < A.gymax);
    }
};

}  // namespace

size_t pool64_lds_bytes(int slots, int max_nwells, int pool_tiles)
{
    return pool::lds_bytes<PoolFP64>(slots, max_nwells, pool_tiles);
}

hipError_t launch_trace_pool64(const PoolArgs &p, int waves, int conf_mode, hipStream_t s)
{
    return pool::launch<PoolFP64>(p, waves, conf_mode, s);
}

}  // namespace oka
