// oka_fp64.hip -- FP64-exact kernels for gfx950.
//
// Every floating-point expression (in this file and oka_fp64_dev.h) is
// evaluated in the reference's operation order with contraction disabled, so
// that for identical inputs the results are bit-identical to the reference's
// numpy scalar arithmetic:
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

#include "oka_fp64_dev.h"
#include "oka_internal.h"
#include "oka_math.h"

#pragma clang fp contract(off)

namespace oka {
namespace {

using namespace fp64;


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
// Both tracers are held to 128 VGPRs (4 waves per SIMD).  The confined one
// spills one 64-bit value, stored before the realization loop and read after
// it, and is 3 % faster than with the 150 the compiler picks.  The unconfined
// one spills 256 bytes per lane but runs 13 % faster than with the 219 VGPRs
// (2 waves per SIMD) the compiler picks: its correctly rounded logarithms
// wait on table loads, which more waves hide.
template <bool kConfined, int kMode>
__global__ void __launch_bounds__(kTraceBlock)
__attribute__((amdgpu_waves_per_eu(4))) k_trace_fp64(TraceArgs a)
{
    constexpr bool kVerts = kMode == kModeVertices;
    extern __shared__ double smem[];
    Well64 *s_wells = reinterpret_cast<Well64 *>(smem);
    __shared__ int s_next;
    __shared__ unsigned long long s_stat[kStatCount];

    const int tid = threadIdx.x;
    const int P = a.npaths;
    const Integ I = make_integ(a);
    const double umbra = a.umbra;
    const double u2 = umbra * umbra;
    const bool raster = !kVerts && a.rasterize;
    uint32_t *lds_tail = reinterpret_cast<uint32_t *>(s_wells + a.nwells);
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
        int ok = 1;
        for (int i = tid; i < a.nwells; i += blockDim.x) {
            Well64 wl;
            wl.x = a.wells_xy[2 * i];
            wl.y = a.wells_xy[2 * i + 1];
            wl.nq = -a.q[(int64_t)r * a.nwells + i];
            wl.pad = 0.0;
            s_wells[i] = wl;
            ok &= plain_coord(wl.x) && plain_coord(wl.y) && plain_strength(wl.nq);
        }
        if (raster)
            ch.begin(a);
        if (tid == 0)
            s_next = blockDim.x;
        const bool plain = __syncthreads_and(ok) != 0;

        const Field64 f = make_field(a, r, s_wells, plain);
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
            pw = pw + (wl[3] * log_rn(r2)) * kC4Pi;
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

size_t trace_fp64_wells_lds(int nwells) { return (size_t)nwells * sizeof(Well64); }

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
double host_log(double x) { return log_rn(x); }

}  // namespace oka
