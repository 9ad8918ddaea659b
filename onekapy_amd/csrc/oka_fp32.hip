// oka_fp32.hip -- FP32 particle tracing for gfx950 (confined or unconfined
// velocity): one 256-lane workgroup per realization, grid-stride over
// realizations.  The step,
// field and rasterizer are in oka_fp32_dev.h.  Results are not bit-identical
// to the FP64 path; the difference in capture probability is reported against
// the FP64-exact mode.
#include <math.h>

#include <type_traits>

#include "oka_fp32_dev.h"
#include "oka_internal.h"

namespace oka {
namespace {

using namespace fp32;

// At most 128 VGPRs, so that 4 waves share a SIMD: the tracer is bound by
// VALU issue, and with 3 waves (the compiler's own choice, 129 VGPRs) the
// SIMDs idle about 12 % of the time on LDS and scalar waits.
template <int kKind, bool kConfined>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_trace_fp32(TraceArgs a)
{
    extern __shared__ float4 smem4[];
    float4 *s_w = smem4;
    uint32_t *s_bm = reinterpret_cast<uint32_t *>(smem4 + a.nwells);
    using Bits = typename std::conditional<kKind == kBmTiled, TileBits, FlatBits>::type;
    __shared__ int s_next;
    __shared__ unsigned long long s_stat[kStatCount];
    __shared__ float s_reg[11];

    const int tid = threadIdx.x;
    const int P = a.npaths;
    Integ32 I;
    I.dur = a.duration;
    I.absdur = fabs(a.duration);
    I.sg = a.duration > 0.0 ? 1.0 : (a.duration < 0.0 ? -1.0 : (a.duration == 0.0 ? 0.0 : (double)NAN));
    I.tol = (float)a.tol;
    I.maxstep = (float)a.maxstep;
    I.cap = a.max_steps;
    const float umbra = (float)a.umbra;
    const float u2 = umbra * umbra;
    const float gdx = (float)a.gdx, gdy = (float)a.gdy;
    const float igdx = (float)(1.0 / a.gdx), igdy = (float)(1.0 / a.gdy);
    const float gx0 = (float)(a.gxmin - a.xo), gx1 = (float)(a.gxmax - a.xo);
    const float gy0 = (float)(a.gymin - a.yo), gy1 = (float)(a.gymax - a.yo);
    Bits bits;
    uint32_t *lds_tail = s_bm;
    if constexpr (kKind == kBmTiled) {
        bits = tile_bits(a, s_bm);
        tile_init(bits);
        lds_tail += tile_lds_words(a.ntiles, a.tile_slots);
    } else if constexpr (kKind == kBmLdsFlat) {
        bits = FlatBits{s_bm, a.gncols};
        lds_tail += a.rasterize ? a.nwords : 0;
    } else {
        bits = FlatBits{a.gbitmap + (int64_t)blockIdx.x * a.nwords, a.gncols};
    }
    unsigned int *s_win = nullptr;
    if (a.rasterize && a.wmarks) {
        s_win = lds_tail;
        for (int c = threadIdx.x; c < a.w_nr * a.w_nc; c += blockDim.x)
            s_win[c] = 0u;
    }

    for (int r = blockIdx.x; r < a.nreal; r += gridDim.x) {
        if (tid < kStatCount)
            s_stat[tid] = 0;
        const double bn = a.b[r] * a.n[r];
        // h = K b^2 / 2, the potential at which the head reaches the top of
        // the aquifer (model.py:344)
        const double h = kConfined ? 1.0 : 0.5 * a.K[r] * (a.b[r] * a.b[r]);
        for (int i = tid; i < a.nwells; i += blockDim.x) {
            const double qw = a.q[(int64_t)r * a.nwells + i];
            s_w[i] = make_float4((float)(a.wells_xy[2 * i] - a.xo),
                                 (float)(a.wells_xy[2 * i + 1] - a.yo),
                                 (float)(qw * kInv2Pi / bn),
                                 kConfined ? 0.0f : (float)(qw * (kInv4Pi * kLn2) / h));
        }
        if (tid == 0) {
            const double *c = a.coef + 6 * (int64_t)r;
            s_reg[0] = (float)(2.0 * c[0] / bn);
            s_reg[1] = (float)(2.0 * c[1] / bn);
            s_reg[2] = (float)(c[2] / bn);
            s_reg[3] = (float)(c[3] / bn);
            s_reg[4] = (float)(c[4] / bn);
            for (int j = 0; j < 6; ++j)
                s_reg[5 + j] = (float)(c[j] / h);
            s_next = blockDim.x;
        }
        if constexpr (kKind != kBmTiled) {   // a tiled bitmap is left empty by its flush
            if (a.rasterize)
                for (int w = tid; w < a.nwords; w += blockDim.x)
                    bits.bm[w] = 0u;
        }
        __syncthreads();

        Field32 f;
        f.w = s_w;
        f.nw = a.nwells;
        f.a2 = s_reg[0];
        f.b2 = s_reg[1];
        f.c = s_reg[2];
        f.d = s_reg[3];
        f.e = s_reg[4];
        if constexpr (!kConfined) {
            f.pa = s_reg[5];
            f.pb = s_reg[6];
            f.pc = s_reg[7];
            f.pd = s_reg[8];
            f.pe = s_reg[9];
            f.pf = s_reg[10];
        }
        // per-trace statistics go to LDS when the trace ends (few registers)
        unsigned long long n_cand = 0;   // candidate nodes of the lane's trace

        Trace32 s;
        Epoch32 ep;
        int p = tid;
        bool have = p < P;
        bool live = false, fresh = false;
        bool a_done = false;   // the disk around the trace's current point is marked
        bool dry = false;      // an evaluation of the trace raised AquiferError
        for (;;) {
            bool ended = false;
            const int64_t k = (int64_t)r * P + p;
            if (!live) {
                if (!have)
                    break;
                if (a.rasterize)
                    ep = find_epoch32(a, k);
                open32(s, I, (float)(a.start_xy[2 * p] - a.xo),
                       (float)(a.start_xy[2 * p + 1] - a.yo));
                live = fresh = true;
                a_done = false;
                dry = false;
            }
            // one field evaluation for every lane: k1 of a fresh trace, else
            // the first stage of an attempted step
            double dtd;
            float dt, px, py, kx, ky;
            first_point(I, s, fresh, dtd, dt, px, py);
            feval32<kConfined>(f, px, py, kx, ky, dry);
            if (fresh) {
                fresh = false;
                if (dry) {   // k1 raised: the trace is its start vertex (capturezone.py:218, 249)
                    s.status = OKA_TRACE_TRUNCATED;
                    live = false;
                    ended = true;
                } else if (!take_k1(I, s, kx, ky)) {
                    live = false;
                    ended = true;
                }
            } else {
                float ox, oy;
                const int res = finish32(LaneEval<kConfined>{f, dry}, I, s, dtd, dt, kx, ky, ox, oy);
                // a stage raised: the step is not accepted and the trace ends
                if (!kConfined && dry)
                    s.status = OKA_TRACE_TRUNCATED;
                if ((res & 1) && a.rasterize)
                    n_cand += (unsigned long long)raster32(bits, ep, gdx, gdy, igdx, igdy, umbra,
                                                           u2, ox, oy, s.x, s.y, !a_done, a_done);
                if (res & 2) {
                    live = false;
                    ended = true;
                }
            }
            if (ended) {
                atomicAdd(&s_stat[kStatAttempted], (unsigned long long)s.att);
                atomicAdd(&s_stat[kStatAccepted], (unsigned long long)s.acc);
                atomicAdd(&s_stat[kStatStatus0 + (s.status & 3)], 1ull);
                atomicAdd(&s_stat[kStatCandidates], n_cand);
                n_cand = 0;
                if (!(s.bx0 > gx0 && s.bx1 < gx1 && s.by0 > gy0 && s.by1 < gy1))
                    atomicAdd(&s_stat[kStatEscaped], 1ull);
                if (a.bbox) {
                    a.bbox[4 * k] = a.xo + (double)s.bx0;
                    a.bbox[4 * k + 1] = a.xo + (double)s.bx1;
                    a.bbox[4 * k + 2] = a.yo + (double)s.by0;
                    a.bbox[4 * k + 3] = a.yo + (double)s.by1;
                }
                if (a.trace_status)
                    a.trace_status[k] = s.status;
                if (a.trace_steps)
                    a.trace_steps[k] = s.att;
                p = atomicAdd(&s_next, 1);
                have = p < P;
            }
        }
        __syncthreads();
        if (a.rasterize) {
            if constexpr (kKind == kBmTiled)
                atomicAdd(&s_stat[kStatVisited], flush_tiles(a, bits, s_win));
            else
                atomicAdd(&s_stat[kStatVisited], flush_bitmap(a, bits.bm, s_win));
        }
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

// The FP32 tracers' segment rasterizer over the segments of one polyline (one
// per lane), in coordinates relative to the grid's centre node.
__global__ void __launch_bounds__(256) k_raster_fp32(RasterArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nseg)
        return;
    const double ox = a.xmin + (double)(a.ncols / 2) * a.dx;
    const double oy = a.ymin + (double)(a.nrows / 2) * a.dy;
    const Epoch32 e{(float)(a.xmin - ox), (float)(a.ymin - oy), a.nrows, a.ncols, 0, 0};
    FlatBits bits{a.bitmap, a.ncols};
    const float umbra = (float)a.umbra;
    bool covered;
    raster32(bits, e, (float)a.dx, (float)a.dy, (float)(1.0 / a.dx), (float)(1.0 / a.dy), umbra,
             umbra * umbra, (float)(a.xy[2 * i] - ox), (float)(a.xy[2 * i + 1] - oy),
             (float)(a.xy[2 * i + 2] - ox), (float)(a.xy[2 * i + 3] - oy), true, covered);
}

}  // namespace

size_t trace_fp32_wells_lds(int nwells) { return (size_t)nwells * sizeof(float4); }

hipError_t launch_trace_fp32(const TraceArgs &a, int grid_blocks, int kind, hipStream_t s)
{
    size_t lds = trace_fp32_wells_lds(a.nwells) + (a.wmarks ? (size_t)a.w_nr * a.w_nc * 4 : 0);
    dim3 grid(grid_blocks), block(kTraceBlock);
    if (!a.rasterize || kind == kBmLdsFlat) {
        lds += a.rasterize ? (size_t)a.nwords * 4 : 0;
        if (a.confined)
            hipLaunchKernelGGL((k_trace_fp32<kBmLdsFlat, true>), grid, block, lds, s, a);
        else
            hipLaunchKernelGGL((k_trace_fp32<kBmLdsFlat, false>), grid, block, lds, s, a);
    } else if (kind == kBmTiled) {
        lds += (size_t)tile_lds_words(a.ntiles, a.tile_slots) * 4;
        if (a.confined)
            hipLaunchKernelGGL((k_trace_fp32<kBmTiled, true>), grid, block, lds, s, a);
        else
            hipLaunchKernelGGL((k_trace_fp32<kBmTiled, false>), grid, block, lds, s, a);
    } else if (a.confined) {
        hipLaunchKernelGGL((k_trace_fp32<kBmHbmFlat, true>), grid, block, lds, s, a);
    } else {
        hipLaunchKernelGGL((k_trace_fp32<kBmHbmFlat, false>), grid, block, lds, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_raster_fp32(const RasterArgs &a, hipStream_t s)
{
    const int64_t blocks = (a.nseg + 255) / 256;
    hipLaunchKernelGGL(k_raster_fp32, dim3((unsigned)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace oka
