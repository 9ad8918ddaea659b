// oka_fp32.hip -- FP32 particle tracing for gfx950 (confined velocity): one
// 256-lane workgroup per realization, grid-stride over realizations.  The step,
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

template <int kKind>
__global__ void __launch_bounds__(kTraceBlock) k_trace_fp32(TraceArgs a)
{
    extern __shared__ float4 smem4[];
    float4 *s_w = smem4;
    uint32_t *s_bm = reinterpret_cast<uint32_t *>(smem4 + a.nwells);
    using Bits = typename std::conditional<kKind == kBmTiled, TileBits, FlatBits>::type;
    __shared__ int s_next;
    __shared__ unsigned long long s_stat[kStatCount];
    __shared__ float s_reg[5];

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
        for (int i = tid; i < a.nwells; i += blockDim.x) {
            const double qw = a.q[(int64_t)r * a.nwells + i];
            s_w[i] = make_float4((float)(a.wells_xy[2 * i] - a.xo),
                                 (float)(a.wells_xy[2 * i + 1] - a.yo),
                                 (float)(qw * kInv2Pi / bn), 0.0f);
        }
        if (tid == 0) {
            const double *c = a.coef + 6 * (int64_t)r;
            s_reg[0] = (float)(2.0 * c[0] / bn);
            s_reg[1] = (float)(2.0 * c[1] / bn);
            s_reg[2] = (float)(c[2] / bn);
            s_reg[3] = (float)(c[3] / bn);
            s_reg[4] = (float)(c[4] / bn);
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
        unsigned long long n_att = 0, n_acc = 0, n_esc = 0, n_cand = 0;
        unsigned long long n_st[4] = {0, 0, 0, 0};

        Trace32 s;
        Epoch32 ep;
        int p = tid;
        bool have = p < P;
        bool live = false;
        for (;;) {
            bool ended = false;
            const int64_t k = (int64_t)r * P + p;
            if (!live) {
                if (!have)
                    break;
                if (a.rasterize)
                    ep = find_epoch32(a, k);
                live = start32(f, I, s, (float)(a.start_xy[2 * p] - a.xo),
                               (float)(a.start_xy[2 * p + 1] - a.yo));
                ended = !live;
            } else {
                float ox, oy;
                const int res = attempt32(f, I, s, ox, oy);
                if ((res & 1) && a.rasterize)
                    n_cand += (unsigned long long)raster32(bits, ep, gdx, gdy, igdx, igdy, umbra,
                                                           u2, ox, oy, s.x, s.y);
                if (res & 2) {
                    live = false;
                    ended = true;
                }
            }
            if (ended) {
                n_att += (unsigned long long)s.att;
                n_acc += (unsigned long long)s.acc;
                n_st[s.status & 3] += 1;
                if (!(s.bx0 > gx0 && s.bx1 < gx1 && s.by0 > gy0 && s.by1 < gy1))
                    n_esc += 1;
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
        atomicAdd(&s_stat[kStatAttempted], n_att);
        atomicAdd(&s_stat[kStatAccepted], n_acc);
        atomicAdd(&s_stat[kStatEscaped], n_esc);
        atomicAdd(&s_stat[kStatCandidates], n_cand);
        for (int j = 0; j < 4; ++j)
            if (n_st[j])
                atomicAdd(&s_stat[kStatStatus0 + j], n_st[j]);
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

}  // namespace

size_t trace_fp32_wells_lds(int nwells) { return (size_t)nwells * sizeof(float4); }

hipError_t launch_trace_fp32(const TraceArgs &a, int grid_blocks, int kind, hipStream_t s)
{
    size_t lds = trace_fp32_wells_lds(a.nwells) + (a.wmarks ? (size_t)a.w_nr * a.w_nc * 4 : 0);
    dim3 grid(grid_blocks), block(kTraceBlock);
    if (!a.rasterize || kind == kBmLdsFlat) {
        lds += a.rasterize ? (size_t)a.nwords * 4 : 0;
        hipLaunchKernelGGL(k_trace_fp32<kBmLdsFlat>, grid, block, lds, s, a);
    } else if (kind == kBmTiled) {
        lds += (size_t)tile_lds_words(a.ntiles, a.tile_slots) * 4;
        hipLaunchKernelGGL(k_trace_fp32<kBmTiled>, grid, block, lds, s, a);
    } else {
        hipLaunchKernelGGL(k_trace_fp32<kBmHbmFlat>, grid, block, lds, s, a);
    }
    return hipGetLastError();
}

}  // namespace oka
