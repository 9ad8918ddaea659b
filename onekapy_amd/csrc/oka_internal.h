// oka_internal.h -- kernel argument blocks and launchers shared by the HIP
// translation units of liboka.so.  Not part of the public ABI (include/oka.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "oka.h"

namespace oka {

constexpr int kTraceBlock = 256;            // lanes per realization workgroup
constexpr int kLdsBitmapMaxWords = 14336;   // 56 KiB: at most 2 realization blocks per CU

// Device-side statistics accumulators (unsigned 64-bit, atomically added).
enum StatSlot {
    kStatAttempted = 0,
    kStatAccepted = 1,
    kStatStatus0 = 2,     // 2..5: status counts
    kStatEscaped = 6,
    kStatVisited = 7,
    kStatCandidates = 8,  // candidate nodes examined by the segment rasterizer
    kStatCount = 9
};

struct TraceArgs {
    // site (device pointers)
    int32_t nwells;
    int32_t npaths;
    const double *wells_xy;     // nwells x 2
    const double *start_xy;     // npaths x 2
    double xo, yo;
    // realizations
    int32_t nreal;
    int32_t confined;
    const double *K, *n, *b;    // nreal
    const double *q;            // nreal x nwells
    const double *coef;         // nreal x 6
    // integrator
    double duration, tol, maxstep, umbra;
    int64_t max_steps;
    // grid: the final lattice and its expansion epochs
    double gxmin, gymin, gxmax, gymax, gdx, gdy;
    int32_t gnrows, gncols;
    int32_t nepochs;            // >= 1 (a fixed grid is a single epoch)
    const oka_epoch *epochs;
    int32_t rasterize;
    int32_t nwords;             // bitmap words per realization
    uint32_t *gbitmap;          // per-block bitmaps in HBM when the LDS cannot hold one
    int32_t *counts;            // gnrows x gncols
    // optional per-trace outputs
    double *bbox;               // ntraces x 4
    int32_t *trace_status;
    int64_t *trace_steps;
    unsigned long long *stats;  // kStatCount
    // vertices mode
    int64_t vcap;
    double *verts;              // ntraces x vcap x 2
    int64_t *nverts;
    // per-realization window marks (oka_trace_marks); wmarks == nullptr: off
    int32_t w_row0, w_col0, w_cell, w_nr, w_nc;
    uint8_t *wmarks;            // nreal x (w_nr * w_nc)
};

// Flush of one realization's visited-node bitmap (probabilityfield.py:357-359):
// one atomic add per set bit into the count grid, and, when a window is
// requested, the per-cell number of visited fine nodes into LDS counters.
// Called by every lane of the workgroup; returns this lane's share of the
// visited count.
__device__ inline unsigned long long flush_bitmap(const TraceArgs &a, const uint32_t *bm,
                                                  unsigned int *s_win)
{
    unsigned long long vis = 0;
    for (int w = threadIdx.x; w < a.nwords; w += blockDim.x) {
        uint32_t bits = bm[w];
        vis += (unsigned long long)__popc(bits);
        while (bits) {
            const int bit = __ffs(bits) - 1;
            bits &= bits - 1u;
            const int idx = w * 32 + bit;
            if (a.counts)
                atomicAdd(&a.counts[idx], 1);
            if (s_win) {
                const int i = idx / a.gncols, j = idx - i * a.gncols;
                const int ci = i - a.w_row0, cj = j - a.w_col0;
                if (ci >= 0 && cj >= 0) {
                    const int r = ci / a.w_cell, c = cj / a.w_cell;
                    if (r < a.w_nr && c < a.w_nc)
                        atomicAdd(&s_win[r * a.w_nc + c], 1u);
                }
            }
        }
    }
    return vis;
}

// Writes the window counters of realization r (capped at 255) and clears them.
__device__ inline void store_window(const TraceArgs &a, int r, unsigned int *s_win)
{
    const int n = a.w_nr * a.w_nc;
    uint8_t *dst = a.wmarks + (int64_t)r * n;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const unsigned int v = s_win[c];
        dst[c] = (uint8_t)(v < 255u ? v : 255u);
        s_win[c] = 0u;
    }
}

struct EvalArgs {
    int32_t kind;
    int32_t nwells;
    const double *wells;        // nwells x 4
    double base, K, n, b, xo, yo;
    double coef[6];
    int64_t npts;
    const double *xy;
    double *out;
    int32_t *status;
};

struct RasterArgs {
    double xmin, ymin, dx, dy, umbra;
    int32_t nrows, ncols;
    int64_t nseg;
    const double *xy;           // (nseg + 1) x 2
    uint32_t *bitmap;           // ceil(nrows*ncols/32) words
};

struct FitArgs {
    int32_t nwells, nobs, nreal;
    const double *wells_xy;     // nwells x 2
    const double *obs;          // nobs x 4
    double xo, yo, base;
    const double *logr2;        // nobs x nwells: log(r^2) observation to well
    const double *K, *b, *q, *z;
    double *ev, *cov, *coef;
    int32_t *status;
};

struct DrawArgs {
    int32_t nwells;
    const oka_dist *dists;      // device: nwells + 3
    int64_t n, first;
    uint64_t seed;
    double *q, *K, *por, *thick, *z;
};

struct FeatureArgs {
    int32_t nwells;
    int64_t n;
    const double *q, *K, *por, *thick, *coef;
    const double *mean, *std;   // device: nwells + 9
    float *x;                   // n x (nwells + 9)
};

// oka_rng.hip
hipError_t launch_draw(const DrawArgs &a, hipStream_t s);
hipError_t launch_features(const FeatureArgs &a, hipStream_t s);
// sums[c] += sum over rows of marks[r, c], sums_sq[c] += sum of marks[r, c]^2
hipError_t launch_mark_moments(const uint8_t *marks, int64_t n, int32_t ncell, uint64_t *sums,
                               uint64_t *sums_sq, hipStream_t s);
uint32_t host_philox_word(uint64_t seed, int64_t r, uint32_t slot, uint32_t sub, int word);

// oka_fp64.hip
hipError_t launch_trace_fp64(const TraceArgs &a, int grid_blocks, bool lds_bitmap,
                             hipStream_t s);
hipError_t launch_vertices_fp64(const TraceArgs &a, hipStream_t s);
hipError_t launch_eval_fp64(const EvalArgs &a, hipStream_t s);
hipError_t launch_raster_fp64(const RasterArgs &a, hipStream_t s);
size_t trace_fp64_lds_bytes(int nwells, int nwords, bool lds_bitmap);

// oka_fp32.hip
hipError_t launch_trace_fp32(const TraceArgs &a, int grid_blocks, bool lds_bitmap,
                             hipStream_t s);
size_t trace_fp32_lds_bytes(int nwells, int nwords, bool lds_bitmap);

// oka_fit.hip
hipError_t launch_fit_logs(const FitArgs &a, double *logr2, hipStream_t s);
hipError_t launch_fit_draw(const FitArgs &a, hipStream_t s);

double host_pow_fifth(double x);
double host_log(double x);

}  // namespace oka
