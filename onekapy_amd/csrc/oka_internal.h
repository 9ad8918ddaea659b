// oka_internal.h -- kernel argument blocks and launchers shared by the HIP
// translation units of liboka.so.  Not part of the public ABI (include/oka.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "oka.h"

namespace oka {

constexpr int kTraceBlock = 256;            // lanes per realization workgroup
constexpr int kLdsBitmapMaxWords = 14336;   // 56 KiB: at most 2 realization blocks per CU
constexpr int kLdsBudget = 57344;           // LDS bytes per trace workgroup (2 per CU)

// Bitmap placement chosen by the launcher (OKA_BITMAP_* minus one).
enum BitmapKind { kBmLdsFlat = 0, kBmTiled = 1, kBmHbmFlat = 2 };
constexpr int kTileEmpty = -1, kTileOverflow = -2;
constexpr int kMinTileSlots = 16;

// Far-well expansions of the FP32 pool (fp32::FarField): the number of
// coefficients each box keeps and evaluates (those above a plan's order are
// zero), which is also the largest order and the one oka_trace_opts.far_order
// = 0 chooses from kFarAutoWells wells on.
constexpr int kFarTerms = 16;
constexpr int kFarAutoWells = 64;
// A plan's boxes hold about kFarNear wells in each 3 x 3 neighbourhood, with
// at most kFarMaxSide boxes a side.
constexpr int kFarNear = 2;
constexpr int kFarMaxSide = 64;

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
    int32_t nwords;             // bitmap words per realization (flat modes)
    uint32_t *gbitmap;          // per-block bitmaps in HBM: flat, or the overflow tiles
    int32_t tiles_x, ntiles;    // tiled mode: 32 x 32-node tiles of the final grid
    int32_t tile_slots;         // tiled mode: LDS slots per workgroup
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
    uint16_t *wmasks;           // nreal x (w_nr * w_nc) node masks (w_cell <= 4), or nullptr
    // FP32 pool: far wells as local expansions of this order (0: every well
    // summed directly); boxes of side far_len from (far_x0, far_y0), in
    // coordinates relative to (xo, yo) (fp32::FarField)
    int32_t far_order, far_nbx, far_nby, far_near_stride;
    float far_x0, far_y0, far_len;
    const int32_t *far_near;        // [far_nbx * far_nby][far_near_stride]: per box
                                    // the near-well count, their indices, then -1
    const int32_t *far_wbox;        // [nwells][2]: each well's box column and row
};

// One marked node of a realization: its count (probabilityfield.py:357-359)
// and, when a window is requested, its cell's LDS word: the count of the
// cell's marked nodes, or with node masks (cells of at most 4 x 4 nodes) the
// bit (row in cell) * cell + (column in cell).
__device__ inline void count_node(const TraceArgs &a, int row, int col, unsigned int *s_win)
{
    if (a.counts)
        atomicAdd(&a.counts[(int64_t)row * a.gncols + col], 1);
    if (s_win) {
        const int ci = row - a.w_row0, cj = col - a.w_col0;
        if (ci >= 0 && cj >= 0) {
            const int r = ci / a.w_cell, c = cj / a.w_cell;
            if (r < a.w_nr && c < a.w_nc) {
                if (a.wmasks)
                    atomicOr(&s_win[r * a.w_nc + c],
                             1u << ((ci - r * a.w_cell) * a.w_cell + (cj - c * a.w_cell)));
                else
                    atomicAdd(&s_win[r * a.w_nc + c], 1u);
            }
        }
    }
}

// The nodes of one tile row: bit b of `bits` is column col0 + b of `row`.
__device__ inline unsigned long long count_bits(const TraceArgs &a, int row, int col0,
                                                uint32_t bits, unsigned int *s_win)
{
    const unsigned long long n = (unsigned long long)__popc(bits);
    while (bits) {
        const int col = col0 + __ffs(bits) - 1;
        bits &= bits - 1u;
        count_node(a, row, col, s_win);
    }
    return n;
}

// n consecutive bits starting at bit b of a word (1 <= n <= 32 - b; for
// n = 32, 2 << 31 wraps to 0 and the mask is all ones).
__device__ __forceinline__ uint32_t span_mask(int n, int b)
{
    return ((2u << (n - 1)) - 1u) << b;
}

// A realization's visited-node bitmap over the whole final grid (LDS or HBM):
// node (row, col) is bit idx % 32 of word idx / 32, idx = row * ncols + col.
struct FlatBits {
    uint32_t *bm;
    int ncols;
    __device__ __forceinline__ bool test(int row, int col) const
    {
        const int idx = row * ncols + col;
        return (bm[idx >> 5] >> (idx & 31)) & 1u;
    }
    __device__ __forceinline__ void set(int row, int col)
    {
        const int idx = row * ncols + col;
        atomicOr(&bm[idx >> 5], 1u << (idx & 31));
    }
    // Row stride of the flat index (the span markers' `base` = row * stride).
    __device__ __forceinline__ int stride() const { return ncols; }
    // Marks columns c0..c1 (inclusive) of one row, whose first node has flat
    // index base = row * ncols, a word at a time; a word whose bits are all
    // set already is only read.
    __device__ __forceinline__ void mark(int /*row*/, int base, int c0, int c1)
    {
        int k = base + c0;
        const int end = base + c1;
        for (;;) {
            const int b = k & 31;
            const int n = min(end - k + 1, 32 - b);
            const uint32_t m = span_mask(n, b);
            uint32_t *w = &bm[k >> 5];
            if ((*w & m) != m)
                atomicOr(w, m);
            k += n;
            if (k > end)
                break;
        }
    }
};

// The bitmap in 32 x 32-node tiles (tile t = (row / 32) * tiles_x + col / 32;
// row r of a tile is one word, column c its bit c).  A tile takes an LDS slot
// the first time one of its nodes is marked; when the slots run out it lives
// in the workgroup's HBM region instead.  Unmarked tiles cost nothing.
struct TileBits {
    int *tmap;          // LDS [ntiles]: slot index, kTileEmpty or kTileOverflow
    int *slot_tile;     // LDS [slots]: tile of a claimed slot, or -1
    uint32_t *slots;    // LDS [slots][32]
    int *nslots;        // LDS: slots handed out (may exceed `slots`)
    uint32_t *hbm;      // [ntiles][32] of this workgroup
    int maxslots, tiles_x, ntiles;

    __device__ __forceinline__ uint32_t *word(int t, int r) const
    {
        const int s = __hip_atomic_load(&tmap[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (s == kTileEmpty)
            return nullptr;
        return s == kTileOverflow ? hbm + (int64_t)t * 32 + r : slots + s * 32 + r;
    }
    __device__ __forceinline__ bool test(int row, int col) const
    {
        const uint32_t *w = word((row >> 5) * tiles_x + (col >> 5), row & 31);
        return w != nullptr && ((*w >> (col & 31)) & 1u);
    }
    // The first lane to claim a tile maps it; a lane that loses the race uses
    // the winner's slot, and its own slot stays unmapped (and zero).
    __device__ __forceinline__ uint32_t *claim(int t)
    {
        const int s = atomicAdd(nslots, 1);
        const int v = s < maxslots ? s : kTileOverflow;
        const int prev = atomicCAS(&tmap[t], kTileEmpty, v);
        const int got = prev == kTileEmpty ? v : prev;
        if (prev == kTileEmpty && v >= 0)
            slot_tile[v] = t;
        return got == kTileOverflow ? hbm + (int64_t)t * 32 : slots + got * 32;
    }
    __device__ __forceinline__ void set(int row, int col)
    {
        const int t = (row >> 5) * tiles_x + (col >> 5);
        uint32_t *w = word(t, row & 31);
        if (w == nullptr)
            w = claim(t) + (row & 31);
        atomicOr(w, 1u << (col & 31));
    }
    __device__ __forceinline__ int stride() const { return 0; }
    // Marks columns c0..c1 (inclusive) of one row, one tile word at a time.
    __device__ __forceinline__ void mark(int row, int /*base*/, int c0, int c1)
    {
        const int tr = (row >> 5) * tiles_x, r = row & 31;
        for (int c = c0; c <= c1;) {
            const int b = c & 31;
            const int n = min(c1 - c + 1, 32 - b);
            const uint32_t m = span_mask(n, b);
            const int t = tr + (c >> 5);
            uint32_t *w = word(t, r);
            if (w == nullptr)
                w = claim(t) + r;
            if ((*w & m) != m)
                atomicOr(w, m);
            c += n;
        }
    }
};

// LDS words a tiled bitmap needs beside its slots.
__host__ __device__ inline int tile_lds_words(int ntiles, int slots)
{
    return ntiles + slots + 1 + slots * 32;
}

// Carves a tiled bitmap out of LDS at `lds` (tile_lds_words words).
__device__ inline TileBits tile_bits(const TraceArgs &a, uint32_t *lds)
{
    TileBits tb;
    tb.tmap = reinterpret_cast<int *>(lds);
    tb.slot_tile = tb.tmap + a.ntiles;
    tb.nslots = tb.slot_tile + a.tile_slots;
    tb.slots = reinterpret_cast<uint32_t *>(tb.nslots + 1);
    tb.hbm = a.gbitmap + (int64_t)blockIdx.x * a.ntiles * 32;
    tb.maxslots = a.tile_slots;
    tb.tiles_x = a.tiles_x;
    tb.ntiles = a.ntiles;
    return tb;
}

// Empty tiled bitmap (the HBM region must already be zero).  Every lane calls
// it; a barrier must follow before first use.
__device__ inline void tile_init(TileBits &tb)
{
    for (int t = threadIdx.x; t < tb.ntiles; t += blockDim.x)
        tb.tmap[t] = kTileEmpty;
    for (int s = threadIdx.x; s < tb.maxslots; s += blockDim.x)
        tb.slot_tile[s] = -1;
    for (int w = threadIdx.x; w < tb.maxslots * 32; w += blockDim.x)
        tb.slots[w] = 0u;
    if (threadIdx.x == 0)
        *tb.nslots = 0;
}

// Flush of one realization's flat bitmap; every lane calls it, after a
// barrier.  Returns this lane's share of the visited count.
__device__ inline unsigned long long flush_bitmap(const TraceArgs &a, const uint32_t *bm,
                                                  unsigned int *s_win)
{
    unsigned long long vis = 0;
    for (int w = threadIdx.x; w < a.nwords; w += blockDim.x) {
        uint32_t bits = bm[w];
        vis += (unsigned long long)__popc(bits);
        while (bits) {
            const int idx = w * 32 + __ffs(bits) - 1;
            bits &= bits - 1u;
            const int row = idx / a.gncols;
            count_node(a, row, idx - row * a.gncols, s_win);
        }
    }
    return vis;
}

// Flush of one realization's tiled bitmap, leaving it empty for the next one.
// Every lane calls it, after a barrier; it ends with a barrier.
__device__ inline unsigned long long flush_tiles(const TraceArgs &a, TileBits &tb,
                                                 unsigned int *s_win)
{
    unsigned long long vis = 0;
    const int handed = *tb.nslots;
    const int ns = handed < tb.maxslots ? handed : tb.maxslots;
    for (int w = threadIdx.x; w < ns * 32; w += blockDim.x) {
        const int t = tb.slot_tile[w >> 5];
        const uint32_t bits = tb.slots[w];
        if (t < 0 || !bits)
            continue;
        tb.slots[w] = 0u;
        vis += count_bits(a, (t / tb.tiles_x) * 32 + (w & 31), (t % tb.tiles_x) * 32, bits, s_win);
    }
    if (handed > tb.maxslots) {   // some tiles went to HBM
        for (int w = threadIdx.x; w < tb.ntiles * 32; w += blockDim.x) {
            const int t = w >> 5;
            if (tb.tmap[t] != kTileOverflow)
                continue;
            const uint32_t bits = tb.hbm[w];
            if (!bits)
                continue;
            tb.hbm[w] = 0u;
            vis += count_bits(a, (t / tb.tiles_x) * 32 + (w & 31), (t % tb.tiles_x) * 32, bits,
                              s_win);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < tb.ntiles; t += blockDim.x)
        tb.tmap[t] = kTileEmpty;
    for (int s = threadIdx.x; s < ns; s += blockDim.x)
        tb.slot_tile[s] = -1;
    if (threadIdx.x == 0)
        *tb.nslots = 0;
    __syncthreads();
    return vis;
}

// Writes the window counters of realization r (capped at 255), or its node
// masks and their counts, and clears them.
__device__ inline void store_window(const TraceArgs &a, int r, unsigned int *s_win)
{
    const int n = a.w_nr * a.w_nc;
    uint8_t *dst = a.wmarks + (int64_t)r * n;
    uint16_t *msk = a.wmasks ? a.wmasks + (int64_t)r * n : nullptr;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const unsigned int v = s_win[c];
        if (msk) {
            msk[c] = (uint16_t)v;
            dst[c] = (uint8_t)__popc(v);
        } else {
            dst[c] = (uint8_t)(v < 255u ? v : 255u);
        }
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
    // q layout: 0 = [nreal x nwells]; else well w of realization r is at
    // q[w * q_well_stride + r] (well-major: a wave's loads are contiguous)
    int64_t q_well_stride;
};

struct DrawArgs {
    int32_t nwells;
    const oka_dist *dists;      // device: nwells + 3
    int64_t n, first;
    uint64_t seed;
    double *q, *K, *por, *thick, *z;
    int32_t antithetic;         // odd realizations reflect their even partner's draws
    int64_t q_well_stride;      // q layout, as FitArgs::q_well_stride
    const uint32_t *sobol;      // QMC: device copy of sobol_directions() (>= qmc_dims x 32)
    const uint32_t *qmc_keys;   // QMC: scramble keys, qmc_dims per replicate (qmc_key())
    int32_t qmc_dims;
    int64_t qmc_npts;           // QMC: points per replicate (0: a single replicate)
};

struct FeatureArgs {
    int32_t nwells;
    int64_t n;
    const double *q, *K, *por, *thick, *coef;
    const double *mean, *std;   // device: nwells + 9
    float *x;                   // n x (nwells + 9)
    int64_t q_well_stride;      // q layout, as FitArgs::q_well_stride
};

// oka_rng.hip
hipError_t launch_draw(const DrawArgs &a, hipStream_t s);
hipError_t launch_features(const FeatureArgs &a, hipStream_t s);
// sums[c] += sum over rows of marks[r, c], sums_sq[c] += sum of marks[r, c]^2
hipError_t launch_mark_moments(const uint8_t *marks, int64_t n, int32_t ncell, uint64_t *sums,
                               uint64_t *sums_sq, hipStream_t s);
uint32_t host_philox_word(uint64_t seed, int64_t r, uint32_t slot, uint32_t sub, int word);

// Sobol' direction numbers of kSobolMaxDims dimensions, 32 per dimension
// (host memory, built once); qmc_dims is the number of dimensions a QMC draw
// of these distributions uses.
constexpr int kSobolMaxDims = 1024;
const uint32_t *sobol_directions();
int qmc_dims(int32_t nwells, const oka_dist *dists);
// The scramble key of dimension d under `seed`: word 0 of Philox4x32-10 at
// counter (d, 2^32 - 1, kQmcKeySlot, 2), a counter the Philox draws never use.
constexpr uint32_t kQmcKeySlot = 0x50B0150u;
uint32_t qmc_key(uint64_t seed, int d);

// Persistent tracer over a table of sites (oka_pool_impl.h; FP32 in
// oka_pool.hip, FP64-exact in oka_pool64.hip).
// A queue item is a chunk of `chunk` particles of one realization.  A
// realization of several chunks is gathered in a realization record in HBM
// (its visited-node tiles and dirty-tile bits), which the wave that merges
// the last chunk flushes.
// The chunks of a realization of at least this many chunks share one
// far-well table, computed once into the realization's record: on the
// 200-well site (4 chunks of 64 particles) the pool launch is 3 % faster than
// with a table per chunk, on the 64-site workload (2 chunks of 128) 5 %
// slower.
constexpr int kFarShareMinChunks = 4;

struct PoolArgs {
    const TraceArgs *sites;     // device [nsites]
    int32_t nsites;
    const int64_t *site_first;  // device [nsites + 1]: first queue item of each site
    const int64_t *real_first;  // device [nsites + 1]: first realization id of each site
    int64_t total;              // queue items over all sites
    unsigned long long *queue;  // next queue item; zero before the launch
    uint32_t *hbm;              // [waves][slots][max_ntiles][32]; zero before the launch
    int32_t slots;              // chunks in flight per wave
    int32_t max_nwells, pool_tiles, max_ntiles;
    int32_t chunk;              // particles per queue item
    int32_t dirty_words;        // ceil(max_ntiles / 32)
    int64_t nrec;               // realization records
    int64_t *rec_owner;         // [nrec]: realization id holding the record, or -1
    int32_t *rec_done;          // [nrec]: chunks merged; zero before the launch
    unsigned long long *rec_table;   // [nrec]: ((g + 1) << 1) | built for the realization
                                     // g whose far-well table the record holds; zero
                                     // before the launch
    uint32_t *rec_bits;         // [nrec][max_ntiles][32]; zero before the launch
    uint32_t *rec_dirty;        // [nrec][dirty_words]; zero before the launch
    float2 *far_beta;           // [waves][slots][far_stride]: far-well coefficients
    float2 *far_beta_rec;       // [nrec][far_stride]: the tables that the chunks of a
                                // realization share, or null (each chunk computes its own)
    int64_t far_stride;         // float2 per slot (the largest nbx * nby * order)
};
// LDS bytes of one wave of the pool in `precision` (OKA_FP32 or
// OKA_FP64_EXACT); conf_mode: 0 every site confined, 1 every site
// unconfined, 2 a choice per site.
size_t pool_lds_bytes(int precision, int slots, int max_nwells, int pool_tiles);
size_t pool64_lds_bytes(int slots, int max_nwells, int pool_tiles);
int pool_max_tiles();
int pool_max_pool_tiles();
hipError_t launch_trace_pool(const PoolArgs &p, int waves, int precision, int conf_mode,
                             hipStream_t s);
hipError_t launch_trace_pool64(const PoolArgs &p, int waves, int conf_mode, hipStream_t s);

// oka_fp64.hip
hipError_t launch_trace_fp64(const TraceArgs &a, int grid_blocks, int kind,
                             hipStream_t s);
hipError_t launch_vertices_fp64(const TraceArgs &a, hipStream_t s);
hipError_t launch_eval_fp64(const EvalArgs &a, hipStream_t s);
hipError_t launch_raster_fp64(const RasterArgs &a, hipStream_t s);
hipError_t launch_raster_fp32(const RasterArgs &a, hipStream_t s);
size_t trace_fp64_wells_lds(int nwells);

// oka_fp32.hip
hipError_t launch_trace_fp32(const TraceArgs &a, int grid_blocks, int kind,
                             hipStream_t s);
size_t trace_fp32_wells_lds(int nwells);

// oka_fit.hip
hipError_t launch_fit_logs(const FitArgs &a, double *logr2, hipStream_t s);
// Over a device table of sites; `first` holds the prefix sums of the item
// counts (nobs x nwells for the logs, nreal for the fits), total the last.
hipError_t launch_fit_logs_multi(const FitArgs *sites, int nsites, const int64_t *first,
                                 int64_t total, hipStream_t s);
hipError_t launch_fit_draw_multi(const FitArgs *sites, int nsites, const int64_t *first,
                                 int64_t total, hipStream_t s);
hipError_t launch_fit_draw(const FitArgs &a, hipStream_t s);

double host_pow_fifth(double x);
double host_log(double x);

}  // namespace oka
