// oka_pool_impl.h -- the persistent tracer with work stealing over live
// particles, generic over the tracing arithmetic (oka_pool.hip: FP32,
// oka_pool64.hip: FP64-exact).
//
// One wave (64 lanes) per workgroup, a persistent grid.  Each wave keeps up to
// `slots` realizations in flight: when a lane's trace ends it takes the next
// particle of the newest open realization, and when that realization has no
// particles left the wave opens another one from a global queue shared by all
// waves (possibly of another site).  A realization is flushed as soon as its
// last particle ends, so a few long traces (particles circling an injection
// well, a 50-year backtrace beside a 1-year one) no longer idle the lanes that
// finished their own work.
//
// Per realization slot, the visited-node bitmap is kept as 32 x 32-node tiles:
// a small LDS hash maps a tile to one of the slot's LDS pool tiles, or to the
// slot's HBM region when the pool is full; tiles that find no hash entry live
// in the HBM region unmapped, and the flush then scans that region.  All
// coordination is within one wave; barriers separate the phases of an
// iteration so that LDS writes by one lane are seen by the others.
//
// The arithmetic policy `Prec` supplies the per-slot data and the per-lane
// trace:
//   Well, Consts        a well record in LDS and the slot's field, integrator
//                       and raster constants (in the slot header);
//   Lane                a lane's trace state;
//   kWavesPerEU         the register budget of the kernel;
//   stage_wells(...)    the lanes stage realization r's wells (returns the
//                       wave-uniform flag that stage_consts receives);
//   stage_consts(...)   lane 0 fills the slot's constants;
//   stage_slot(...)     the lanes' work on the staged slot (after a barrier),
//                       such as per-realization tables that the lanes then read;
//   table_shared(k)     whether that table may be shared by the chunks of a
//                       realization (computed once, in its record);
//   table_to(k, beta)   lane 0 points the slot's constants at the table beta;
//   open(...)           a lane takes particle p of the slot's realization;
//   advance<kConf>(...) one iteration of the lane's trace (one field
//                       evaluation for a fresh trace, an attempted step
//                       otherwise), rasterizing an accepted segment; returns
//                       bit 1 when the trace has ended;
//   summary(...)        the ended trace's steps, candidate nodes, status,
//                       bounding box (absolute coordinates) and whether it
//                       left the grid (in the policy's own arithmetic, as
//                       its workgroup tracer decides it).
// Each trace's arithmetic is lane-local, so the schedule never changes its
// bits: the pool gives the workgroup tracer's counts, statuses and statistics.
#pragma once

#include <math.h>

#include "oka_internal.h"

namespace oka {
namespace pool {

constexpr int kWave = 64;
constexpr int kHashBits = 7;
constexpr int kHash = 1 << kHashBits;                 // map entries per slot
constexpr int kValBits = 11;
constexpr uint32_t kValMask = (1u << kValBits) - 1u;  // pool tile, or kValHbm
constexpr uint32_t kValHbm = kValMask;
constexpr uint32_t kValNone = 0xffffffffu;              // no map entry (find)
constexpr uint32_t kKeyEmpty = 0xffffffffu;

// Which velocity fields a launch traces: every slot confined, every slot
// unconfined, or a per-slot choice (sites of both kinds in one batch).
enum ConfMode { kConfAll = 0, kUnconfAll = 1, kConfMixed = 2 };

template <class Prec>
struct __align__(16) SlotHdr {   // a multiple of 16 bytes: the well records follow
    unsigned long long att, acc, esc, cand, vis, st[4];
    int64_t r;          // realization within its site
    int64_t g;          // realization id over all sites
    int state;          // 0 free, 1 open, 2 waiting for its realization record,
                        // 3 waiting for its record's shared table
    int site;
    int next, np, rem;  // next particle to hand out, end of the chunk, not yet ended
    int nchunks, rec;   // chunks of the realization; its record when nchunks > 1
    int pool_next;      // pool tiles handed out (may exceed the pool)
    int hash_full;      // a tile found no map entry
    int nw, raster, tiles_x, confined;
    // the slot's field, integrator and raster constants, read by the lanes
    // when they need them rather than kept in registers per lane
    typename Prec::Consts k;
};

__device__ __forceinline__ uint32_t tile_hash(uint32_t t)
{
    return (t * 2654435761u) >> (32 - kHashBits);
}

// One slot's tiled bitmap, as seen by one lane.
struct PoolBits {
    uint32_t *keys;     // LDS [kHash]: (tile << kValBits) | value, or kKeyEmpty
    uint32_t *pool;     // LDS [pool_tiles][32]
    int *pool_next, *hash_full;
    int pool_tiles;
    uint32_t *hbm;      // [ntiles][32]
    int tiles_x;
    // the last tile this lane found mapped and its value (a mapping lasts
    // until the flush); kept by the lane between steps
    uint32_t last_t, last_v;

    __device__ __forceinline__ uint32_t *tile_words(uint32_t t, uint32_t v) const
    {
        return v == kValHbm ? hbm + (int64_t)t * 32 : pool + v * 32;
    }
    // Tile t's value (pool tile or kValHbm), or kValNone when unmapped.
    __device__ __forceinline__ uint32_t find(uint32_t t) const
    {
        uint32_t h = tile_hash(t);
        for (int k = 0; k < kHash; ++k) {
            const uint32_t e = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
            if (e == kKeyEmpty)
                return kValNone;
            if ((e >> kValBits) == t)
                return e & kValMask;
            h = (h + 1) & (kHash - 1);
        }
        return kValNone;
    }
    // Maps tile t (or finds the mapping another lane made first); kValHbm
    // when it lives in the HBM region, unmapped if the map is full.
    __device__ __forceinline__ uint32_t claim(uint32_t t)
    {
        const int v0 = atomicAdd(pool_next, 1);
        const uint32_t v = v0 < pool_tiles ? (uint32_t)v0 : kValHbm;
        const uint32_t want = (t << kValBits) | v;
        uint32_t h = tile_hash(t);
        for (int k = 0; k < kHash; ++k) {
            const uint32_t prev = atomicCAS(&keys[h], kKeyEmpty, want);
            if (prev == kKeyEmpty)
                return v;
            if ((prev >> kValBits) == t)   // another lane mapped it first
                return prev & kValMask;
            h = (h + 1) & (kHash - 1);
        }
        atomicOr(hash_full, 1);   // unmapped: the flush scans the HBM region
        return kValHbm;
    }
    __device__ __forceinline__ uint32_t lookup(uint32_t t)
    {
        if (t == last_t)
            return last_v;
        const uint32_t v = find(t);
        if (v != kValNone) {
            last_t = t;
            last_v = v;
        }
        return v;
    }
    __device__ __forceinline__ bool test(int row, int col)
    {
        const uint32_t t = (uint32_t)((row >> 5) * tiles_x + (col >> 5));
        uint32_t v = lookup(t);
        if (v == kValNone) {
            if (!__hip_atomic_load(hash_full, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
                return false;
            v = kValHbm;
        }
        return (tile_words(t, v)[row & 31] >> (col & 31)) & 1u;
    }
    __device__ __forceinline__ void set(int row, int col)
    {
        const uint32_t t = (uint32_t)((row >> 5) * tiles_x + (col >> 5));
        uint32_t v = lookup(t);
        if (v == kValNone) {
            v = claim(t);
            last_t = t;
            last_v = v;
        }
        atomicOr(&tile_words(t, v)[row & 31], 1u << (col & 31));
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
            const uint32_t t = (uint32_t)(tr + (c >> 5));
            uint32_t v = lookup(t);
            if (v == kValNone) {
                v = claim(t);
                last_t = t;
                last_v = v;
            }
            uint32_t *w = tile_words(t, v) + r;
            if ((*w & m) != m)
                atomicOr(w, m);
            c += n;
        }
    }
};

template <class Prec>
struct PoolLds {
    SlotHdr<Prec> *hdr;             // [slots]
    typename Prec::Well *wells;     // [slots][max_nwells]
    uint32_t *keys;                 // [slots][kHash]
    uint32_t *pool;                 // [slots][pool_tiles][32]
};

template <class Prec>
__device__ __forceinline__ PoolLds<Prec> carve(const PoolArgs &P, void *lds)
{
    PoolLds<Prec> L;
    L.hdr = reinterpret_cast<SlotHdr<Prec> *>(lds);
    L.wells = reinterpret_cast<typename Prec::Well *>(L.hdr + P.slots);
    L.keys = reinterpret_cast<uint32_t *>(L.wells + (size_t)P.slots * P.max_nwells);
    L.pool = L.keys + (size_t)P.slots * kHash;
    return L;
}

template <class Prec>
__device__ __forceinline__ PoolBits slot_bits(const PoolArgs &P, const PoolLds<Prec> &L, int s,
                                              int tiles_x)
{
    PoolBits b;
    b.keys = L.keys + (size_t)s * kHash;
    b.pool = L.pool + (size_t)s * P.pool_tiles * 32;
    b.pool_next = &L.hdr[s].pool_next;
    b.hash_full = &L.hdr[s].hash_full;
    b.pool_tiles = P.pool_tiles;
    b.hbm = P.hbm + ((int64_t)blockIdx.x * P.slots + s) * P.max_ntiles * 32;
    b.tiles_x = tiles_x;
    b.last_t = kKeyEmpty;
    b.last_v = kValNone;
    return b;
}

// Site and realization of queue item `item` (site_first: prefix sums of nreal).
__device__ __forceinline__ int find_site(const PoolArgs &P, int64_t item)
{
    int lo = 0, hi = P.nsites - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (P.site_first[mid] <= item)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Claims realization record h.rec for h.g (lane 0; the result is broadcast).
template <class Prec>
__device__ __forceinline__ bool claim_record(const PoolArgs &P, const SlotHdr<Prec> &h)
{
    int ok = 0;
    if (threadIdx.x == 0) {
        const unsigned long long prev = atomicCAS((unsigned long long *)&P.rec_owner[h.rec],
                                                  ~0ull, (unsigned long long)h.g);
        ok = prev == ~0ull || prev == (unsigned long long)h.g;
    }
    return __shfl(ok, 0) != 0;
}

__device__ __forceinline__ bool shares_table(const PoolArgs &P, int nchunks)
{
    return nchunks >= kFarShareMinChunks && P.far_beta_rec != nullptr;
}

// A slot whose realization holds its record, when the realization's chunks
// share one table (shares_table): the first chunk here computes it
// into the record and publishes it; the others use it once it is published
// (state 3 until then).  The record's table word names the realization
// whose table the record holds, ((g + 1) << 1) | published.  While g holds
// the record only g's chunks write that word (the previous holder's chunks
// have all ended), so it goes from another realization's value to g's and
// then to g's published value.  The builder computes the table without
// waiting on anything, so a waiting chunk's wait ends.  Returns the slot's
// state, 1 or 3.  Wave-uniform; ends with a barrier.
template <class Prec>
__device__ int open_table(const PoolArgs &P, const PoolLds<Prec> &L, int s)
{
    SlotHdr<Prec> &h = L.hdr[s];
    if (!shares_table(P, h.nchunks) || !Prec::table_shared(h.k))
        return 1;
    const unsigned long long key = (unsigned long long)(h.g + 1) << 1;
    int code = 0;   // 0: build it, 1: published, 3: being built by another chunk
    if (threadIdx.x == 0) {
        Prec::table_to(h.k, P.far_beta_rec + (int64_t)h.rec * P.far_stride);
        unsigned long long cur = __hip_atomic_load(&P.rec_table[h.rec], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
        if (cur != key && cur != (key | 1ull)) {   // the previous holder's table
            const unsigned long long prev = atomicCAS(&P.rec_table[h.rec], cur, key);
            // another chunk of g got there first; any other value cannot occur,
            // and computing the table again would write the same words
            cur = prev == cur ? 0ull : prev;
        }
        code = cur == (key | 1ull) ? 1 : (cur == key ? 3 : 0);
    }
    code = __shfl(code, 0);
    __syncthreads();   // the lanes see the slot's table pointer
    if (code == 0) {
        const typename Prec::Well *w = L.wells + (size_t)s * P.max_nwells;
        Prec::stage_slot(h.k, w, h.nw);
        __threadfence();   // each lane's coefficients, device-wide, before the flag
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_store(&P.rec_table[h.rec], key | 1ull, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT);
        return 1;
    }
    if (code == 1)
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // the table after its flag
    return code;
}

// Whether the shared table that slot s waits for (state 3) is published.
// Wave-uniform.
template <class Prec>
__device__ bool table_published(const PoolArgs &P, const SlotHdr<Prec> &h)
{
    int ok = 0;
    if (threadIdx.x == 0)
        ok = __hip_atomic_load(&P.rec_table[h.rec], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) ==
             (((unsigned long long)(h.g + 1) << 1) | 1ull);
    ok = __shfl(ok, 0);
    if (ok)
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // every lane reads after the flag
    return ok != 0;
}

// Opens slot s for queue item `item`: wells, constants, header.  A chunk of a
// realization of several chunks also claims the realization's record (state
// 2 while an older realization still holds it) and then its shared table
// (open_table).  Wave-uniform; ends with a barrier.
template <class Prec>
__device__ void open_slot(const PoolArgs &P, const PoolLds<Prec> &L, int s, int64_t item)
{
    const int site = find_site(P, item);
    const TraceArgs &A = P.sites[site];
    const int nch = (A.npaths + P.chunk - 1) / P.chunk;
    const int64_t local = item - P.site_first[site];
    const int64_t r = nch > 0 ? local / nch : 0;
    const int ch = nch > 0 ? (int)(local - r * nch) : 0;
    typename Prec::Well *w = L.wells + (size_t)s * P.max_nwells;
    const bool flag = Prec::stage_wells(w, A, r);
    if (threadIdx.x == 0) {
        SlotHdr<Prec> &h = L.hdr[s];
        Prec::stage_consts(h.k, P, A, r, s, flag);
        h.att = h.acc = h.esc = h.cand = h.vis = 0;
        for (int j = 0; j < 4; ++j)
            h.st[j] = 0;
        h.r = r;
        h.g = P.real_first[site] + r;
        h.site = site;
        h.next = ch * P.chunk;
        h.np = min(A.npaths, h.next + P.chunk);
        h.rem = h.np - h.next;
        h.nchunks = nch;
        h.rec = nch > 1 ? (int)(h.g % P.nrec) : -1;
        h.pool_next = 0;
        h.hash_full = 0;
        h.nw = A.nwells;
        h.raster = A.rasterize != 0;
        h.tiles_x = A.tiles_x;
        h.confined = A.confined != 0;
    }
    __syncthreads();
    const bool chunked = L.hdr[s].nchunks > 1;
    const bool shared = shares_table(P, L.hdr[s].nchunks) && Prec::table_shared(L.hdr[s].k);
    if (!shared && Prec::stage_slot(L.hdr[s].k, w, A.nwells))
        __syncthreads();
    const int state = !chunked ? 1 : (claim_record(P, L.hdr[s]) ? open_table(P, L, s) : 2);
    if (threadIdx.x == 0)
        L.hdr[s].state = state;
    __syncthreads();
}

// Flushes slot s into its site's counts and statistics and frees it.
// Wave-uniform; ends with a barrier.
template <class Prec>
__device__ void flush_slot(const PoolArgs &P, const PoolLds<Prec> &L, int s)
{
    SlotHdr<Prec> &h = L.hdr[s];
    const TraceArgs &A = P.sites[h.site];
    const PoolBits b = slot_bits(P, L, s, A.tiles_x);
    unsigned long long vis = 0;
    if (A.rasterize) {
        for (int k = threadIdx.x; k < kHash * 32; k += kWave) {
            const uint32_t e = b.keys[k >> 5];
            if (e == kKeyEmpty)
                continue;
            const uint32_t t = e >> kValBits;
            uint32_t *w = b.tile_words(t, e & kValMask) + (k & 31);
            const uint32_t bits = *w;
            if (!bits)
                continue;
            *w = 0u;
            vis += count_bits(A, (int)(t / A.tiles_x) * 32 + (k & 31), (int)(t % A.tiles_x) * 32,
                              bits, nullptr);
        }
        if (h.hash_full) {   // unmapped tiles: scan the whole HBM region
            for (int k = threadIdx.x; k < A.ntiles * 32; k += kWave) {
                const uint32_t bits = b.hbm[k];
                if (!bits)
                    continue;
                b.hbm[k] = 0u;
                const int t = k >> 5;
                vis += count_bits(A, (t / A.tiles_x) * 32 + (k & 31), (t % A.tiles_x) * 32, bits,
                                  nullptr);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        vis += __shfl_xor(vis, off);
    __syncthreads();
    for (int k = threadIdx.x; k < kHash; k += kWave)
        b.keys[k] = kKeyEmpty;
    if (threadIdx.x == 0) {
        if (A.stats) {
            atomicAdd(&A.stats[kStatAttempted], h.att);
            atomicAdd(&A.stats[kStatAccepted], h.acc);
            atomicAdd(&A.stats[kStatEscaped], h.esc);
            atomicAdd(&A.stats[kStatCandidates], h.cand);
            atomicAdd(&A.stats[kStatVisited], vis);
            for (int j = 0; j < 4; ++j)
                atomicAdd(&A.stats[kStatStatus0 + j], h.st[j]);
        }
        h.pool_next = 0;
        h.hash_full = 0;
        h.state = 0;
    }
    __syncthreads();
}

// A finished chunk of a realization of several chunks: ORs the chunk's marked
// words into the realization record and sets its dirty-tile bits; the wave
// that merges the last chunk flushes the record into the counts and frees
// it.  Wave-uniform; ends with a barrier.
template <class Prec>
__device__ void merge_slot(const PoolArgs &P, const PoolLds<Prec> &L, int s)
{
    SlotHdr<Prec> &h = L.hdr[s];
    const TraceArgs &A = P.sites[h.site];
    const PoolBits b = slot_bits(P, L, s, A.tiles_x);
    uint32_t *rb = P.rec_bits + (int64_t)h.rec * P.max_ntiles * 32;
    uint32_t *rd = P.rec_dirty + (int64_t)h.rec * P.dirty_words;
    if (A.rasterize) {
        for (int k = threadIdx.x; k < kHash * 32; k += kWave) {
            const uint32_t e = b.keys[k >> 5];
            if (e == kKeyEmpty)
                continue;
            const uint32_t t = e >> kValBits;
            uint32_t *w = b.tile_words(t, e & kValMask) + (k & 31);
            const uint32_t bits = *w;
            if (!bits)
                continue;
            *w = 0u;
            atomicOr(&rb[(int64_t)t * 32 + (k & 31)], bits);
            atomicOr(&rd[t >> 5], 1u << (t & 31));
        }
        if (h.hash_full) {
            for (int k = threadIdx.x; k < A.ntiles * 32; k += kWave) {
                const uint32_t bits = b.hbm[k];
                if (!bits)
                    continue;
                b.hbm[k] = 0u;
                atomicOr(&rb[k], bits);
                atomicOr(&rd[k >> 10], 1u << ((k >> 5) & 31));
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kHash; k += kWave)
        b.keys[k] = kKeyEmpty;
    if (threadIdx.x == 0) {
        if (A.stats) {
            atomicAdd(&A.stats[kStatAttempted], h.att);
            atomicAdd(&A.stats[kStatAccepted], h.acc);
            atomicAdd(&A.stats[kStatEscaped], h.esc);
            atomicAdd(&A.stats[kStatCandidates], h.cand);
            for (int j = 0; j < 4; ++j)
                atomicAdd(&A.stats[kStatStatus0 + j], h.st[j]);
        }
        h.pool_next = 0;
        h.hash_full = 0;
    }
    // publish the chunk's marks before counting it
    __threadfence();
    int last = 0;
    if (threadIdx.x == 0)
        last = atomicAdd(&P.rec_done[h.rec], 1) == h.nchunks - 1;
    last = __shfl(last, 0);
    if (last) {
        __threadfence();
        unsigned long long vis = 0;
        if (A.rasterize) {
            const int dwords = (A.ntiles + 31) / 32;
            for (int dw = 0; dw < dwords; ++dw) {
                uint32_t dbits = __hip_atomic_load(&rd[dw], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                while (dbits) {
                    const int t = dw * 32 + __ffs(dbits) - 1;
                    dbits &= dbits - 1u;
                    if (threadIdx.x < 32) {
                        uint32_t *w = &rb[(int64_t)t * 32 + threadIdx.x];
                        const uint32_t bits = __hip_atomic_load(w, __ATOMIC_RELAXED,
                                                                __HIP_MEMORY_SCOPE_AGENT);
                        if (bits) {
                            __hip_atomic_store(w, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            vis += count_bits(A, (t / A.tiles_x) * 32 + (int)threadIdx.x,
                                              (t % A.tiles_x) * 32, bits, nullptr);
                        }
                    }
                }
                if (threadIdx.x == 0)
                    __hip_atomic_store(&rd[dw], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        for (int off = 32; off > 0; off >>= 1)
            vis += __shfl_xor(vis, off);
        if (threadIdx.x == 0) {
            if (A.stats)
                atomicAdd(&A.stats[kStatVisited], vis);
            __hip_atomic_store(&P.rec_done[h.rec], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // the record is clean before another realization may claim it
        __threadfence();
        if (threadIdx.x == 0)
            atomicExch((unsigned long long *)&P.rec_owner[h.rec], ~0ull);
    }
    if (threadIdx.x == 0)
        h.state = 0;
    __syncthreads();
}

template <class Prec, int kConf>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(Prec::kWavesPerEU)))
k_trace_pool(PoolArgs P)
{
    extern __shared__ __align__(16) unsigned char pool_lds[];
    const PoolLds<Prec> L = carve<Prec>(P, pool_lds);
    const int lane = threadIdx.x;
    for (int s = 0; s < P.slots; ++s) {
        if (lane == 0) {
            L.hdr[s].state = 0;
            L.hdr[s].pool_next = 0;
            L.hdr[s].hash_full = 0;
        }
        for (int k = lane; k < kHash; k += kWave)
            L.keys[(size_t)s * kHash + k] = kKeyEmpty;
    }
    for (int k = lane; k < P.slots * P.pool_tiles * 32; k += kWave)
        L.pool[k] = 0u;
    __syncthreads();

    // the lane's trace; its slot's constants stay in the slot header
    bool live = false;
    int slot = 0, pidx = 0;
    typename Prec::Lane ln;
    // wave-uniform
    bool queue_done = false;
    int intake = -1;

    for (;;) {
        // A. idle lanes take particles from the intake slot, opening slots as needed
        bool ended = false;
        unsigned long long need = __ballot(!live);
        while (need != 0ull) {
            if (intake >= 0 && L.hdr[intake].state >= 2) {   // its record or table is busy
                int st = L.hdr[intake].state;
                if (st == 2)
                    st = claim_record(P, L.hdr[intake]) ? open_table(P, L, intake) : 2;
                else if (table_published(P, L.hdr[intake]))
                    st = 1;
                __syncthreads();   // every lane has read the state before lane 0 writes it
                if (threadIdx.x == 0)
                    L.hdr[intake].state = st;
                __syncthreads();
                if (st != 1)
                    break;
            }
            if (intake >= 0 && L.hdr[intake].next < L.hdr[intake].np) {
                const int next = L.hdr[intake].next;
                const int avail = L.hdr[intake].np - next;
                const int rank = __popcll(need & ((1ull << lane) - 1ull));
                const int took = min(__popcll(need), avail);
                if (((need >> lane) & 1ull) && rank < avail) {
                    slot = intake;
                    pidx = next + rank;
                    const SlotHdr<Prec> &h = L.hdr[slot];
                    Prec::open(ln, h.k, P.sites[h.site], h.r, pidx, h.raster != 0);
                    live = true;
                }
                __syncthreads();
                if (lane == 0)
                    L.hdr[intake].next = next + took;
                __syncthreads();
                need = __ballot(!live && !ended);
                continue;
            }
            if (queue_done)
                break;
            int fs = -1;
            for (int s = 0; s < P.slots; ++s)
                if (fs < 0 && L.hdr[s].state == 0)
                    fs = s;
            if (fs < 0)
                break;
            unsigned long long item = 0ull;
            if (lane == 0)
                item = atomicAdd(P.queue, 1ull);
            item = __shfl(item, 0);
            if (item >= (unsigned long long)P.total) {
                queue_done = true;
                break;
            }
            open_slot(P, L, fs, (int64_t)item);
            intake = fs;
        }

        // B. one iteration of every live lane's trace
        if (live) {
            const SlotHdr<Prec> &h = L.hdr[slot];
            PoolBits bits = slot_bits(P, L, slot, h.tiles_x);
            const typename Prec::Well *w = L.wells + (size_t)slot * P.max_nwells;
            int res;
            if (kConf == kConfAll || (kConf == kConfMixed && h.confined))
                res = Prec::template advance<true>(ln, h.k, w, h.nw, bits, h.raster != 0);
            else
                res = Prec::template advance<false>(ln, h.k, w, h.nw, bits, h.raster != 0);
            if (res & 2) {
                live = false;
                ended = true;
            }
        }
        __syncthreads();

        // C. ended traces: statistics, per-trace outputs, the slot's count down
        if (ended) {
            SlotHdr<Prec> &h = L.hdr[slot];
            const TraceArgs &A = P.sites[h.site];
            long long att, acc;
            unsigned long long cand;
            int status;
            bool escaped;
            double bb[4];
            Prec::summary(ln, A, att, acc, cand, status, bb, escaped);
            atomicAdd(&h.att, (unsigned long long)att);
            atomicAdd(&h.acc, (unsigned long long)acc);
            atomicAdd(&h.cand, cand);
            atomicAdd(&h.st[status & 3], 1ull);
            if (escaped)
                atomicAdd(&h.esc, 1ull);
            const int64_t k = h.r * A.npaths + pidx;
            if (A.bbox)
                for (int j = 0; j < 4; ++j)
                    A.bbox[4 * k + j] = bb[j];
            if (A.trace_status)
                A.trace_status[k] = status;
            if (A.trace_steps)
                A.trace_steps[k] = att;
            atomicSub(&h.rem, 1);
        }
        __syncthreads();

        // D. flush finished realizations
        bool open = false;
        for (int s = 0; s < P.slots; ++s) {
            const SlotHdr<Prec> &h = L.hdr[s];
            if (h.state == 1 && h.next == h.np && h.rem == 0) {
                if (h.nchunks > 1)
                    merge_slot(P, L, s);
                else
                    flush_slot(P, L, s);
                if (intake == s)
                    intake = -1;
            }
            open = open || L.hdr[s].state != 0;
        }
        if (queue_done && !open && __ballot(live) == 0ull)
            break;
    }
}

// LDS of one wave: per slot a header, the well records, the tile map and the
// pool tiles.
template <class Prec>
size_t lds_bytes(int slots, int max_nwells, int pool_tiles)
{
    static_assert(sizeof(SlotHdr<Prec>) % 16 == 0, "slot headers keep the wells 16-byte aligned");
    return (size_t)slots * (sizeof(SlotHdr<Prec>) + (size_t)max_nwells * sizeof(typename Prec::Well) +
                            (size_t)kHash * 4 + (size_t)pool_tiles * 128);
}

template <class Prec>
hipError_t launch(const PoolArgs &p, int waves, int conf_mode, hipStream_t s)
{
    const size_t lds = lds_bytes<Prec>(p.slots, p.max_nwells, p.pool_tiles);
    if (conf_mode == kConfAll)
        hipLaunchKernelGGL((k_trace_pool<Prec, kConfAll>), dim3(waves), dim3(kWave), lds, s, p);
    else if (conf_mode == kUnconfAll)
        hipLaunchKernelGGL((k_trace_pool<Prec, kUnconfAll>), dim3(waves), dim3(kWave), lds, s, p);
    else
        hipLaunchKernelGGL((k_trace_pool<Prec, kConfMixed>), dim3(waves), dim3(kWave), lds, s, p);
    return hipGetLastError();
}

}  // namespace pool
}  // namespace oka
