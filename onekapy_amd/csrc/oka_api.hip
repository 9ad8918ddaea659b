// oka_api.hip -- the extern "C" entry points of liboka.so (include/oka.h):
// context and stream handling, staging of host arrays, launch configuration,
// error reporting, and the host-side expansion-history scan.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "oka_ctx.h"
#include "oka_internal.h"

using namespace oka;

// Scratch memory of oka_trace_chronicle_noalloc: the launch's statistics
// accumulators, the whole-grid epoch and the HBM bitmap region of grids too
// large for LDS.  Each GPU the library runs on holds its own copy until the
// process ends.  In host code `scratch` is only a stand-in for the symbol:
// kernels get the device address that hipGetSymbolAddress returns for the
// current device (oka_ctx::scratch).
static __device__ __align__(256) uint32_t scratch[OKA_NOALLOC_SCRATCH_BYTES / 4];

namespace {

constexpr size_t kScrStats = 0;        // kStatCount x u64
constexpr size_t kScrEpoch = 128;      // one oka_epoch
constexpr size_t kScrBitmap = 256;     // the rest: bitmap words
static_assert(sizeof(unsigned long long) * oka::kStatCount <= kScrEpoch, "statistics fit");
static_assert(kScrEpoch + sizeof(oka_epoch) <= kScrBitmap, "the epoch fits");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
bool g_scratch_done[kMaxDevices];
void *g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];

// The device address of `scratch` on `device`, which must be current; looked
// up once per device.
void *scratch_address(int device, hipError_t *err)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    if (!g_scratch_done[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
        g_scratch_err[device] = e;
        g_scratch_done[device] = true;
    }
    *err = g_scratch_err[device];
    return g_scratch_ptr[device];
}

}  // namespace

namespace oka {

int fail(oka_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        char tmp[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(tmp, sizeof(tmp), fmt, ap);
        va_end(ap);
        c->err = tmp;
    }
    return code;
}

int ensure(oka_ctx *c, Slot s, size_t bytes, void **out)
{
    DevBuf &b = c->buf[s];
    if (bytes == 0)
        bytes = 8;
    if (b.cap < bytes) {
        if (b.p) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
            HIPCHK(c, hipFree(b.p));
            b.p = nullptr;
            b.cap = 0;
        }
        size_t want = bytes + bytes / 4;
        HIPCHK(c, hipMalloc(&b.p, want));
        b.cap = want;
    }
    *out = b.p;
    return OKA_OK;
}

// Device view of an input array: the caller's device pointer, or a staged copy.
int stage_in(oka_ctx *c, Slot s, const void *src, size_t bytes, uint32_t flags, const void **out)
{
    if (!src) {
        *out = nullptr;
        return OKA_OK;
    }
    if (flags & OKA_PTR_DEVICE) {
        *out = src;
        return OKA_OK;
    }
    void *d;
    int rc = ensure(c, s, bytes, &d);
    if (rc)
        return rc;
    if (bytes)
        HIPCHK(c, hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, c->stream));
    *out = d;
    return OKA_OK;
}

// Device view of an output array (optionally preloaded with its host contents).
int stage_out(oka_ctx *c, Slot s, void *dst, size_t bytes, uint32_t flags, bool preload,
              void **out)
{
    if (!dst) {
        *out = nullptr;
        return OKA_OK;
    }
    if (flags & OKA_PTR_DEVICE) {
        *out = dst;
        return OKA_OK;
    }
    int rc = ensure(c, s, bytes, out);
    if (rc)
        return rc;
    if (preload && bytes)
        HIPCHK(c, hipMemcpyAsync(*out, dst, bytes, hipMemcpyHostToDevice, c->stream));
    return OKA_OK;
}

int copy_back(oka_ctx *c, void *dst, const void *dev, size_t bytes, uint32_t flags)
{
    if (!dst || (flags & OKA_PTR_DEVICE) || bytes == 0)
        return OKA_OK;
    HIPCHK(c, hipMemcpyAsync(dst, dev, bytes, hipMemcpyDeviceToHost, c->stream));
    return OKA_OK;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// hist[k] += counts equal to k.  The empty nodes (usually most of a grid) are
// counted per lane and added once per wave; *bad is set when a count lies
// outside [0, max_count].
__global__ void k_count_histogram(const int32_t *counts, int64_t n, int32_t max_count,
                                  unsigned long long *hist, int32_t *bad)
{
    unsigned long long zeros = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = counts[i];
        if (v < 0 || v > max_count)
            atomicOr(bad, 1);
        else if (v == 0)
            zeros += 1;
        else
            atomicAdd(&hist[v], 1ull);
    }
    for (int off = 32; off > 0; off >>= 1)
        zeros += __shfl_xor(zeros, off);
    if ((threadIdx.x & 63) == 0 && zeros)
        atomicAdd(&hist[0], zeros);
}

// Writes one epoch into device memory (a kernel, so that callers may capture it).
__global__ void k_store_epoch(oka_epoch e, oka_epoch *dst)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *dst = e;
}

__global__ void k_bits_to_u8(const uint32_t *bits, uint8_t *out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && ((bits[i >> 5] >> (i & 31)) & 1u))
        out[i] = 1;
}

// Zero-fill without hipMemsetAsync, so that the call can be graph-captured.
int fill_u32v(oka_ctx *c, void *p, uint32_t value, int64_t nwords)
{
    if (nwords <= 0)
        return OKA_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, c->stream,
                       (uint32_t *)p, value, nwords);
    HIPCHK(c, hipGetLastError());
    return OKA_OK;
}

int fill_u32(oka_ctx *c, void *p, int64_t nwords) { return fill_u32v(c, p, 0u, nwords); }



int pool_shape(oka_ctx *c, const TraceArgs *sites, int nsites, const oka_trace_opts *o,
               int precision, PoolShape *out)
{
    PoolShape &sh = *out;
    if (precision != OKA_FP32 && precision != OKA_FP64_EXACT)
        return fail(c, OKA_ERR_ARG, "unknown precision %d", precision);
    sh.precision = precision;
    sh.chunk = o->pool_chunk > 0 ? o->pool_chunk : 64;
    sh.nsites = nsites;
    // first[]: queue items, then realizations, per site (2 x (nsites + 1))
    sh.first.assign(2 * ((size_t)nsites + 1), 0);
    int64_t total = 0, nreal = 0, chunked = 0;
    int nconf = 0;
    bool far_share = false;   // a site whose realizations share far-well tables
    sh.max_nw = 1;
    sh.max_tiles = 1;
    for (int i = 0; i < nsites; ++i) {
        const int nch = (sites[i].npaths + sh.chunk - 1) / sh.chunk;
        sh.first[i] = total;
        sh.first[nsites + 1 + i] = nreal;
        total += (int64_t)sites[i].nreal * nch;
        nreal += sites[i].nreal;
        if (nch > 1)
            chunked += sites[i].nreal;
        sh.max_nw = std::max(sh.max_nw, (int)sites[i].nwells);
        if (sites[i].rasterize)
            sh.max_tiles = std::max(sh.max_tiles, (int)sites[i].ntiles);
        nconf += sites[i].confined ? 1 : 0;
        if (sites[i].far_order > 0) {
            sh.far_stride = std::max<int64_t>(
                sh.far_stride, (int64_t)sites[i].far_nbx * sites[i].far_nby * kFarTerms);
            far_share = far_share || nch >= kFarShareMinChunks;
        }
    }
    sh.conf_mode = nconf == nsites ? 0 : (nconf == 0 ? 1 : 2);
    sh.first[nsites] = total;
    sh.first[2 * (size_t)nsites + 1] = nreal;
    sh.total = total;
    if (sh.max_tiles > pool_max_tiles())
        return fail(c, OKA_ERR_ARG, "grid too large for the pool schedule (%d tiles)",
                    sh.max_tiles);
    sh.slots = o->pool_slots > 0 ? std::min(o->pool_slots, 16) : 2;
    // the FP64 kernel's registers allow 2 waves per SIMD, the FP32 one's 4
    const int max_wpc = precision == OKA_FP64_EXACT ? 8 : 16;
    if (o->pool_waves_per_cu > 0) {
        sh.wpc = std::min(o->pool_waves_per_cu, 32);
    } else {
        // default: the most waves per CU (of 16, 12, 8, 6, 4, no more than
        // the registers let run) whose LDS share still holds the slots'
        // wells and maps and 8 pool tiles per slot
        const int choice[] = {16, 12, 8, 6, 4};
        sh.wpc = 4;
        for (int w : choice) {
            if (w <= max_wpc && pool_lds_bytes(precision, sh.slots, sh.max_nw, 8) <=
                                    std::min(c->lds_per_cu / (size_t)w, c->max_lds)) {
                sh.wpc = w;
                break;
            }
        }
    }
    const size_t budget = std::min(c->lds_per_cu / sh.wpc, c->max_lds);
    const size_t fixed = pool_lds_bytes(precision, sh.slots, sh.max_nw, 0);
    if (fixed > budget)
        return fail(c, OKA_ERR_ARG, "%d wells x %d slots do not fit in LDS at %d waves per CU",
                    sh.max_nw, sh.slots, sh.wpc);
    sh.pool_tiles = (int)std::min<size_t>((budget - fixed) / ((size_t)sh.slots * 128),
                                          pool_max_pool_tiles());
    sh.waves = c->num_cu * sh.wpc;
    sh.hbm_words = (int64_t)sh.waves * sh.slots * sh.max_tiles * 32;
    // realization records, at most 32 GiB; per record: owner (2 words), done
    // and a pad word, the shared far-well table's state (2 words), the tiles,
    // dirty bits.  Realization g uses record
    // g % nrec, and a chunk waits while an older realization still holds it.
    // A realization holds its record until its longest trace ends, which on
    // sites with traces of very different lengths takes as long as the queue
    // takes to advance by a few times the chunks in flight: with twice the
    // slots' count of records, half the pool's iterations on a 200-well site
    // found their intake waiting; with four times, a few in a thousand.
    sh.dirty_words = (sh.max_tiles + 31) / 32;
    const int64_t per_rec = (int64_t)sh.max_tiles * 32 + sh.dirty_words + 6;
    sh.nrec = 0;
    if (chunked > 0) {
        sh.nrec = std::min<int64_t>(chunked, std::max<int64_t>(4LL * sh.waves * sh.slots, 64));
        sh.nrec = std::max<int64_t>(1, std::min<int64_t>(sh.nrec, (8LL << 30) / per_rec));
    }
    sh.rec_words = sh.nrec * per_rec;
    // the chunks of a realization share one far-well table in its record
    // (kFarShareMinChunks) when those tables take at most 4 GiB; else each
    // chunk computes its own
    const int64_t rec_table_bytes = sh.nrec * sh.far_stride * (int64_t)sizeof(float2);
    sh.far_rec = far_share && rec_table_bytes <= (4LL << 30) ? sh.nrec : 0;
    return OKA_OK;
}

int far_requested(const TraceArgs &a, int precision, int far_order)
{
    if (precision != OKA_FP32 || !a.confined || a.nwells < 2 || far_order < 0)
        return 0;
    if (far_order > 0)
        return std::min(far_order, kFarTerms);
    return a.nwells >= kFarAutoWells ? kFarTerms : 0;
}

// Boxes of side L over the wells and an eighth of their extent around them
// (in the FP32 tracers' coordinates, relative to the target), sized so that
// the 3 x 3 boxes around a point hold about kFarNear wells, with at most
// kFarMaxSide boxes a side.
// Every well of a box's 3 x 3 neighbourhood is summed directly; the rest are
// at least 1.5 L from the box centre while the point is within L / sqrt(2),
// so the expansion's terms fall as (0.47)^k.  In the automatic mode a plan
// whose near lists average more than half the wells is not used.
void far_plan(const TraceArgs &a, const double *wxy, int precision, int far_order,
              FarPlan *out)
{
    FarPlan &P = *out;
    P = FarPlan();
    const int order = far_requested(a, precision, far_order);
    const int nw = a.nwells;
    if (order == 0 || !wxy)
        return;
    std::vector<double> wx(nw), wy(nw);
    double x0 = HUGE_VAL, x1 = -HUGE_VAL, y0 = HUGE_VAL, y1 = -HUGE_VAL;
    for (int i = 0; i < nw; ++i) {
        wx[i] = (double)(float)(wxy[2 * i] - a.xo);
        wy[i] = (double)(float)(wxy[2 * i + 1] - a.yo);
        if (!std::isfinite(wx[i]) || !std::isfinite(wy[i]))
            return;
        x0 = std::min(x0, wx[i]);
        x1 = std::max(x1, wx[i]);
        y0 = std::min(y0, wy[i]);
        y1 = std::max(y1, wy[i]);
    }
    // an eighth of the extent beyond the wells on each side; the plan does not
    // depend on the grid, so a chronicle's passes trace the same paths
    const double mx = 0.125 * (x1 - x0), my = 0.125 * (y1 - y0);
    x0 -= mx;
    x1 += mx;
    y0 -= my;
    y1 += my;
    const double W = x1 - x0, H = y1 - y0;
    double L = std::sqrt((double)kFarNear * W * H / (9.0 * nw));
    const double side = kFarMaxSide - 0.5;
    L = std::max({L, W / side, H / side, 1e-3 * std::max(W, H)});
    if (!(L > 0.0) || !std::isfinite(L))
        return;
    const int nbx = (int)std::floor(W / L) + 1, nby = (int)std::floor(H / L) + 1;
    const int nb = nbx * nby;
    // centre the boxes on the bounds: every well lies at least a few
    // rounding steps inside
    const float fx0 = (float)(x0 - 0.5 * (nbx * L - W));
    const float fy0 = (float)(y0 - 0.5 * (nby * L - H));
    const float fl = (float)L;
    P.wbox.resize(2 * (size_t)nw);
    for (int i = 0; i < nw; ++i) {
        const int bx = (int)std::floor((wx[i] - fx0) / fl);
        const int by = (int)std::floor((wy[i] - fy0) / fl);
        P.wbox[2 * i] = std::min(std::max(bx, 0), nbx - 1);
        P.wbox[2 * i + 1] = std::min(std::max(by, 0), nby - 1);
    }
    // near rows of a fixed length (a multiple of 4, read 4 words at a time):
    // the count, the indices in index order, then -1
    std::vector<std::vector<int32_t>> lists(nb);
    int64_t nnear = 0;
    int most = 0;
    for (int b = 0; b < nb; ++b) {
        const int bx = b % nbx, by = b / nbx;
        for (int i = 0; i < nw; ++i)
            if (std::abs(P.wbox[2 * i] - bx) <= 1 && std::abs(P.wbox[2 * i + 1] - by) <= 1)
                lists[b].push_back(i);
        nnear += (int64_t)lists[b].size();
        most = std::max(most, (int)lists[b].size());
    }
    P.near_stride = (most + 1 + 3) / 4 * 4;
    P.near.assign((size_t)nb * P.near_stride, -1);
    for (int b = 0; b < nb; ++b) {
        int32_t *row = P.near.data() + (size_t)b * P.near_stride;
        row[0] = (int32_t)lists[b].size();
        std::copy(lists[b].begin(), lists[b].end(), row + 1);
    }
    if (far_order == 0 && 2 * nnear > (int64_t)nb * nw) {
        P = FarPlan();
        return;
    }
    P.order = order;
    P.nbx = nbx;
    P.nby = nby;
    P.x0 = fx0;
    P.y0 = fy0;
    P.len = fl;
}

void far_args(const FarPlan &P, const int32_t *near, const int32_t *wbox, TraceArgs &a)
{
    a.far_order = P.order;
    a.far_nbx = P.nbx;
    a.far_nby = P.nby;
    a.far_x0 = P.x0;
    a.far_y0 = P.y0;
    a.far_len = P.len;
    a.far_near_stride = P.near_stride;
    a.far_near = near;
    a.far_wbox = wbox;
}

PoolArgs pool_args(const PoolShape &sh, void *sites, void *first, void *queue, void *hbm,
                   void *rec, void *far_beta)
{
    PoolArgs p{};
    p.far_beta = (float2 *)far_beta;
    p.far_stride = sh.far_stride;
    if (sh.far_rec > 0)
        p.far_beta_rec = p.far_beta + (int64_t)sh.waves * sh.slots * sh.far_stride;
    p.sites = (const TraceArgs *)sites;
    p.nsites = sh.nsites;
    p.site_first = (const int64_t *)first;
    p.real_first = (const int64_t *)first + sh.nsites + 1;
    p.total = sh.total;
    p.queue = (unsigned long long *)queue;
    p.hbm = (uint32_t *)hbm;
    p.slots = sh.slots;
    p.max_nwells = sh.max_nw;
    p.pool_tiles = sh.pool_tiles;
    p.max_ntiles = sh.max_tiles;
    p.chunk = sh.chunk;
    p.dirty_words = sh.dirty_words;
    p.nrec = std::max<int64_t>(sh.nrec, 1);
    if (sh.nrec > 0) {
        p.rec_owner = (int64_t *)rec;
        p.rec_done = (int32_t *)(p.rec_owner + sh.nrec);
        p.rec_table = (unsigned long long *)(p.rec_done + 2 * sh.nrec);
        p.rec_bits = (uint32_t *)(p.rec_table + sh.nrec);
        p.rec_dirty = p.rec_bits + sh.nrec * (int64_t)sh.max_tiles * 32;
    }
    return p;
}

int pool_launch(oka_ctx *c, const PoolShape &sh, const PoolArgs &p)
{
    TRY(fill_u32(c, p.queue, 2));
    TRY(fill_u32(c, p.hbm, sh.hbm_words));
    if (sh.nrec > 0) {
        TRY(fill_u32v(c, p.rec_owner, 0xffffffffu, 2 * sh.nrec));
        TRY(fill_u32(c, p.rec_done, sh.rec_words - 2 * sh.nrec));
    }
    if (sh.total > 0)
        HIPCHK(c, launch_trace_pool(p, sh.waves, sh.precision, sh.conf_mode, c->stream));
    return OKA_OK;
}

}  // namespace oka

namespace {

int check_opts(oka_ctx *c, const oka_trace_opts *o)
{
    if (!o)
        return fail(c, OKA_ERR_ARG, "trace options are NULL");
    if (o->npaths < 0)
        return fail(c, OKA_ERR_ARG, "npaths < 0");
    if (o->npaths > 0 && !o->start_xy)
        return fail(c, OKA_ERR_ARG, "start_xy is NULL");
    if (o->precision != OKA_FP64_EXACT && o->precision != OKA_FP32)
        return fail(c, OKA_ERR_ARG, "unknown precision %d", o->precision);
    if (o->schedule < OKA_SCHED_AUTO || o->schedule > OKA_SCHED_POOL)
        return fail(c, OKA_ERR_ARG, "unknown schedule %d", o->schedule);
    if (o->far_order > kFarTerms)
        return fail(c, OKA_ERR_ARG, "far_order %d above %d", o->far_order, kFarTerms);
    return OKA_OK;
}

constexpr int64_t kDefaultMaxSteps = 1000000;

}  // namespace

extern "C" {

int oka_abi_version(void) { return OKA_ABI_VERSION; }

int oka_count_histogram(oka_ctx *c, int64_t n, const int32_t *counts, int32_t max_count,
                        int64_t *hist, uint32_t flags)
{
    if (!c || !hist || n < 0 || max_count < 0 || (n > 0 && !counts))
        return OKA_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    const void *dc;
    TRY(stage_in(c, kSlotCounts, counts, sizeof(int32_t) * n, flags, &dc));
    void *dh, *db;
    TRY(stage_out(c, kSlotOut, hist, sizeof(int64_t) * ((size_t)max_count + 1), flags, true, &dh));
    TRY(ensure(c, kSlotEStatus, sizeof(int32_t), &db));
    TRY(fill_u32(c, db, 1));
    if (n > 0) {
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, (int64_t)c->num_cu * 16);
        hipLaunchKernelGGL(k_count_histogram, dim3((unsigned)blocks), dim3(256), 0, c->stream,
                           (const int32_t *)dc, n, max_count, (unsigned long long *)dh,
                           (int32_t *)db);
        HIPCHK(c, hipGetLastError());
    }
    int32_t bad = 0;
    HIPCHK(c, hipMemcpyAsync(&bad, db, sizeof(bad), hipMemcpyDeviceToHost, c->stream));
    TRY(copy_back(c, hist, dh, sizeof(int64_t) * ((size_t)max_count + 1), flags));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (bad)
        return fail(c, OKA_ERR_RANGE, "a count lies outside [0, %d]", max_count);
    return OKA_OK;
}

int oka_ctx_device_info(oka_ctx *c, oka_device_info *out)
{
    if (!c || !out)
        return OKA_ERR_ARG;
    out->num_cu = c->num_cu;
    out->warp_size = c->warp_size;
    out->lds_per_workgroup = (int64_t)c->max_lds;
    out->lds_per_cu = (int64_t)c->lds_per_cu;
    return OKA_OK;
}

int oka_ctx_create(oka_ctx **out, int device)
{
    if (!out)
        return OKA_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return OKA_ERR_DEVICE;
    if (device < 0 || device >= n)
        return OKA_ERR_ARG;
    oka_ctx *c = new oka_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return OKA_ERR_DEVICE;
    }
    c->stream = c->own;
    c->num_cu = prop.multiProcessorCount;
    c->max_lds = prop.sharedMemPerBlock;
    c->warp_size = prop.warpSize;
    int lds_cu = 0;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                              device) != hipSuccess)
        lds_cu = 0;
    c->lds_per_cu = lds_cu > 0 ? (size_t)lds_cu : prop.sharedMemPerBlock;
    c->scratch = scratch_address(device, &c->scratch_err);
    *out = c;
    return OKA_OK;
}

void oka_ctx_destroy(oka_ctx *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &b : c->buf)
        if (b.p)
            (void)hipFree(b.p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
}

const char *oka_last_error(const oka_ctx *c) { return c ? c->err.c_str() : "null context"; }

int oka_ctx_set_stream(oka_ctx *c, void *s)
{
    if (!c)
        return OKA_ERR_ARG;
    c->stream = s ? (hipStream_t)s : c->own;
    return OKA_OK;
}

void *oka_ctx_stream(oka_ctx *c) { return c ? (void *)c->stream : nullptr; }

int oka_ctx_synchronize(oka_ctx *c)
{
    if (!c)
        return OKA_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OKA_OK;
}

double oka_pow_fifth(double x) { return host_pow_fifth(x); }
double oka_log(double x) { return host_log(x); }

int oka_model_eval(oka_ctx *c, const oka_model *m, int32_t kind, int64_t npts, const double *xy,
                   double *out, int32_t *status, uint32_t flags)
{
    if (!c || !m)
        return OKA_ERR_ARG;
    if (kind < OKA_EVAL_POTENTIAL || kind > OKA_EVAL_VELOCITY_CONFINED)
        return fail(c, OKA_ERR_ARG, "unknown evaluation kind %d", kind);
    if (npts < 0 || m->nwells < 0 || (m->nwells > 0 && !m->wells))
        return fail(c, OKA_ERR_ARG, "bad points or wells");
    if (npts == 0)
        return OKA_OK;
    if (!xy || !out)
        return fail(c, OKA_ERR_ARG, "xy/out are NULL");
    HIPCHK(c, hipSetDevice(c->device));
    const int width = (kind == OKA_EVAL_DISCHARGE || kind == OKA_EVAL_VELOCITY ||
                       kind == OKA_EVAL_VELOCITY_CONFINED) ? 2 : 1;
    EvalArgs a{};
    a.kind = kind;
    a.nwells = m->nwells;
    a.base = m->base;
    a.K = m->conductivity;
    a.n = m->porosity;
    a.b = m->thickness;
    a.xo = m->xo;
    a.yo = m->yo;
    for (int j = 0; j < 6; ++j)
        a.coef[j] = m->coef[j];
    a.npts = npts;
    const void *d;
    TRY(stage_in(c, kSlotWells, m->wells, sizeof(double) * 4 * m->nwells, 0, &d));
    a.wells = (const double *)d;
    TRY(stage_in(c, kSlotXY, xy, sizeof(double) * 2 * npts, flags, &d));
    a.xy = (const double *)d;
    void *o;
    TRY(stage_out(c, kSlotOut, out, sizeof(double) * width * npts, flags, false, &o));
    a.out = (double *)o;
    TRY(stage_out(c, kSlotEStatus, status, sizeof(int32_t) * npts, flags, false, &o));
    a.status = (int32_t *)o;
    HIPCHK(c, launch_eval_fp64(a, c->stream));
    TRY(copy_back(c, out, a.out, sizeof(double) * width * npts, flags));
    TRY(copy_back(c, status, a.status, sizeof(int32_t) * npts, flags));
    if (!(flags & OKA_ASYNC) || !(flags & OKA_PTR_DEVICE))
        HIPCHK(c, hipStreamSynchronize(c->stream));
    return OKA_OK;
}

int oka_fit_draw(oka_ctx *c, const oka_site *site, int32_t nreal, const double *K,
                 const double *b, const double *q, const double *z, double *ev, double *cov,
                 double *coef, int32_t *status, uint32_t flags)
{
    if (!c || !site)
        return OKA_ERR_ARG;
    if (nreal < 0 || site->nwells < 0 || site->nobs < 0)
        return fail(c, OKA_ERR_ARG, "negative sizes");
    if (nreal == 0)
        return OKA_OK;
    if (!K || !b || (site->nwells > 0 && (!q || !site->wells_xy)) || !site->obs)
        return fail(c, OKA_ERR_ARG, "missing fit inputs");
    if (site->nobs < 6)
        return fail(c, OKA_ERR_LINALG, "fewer than 6 observations: the fit is underdetermined");
    HIPCHK(c, hipSetDevice(c->device));
    FitArgs a{};
    a.nwells = site->nwells;
    a.nobs = site->nobs;
    a.nreal = nreal;
    a.xo = site->xo;
    a.yo = site->yo;
    a.base = site->base;
    const void *d;
    TRY(stage_in(c, kSlotWells, site->wells_xy, sizeof(double) * 2 * site->nwells, flags, &d));
    a.wells_xy = (const double *)d;
    TRY(stage_in(c, kSlotObs, site->obs, sizeof(double) * 4 * site->nobs, flags, &d));
    a.obs = (const double *)d;
    TRY(stage_in(c, kSlotK, K, sizeof(double) * nreal, flags, &d));
    a.K = (const double *)d;
    TRY(stage_in(c, kSlotB, b, sizeof(double) * nreal, flags, &d));
    a.b = (const double *)d;
    TRY(stage_in(c, kSlotQ, q, sizeof(double) * nreal * site->nwells, flags, &d));
    a.q = (const double *)d;
    TRY(stage_in(c, kSlotZ, z, sizeof(double) * 6 * nreal, flags, &d));
    a.z = (const double *)d;
    void *o;
    TRY(ensure(c, kSlotLogs, sizeof(double) * site->nobs * std::max(site->nwells, 1), &o));
    a.logr2 = (const double *)o;
    TRY(stage_out(c, kSlotEv, ev, sizeof(double) * 6 * nreal, flags, false, &o));
    a.ev = (double *)o;
    TRY(stage_out(c, kSlotCov, cov, sizeof(double) * 36 * nreal, flags, false, &o));
    a.cov = (double *)o;
    TRY(stage_out(c, kSlotCoef, coef, sizeof(double) * 6 * nreal, flags, false, &o));
    a.coef = (double *)o;
    // statuses are always produced, so that the call can report the first failure
    int32_t *dst = status;
    uint32_t sflags = flags;
    std::vector<int32_t> host_status;
    if (!status) {
        host_status.resize(nreal);
        dst = host_status.data();
        sflags = 0;
    }
    TRY(stage_out(c, kSlotFStatus, dst, sizeof(int32_t) * nreal, sflags, false, &o));
    a.status = (int32_t *)o;
    HIPCHK(c, launch_fit_logs(a, (double *)a.logr2, c->stream));
    HIPCHK(c, launch_fit_draw(a, c->stream));
    TRY(copy_back(c, ev, a.ev, sizeof(double) * 6 * nreal, flags));
    TRY(copy_back(c, cov, a.cov, sizeof(double) * 36 * nreal, flags));
    TRY(copy_back(c, coef, a.coef, sizeof(double) * 6 * nreal, flags));
    TRY(copy_back(c, dst, a.status, sizeof(int32_t) * nreal, sflags));
    if ((flags & OKA_ASYNC) && (sflags & OKA_PTR_DEVICE))
        return OKA_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int32_t *hs = dst;
    std::vector<int32_t> tmp;
    if (sflags & OKA_PTR_DEVICE) {
        tmp.resize(nreal);
        HIPCHK(c, hipMemcpy(tmp.data(), a.status, sizeof(int32_t) * nreal, hipMemcpyDeviceToHost));
        hs = tmp.data();
    }
    for (int32_t r = 0; r < nreal; ++r) {
        if (hs[r] == OKA_ERR_RANGE)
            return fail(c, OKA_ERR_RANGE, "model.fit_coeficient: elevation < base (realization %d)", r);
        if (hs[r] == OKA_ERR_LINALG)
            return fail(c, OKA_ERR_LINALG,
                        "normal matrix or covariance not positive definite (realization %d)", r);
    }
    return OKA_OK;
}

int oka_grid_history(const oka_grid *init, int64_t ntraces, const double *bbox,
                     oka_grid *final_grid, oka_epoch *epochs, int64_t max_epochs,
                     int64_t *nepochs)
{
    // probabilityfield.py:196-261 applied to each trace's bbox in order.
    if (!init || !final_grid || !nepochs || (ntraces > 0 && !bbox) || ntraces < 0)
        return OKA_ERR_ARG;
    if (!(init->dx > 0.0) || !(init->dy > 0.0))
        return OKA_ERR_RANGE;
    struct St {
        double xmin, xmax, ymin, ymax;
        int64_t nrows, ncols, cl, rb;   // cl/rb: cumulative left/bottom growth
    } s{init->xmin, init->xmax, init->ymin, init->ymax, init->nrows, init->ncols, 0, 0};
    const double dx = init->dx, dy = init->dy;
    std::vector<int64_t> first;
    std::vector<St> states;
    bool empty = init->nrows == 0 || init->ncols == 0;
    if (!empty) {
        first.push_back(0);
        states.push_back(s);
    }
    for (int64_t k = 0; k < ntraces; ++k) {
        const double bx0 = bbox[4 * k], bx1 = bbox[4 * k + 1];
        const double by0 = bbox[4 * k + 2], by1 = bbox[4 * k + 3];
        if (bx0 > bx1 || by0 > by1)
            return OKA_ERR_RANGE;
        bool changed = false;
        if (empty) {
            s.xmin = bx0 - dx;
            s.ymin = by0 - dy;
            const double nc = ceil((bx1 - s.xmin) / dx) + 2.0;
            const double nr = ceil((by1 - s.ymin) / dy) + 2.0;
            s.ncols = nc > 3.0 ? (int64_t)nc : 3;
            s.nrows = nr > 3.0 ? (int64_t)nr : 3;
            s.xmax = s.xmin + (double)(s.ncols - 1) * dx;
            s.ymax = s.ymin + (double)(s.nrows - 1) * dy;
            empty = false;
            changed = true;
        } else {
            while (bx0 <= s.xmin) { s.xmin -= dx; s.ncols += 1; s.cl += 1; changed = true; }
            while (bx1 >= s.xmax) { s.xmax += dx; s.ncols += 1; changed = true; }
            while (by0 <= s.ymin) { s.ymin -= dy; s.nrows += 1; s.rb += 1; changed = true; }
            while (by1 >= s.ymax) { s.ymax += dy; s.nrows += 1; changed = true; }
        }
        if (changed) {
            if (!first.empty() && first.back() == k) {
                states.back() = s;
            } else {
                first.push_back(k);
                states.push_back(s);
            }
        }
    }
    if (s.nrows > INT32_MAX || s.ncols > INT32_MAX)
        return OKA_ERR_RANGE;
    final_grid->xmin = s.xmin;
    final_grid->ymin = s.ymin;
    final_grid->xmax = s.xmax;
    final_grid->ymax = s.ymax;
    final_grid->dx = dx;
    final_grid->dy = dy;
    final_grid->nrows = (int32_t)s.nrows;
    final_grid->ncols = (int32_t)s.ncols;
    *nepochs = (int64_t)first.size();
    if (!epochs)
        return OKA_OK;
    if (max_epochs < (int64_t)first.size())
        return OKA_ERR_CAPACITY;
    for (size_t e = 0; e < first.size(); ++e) {
        const St &t = states[e];
        epochs[e].first_trace = first[e];
        epochs[e].xmin = t.xmin;
        epochs[e].ymin = t.ymin;
        epochs[e].nrows = (int32_t)t.nrows;
        epochs[e].ncols = (int32_t)t.ncols;
        epochs[e].row_off = (int32_t)(s.rb - t.rb);
        epochs[e].col_off = (int32_t)(s.cl - t.cl);
    }
    return OKA_OK;
}

static int trace_common(oka_ctx *c, const oka_site *site, int32_t nreal, const double *K,
                        const double *n, const double *b, const double *q, const double *coef,
                        const oka_trace_opts *o, uint32_t flags, TraceArgs &a)
{
    TRY(check_opts(c, o));
    if (!site || nreal < 0 || site->nwells < 0)
        return fail(c, OKA_ERR_ARG, "bad site or realization count");
    if (nreal > 0 && (!K || !n || !b || !coef || (site->nwells > 0 && (!q || !site->wells_xy))))
        return fail(c, OKA_ERR_ARG, "missing realization inputs");
    HIPCHK(c, hipSetDevice(c->device));
    a.nwells = site->nwells;
    a.npaths = o->npaths;
    a.xo = site->xo;
    a.yo = site->yo;
    a.nreal = nreal;
    a.confined = o->confined ? 1 : 0;
    a.duration = o->duration;
    a.tol = o->tol;
    a.maxstep = o->maxstep;
    a.umbra = o->umbra;
    a.max_steps = o->max_steps > 0 ? o->max_steps : kDefaultMaxSteps;
    const void *d;
    TRY(stage_in(c, kSlotWells, site->wells_xy, sizeof(double) * 2 * site->nwells, flags, &d));
    a.wells_xy = (const double *)d;
    TRY(stage_in(c, kSlotStart, o->start_xy, sizeof(double) * 2 * o->npaths, flags, &d));
    a.start_xy = (const double *)d;
    TRY(stage_in(c, kSlotK, K, sizeof(double) * nreal, flags, &d));
    a.K = (const double *)d;
    TRY(stage_in(c, kSlotN, n, sizeof(double) * nreal, flags, &d));
    a.n = (const double *)d;
    TRY(stage_in(c, kSlotB, b, sizeof(double) * nreal, flags, &d));
    a.b = (const double *)d;
    TRY(stage_in(c, kSlotQ, q, sizeof(double) * nreal * site->nwells, flags, &d));
    a.q = (const double *)d;
    TRY(stage_in(c, kSlotCoef, coef, sizeof(double) * 6 * nreal, flags, &d));
    a.coef = (const double *)d;
    return OKA_OK;
}

static int chronicle_finish(oka_ctx *c, const TraceArgs &a, int32_t rasterize, int64_t nnodes,
                     int64_t ntr, int32_t *counts, double *bbox, int32_t *trace_status,
                     int64_t *trace_steps, oka_trace_stats *stats, uint32_t flags,
                     const oka_window *win, uint8_t *marks, uint16_t *masks);

// The far-well plan of a single-site pool launch, in the context's slots.
int stage_far(oka_ctx *c, const oka_site *site, const oka_trace_opts *o, uint32_t flags,
              TraceArgs &a)
{
    if (far_requested(a, o->precision, o->far_order) == 0)
        return OKA_OK;
    std::vector<double> wh;
    const double *wxy = site->wells_xy;
    if (flags & OKA_PTR_DEVICE) {
        wh.resize(2 * (size_t)a.nwells);
        HIPCHK(c, hipMemcpy(wh.data(), wxy, sizeof(double) * wh.size(), hipMemcpyDeviceToHost));
        wxy = wh.data();
    }
    FarPlan P;
    far_plan(a, wxy, o->precision, o->far_order, &P);
    if (P.order == 0)
        return OKA_OK;
    void *d[2];
    const std::vector<int32_t> *v[2] = {&P.near, &P.wbox};
    const Slot slot[2] = {kSlotFarNear, kSlotFarWbox};
    for (int i = 0; i < 2; ++i) {
        TRY(ensure(c, slot[i], sizeof(int32_t) * std::max<size_t>(v[i]->size(), 1), &d[i]));
        if (!v[i]->empty())
            HIPCHK(c, hipMemcpy(d[i], v[i]->data(), sizeof(int32_t) * v[i]->size(),
                                hipMemcpyHostToDevice));
    }
    far_args(P, (const int32_t *)d[0], (const int32_t *)d[1], a);
    return OKA_OK;
}

// Launches the persistent pool tracer over `nsites` sites (host copies of
// their argument blocks, with device pointers inside), with its buffers in the
// context's slots; times it with the context events when `timed`.
int run_pool(oka_ctx *c, const TraceArgs *sites, int nsites, const oka_trace_opts *o, bool timed)
{
    PoolShape sh;
    TRY(pool_shape(c, sites, nsites, o, o->precision, &sh));
    void *p_sites, *p_first, *p_queue, *p_hbm, *p_rec = nullptr, *p_beta = nullptr;
    if (sh.far_stride > 0)
        TRY(ensure(c, kSlotFarBeta, sizeof(float2) * pool_beta_count(sh) * sh.far_stride,
                   &p_beta));
    TRY(ensure(c, kSlotPoolSites, sizeof(TraceArgs) * nsites, &p_sites));
    TRY(ensure(c, kSlotPoolFirst, sizeof(int64_t) * sh.first.size(), &p_first));
    TRY(ensure(c, kSlotPoolQueue, sizeof(unsigned long long), &p_queue));
    TRY(ensure(c, kSlotPoolHbm, sizeof(uint32_t) * sh.hbm_words, &p_hbm));
    if (sh.rec_words > 0)
        TRY(ensure(c, kSlotPoolRec, sizeof(uint32_t) * sh.rec_words, &p_rec));
    HIPCHK(c, hipMemcpyAsync(p_sites, sites, sizeof(TraceArgs) * nsites, hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(p_first, sh.first.data(), sizeof(int64_t) * sh.first.size(),
                             hipMemcpyHostToDevice, c->stream));
    const PoolArgs p = pool_args(sh, p_sites, p_first, p_queue, p_hbm, p_rec, p_beta);
    if (timed)
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    TRY(pool_launch(c, sh, p));
    if (timed)
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    return OKA_OK;
}

// Where a realization's visited-node bitmap lives (OKA_BITMAP_*, see
// include/oka.h): a flat LDS bitmap when the whole grid fits in the LDS
// budget beside the wells and window counters, else 32 x 32-node tiles in LDS
// slots, else a flat bitmap in HBM.
int choose_bitmap(oka_ctx *c, const oka_trace_opts *o, size_t base_lds, TraceArgs &a, int *kind)
{
    const int mode = o->bitmap_mode;
    if (mode < OKA_BITMAP_AUTO || mode > OKA_BITMAP_HBM_FLAT)
        return fail(c, OKA_ERR_ARG, "unknown bitmap mode %d", mode);
    const size_t budget = std::min<size_t>(kLdsBudget, c->max_lds);
    const bool flat_fits = base_lds + (size_t)a.nwords * 4 <= budget;
    a.tiles_x = (a.gncols + 31) / 32;
    const int64_t ntiles = (int64_t)a.tiles_x * ((a.gnrows + 31) / 32);
    int64_t slots = 0;
    if (base_lds + (size_t)(ntiles + 1) * 4 < budget)
        slots = (int64_t)(budget - base_lds - (size_t)(ntiles + 1) * 4) / (4 * 33);
    if (o->tile_slots > 0)
        slots = std::min<int64_t>(slots, o->tile_slots);
    a.ntiles = (int32_t)ntiles;
    a.tile_slots = (int32_t)slots;
    if (mode == OKA_BITMAP_LDS_FLAT && !flat_fits)
        return fail(c, OKA_ERR_ARG, "the grid's bitmap (%d words) does not fit in LDS", a.nwords);
    if (mode == OKA_BITMAP_TILED && slots < 1)
        return fail(c, OKA_ERR_ARG, "no LDS left for tile slots (%lld tiles)", (long long)ntiles);
    if (mode == OKA_BITMAP_LDS_FLAT || (mode == OKA_BITMAP_AUTO && flat_fits))
        *kind = kBmLdsFlat;
    else if (mode == OKA_BITMAP_TILED ||
             (mode == OKA_BITMAP_AUTO && slots >= kMinTileSlots))
        *kind = kBmTiled;
    else
        *kind = kBmHbmFlat;
    return OKA_OK;
}

int chronicle_impl(oka_ctx *c, const oka_site *site, int32_t nreal, const double *K,
                          const double *n, const double *b, const double *q, const double *coef,
                          const oka_trace_opts *o, const oka_grid *g, int32_t nepochs,
                          const oka_epoch *epochs, int32_t rasterize, int32_t *counts,
                          double *bbox, int32_t *trace_status, int64_t *trace_steps,
                          oka_trace_stats *stats, uint32_t flags, const oka_window *win,
                          uint8_t *marks, uint16_t *masks = nullptr)
{
    if (!c)
        return OKA_ERR_ARG;
    if (!g)
        return fail(c, OKA_ERR_ARG, "grid is NULL");
    if (rasterize && (g->nrows <= 0 || g->ncols <= 0 || (!counts && !marks)))
        return fail(c, OKA_ERR_ARG, "rasterization needs a non-empty grid and an output");
    if (marks && (!win || win->cell < 1 || win->cell > 15 || win->nrows <= 0 ||
                  win->ncols <= 0 || (int64_t)win->nrows * win->ncols > 16384))
        return fail(c, OKA_ERR_ARG, "bad window (cell 1..15, at most 16384 cells)");
    if (rasterize && (!(g->dx > 0.0) || !(g->dy > 0.0)))
        return fail(c, OKA_ERR_RANGE, "grid spacing must be > 0");
    if (nepochs < 0 || (nepochs > 0 && !epochs))
        return fail(c, OKA_ERR_ARG, "bad epochs");
    TraceArgs a{};
    TRY(trace_common(c, site, nreal, K, n, b, q, coef, o, flags, a));
    const int64_t ntr = (int64_t)nreal * o->npaths;
    if (ntr == 0) {
        if (stats)
            memset(stats, 0, sizeof(*stats));
        return OKA_OK;
    }
    a.gxmin = g->xmin;
    a.gymin = g->ymin;
    a.gxmax = g->xmax;
    a.gymax = g->ymax;
    a.gdx = g->dx;
    a.gdy = g->dy;
    a.gnrows = g->nrows;
    a.gncols = g->ncols;
    a.rasterize = rasterize ? 1 : 0;
    const int64_t nnodes = (int64_t)g->nrows * g->ncols;
    a.nwords = rasterize ? (int32_t)((nnodes + 31) / 32) : 0;
    // epochs: the caller's, or the whole grid as one epoch
    oka_epoch whole{0, g->xmin, g->ymin, g->nrows, g->ncols, 0, 0};
    const void *d;
    if (nepochs == 0) {
        TRY(stage_in(c, kSlotEpochs, &whole, sizeof(whole), 0, &d));
        a.nepochs = 1;
    } else {
        TRY(stage_in(c, kSlotEpochs, epochs, sizeof(oka_epoch) * nepochs, flags, &d));
        a.nepochs = nepochs;
    }
    a.epochs = (const oka_epoch *)d;
    void *o2;
    TRY(stage_out(c, kSlotCounts, rasterize ? counts : nullptr, sizeof(int32_t) * nnodes, flags,
                  true, &o2));
    a.counts = (int32_t *)o2;
    TRY(stage_out(c, kSlotBbox, bbox, sizeof(double) * 4 * ntr, flags, false, &o2));
    a.bbox = (double *)o2;
    TRY(stage_out(c, kSlotTStatus, trace_status, sizeof(int32_t) * ntr, flags, false, &o2));
    a.trace_status = (int32_t *)o2;
    TRY(stage_out(c, kSlotTSteps, trace_steps, sizeof(int64_t) * ntr, flags, false, &o2));
    a.trace_steps = (int64_t *)o2;
    TRY(ensure(c, kSlotStats, sizeof(unsigned long long) * kStatCount, &o2));
    a.stats = (unsigned long long *)o2;
    TRY(fill_u32(c, a.stats, 2 * kStatCount));  // 64-bit slots
    size_t win_lds = 0;
    if (marks) {
        a.w_row0 = win->row0;
        a.w_col0 = win->col0;
        a.w_cell = win->cell;
        a.w_nr = win->nrows;
        a.w_nc = win->ncols;
        win_lds = (size_t)win->nrows * win->ncols * 4;
        TRY(stage_out(c, kSlotMarks, marks, (size_t)nreal * win->nrows * win->ncols, flags, false,
                      &o2));
        a.wmarks = (uint8_t *)o2;
        if (masks) {
            TRY(stage_out(c, kSlotMasks, masks,
                          sizeof(uint16_t) * nreal * win->nrows * win->ncols, flags, false, &o2));
            a.wmasks = (uint16_t *)o2;
        }
    }

    const bool timed = stats != nullptr;
    const bool fp64 = o->precision == OKA_FP64_EXACT;
    // AUTO runs the pool where it has what the workgroup schedule lacks: the
    // far-well expansions of an FP32 confined site of many wells (3.7x on a
    // 200-well site).  Elsewhere the workgroup schedule is as fast or faster,
    // with the same counts, and a forced bitmap placement or tile-slot count
    // asks for it.
    bool pool = o->schedule == OKA_SCHED_POOL;
    if (o->schedule == OKA_SCHED_AUTO && !marks && o->bitmap_mode == OKA_BITMAP_AUTO &&
        o->tile_slots == 0 && far_requested(a, o->precision, o->far_order) > 0) {
        // the pool must fit: its tile index, and the LDS share of the fewest
        // waves per CU it would choose (or the caller's shape) for the slots
        const int tx = (a.gncols + 31) / 32;
        const int64_t nt = (int64_t)tx * ((a.gnrows + 31) / 32);
        const int slots = o->pool_slots > 0 ? std::min(o->pool_slots, 16) : 2;
        const int wpc = o->pool_waves_per_cu > 0 ? std::min(o->pool_waves_per_cu, 32) : 4;
        const size_t share = std::min(c->lds_per_cu / (size_t)wpc, c->max_lds);
        if (nt <= pool_max_tiles() && pool_lds_bytes(o->precision, slots, a.nwells, 8) <= share) {
            TRY(stage_far(c, site, o, flags, a));
            pool = a.far_order > 0;   // the plan was taken
        }
    }
    if (pool) {
        if (marks)
            return fail(c, OKA_ERR_ARG, "window marks need the workgroup schedule");
        a.tiles_x = (a.gncols + 31) / 32;
        a.ntiles = a.tiles_x * ((a.gnrows + 31) / 32);
        if (o->schedule == OKA_SCHED_POOL)
            TRY(stage_far(c, site, o, flags, a));
        TRY(run_pool(c, &a, 1, o, timed));
        return chronicle_finish(c, a, rasterize, nnodes, ntr, counts, bbox, trace_status,
                                trace_steps, stats, flags, win, marks, masks);
    }

    // launch shape: one realization per 256-lane workgroup, grid-stride
    const size_t base_lds =
        (fp64 ? trace_fp64_wells_lds(a.nwells) : trace_fp32_wells_lds(a.nwells)) + win_lds;
    if (base_lds > c->max_lds)
        return fail(c, OKA_ERR_ARG, "too many wells for LDS (%d)", a.nwells);
    int kind = kBmLdsFlat;
    int blocks = nreal;
    if (rasterize) {
        TRY(choose_bitmap(c, o, base_lds, a, &kind));
        if (kind != kBmLdsFlat) {
            // per-workgroup HBM regions: bound the resident set to what the
            // LDS lets run at once
            const size_t lds = base_lds + (kind == kBmTiled
                                               ? (size_t)tile_lds_words(a.ntiles, a.tile_slots) * 4
                                               : 0);
            const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, c->lds_per_cu / lds));
            blocks = std::min(nreal, c->num_cu * per_cu);
            const size_t words = kind == kBmTiled ? (size_t)a.ntiles * 32 : (size_t)a.nwords;
            TRY(ensure(c, kSlotBitmap, sizeof(uint32_t) * words * blocks, &o2));
            a.gbitmap = (uint32_t *)o2;
            if (kind == kBmTiled)   // overflow tiles start empty; each flush leaves them so
                TRY(fill_u32(c, a.gbitmap, (int64_t)words * blocks));
        }
    }
    if (timed)
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (fp64)
        HIPCHK(c, launch_trace_fp64(a, blocks, kind, c->stream));
    else
        HIPCHK(c, launch_trace_fp32(a, blocks, kind, c->stream));
    if (timed)
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    return chronicle_finish(c, a, rasterize, nnodes, ntr, counts, bbox, trace_status, trace_steps,
                            stats, flags, win, marks, masks);
}

// Copies a trace launch's outputs back (host buffers) and reads its statistics.
int chronicle_finish(oka_ctx *c, const TraceArgs &a, int32_t rasterize, int64_t nnodes,
                     int64_t ntr, int32_t *counts, double *bbox, int32_t *trace_status,
                     int64_t *trace_steps, oka_trace_stats *stats, uint32_t flags,
                     const oka_window *win, uint8_t *marks, uint16_t *masks)
{
    const int32_t nreal = a.nreal;
    TRY(copy_back(c, rasterize ? counts : nullptr, a.counts, sizeof(int32_t) * nnodes, flags));
    TRY(copy_back(c, bbox, a.bbox, sizeof(double) * 4 * ntr, flags));
    TRY(copy_back(c, trace_status, a.trace_status, sizeof(int32_t) * ntr, flags));
    TRY(copy_back(c, trace_steps, a.trace_steps, sizeof(int64_t) * ntr, flags));
    if (marks)
        TRY(copy_back(c, marks, a.wmarks, (size_t)nreal * win->nrows * win->ncols, flags));
    if (masks)
        TRY(copy_back(c, masks, a.wmasks, sizeof(uint16_t) * nreal * win->nrows * win->ncols,
                      flags));
    if ((flags & OKA_ASYNC) && (flags & OKA_PTR_DEVICE) && !stats)
        return OKA_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (stats) {
        unsigned long long h[kStatCount];
        HIPCHK(c, hipMemcpy(h, a.stats, sizeof(h), hipMemcpyDeviceToHost));
        stats->attempted = (int64_t)h[kStatAttempted];
        stats->accepted = (int64_t)h[kStatAccepted];
        for (int j = 0; j < 4; ++j)
            stats->status_count[j] = (int64_t)h[kStatStatus0 + j];
        stats->escaped = (int64_t)h[kStatEscaped];
        stats->visited = (int64_t)h[kStatVisited];
        stats->candidates = (int64_t)h[kStatCandidates];
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        stats->kernel_ms = ms;
    }
    return OKA_OK;
}

int oka_trace_chronicle(oka_ctx *c, const oka_site *site, int32_t nreal, const double *K,
                        const double *n, const double *b, const double *q, const double *coef,
                        const oka_trace_opts *o, const oka_grid *g, int32_t nepochs,
                        const oka_epoch *epochs, int32_t rasterize, int32_t *counts,
                        double *bbox, int32_t *trace_status, int64_t *trace_steps,
                        oka_trace_stats *stats, uint32_t flags)
{
    if (c && rasterize && !counts)
        return fail(c, OKA_ERR_ARG, "rasterization needs counts");
    return chronicle_impl(c, site, nreal, K, n, b, q, coef, o, g, nepochs, epochs, rasterize,
                          counts, bbox, trace_status, trace_steps, stats, flags, nullptr,
                          nullptr);
}

int oka_trace_chronicle_noalloc(oka_ctx *c, const oka_site *site, int32_t nreal, const double *K,
                                const double *n, const double *b, const double *q,
                                const double *coef, const oka_trace_opts *o, const oka_grid *g,
                                int32_t nepochs, const oka_epoch *epochs, int32_t rasterize,
                                int32_t *counts, double *bbox, int32_t *trace_status,
                                int64_t *trace_steps, oka_trace_stats *stats, uint32_t flags)
{
    if (!c)
        return OKA_ERR_ARG;
    if (!(flags & OKA_PTR_DEVICE))
        return fail(c, OKA_ERR_ARG, "oka_trace_chronicle_noalloc takes device arrays only "
                                    "(flags need OKA_PTR_DEVICE)");
    if (!c->scratch)
        return fail(c, OKA_ERR_DEVICE, "no scratch address on device %d: %s", c->device,
                    hipGetErrorString(c->scratch_err));
    if (!g)
        return fail(c, OKA_ERR_ARG, "grid is NULL");
    if (rasterize && (g->nrows <= 0 || g->ncols <= 0 || !counts))
        return fail(c, OKA_ERR_ARG, "rasterization needs a non-empty grid and counts");
    if (rasterize && (!(g->dx > 0.0) || !(g->dy > 0.0)))
        return fail(c, OKA_ERR_RANGE, "grid spacing must be > 0");
    if (nepochs < 0 || (nepochs > 0 && !epochs))
        return fail(c, OKA_ERR_ARG, "bad epochs");
    if (o && o->schedule == OKA_SCHED_POOL)
        return fail(c, OKA_ERR_ARG, "oka_trace_chronicle_noalloc runs the workgroup schedule only");
    TraceArgs a{};
    TRY(trace_common(c, site, nreal, K, n, b, q, coef, o, flags, a));
    if (stats)
        memset(stats, 0, sizeof(*stats));
    const int64_t ntr = (int64_t)nreal * o->npaths;
    if (ntr == 0)
        return OKA_OK;
    a.gxmin = g->xmin;
    a.gymin = g->ymin;
    a.gxmax = g->xmax;
    a.gymax = g->ymax;
    a.gdx = g->dx;
    a.gdy = g->dy;
    a.gnrows = g->nrows;
    a.gncols = g->ncols;
    a.rasterize = rasterize ? 1 : 0;
    const int64_t nnodes = (int64_t)g->nrows * g->ncols;
    a.nwords = rasterize ? (int32_t)((nnodes + 31) / 32) : 0;
    unsigned char *scr = static_cast<unsigned char *>(c->scratch);
    a.stats = reinterpret_cast<unsigned long long *>(scr + kScrStats);
    TRY(fill_u32(c, a.stats, 2 * kStatCount));   // 64-bit slots
    if (nepochs == 0) {
        const oka_epoch whole{0, g->xmin, g->ymin, g->nrows, g->ncols, 0, 0};
        oka_epoch *dst = reinterpret_cast<oka_epoch *>(scr + kScrEpoch);
        hipLaunchKernelGGL(k_store_epoch, dim3(1), dim3(64), 0, c->stream, whole, dst);
        HIPCHK(c, hipGetLastError());
        a.nepochs = 1;
        a.epochs = dst;
    } else {
        a.nepochs = nepochs;
        a.epochs = epochs;
    }
    a.counts = rasterize ? counts : nullptr;
    a.bbox = bbox;
    a.trace_status = trace_status;
    a.trace_steps = trace_steps;

    const bool fp64 = o->precision == OKA_FP64_EXACT;
    const size_t base_lds = fp64 ? trace_fp64_wells_lds(a.nwells) : trace_fp32_wells_lds(a.nwells);
    if (base_lds > c->max_lds)
        return fail(c, OKA_ERR_ARG, "too many wells for LDS (%d)", a.nwells);
    int kind = kBmLdsFlat;
    int blocks = nreal;
    if (rasterize) {
        TRY(choose_bitmap(c, o, base_lds, a, &kind));
        if (kind != kBmLdsFlat) {
            // per-workgroup HBM bitmaps in the scratch region: as many
            // workgroups in flight as the LDS lets run and the region holds
            const size_t lds = base_lds + (kind == kBmTiled
                                               ? (size_t)tile_lds_words(a.ntiles, a.tile_slots) * 4
                                               : 0);
            const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, c->lds_per_cu / lds));
            const int64_t per_block = kind == kBmTiled ? (int64_t)a.ntiles * 32 : (int64_t)a.nwords;
            const int64_t room = (int64_t)((OKA_NOALLOC_SCRATCH_BYTES - kScrBitmap) / 4) / per_block;
            if (room < 1)
                return fail(c, OKA_ERR_CAPACITY,
                            "a %d x %d grid's bitmap (%lld words) exceeds the no-allocation "
                            "scratch (%u bytes)", g->nrows, g->ncols, (long long)per_block,
                            (unsigned)OKA_NOALLOC_SCRATCH_BYTES);
            blocks = (int)std::min<int64_t>(std::min<int64_t>(nreal, (int64_t)c->num_cu * per_cu),
                                            room);
            a.gbitmap = reinterpret_cast<uint32_t *>(scr + kScrBitmap);
            if (kind == kBmTiled)   // overflow tiles start empty; each flush leaves them so
                TRY(fill_u32(c, a.gbitmap, per_block * blocks));
        }
    }
    if (stats)
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (fp64)
        HIPCHK(c, launch_trace_fp64(a, blocks, kind, c->stream));
    else
        HIPCHK(c, launch_trace_fp32(a, blocks, kind, c->stream));
    if (!stats) {
        if (!(flags & OKA_ASYNC))
            HIPCHK(c, hipStreamSynchronize(c->stream));
        return OKA_OK;
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unsigned long long h[kStatCount];
    HIPCHK(c, hipMemcpy(h, a.stats, sizeof(h), hipMemcpyDeviceToHost));
    stats->attempted = (int64_t)h[kStatAttempted];
    stats->accepted = (int64_t)h[kStatAccepted];
    for (int j = 0; j < 4; ++j)
        stats->status_count[j] = (int64_t)h[kStatStatus0 + j];
    stats->escaped = (int64_t)h[kStatEscaped];
    stats->visited = (int64_t)h[kStatVisited];
    stats->candidates = (int64_t)h[kStatCandidates];
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    stats->kernel_ms = ms;
    return OKA_OK;
}

int oka_trace_marks(oka_ctx *c, const oka_site *site, int32_t nreal, const double *K,
                    const double *n, const double *b, const double *q, const double *coef,
                    const oka_trace_opts *o, const oka_grid *g, const oka_window *win,
                    uint8_t *marks, uint16_t *node_masks, int32_t *counts,
                    oka_trace_stats *stats, uint32_t flags)
{
    if (c && !marks)
        return fail(c, OKA_ERR_ARG, "marks is NULL");
    if (c && node_masks && win && win->cell > 4)
        return fail(c, OKA_ERR_ARG, "node masks need cells of at most 4 x 4 nodes (cell %d)",
                    win->cell);
    return chronicle_impl(c, site, nreal, K, n, b, q, coef, o, g, 0, nullptr, 1, counts,
                          nullptr, nullptr, nullptr, stats, flags, win, marks, node_masks);
}

int oka_trace_vertices(oka_ctx *c, const oka_site *site, int32_t nreal, const double *K,
                       const double *n, const double *b, const double *q, const double *coef,
                       const oka_trace_opts *o, int64_t cap, double *verts, int64_t *nverts,
                       int32_t *trace_status, int64_t *trace_steps, uint32_t flags)
{
    if (!c)
        return OKA_ERR_ARG;
    if (cap < 0 || (cap > 0 && !verts) || !nverts)
        return fail(c, OKA_ERR_ARG, "bad vertex buffers");
    TraceArgs a{};
    TRY(trace_common(c, site, nreal, K, n, b, q, coef, o, flags, a));
    if (o->precision != OKA_FP64_EXACT)
        return fail(c, OKA_ERR_ARG, "vertices are produced by the FP64-exact tracer only");
    const int64_t ntr = (int64_t)nreal * o->npaths;
    if (ntr == 0)
        return OKA_OK;
    a.gxmin = -INFINITY;
    a.gymin = -INFINITY;
    a.gxmax = INFINITY;
    a.gymax = INFINITY;
    a.nepochs = 1;
    a.vcap = cap;
    void *o2;
    TRY(stage_out(c, kSlotVerts, verts, sizeof(double) * 2 * cap * ntr, flags, false, &o2));
    a.verts = (double *)o2;
    TRY(stage_out(c, kSlotNVerts, nverts, sizeof(int64_t) * ntr, flags, false, &o2));
    a.nverts = (int64_t *)o2;
    TRY(stage_out(c, kSlotTStatus, trace_status, sizeof(int32_t) * ntr, flags, false, &o2));
    a.trace_status = (int32_t *)o2;
    TRY(stage_out(c, kSlotTSteps, trace_steps, sizeof(int64_t) * ntr, flags, false, &o2));
    a.trace_steps = (int64_t *)o2;
    HIPCHK(c, launch_vertices_fp64(a, c->stream));
    TRY(copy_back(c, verts, a.verts, sizeof(double) * 2 * cap * ntr, flags));
    TRY(copy_back(c, nverts, a.nverts, sizeof(int64_t) * ntr, flags));
    TRY(copy_back(c, trace_status, a.trace_status, sizeof(int32_t) * ntr, flags));
    TRY(copy_back(c, trace_steps, a.trace_steps, sizeof(int64_t) * ntr, flags));
    if (!(flags & OKA_ASYNC) || !(flags & OKA_PTR_DEVICE))
        HIPCHK(c, hipStreamSynchronize(c->stream));
    return OKA_OK;
}

int oka_rasterize(oka_ctx *c, const oka_grid *g, int64_t nverts, const double *xy, double umbra,
                  uint8_t *rgrid, uint32_t flags)
{
    if (!c || !g)
        return OKA_ERR_ARG;
    if (g->nrows <= 0 || g->ncols <= 0 || !rgrid)
        return fail(c, OKA_ERR_ARG, "empty grid");
    if (nverts < 2)
        return OKA_OK;
    if (!xy)
        return fail(c, OKA_ERR_ARG, "xy is NULL");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t nnodes = (int64_t)g->nrows * g->ncols;
    const int64_t nwords = (nnodes + 31) / 32;
    RasterArgs a{};
    a.xmin = g->xmin;
    a.ymin = g->ymin;
    a.dx = g->dx;
    a.dy = g->dy;
    a.umbra = umbra;
    a.nrows = g->nrows;
    a.ncols = g->ncols;
    a.nseg = nverts - 1;
    const void *d;
    TRY(stage_in(c, kSlotXY, xy, sizeof(double) * 2 * nverts, flags, &d));
    a.xy = (const double *)d;
    void *bm;
    TRY(ensure(c, kSlotBitmap, sizeof(uint32_t) * nwords, &bm));
    a.bitmap = (uint32_t *)bm;
    TRY(fill_u32(c, bm, nwords));
    HIPCHK(c, (flags & OKA_RASTER_FP32) ? launch_raster_fp32(a, c->stream)
                                        : launch_raster_fp64(a, c->stream));
    void *rg;
    TRY(stage_out(c, kSlotRgrid, rgrid, (size_t)nnodes, flags, true, &rg));
    hipLaunchKernelGGL(k_bits_to_u8, dim3((unsigned)((nnodes + 255) / 256)), dim3(256), 0,
                       c->stream, (const uint32_t *)bm, (uint8_t *)rg, nnodes);
    HIPCHK(c, hipGetLastError());
    TRY(copy_back(c, rgrid, rg, (size_t)nnodes, flags));
    if (!(flags & OKA_ASYNC) || !(flags & OKA_PTR_DEVICE))
        HIPCHK(c, hipStreamSynchronize(c->stream));
    return OKA_OK;
}

}  // extern "C"
