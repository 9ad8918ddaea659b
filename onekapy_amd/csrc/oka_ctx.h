// oka_ctx.h -- the context object and host-side staging helpers shared by
// the entry-point translation units of liboka.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "oka_internal.h"

namespace oka {

// Device buffer slots owned by a context; grown on demand, freed at destroy.
enum Slot {
    kSlotWells, kSlotStart, kSlotK, kSlotN, kSlotB, kSlotQ, kSlotCoef, kSlotCounts,
    kSlotBbox, kSlotTStatus, kSlotTSteps, kSlotStats, kSlotEpochs, kSlotBitmap,
    kSlotVerts, kSlotNVerts, kSlotObs, kSlotLogs, kSlotZ, kSlotEv, kSlotCov, kSlotFStatus,
    kSlotXY, kSlotOut, kSlotEStatus, kSlotRgrid, kSlotMarks, kSlotMlpX, kSlotMlpMarks,
    kSlotMlpOrder, kSlotMlpOut, kSlotMlpOut2, kSlotMlpLoss, kSlotCvDists, kSlotCvSite,
    kSlotCvDraw, kSlotPoolSites, kSlotPoolFirst, kSlotPoolQueue, kSlotPoolHbm, kSlotPoolRec, kSlotQmc,
    kSlotFarNear, kSlotFarWbox, kSlotFarBeta, kSlotMasks, kSlotNodeProbs, kSlotNodeMasks,
    kSlotNodeMarks, kSlotNodeIn, kSlotNodeX, kSlotNodeCounts, kSlotNodePart, kSlotNodeAcc,
    kSlotCount
};

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

}  // namespace oka

struct oka_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int num_cu = 0;
    size_t max_lds = 0;      // per workgroup
    size_t lds_per_cu = 0;
    int warp_size = 64;
    std::string err;
    oka::DevBuf buf[oka::kSlotCount];
    // device address of this device's copy of the no-allocation scratch
    // array (oka_trace_chronicle_noalloc), looked up at create; null if the
    // lookup failed, with its HIP error in scratch_err
    void *scratch = nullptr;
    hipError_t scratch_err = hipSuccess;
};

namespace oka {

int fail(oka_ctx *c, int code, const char *fmt, ...);
int ensure(oka_ctx *c, Slot s, size_t bytes, void **out);
int stage_in(oka_ctx *c, Slot s, const void *src, size_t bytes, uint32_t flags, const void **out);
int stage_out(oka_ctx *c, Slot s, void *dst, size_t bytes, uint32_t flags, bool preload,
              void **out);
int copy_back(oka_ctx *c, void *dst, const void *dev, size_t bytes, uint32_t flags);
int fill_u32(oka_ctx *c, void *p, int64_t nwords);
int fill_u32v(oka_ctx *c, void *p, uint32_t value, int64_t nwords);

// The far-well plan of one site (fp32::FarField): its box grid, each box's
// near wells and each well's box, in the FP32 tracers' target-relative
// coordinates; order 0 when the site sums every well directly.
struct FarPlan {
    int order = 0, nbx = 0, nby = 0, near_stride = 0;
    float x0 = 0.0f, y0 = 0.0f, len = 0.0f;
    std::vector<int32_t> near, wbox;   // TraceArgs::far_near, far_wbox
};
// The order oka_trace_opts.far_order asks for site `a` in `precision`
// (0: none), and the plan of the site (wells_host: its wells, nwells x 2,
// host memory); far_args points `a` at the plan's device copies.
int far_requested(const TraceArgs &a, int precision, int far_order);
void far_plan(const TraceArgs &a, const double *wells_host, int precision, int far_order,
              FarPlan *out);
void far_args(const FarPlan &P, const int32_t *near, const int32_t *wbox, TraceArgs &a);

// Shape of a persistent pool launch (oka_pool.hip) over a table of sites.
struct PoolShape {
    int precision = OKA_FP32, conf_mode = 0;   // the arithmetic and velocity fields traced
    int nsites = 0, chunk = 64, slots = 4, wpc = 8, waves = 0, pool_tiles = 0;
    int max_nw = 1, max_tiles = 1, dirty_words = 1;
    int64_t total = 0, nrec = 0, hbm_words = 0, rec_words = 0;
    int64_t far_stride = 0;       // float2 of far-well coefficients per slot
    int64_t far_rec = 0;          // records holding a far-well table their chunks share
    std::vector<int64_t> first;   // items, then realizations, per site: 2 x (nsites + 1)
};
// Far-well tables of a launch: one per slot, then one per sharing record.
inline int64_t pool_beta_count(const PoolShape &sh)
{
    return sh.far_stride > 0 ? (int64_t)sh.waves * sh.slots + sh.far_rec : 0;
}
int pool_shape(oka_ctx *c, const TraceArgs *sites, int nsites, const oka_trace_opts *o,
               int precision, PoolShape *out);
// Device buffers: sites [nsites] TraceArgs, first [2 (nsites + 1)] int64
// (uploaded by the caller), queue (8 bytes), hbm [hbm_words], rec [rec_words],
// far_beta [pool_beta_count * far_stride] float2 (when far_stride > 0).
PoolArgs pool_args(const PoolShape &sh, void *sites, void *first, void *queue, void *hbm,
                   void *rec, void *far_beta);
// Clears the launch's queue, bitmaps and records and launches it: kernel
// launches only, so a caller may capture it in a graph.
int pool_launch(oka_ctx *c, const PoolShape &sh, const PoolArgs &p);

}  // namespace oka

#define HIPCHK(c, expr)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return ::oka::fail((c), OKA_ERR_DEVICE, "%s: %s (%s:%d)", #expr,                  \
                               hipGetErrorString(e_), __FILE__, __LINE__);                    \
    } while (0)

#define TRY(expr)                                                                             \
    do {                                                                                      \
        int rc_ = (expr);                                                                     \
        if (rc_)                                                                              \
            return rc_;                                                                       \
    } while (0)
