// oka_field.hip -- the capture-probability field after the run: Pr =
// pgrid / total_weight, its optional Gaussian smoothing, and its descending
// sort for the impact curve (oneka/visualize.py:228-233, 382-385).
//
// FP64 with contraction off: the smoothing restates the separable 1-D
// correlation of scipy.ndimage.gaussian_filter(Z, sigma, mode='constant',
// cval=0.0) (visualize.py:233) in its own operation order -- for a symmetric
// kernel of radius r with weights w[0..2r] (centre w[r]), each output is
//   acc = x[0] * w[r];  then for j = r, r - 1, ..., 1:
//   acc = acc + (x[-j] + x[+j]) * w[r - j]
// with zeros outside the grid -- along axis 0 (rows), then along axis 1.
// The weights are the caller's (the host computes them as scipy does).
#include <math.h>

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "oka_ctx.h"
#include "oka_internal.h"

#pragma clang fp contract(off)

using namespace oka;

namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock) k_probability(const double *pgrid, double total, double *z,
                                                        int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        z[i] = pgrid[i] / total;   // visualize.py:230, 382
}

// One pass of the symmetric correlation along rows (axis 0: neighbours i +- j
// in column c) or columns (axis 1: neighbours c +- j in row i).
template <int kAxis>
__global__ void __launch_bounds__(kBlock) k_correlate(const double *in, double *out, int nrows,
                                                      int ncols, const double *w, int radius)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nrows * ncols)
        return;
    const int i = (int)(idx / ncols), c = (int)(idx - (int64_t)i * ncols);
    const int pos = kAxis == 0 ? i : c, len = kAxis == 0 ? nrows : ncols;
    const int64_t step = kAxis == 0 ? ncols : 1;
    double acc = in[idx] * w[radius];
    for (int j = radius; j >= 1; --j) {
        const double lo = pos - j >= 0 ? in[idx - j * step] : 0.0;
        const double hi = pos + j < len ? in[idx + j * step] : 0.0;
        acc = acc + (lo + hi) * w[radius - j];
    }
    out[idx] = acc;
}

unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

extern "C" {

int oka_field_smooth(oka_ctx *c, int32_t nrows, int32_t ncols, const double *pgrid,
                     double total_weight, int32_t radius, const double *weights, double *out,
                     uint32_t flags)
{
    if (!c)
        return OKA_ERR_ARG;
    if (nrows < 0 || ncols < 0 || radius < 0 || (radius > 0 && !weights))
        return fail(c, OKA_ERR_ARG, "bad field or kernel sizes");
    const int64_t n = (int64_t)nrows * ncols;
    if (n == 0)
        return OKA_OK;
    if (!pgrid || !out)
        return fail(c, OKA_ERR_ARG, "pgrid/out are NULL");
    if (!(total_weight > 0.0))
        return fail(c, OKA_ERR_RANGE, "total_weight must be > 0 (an empty capture zone)");
    HIPCHK(c, hipSetDevice(c->device));
    const void *dp;
    TRY(stage_in(c, kSlotXY, pgrid, sizeof(double) * n, flags, &dp));
    void *dz;
    TRY(stage_out(c, kSlotOut, out, sizeof(double) * n, flags, false, &dz));
    hipLaunchKernelGGL(k_probability, dim3(grid_for(n)), dim3(kBlock), 0, c->stream,
                       (const double *)dp, total_weight, (double *)dz, n);
    HIPCHK(c, hipGetLastError());
    if (radius > 0) {
        const void *dw;
        TRY(stage_in(c, kSlotWells, weights, sizeof(double) * (2 * (size_t)radius + 1), 0, &dw));
        void *tmp;
        TRY(ensure(c, kSlotBitmap, sizeof(double) * n, &tmp));
        hipLaunchKernelGGL((k_correlate<0>), dim3(grid_for(n)), dim3(kBlock), 0, c->stream,
                           (const double *)dz, (double *)tmp, nrows, ncols, (const double *)dw,
                           radius);
        HIPCHK(c, hipGetLastError());
        hipLaunchKernelGGL((k_correlate<1>), dim3(grid_for(n)), dim3(kBlock), 0, c->stream,
                           (const double *)tmp, (double *)dz, nrows, ncols, (const double *)dw,
                           radius);
        HIPCHK(c, hipGetLastError());
    }
    TRY(copy_back(c, out, dz, sizeof(double) * n, flags));
    if (!(flags & OKA_ASYNC) || !(flags & OKA_PTR_DEVICE))
        HIPCHK(c, hipStreamSynchronize(c->stream));
    return OKA_OK;
}

int oka_field_sorted_probabilities(oka_ctx *c, int64_t n, const double *pgrid, double total_weight,
                                   double *pr, uint32_t flags)
{
    if (!c)
        return OKA_ERR_ARG;
    if (n < 0 || (n > 0 && (!pgrid || !pr)))
        return fail(c, OKA_ERR_ARG, "bad field");
    if (n == 0)
        return OKA_OK;
    if (n > INT32_MAX)
        return fail(c, OKA_ERR_ARG, "more than 2^31 nodes");
    if (!(total_weight > 0.0))
        return fail(c, OKA_ERR_RANGE, "total_weight must be > 0 (an empty capture zone)");
    HIPCHK(c, hipSetDevice(c->device));
    const void *dp;
    TRY(stage_in(c, kSlotXY, pgrid, sizeof(double) * n, flags, &dp));
    void *dz, *dout;
    TRY(ensure(c, kSlotBitmap, sizeof(double) * n, &dz));
    TRY(stage_out(c, kSlotOut, pr, sizeof(double) * n, flags, false, &dout));
    hipLaunchKernelGGL(k_probability, dim3(grid_for(n)), dim3(kBlock), 0, c->stream,
                       (const double *)dp, total_weight, (double *)dz, n);
    HIPCHK(c, hipGetLastError());
    size_t temp = 0;
    HIPCHK(c, hipcub::DeviceRadixSort::SortKeysDescending(nullptr, temp, (const double *)dz,
                                                          (double *)dout, (int)n, 0, 64,
                                                          c->stream));
    void *dtemp;
    TRY(ensure(c, kSlotMarks, temp, &dtemp));
    HIPCHK(c, hipcub::DeviceRadixSort::SortKeysDescending(dtemp, temp, (const double *)dz,
                                                          (double *)dout, (int)n, 0, 64,
                                                          c->stream));
    TRY(copy_back(c, pr, dout, sizeof(double) * n, flags));
    if (!(flags & OKA_ASYNC) || !(flags & OKA_PTR_DEVICE))
        HIPCHK(c, hipStreamSynchronize(c->stream));
    return OKA_OK;
}

}  // extern "C"
