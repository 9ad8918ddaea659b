// oka_rng.hip -- realization draws on the device and the surrogate's features.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC 2011) is a counter-based generator: the 128-bit output for
// counter (realization, slot, sub) and key (seed) needs no state, so every
// realization's draws are reproducible on their own and the launch shape does
// not matter.
#include <math.h>

#include "oka_internal.h"

namespace oka {
namespace {

struct U4 {
    uint32_t x, y, z, w;
};

__host__ __device__ inline uint32_t mulhi32(uint32_t a, uint32_t b)
{
    return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
    for (int i = 0; i < 10; ++i) {
        const uint32_t hi0 = mulhi32(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        const uint32_t hi1 = mulhi32(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Two uniforms in (0, 1) with 53 random bits each.
__device__ inline void uniforms(uint64_t seed, int64_t r, uint32_t slot, uint32_t sub, double &u1,
                                double &u2)
{
    const U4 o = philox4x32_10(U4{(uint32_t)r, (uint32_t)((uint64_t)r >> 32), slot, sub},
                               (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t a = ((uint64_t)o.x << 21) ^ (uint64_t)(o.y >> 11);
    const uint64_t b = ((uint64_t)o.z << 21) ^ (uint64_t)(o.w >> 11);
    u1 = ((double)(a & ((1ull << 53) - 1)) + 0.5) * 0x1p-53;
    u2 = ((double)(b & ((1ull << 53) - 1)) + 0.5) * 0x1p-53;
}

// With `flip` the draw is the antithetic partner of realization r's: every
// uniform u becomes 1 - u (exact: u = (k + 1/2) 2^-53 maps to the lattice
// point of 2^53 - 1 - k) and the lognormal's normal g becomes -g, so the
// marginal distribution is the same and the pair is negatively correlated.
__device__ inline double sample(const oka_dist &d, uint64_t seed, int64_t r, uint32_t slot,
                                bool flip)
{
    double u1, u2;
    uniforms(seed, r, slot, 0, u1, u2);
    const double ui = flip ? 1.0 - u1 : u1;   // the inverse-CDF variate
    double v, us = u2;
    switch (d.kind) {
    case OKA_DIST_UNIFORM:
        v = d.a + (d.b - d.a) * ui;
        break;
    case OKA_DIST_TRIANGULAR: {
        // inverse CDF of triangular(lo = a, mode = b, hi = c)
        const double fc = (d.b - d.a) / (d.c - d.a);
        v = ui < fc ? d.a + sqrt(ui * (d.c - d.a) * (d.b - d.a))
                    : d.c - sqrt((1.0 - ui) * (d.c - d.a) * (d.c - d.b));
        break;
    }
    case OKA_DIST_LOGNORMAL: {
        const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        v = exp(d.a + d.b * (flip ? -g : g));
        double w;
        uniforms(seed, r, slot, 1, us, w);
        break;
    }
    default:
        v = d.a;
        break;
    }
    if (flip)
        us = 1.0 - us;
    if (d.p_negative > 0.0 && us < d.p_negative)
        v = -v;
    return v;
}

__global__ void __launch_bounds__(256) k_draw(DrawArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n)
        return;
    // antithetic pairs (2k, 2k + 1): the odd member reuses counter 2k, reflected
    const bool flip = a.antithetic && (a.first + i) % 2 == 1;
    const int64_t r = flip ? a.first + i - 1 : a.first + i;
    const int nw = a.nwells;
    double *q = a.q + (a.q_well_stride ? i : i * nw);
    const int64_t qs = a.q_well_stride ? a.q_well_stride : 1;
    for (int w = 0; w < nw; ++w)
        q[w * qs] = sample(a.dists[w], a.seed, r, (uint32_t)w, flip);
    a.K[i] = sample(a.dists[nw], a.seed, r, (uint32_t)nw, flip);
    a.por[i] = sample(a.dists[nw + 1], a.seed, r, (uint32_t)nw + 1, flip);
    a.thick[i] = sample(a.dists[nw + 2], a.seed, r, (uint32_t)nw + 2, flip);
    const double sg = flip ? -1.0 : 1.0;
    for (int j = 0; j < 3; ++j) {
        double u1, u2;
        uniforms(a.seed, r, (uint32_t)(nw + 3 + j), 0, u1, u2);
        const double rad = sqrt(-2.0 * log(u1));
        double sn, cs;
        sincos(6.283185307179586 * u2, &sn, &cs);
        a.z[i * 6 + 2 * j] = sg * (rad * cs);
        a.z[i * 6 + 2 * j + 1] = sg * (rad * sn);
    }
}

// x[i] = (f - mean) / std with f = (q_1..q_nw, K, porosity, thickness, A..F).
__global__ void __launch_bounds__(256) k_features(FeatureArgs a)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nin = a.nwells + 9;
    if (t >= a.n * nin)
        return;
    const int64_t i = t / nin;
    const int j = (int)(t % nin);
    double f;
    if (j < a.nwells)
        f = a.q_well_stride ? a.q[j * a.q_well_stride + i] : a.q[i * a.nwells + j];
    else if (j == a.nwells)
        f = a.K[i];
    else if (j == a.nwells + 1)
        f = a.por[i];
    else if (j == a.nwells + 2)
        f = a.thick[i];
    else
        f = a.coef[i * 6 + (j - a.nwells - 3)];
    a.x[t] = (float)((f - a.mean[j]) / a.std[j]);
}

// Exact integer moments of the window marks: workgroup (x, y) sums cells
// 256x .. 256x + 255 over rows y * rows_per .. + rows_per (partial sums stay
// below 2^32: rows_per * 255^2 < 2^32).
__global__ void __launch_bounds__(256) k_mark_moments(const uint8_t *marks, int64_t n, int ncell,
                                                      int rows_per, unsigned long long *sums,
                                                      unsigned long long *sums_sq)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncell)
        return;
    const int64_t r0 = (int64_t)blockIdx.y * rows_per;
    const int64_t r1 = r0 + rows_per < n ? r0 + rows_per : n;
    uint32_t s = 0, s2 = 0;
    for (int64_t r = r0; r < r1; ++r) {
        const uint32_t v = marks[r * ncell + c];
        s += v;
        s2 += v * v;
    }
    atomicAdd(&sums[c], (unsigned long long)s);
    if (sums_sq)
        atomicAdd(&sums_sq[c], (unsigned long long)s2);
}

}  // namespace

hipError_t launch_mark_moments(const uint8_t *marks, int64_t n, int32_t ncell, uint64_t *sums,
                               uint64_t *sums_sq, hipStream_t s)
{
    constexpr int kRowsPer = 1024;
    if (n <= 0 || ncell <= 0)
        return hipSuccess;
    const dim3 grid((unsigned)((ncell + 255) / 256), (unsigned)((n + kRowsPer - 1) / kRowsPer));
    hipLaunchKernelGGL(k_mark_moments, grid, dim3(256), 0, s, marks, n, ncell, kRowsPer,
                       (unsigned long long *)sums, (unsigned long long *)sums_sq);
    return hipGetLastError();
}

hipError_t launch_draw(const DrawArgs &a, hipStream_t s)
{
    if (a.n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_draw, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_features(const FeatureArgs &a, hipStream_t s)
{
    const int64_t n = a.n * (a.nwells + 9);
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

uint32_t host_philox_word(uint64_t seed, int64_t r, uint32_t slot, uint32_t sub, int word)
{
    const U4 o = philox4x32_10(U4{(uint32_t)r, (uint32_t)((uint64_t)r >> 32), slot, sub},
                               (uint32_t)seed, (uint32_t)(seed >> 32));
    return word == 0 ? o.x : word == 1 ? o.y : word == 2 ? o.z : o.w;
}

}  // namespace oka
