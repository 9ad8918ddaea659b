// oka_rng.hip -- realization draws on the device and the surrogate's features.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC 2011) is a counter-based generator: the 128-bit output for
// counter (realization, slot, sub) and key (seed) needs no state, so every
// realization's draws are reproducible on their own and the launch shape does
// not matter.
//
// Randomized quasi-Monte Carlo (OKA_QMC): realization j takes its uniforms
// from point j of a Sobol' sequence (Bratley & Fox, ACM TOMS 14, 1988; the
// direction-number recurrence as in Joe & Kuo, ACM TOMS 29, 2003) under a
// nested (Owen-type) scramble keyed by the seed: the hash-based scramble of
// Burley ("Practical hash-based Owen scrambling", JCGT 9(4), 2020), where
// every bit of a coordinate flips as a function of the more significant bits
// only.  Each scrambled point is uniform on the unit cube (a random 32-bit
// offset enters before a bijection), the nets' stratification survives, and
// independent seeds give independent unbiased estimates.
#include <math.h>

#include <mutex>
#include <vector>

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

// The value of distribution d from its uniforms: u (the inverse-CDF variate,
// or the Box-Muller radius variate of a lognormal), w (the lognormal's angle
// variate) and us (the sign variate); neg negates the lognormal's normal.
__device__ inline double dist_value(const oka_dist &d, double u, double w, double us, bool neg)
{
    double v;
    switch (d.kind) {
    case OKA_DIST_UNIFORM:
        v = d.a + (d.b - d.a) * u;
        break;
    case OKA_DIST_TRIANGULAR: {
        // inverse CDF of triangular(lo = a, mode = b, hi = c)
        const double fc = (d.b - d.a) / (d.c - d.a);
        v = u < fc ? d.a + sqrt(u * (d.c - d.a) * (d.b - d.a))
                   : d.c - sqrt((1.0 - u) * (d.c - d.a) * (d.c - d.b));
        break;
    }
    case OKA_DIST_LOGNORMAL: {
        const double g = sqrt(-2.0 * log(u)) * cos(6.283185307179586 * w);
        v = exp(d.a + d.b * (neg ? -g : g));
        break;
    }
    default:
        v = d.a;
        break;
    }
    if (d.p_negative > 0.0 && us < d.p_negative)
        v = -v;
    return v;
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
    const bool ln = d.kind == OKA_DIST_LOGNORMAL;
    double us = u2;
    if (ln) {
        double w;
        uniforms(seed, r, slot, 1, us, w);
    }
    if (flip)
        us = 1.0 - us;
    return dist_value(d, ln || !flip ? u1 : 1.0 - u1, u2, us, flip);
}

// Nested scramble of a 32-bit coordinate: in bit-reversed order, a keyed
// offset and four steps x ^= x * c (c even), each of which flips a bit as a
// function of the bits below it only (Burley 2020, the Laine-Karras form).
__host__ __device__ inline uint32_t nested_scramble(uint32_t x, uint32_t key)
{
    x = __builtin_bitreverse32(x);
    x += key;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return __builtin_bitreverse32(x);
}

// Uniform in (0, 1) of dimension d of Sobol' point j: the XOR of the
// direction numbers of j's set bits, scrambled with key (qmc_key()).
__device__ inline double qmc_u(const DrawArgs &a, uint32_t j, int d, const uint32_t *keys)
{
    const uint32_t *v = a.sobol + 32 * d;
    uint32_t x = 0;
    for (uint32_t b = j; b != 0u; b &= b - 1u)
        x ^= v[__builtin_ctz(b)];
    return ((double)nested_scramble(x, keys[d]) + 0.5) * 0x1p-32;
}

// Realization first + i of the QMC draws.  Dimensions, most influential
// first: the inverse-CDF variates of K, porosity and thickness, the 6
// uniforms of the MVN draw's normals, those of the discharges (wells in
// order), then the lognormals' angle variates and the sign variates (slots
// K, porosity, thickness, then the wells).  A constant takes no dimension.
__device__ inline void draw_qmc(const DrawArgs &a, int64_t i)
{
    // realization first + i: point j of replicate rep (one replicate unless
    // qmc_npts is set), scrambled with that replicate's keys
    const int64_t gi = a.first + i;
    const int64_t rep = a.qmc_npts ? gi / a.qmc_npts : 0;
    const uint32_t j = (uint32_t)(a.qmc_npts ? gi % a.qmc_npts : gi);
    const uint32_t *keys = a.qmc_keys + rep * a.qmc_dims;
    const int nw = a.nwells;
    int ncore = 0, nq = 0;
    for (int s = 0; s < 3; ++s)
        ncore += a.dists[nw + s].kind != OKA_DIST_CONST;
    for (int w = 0; w < nw; ++w)
        nq += a.dists[w].kind != OKA_DIST_CONST;
    int prim = 0, sec = ncore + 6 + nq;   // next primary and secondary dimension
    auto value = [&](int slot) {
        const oka_dist &d = a.dists[slot];
        double u = 0.5, w = 0.5, us = 1.0;
        if (d.kind != OKA_DIST_CONST)
            u = qmc_u(a, j, prim++, keys);
        if (d.kind == OKA_DIST_LOGNORMAL)
            w = qmc_u(a, j, sec++, keys);
        if (d.p_negative > 0.0)
            us = qmc_u(a, j, sec++, keys);
        return dist_value(d, u, w, us, false);
    };
    a.K[i] = value(nw);
    a.por[i] = value(nw + 1);
    a.thick[i] = value(nw + 2);
    for (int p = 0; p < 3; ++p) {
        const double u1 = qmc_u(a, j, prim, keys), u2 = qmc_u(a, j, prim + 1, keys);
        prim += 2;
        const double rad = sqrt(-2.0 * log(u1));
        double sn, cs;
        sincos(6.283185307179586 * u2, &sn, &cs);
        a.z[i * 6 + 2 * p] = rad * cs;
        a.z[i * 6 + 2 * p + 1] = rad * sn;
    }
    double *q = a.q + (a.q_well_stride ? i : i * nw);
    const int64_t qs = a.q_well_stride ? a.q_well_stride : 1;
    for (int w = 0; w < nw; ++w)
        q[w * qs] = value(w);
}

__global__ void __launch_bounds__(256) k_draw(DrawArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n)
        return;
    if (a.sobol) {
        draw_qmc(a, i);
        return;
    }
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
    if (a.sobol && (a.qmc_dims <= 0 || !a.qmc_keys))
        return hipErrorInvalidValue;
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

namespace {

// x^e mod p over GF(2), p of degree deg (bit deg set), deg >= 2.
uint64_t gf2_powmod_x(uint64_t e, uint64_t p, int deg)
{
    auto mulmod = [&](uint64_t a, uint64_t b) {
        uint64_t r = 0;
        for (; b; b >>= 1) {
            if (b & 1)
                r ^= a;
            a <<= 1;
            if ((a >> deg) & 1)
                a ^= p;
        }
        return r;
    };
    uint64_t r = 1, base = 2;
    for (; e; e >>= 1) {
        if (e & 1)
            r = mulmod(r, base);
        base = mulmod(base, base);
    }
    return r;
}

// Primitive: x has order 2^deg - 1 modulo p.
bool gf2_primitive(uint64_t p, int deg)
{
    const uint64_t n = (1ull << deg) - 1;
    if (gf2_powmod_x(n, p, deg) != 1)
        return false;
    uint64_t m = n;
    for (uint64_t f = 2; f * f <= m; ++f) {
        if (m % f)
            continue;
        while (m % f == 0)
            m /= f;
        if (gf2_powmod_x(n / f, p, deg) == 1)
            return false;
    }
    return m == 1 || gf2_powmod_x(n / m, p, deg) != 1;
}

uint64_t splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

std::vector<uint32_t> make_sobol_directions()
{
    std::vector<uint32_t> v((size_t)kSobolMaxDims * 32);
    for (int k = 0; k < 32; ++k)
        v[k] = 1u << (31 - k);   // dimension 0: van der Corput
    int d = 1;
    for (int deg = 1; deg <= 16 && d < kSobolMaxDims; ++deg) {
        for (uint64_t p = (1ull << deg) | 1ull; p < (2ull << deg) && d < kSobolMaxDims; p += 2) {
            if (deg > 1 && !gf2_primitive(p, deg))
                continue;
            uint32_t *vd = v.data() + (size_t)32 * d;
            for (int k = 1; k <= 32; ++k) {
                uint32_t x;
                if (k <= deg) {
                    // initial direction number m_k: odd, below 2^k
                    const uint64_t h = splitmix64(((uint64_t)d << 32) | (uint64_t)k);
                    const uint32_t m = 2u * (uint32_t)(h & ((1ull << (k - 1)) - 1)) + 1u;
                    x = m << (32 - k);
                } else {
                    // v_k = v_{k-s} ^ (v_{k-s} >> s) ^ sum_i a_i v_{k-i},
                    // p = x^s + a_1 x^(s-1) + ... + a_(s-1) x + 1
                    x = vd[k - deg - 1] ^ (vd[k - deg - 1] >> deg);
                    for (int i = 1; i < deg; ++i)
                        if ((p >> (deg - i)) & 1)
                            x ^= vd[k - i - 1];
                }
                vd[k - 1] = x;
            }
            ++d;
        }
    }
    return v;
}

}  // namespace

const uint32_t *sobol_directions()
{
    static std::vector<uint32_t> table;
    static std::once_flag once;
    std::call_once(once, [] { table = make_sobol_directions(); });
    return table.data();
}

int qmc_dims(int32_t nwells, const oka_dist *dists)
{
    int d = 6;   // the MVN draw's normals
    for (int s = 0; s < nwells + 3; ++s) {
        d += dists[s].kind != OKA_DIST_CONST;
        d += dists[s].kind == OKA_DIST_LOGNORMAL;
        d += dists[s].p_negative > 0.0;
    }
    return d;
}

uint32_t qmc_key(uint64_t seed, int d)
{
    const U4 o = philox4x32_10(U4{(uint32_t)d, 0xFFFFFFFFu, kQmcKeySlot, 2u}, (uint32_t)seed,
                               (uint32_t)(seed >> 32));
    return o.x;
}

uint32_t host_philox_word(uint64_t seed, int64_t r, uint32_t slot, uint32_t sub, int word)
{
    const U4 o = philox4x32_10(U4{(uint32_t)r, (uint32_t)((uint64_t)r >> 32), slot, sub},
                               (uint32_t)seed, (uint32_t)(seed >> 32));
    return word == 0 ? o.x : word == 1 ? o.y : word == 2 ? o.z : o.w;
}

}  // namespace oka
