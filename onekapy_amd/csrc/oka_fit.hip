// oka_fit.hip -- batched FP64 regional-flow fit and multivariate-normal draw.
//
// Per realization r (one lane each): the weighted least-squares rows of
// Model.construct_fit (oneka/model.py:540-567) with conductivity K[r],
// thickness b[r] and discharges q[r, :]; the fit of compute_fit
// (model.py:598-604) solved through column-equilibrated normal equations, a
// 6x6 Cholesky factorization and one step of iterative refinement; the
// covariance inv(WA^T WA); and the draw coef = ev + L z of
// oneka/stochastic.py:239-241 with L the lower Cholesky factor of the
// covariance.  The realization-invariant part of the wells' potential at the
// observations, log(r^2) per (observation, well), is computed once per call by
// k_fit_logs.
#include <math.h>

#include "oka_internal.h"
#include "oka_math.h"

#pragma clang fp contract(off)

namespace oka {
namespace {

constexpr double kC4Pi = 0.07957747154594767;  // model.py:264

__global__ void __launch_bounds__(256) k_fit_logs(FitArgs a, double *logr2)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)a.nobs * a.nwells)
        return;
    const int i = (int)(t / a.nwells), w = (int)(t % a.nwells);
    const double dx = a.obs[4 * i] - a.wells_xy[2 * w];
    const double dy = a.obs[4 * i + 1] - a.wells_xy[2 * w + 1];
    logr2[t] = log_rn(dx * dx + dy * dy);
}

// Observations whose wells' potentials one pass over the realization's
// discharges computes: each discharge is loaded once per block rather than
// once per observation (the fit of 10^6 realizations of the 200-well,
// 200-observation site: 402 ms unblocked, 106 ms in blocks of 4, 55 ms in
// blocks of 8, the same bits).
constexpr int kObsBlock = 8;

// The wells' potential at observations i0 .. i0 + kObsBlock - 1 (those below
// nobs), each summed over the wells in index order (model.py:258-264).
__device__ __forceinline__ void well_potentials(const FitArgs &a, int i0, const double *q,
                                                int64_t qs, double (&pw)[kObsBlock])
{
#pragma unroll
    for (int j = 0; j < kObsBlock; ++j)
        pw[j] = 0.0;
    const double *L = a.logr2 + (int64_t)i0 * a.nwells;
    const int nb = min(kObsBlock, a.nobs - i0);
    for (int w = 0; w < a.nwells; ++w) {
        const double qw = q[w * qs];
#pragma unroll
        for (int j = 0; j < kObsBlock; ++j)
            if (j < nb)
                pw[j] = pw[j] + (qw * L[(int64_t)j * a.nwells + w]) * kC4Pi;
    }
}

// One weighted row (model.py:545-565), given the wells' potential pw at the
// observation.  Returns 0 or OKA_ERR_RANGE.
__device__ __forceinline__ int fit_row(const FitArgs &a, int i, double K, double b, double pw,
                                       double *row, double &wb)
{
    const double x = a.obs[4 * i], y = a.obs[4 * i + 1];
    const double z_ev = a.obs[4 * i + 2], z_std = a.obs[4 * i + 3];
    const double head = z_ev - a.base;
    double pot_ev, pot_std;
    if (head >= b) {
        pot_ev = (K * b) * (head - 0.5 * b);
        pot_std = (K * b) * z_std;
    } else if (head > 0.0) {
        pot_ev = (0.5 * K) * (head * head + z_std * z_std);
        pot_std = (K * head) * z_std;
    } else {
        return OKA_ERR_RANGE;
    }
    const double dx = x - a.xo, dy = y - a.yo;
    row[0] = (dx * dx) / pot_std;
    row[1] = (dy * dy) / pot_std;
    row[2] = (dx * dy) / pot_std;
    row[3] = dx / pot_std;
    row[4] = dy / pot_std;
    row[5] = 1.0 / pot_std;
    wb = (pot_ev - pw) / pot_std;
    return 0;
}

// Lower Cholesky factor of a symmetric 6x6 (row-major, lower part read).
__device__ __forceinline__ bool chol6(const double *A, double *L)
{
#pragma unroll
    for (int t = 0; t < 36; ++t)
        L[t] = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double d = A[j * 6 + j];
#pragma unroll
        for (int k = 0; k < j; ++k)
            d = d - L[j * 6 + k] * L[j * 6 + k];
        if (!(d > 0.0))
            return false;
        const double ljj = sqrt(d);
        L[j * 6 + j] = ljj;
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
            double v = A[i * 6 + j];
#pragma unroll
            for (int k = 0; k < j; ++k)
                v = v - L[i * 6 + k] * L[j * 6 + k];
            L[i * 6 + j] = v / ljj;
        }
    }
    return true;
}

// out = S (L L^T)^{-1} S g with S = diag(s).
__device__ __forceinline__ void chol6_solve(const double *L, const double *s, const double *g,
                                            double *out)
{
    double w[6], y[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double v = s[j] * g[j];
#pragma unroll
        for (int k = 0; k < j; ++k)
            v = v - L[j * 6 + k] * w[k];
        w[j] = v / L[j * 6 + j];
    }
#pragma unroll
    for (int j = 5; j >= 0; --j) {
        double v = w[j];
#pragma unroll
        for (int k = j + 1; k < 6; ++k)
            v = v - L[k * 6 + j] * y[k];
        y[j] = v / L[j * 6 + j];
    }
#pragma unroll
    for (int j = 0; j < 6; ++j)
        out[j] = s[j] * y[j];
}

// Fit and MVN draw of realization r of the site `a`.
__device__ __forceinline__ void fit_one(const FitArgs &a, int64_t r)
{
    const double K = a.K[r], b = a.b[r];
    // the realization's discharges: q[w * qs]
    const int64_t qs = a.q_well_stride ? a.q_well_stride : 1;
    const double *q = a.q + (a.q_well_stride ? r : (int64_t)r * a.nwells);
    double N[36], rhs[6], row[6], wb;
#pragma unroll
    for (int t = 0; t < 36; ++t)
        N[t] = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j)
        rhs[j] = 0.0;
    int st = 0;
    for (int i0 = 0; i0 < a.nobs && st == 0; i0 += kObsBlock) {
        double pw[kObsBlock];
        well_potentials(a, i0, q, qs, pw);
#pragma unroll
        for (int o = 0; o < kObsBlock; ++o) {
            if (i0 + o >= a.nobs)
                break;
            st = fit_row(a, i0 + o, K, b, pw[o], row, wb);
            if (st)
                break;
#pragma unroll
            for (int j = 0; j < 6; ++j) {
#pragma unroll
                for (int k = 0; k <= j; ++k)
                    N[j * 6 + k] = N[j * 6 + k] + row[j] * row[k];
                rhs[j] = rhs[j] + row[j] * wb;
            }
        }
    }
    double ev[6], L[36], s[6];
    if (st == 0) {
#pragma unroll
        for (int j = 0; j < 6; ++j)
            s[j] = 1.0 / sqrt(N[j * 6 + j]);
        double Ns[36];
#pragma unroll
        for (int j = 0; j < 6; ++j)
#pragma unroll
            for (int k = 0; k < 6; ++k)
                Ns[j * 6 + k] = k <= j ? (s[j] * N[j * 6 + k]) * s[k] : 0.0;
        if (!chol6(Ns, L))
            st = OKA_ERR_LINALG;
    }
    if (st == 0) {
        chol6_solve(L, s, rhs, ev);
        double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, d[6];
        for (int i0 = 0; i0 < a.nobs; i0 += kObsBlock) {
            double pw[kObsBlock];
            well_potentials(a, i0, q, qs, pw);
#pragma unroll
            for (int o = 0; o < kObsBlock; ++o) {
                if (i0 + o >= a.nobs)
                    break;
                fit_row(a, i0 + o, K, b, pw[o], row, wb);
                double res = wb;
#pragma unroll
                for (int j = 0; j < 6; ++j)
                    res = res - row[j] * ev[j];
#pragma unroll
                for (int j = 0; j < 6; ++j)
                    g[j] = g[j] + row[j] * res;
            }
        }
        chol6_solve(L, s, g, d);
#pragma unroll
        for (int j = 0; j < 6; ++j)
            ev[j] = ev[j] + d[j];
    }
    double cov[36];
    if (st == 0) {
        double Li[36];
#pragma unroll
        for (int t = 0; t < 36; ++t)
            Li[t] = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            Li[j * 6 + j] = 1.0 / L[j * 6 + j];
#pragma unroll
            for (int i = j + 1; i < 6; ++i) {
                double v = 0.0;
#pragma unroll
                for (int k = j; k < i; ++k)
                    v = v - L[i * 6 + k] * Li[k * 6 + j];
                Li[i * 6 + j] = v / L[i * 6 + i];
            }
        }
#pragma unroll
        for (int j = 0; j < 6; ++j) {
#pragma unroll
            for (int k = 0; k <= j; ++k) {
                double v = 0.0;
#pragma unroll
                for (int i = j; i < 6; ++i)
                    v = v + Li[i * 6 + j] * Li[i * 6 + k];
                const double c = (s[j] * v) * s[k];
                cov[j * 6 + k] = c;
                cov[k * 6 + j] = c;
            }
        }
    }
    double coef[6];
    if (st == 0) {
        if (a.z) {
            double Lc[36];
            if (!chol6(cov, Lc)) {
                st = OKA_ERR_LINALG;
            } else {
                const double *z = a.z + 6 * (int64_t)r;
#pragma unroll
                for (int j = 0; j < 6; ++j) {
                    double acc = 0.0;
#pragma unroll
                    for (int k = 0; k <= j; ++k)
                        acc = acc + Lc[j * 6 + k] * z[k];
                    coef[j] = ev[j] + acc;
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < 6; ++j)
                coef[j] = ev[j];
        }
    }
    if (st != 0) {
#pragma unroll
        for (int j = 0; j < 6; ++j)
            ev[j] = coef[j] = (double)NAN;
#pragma unroll
        for (int t = 0; t < 36; ++t)
            cov[t] = (double)NAN;
    }
    if (a.ev)
#pragma unroll
        for (int j = 0; j < 6; ++j)
            a.ev[6 * (int64_t)r + j] = ev[j];
    if (a.cov)
#pragma unroll
        for (int t = 0; t < 36; ++t)
            a.cov[36 * (int64_t)r + t] = cov[t];
    if (a.coef)
#pragma unroll
        for (int j = 0; j < 6; ++j)
            a.coef[6 * (int64_t)r + j] = coef[j];
    if (a.status)
        a.status[r] = st;
}

__global__ void __launch_bounds__(64) k_fit_draw(FitArgs a)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < a.nreal)
        fit_one(a, r);
}

// Site of work item t of a table whose item counts have prefix sums `first`.
__device__ __forceinline__ int table_site(const int64_t *first, int nsites, int64_t t)
{
    int lo = 0, hi = nsites - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// log(r^2) of every (observation, well) pair of every site of a table.
__global__ void __launch_bounds__(256) k_fit_logs_multi(const FitArgs *sites, int nsites,
                                                        const int64_t *first, int64_t total)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total)
        return;
    const int site = table_site(first, nsites, t);
    const FitArgs &a = sites[site];
    const int64_t k = t - first[site];
    const int i = (int)(k / a.nwells), w = (int)(k % a.nwells);
    const double dx = a.obs[4 * i] - a.wells_xy[2 * w];
    const double dy = a.obs[4 * i + 1] - a.wells_xy[2 * w + 1];
    ((double *)a.logr2)[k] = log_rn(dx * dx + dy * dy);
}

// Fit and MVN draw of every realization of every site of a table.
__global__ void __launch_bounds__(64) k_fit_draw_multi(const FitArgs *sites, int nsites,
                                                       const int64_t *first, int64_t total)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total)
        return;
    const int site = table_site(first, nsites, t);
    fit_one(sites[site], t - first[site]);
}

}  // namespace

hipError_t launch_fit_logs(const FitArgs &a, double *logr2, hipStream_t s)
{
    const int64_t n = (int64_t)a.nobs * a.nwells;
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_fit_logs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, logr2);
    return hipGetLastError();
}

hipError_t launch_fit_logs_multi(const FitArgs *sites, int nsites, const int64_t *first,
                                int64_t total, hipStream_t s)
{
    if (total <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_fit_logs_multi, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       sites, nsites, first, total);
    return hipGetLastError();
}

hipError_t launch_fit_draw_multi(const FitArgs *sites, int nsites, const int64_t *first,
                                int64_t total, hipStream_t s)
{
    if (total <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_fit_draw_multi, dim3((unsigned)((total + 63) / 64)), dim3(64), 0, s,
                       sites, nsites, first, total);
    return hipGetLastError();
}

hipError_t launch_fit_draw(const FitArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_fit_draw, dim3((unsigned)((a.nreal + 63) / 64)), dim3(64), 0, s, a);
    return hipGetLastError();
}

}  // namespace oka
