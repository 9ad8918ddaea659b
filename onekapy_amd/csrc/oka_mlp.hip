// oka_mlp.hip -- the surrogate network of the learned control variate:
// BF16 MFMA (v_mfma_f32_16x16x32_bf16) matrix products with FP32
// accumulation and fused epilogues, BCE loss, bias gradients and Adam.
//
// Every product is "NT": C[M,N] = A[M,K] . B[N,K]^T with both operands
// row-major and K-contiguous, so each MFMA operand is one 16-byte LDS read.
// The forward and backward epilogues write activations and gradients both
// row-major and transposed (BF16), which makes every backward product NT too:
//   forward  Z = X W^T + b                    A = X,    B = W
//   dX       G W   = G (W^T)^T                A = G,    B = W^T copy
//   dW       G^T X = (G^T) (X^T)^T            A = G^T,  B = X^T copy
#include <math.h>
#include <stdint.h>

#include "oka_internal.h"
#include "oka_mlp.h"

namespace oka {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u)
        return (uint16_t)((u >> 16) | 0x40u);           // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);                    // round to nearest even
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float gelu(float z)
{
    return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_grad(float z)
{
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * __expf(-0.5f * z * z);
    return cdf + z * pdf;
}

__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }

constexpr int kTile = 64;   // block tile M and N
constexpr int kKStep = 32;  // K per MFMA
constexpr int kPad = 8;     // LDS row padding (bf16 elements)

template <int kEpi>
__global__ void __launch_bounds__(256) k_gemm_nt(GemmArgs g)
{
    __shared__ __align__(16) uint16_t sA[kTile][kKStep + kPad];
    __shared__ __align__(16) uint16_t sB[kTile][kKStep + kPad];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int lr = tid >> 2, lc = (tid & 3) * 8;   // tile loader: row, column chunk
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
            acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // split-K (kEpiAtomicF32): this block sums K in [kbeg, kend)
    const int kbeg = kEpi == kEpiAtomicF32 ? (int)blockIdx.z * g.ksplit : 0;
    const int kend = kEpi == kEpiAtomicF32 ? kbeg + g.ksplit : g.K;
    const uint16_t *pa = g.A + (int64_t)(m0 + lr) * g.lda + lc;
    const uint16_t *pb = g.B + (int64_t)(n0 + lr) * g.ldb + lc;
    uint4 ra = *reinterpret_cast<const uint4 *>(pa + kbeg);
    uint4 rb = *reinterpret_cast<const uint4 *>(pb + kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += kKStep) {
        *reinterpret_cast<uint4 *>(&sA[lr][lc]) = ra;
        *reinterpret_cast<uint4 *>(&sB[lr][lc]) = rb;
        __syncthreads();
        if (k0 + kKStep < kend) {   // prefetch the next K slice while computing
            ra = *reinterpret_cast<const uint4 *>(pa + k0 + kKStep);
            rb = *reinterpret_cast<const uint4 *>(pb + k0 + kKStep);
        }
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            a[i] = *reinterpret_cast<const bf16x8 *>(&sA[wm + i * 16 + (lane & 15)][(lane >> 4) * 8]);
            b[i] = *reinterpret_cast<const bf16x8 *>(&sB[wn + i * 16 + (lane & 15)][(lane >> 4) * 8]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }

    // epilogue: lane holds C[m = mb + 4*(lane/16) + e][n = nb + lane%16], e = 0..3
    double part_s = 0.0, part_d = 0.0, part_l = 0.0;
    const int64_t step = g.dstate ? g.dstate[0] : 0;
    const int64_t row_base = g.dstate ? step * g.step_rows : g.row0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int n = n0 + wn + j * 16 + (lane & 15);
        const float bias = (kEpi == kEpiGeluFwd || kEpi == kEpiLogits || kEpi == kEpiLossGrad ||
                            kEpi == kEpiAccum || kEpi == kEpiProbs) ? g.bias[n] : 0.f;
        float col_s = 0.f, col_d = 0.f, col_s2 = 0.f, col_d2 = 0.f;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int mb = m0 + wm + i * 16 + (lane >> 4) * 4;
            uint16_t hv[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = mb + e;
                const float v = acc[i][j][e];
                if (kEpi == kEpiGeluFwd) {
                    const float z = v + bias;
                    g.Cf[(int64_t)m * g.ldc + n] = z;
                    hv[e] = f2bf(gelu(z));
                    g.Ch[(int64_t)m * g.N + n] = hv[e];
                } else if (kEpi == kEpiLogits) {
                    g.Cf[(int64_t)m * g.ldc + n] = v + bias;
                } else if (kEpi == kEpiGeluBwd) {
                    hv[e] = f2bf(v * gelu_grad(g.Zsaved[(int64_t)m * g.N + n]));
                    g.Ch[(int64_t)m * g.N + n] = hv[e];
                } else if (kEpi == kEpiStoreF32) {
                    g.Cf[(int64_t)m * g.ldc + n] = v;
                } else if (kEpi == kEpiAtomicF32) {
                    atomicAdd(&g.Cf[(int64_t)m * g.ldc + n], v);
                } else if (kEpi == kEpiLossGrad) {
                    const float z = v + bias;
                    const int64_t row = g.rowidx ? g.rowidx[row_base + m] : row_base + m;
                    const float y = (float)g.marks[row * g.N + n] * g.tscale;
                    hv[e] = f2bf((sigmoid(z) - y) * g.gscale);
                    g.Ch[(int64_t)m * g.N + n] = hv[e];
                    part_l += (double)(fmaxf(z, 0.f) - z * y + log1pf(__expf(-fabsf(z))));
                } else if (kEpi == kEpiAccum || kEpi == kEpiProbs) {
                    if (m < g.mvalid) {
                        const float s = sigmoid(v + bias);
                        if (kEpi == kEpiProbs) {
                            g.Cf[(int64_t)m * g.ldc + n] = s;
                        } else {
                            col_s += s;
                            col_s2 += s * s;
                            if (g.marks) {
                                const int64_t row = g.row0 + m;
                                const float dv = (float)g.marks[row * g.N + n] * g.tscale - s;
                                col_d += dv;
                                col_d2 += dv * dv;
                            }
                        }
                    }
                }
            }
            if (kEpi == kEpiGeluFwd || kEpi == kEpiGeluBwd || kEpi == kEpiLossGrad) {
                // transposed copy: 4 consecutive rows of column n are contiguous in C^T
                uint2 packed;
                packed.x = (uint32_t)hv[0] | ((uint32_t)hv[1] << 16);
                packed.y = (uint32_t)hv[2] | ((uint32_t)hv[3] << 16);
                *reinterpret_cast<uint2 *>(&g.Cht[(int64_t)n * g.M + mb]) = packed;
            }
        }
        if (kEpi == kEpiAccum) {
            // lanes l, l^16, l^32, l^48 hold the same column
            col_s += __shfl_xor(col_s, 16);
            col_s += __shfl_xor(col_s, 32);
            col_s2 += __shfl_xor(col_s2, 16);
            col_s2 += __shfl_xor(col_s2, 32);
            col_d += __shfl_xor(col_d, 16);
            col_d += __shfl_xor(col_d, 32);
            col_d2 += __shfl_xor(col_d2, 16);
            col_d2 += __shfl_xor(col_d2, 32);
            if (lane < 16) {
                atomicAdd(&g.sums[n], (double)col_s);
                if (g.sums_sq)
                    atomicAdd(&g.sums_sq[n], (double)col_s2);
                if (g.marks && g.diffs)
                    atomicAdd(&g.diffs[n], (double)col_d);
                if (g.marks && g.diffs_sq)
                    atomicAdd(&g.diffs_sq[n], (double)col_d2);
            }
        }
    }
    if (kEpi == kEpiLossGrad) {
        for (int off = 32; off > 0; off >>= 1)
            part_l += __shfl_xor(part_l, off);
        if (lane == 0)
            atomicAdd(g.loss + step, part_l);
    }
    (void)part_s;
    (void)part_d;
}

// x (FP32, row stride n_in) of rows rowidx[row0 + m] (or row0 + m) -> X0
// (BF16 [M, kin], zero padded) and X0^T ([kin, M]); rows >= mvalid are zero.
__global__ void __launch_bounds__(256) k_prep_input(const float *x, int n_in, int kin,
                                                    const int32_t *rowidx, int64_t row0, int M,
                                                    int mvalid, uint16_t *x0, uint16_t *x0t,
                                                    const int64_t *dstate)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)M * kin)
        return;
    const int m = (int)(t / kin), k = (int)(t % kin);
    if (dstate)
        row0 = dstate[0] * M;
    float v = 0.f;
    if (m < mvalid && k < n_in) {
        const int64_t row = rowidx ? rowidx[row0 + m] : row0 + m;
        v = x[row * n_in + k];
    }
    const uint16_t h = f2bf(v);
    x0[(int64_t)m * kin + k] = h;
    x0t[(int64_t)k * M + m] = h;
}

// db[n] = sum over m of G^T[n, m] (one workgroup per n).
__global__ void __launch_bounds__(256) k_bias_grad(const uint16_t *gt, int M, float *db)
{
    __shared__ float s[256];
    const int n = blockIdx.x;
    float acc = 0.f;
    for (int m = threadIdx.x; m < M; m += blockDim.x)
        acc += bf2f(gt[(int64_t)n * M + m]);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        db[n] = s[0];
}

__global__ void __launch_bounds__(256) k_adam(float *p, const float *grad, float *m, float *v,
                                              int64_t n, float lr, float b1, float b2, float eps,
                                              int64_t t, const int64_t *dstate)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    if (dstate)
        t = dstate[1] + 1;
    // bias corrections 1/(1 - beta^t)
    const float c1 = (float)(1.0 / (1.0 - exp((double)t * log((double)b1))));
    const float c2 = (float)(1.0 / (1.0 - exp((double)t * log((double)b2))));
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// FP32 master W [rows, cols] -> BF16 W and W^T, one 32 x 32 tile per
// workgroup through LDS so that both stores are coalesced.
__global__ void __launch_bounds__(256) k_refresh_bf16(const float *w, int rows, int cols,
                                                      uint16_t *wh, uint16_t *wt)
{
    __shared__ uint16_t tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) {
            const uint16_t h = f2bf(w[(int64_t)r * cols + c]);
            wh[(int64_t)r * cols + c] = h;
            tile[k][tx] = h;
        }
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols)
            wt[(int64_t)c * rows + r] = tile[tx][k];
    }
}

// G = acc * gelu'(Z) as BF16, row-major and transposed (after a split-K product).
__global__ void __launch_bounds__(256) k_gelu_bwd(const float *acc, const float *z, int M, int N,
                                                  uint16_t *g, uint16_t *gt)
{
    __shared__ uint16_t tile[32][33];
    const int n0 = blockIdx.x * 32, m0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8) {
        const int m = m0 + k, n = n0 + tx;
        if (m < M && n < N) {
            const int64_t i = (int64_t)m * N + n;
            const uint16_t h = f2bf(acc[i] * gelu_grad(z[i]));
            g[i] = h;
            tile[k][tx] = h;
        }
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int n = n0 + k, m = m0 + tx;
        if (m < M && n < N)
            gt[(int64_t)n * M + m] = tile[tx][k];
    }
}

__global__ void k_step_advance(int64_t *dstate)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        dstate[0] += 1;
        dstate[1] += 1;
    }
}

__global__ void k_fill_f32(float *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0.f;
}

__global__ void k_zero_f64(double *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0.0;
}

}  // namespace mlp

using namespace mlp;

hipError_t launch_gemm(int epi, const GemmArgs &g, hipStream_t s, int splits)
{
    dim3 grid(g.N / kTile, g.M / kTile, epi == kEpiAtomicF32 ? splits : 1), block(256);
    switch (epi) {
    case kEpiGeluFwd: hipLaunchKernelGGL(k_gemm_nt<kEpiGeluFwd>, grid, block, 0, s, g); break;
    case kEpiLogits: hipLaunchKernelGGL(k_gemm_nt<kEpiLogits>, grid, block, 0, s, g); break;
    case kEpiGeluBwd: hipLaunchKernelGGL(k_gemm_nt<kEpiGeluBwd>, grid, block, 0, s, g); break;
    case kEpiStoreF32: hipLaunchKernelGGL(k_gemm_nt<kEpiStoreF32>, grid, block, 0, s, g); break;
    case kEpiLossGrad: hipLaunchKernelGGL(k_gemm_nt<kEpiLossGrad>, grid, block, 0, s, g); break;
    case kEpiAccum: hipLaunchKernelGGL(k_gemm_nt<kEpiAccum>, grid, block, 0, s, g); break;
    case kEpiProbs: hipLaunchKernelGGL(k_gemm_nt<kEpiProbs>, grid, block, 0, s, g); break;
    case kEpiAtomicF32: hipLaunchKernelGGL(k_gemm_nt<kEpiAtomicF32>, grid, block, 0, s, g); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_prep_input(const float *x, int n_in, int kin, const int32_t *rowidx,
                             int64_t row0, int M, int mvalid, uint16_t *x0, uint16_t *x0t,
                             hipStream_t s, const int64_t *dstate)
{
    const int64_t n = (int64_t)M * kin;
    hipLaunchKernelGGL(k_prep_input, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n_in,
                       kin, rowidx, row0, M, mvalid, x0, x0t, dstate);
    return hipGetLastError();
}

hipError_t launch_bias_grad(const uint16_t *gt, int N, int M, float *db, hipStream_t s)
{
    hipLaunchKernelGGL(k_bias_grad, dim3(N), dim3(256), 0, s, gt, M, db);
    return hipGetLastError();
}

hipError_t launch_adam(float *p, const float *grad, float *m, float *v, int64_t n, float lr,
                       float b1, float b2, float eps, int64_t t, const int64_t *dstate,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, grad, m, v,
                       n, lr, b1, b2, eps, t, dstate);
    return hipGetLastError();
}

hipError_t launch_refresh_bf16(const float *w, int rows, int cols, uint16_t *wh, uint16_t *wt,
                               hipStream_t s)
{
    dim3 grid((cols + 31) / 32, (rows + 31) / 32);
    hipLaunchKernelGGL(k_refresh_bf16, grid, dim3(256), 0, s, w, rows, cols, wh, wt);
    return hipGetLastError();
}

hipError_t launch_gelu_bwd(const float *acc, const float *z, int M, int N, uint16_t *g,
                           uint16_t *gt, hipStream_t s)
{
    dim3 grid((N + 31) / 32, (M + 31) / 32);
    hipLaunchKernelGGL(k_gelu_bwd, grid, dim3(256), 0, s, acc, z, M, N, g, gt);
    return hipGetLastError();
}

hipError_t launch_step_advance(int64_t *dstate, hipStream_t s)
{
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, s, dstate);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float *p, int64_t n, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_zero_f64(double *p, int64_t n, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace oka
