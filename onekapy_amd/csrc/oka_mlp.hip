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
// A dW product also forms the bias gradient (the row sums of G^T) from the
// A slices its first column of tiles loads, and one launch may carry two
// independent products (dW of a layer beside dX of the same layer).
#include <math.h>
#include <stdint.h>

#include "oka_internal.h"
#include "oka_mlp.h"

namespace oka {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32; NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    return __builtin_bit_cast(uint16_t, (__bf16)f);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// sum of the 8 BF16 values packed in a 16-byte word
__device__ __forceinline__ float sum8(uint4 w)
{
    const uint32_t u[4] = {w.x, w.y, w.z, w.w};
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        s += __uint_as_float(u[i] << 16) + __uint_as_float(u[i] & 0xffff0000u);
    return s;
}

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

// Logistic function with the hardware reciprocal (about 1 ulp; exp(-z) = inf
// gives 0 and exp(-z) = 0 gives 1).
__device__ __forceinline__ float sigmoid(float z)
{
    return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

constexpr int kTile = 64;       // the smaller block tile (M and N); products are multiples of it
constexpr int kBK = 64;         // K per LDS stage (two MFMA K steps of 32)
constexpr int kLd = kBK + 8;    // LDS row stride in BF16 elements (16-byte pad)

// Block tile shapes: kBig = 0 is 64 x 64 (each of the 4 waves 32 x 32, 2 x 2
// MFMA tiles), kBig = 1 is 128 x 128 (each wave 64 x 64, 4 x 4 MFMA tiles:
// twice the MFMAs per LDS fragment and half the operand traffic per flop).
template <int kBig>
struct Shape {
    static constexpr int TM = kBig ? 4 : 2, TN = kBig ? 4 : 2;   // MFMA tiles per wave
    static constexpr int BM = 32 * TM, BN = 32 * TN;             // block tile
    static constexpr int RA = BM / 32, RB = BN / 32;             // rows loaded per lane
};

template <int kBig>
struct GemmSmem {
    static_assert(Shape<kBig>::BM == Shape<kBig>::BN, "square block tiles");
    union {
        struct {
            uint16_t a[2][Shape<kBig>::BM][kLd];  // double-buffered A and B slices
            uint16_t b[2][Shape<kBig>::BN][kLd];
        };
        // after the main loop: the BF16 result tile, row-major [0] and transposed [1]
        uint16_t out[2][Shape<kBig>::BM][Shape<kBig>::BN + 8];
    };
    double red[4];
    float col[4][Shape<kBig>::BN];        // kEpiAccum: the workgroup's column sums
};

__device__ __forceinline__ uint4 ld16(const uint16_t *p) { return *reinterpret_cast<const uint4 *>(p); }
__device__ __forceinline__ void st16(uint16_t *p, uint4 v) { *reinterpret_cast<uint4 *>(p) = v; }

template <int kBig>
struct Stage {
    uint4 a[Shape<kBig>::RA], b[Shape<kBig>::RB];   // rows lr + 32 r of the A and B slices
};

// One BM x BN tile of C (tile index = (z * M/BM + row tile) * N/BN + column
// tile).  256 lanes: wave w computes the quadrant (w / 2, w % 2); lane t
// loads 16 bytes of rows t/8 + 32 r of each operand per K stage of 64.
// Global loads run two stages ahead of the MFMAs (two register sets) and LDS
// is double buffered, with one barrier per stage.
template <int kEpi, int kBig>
__device__ __forceinline__ void gemm_tile(const GemmArgs &g, int tile, GemmSmem<kBig> &sm)
{
    using S = Shape<kBig>;
    constexpr int TM = S::TM, TN = S::TN;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int gx = g.N / S::BN, gxy = gx * (g.M / S::BM);
    const int bz = tile / gxy;
    const int rem = tile - bz * gxy;
    const int m0 = (rem / gx) * S::BM, n0 = (rem - (rem / gx) * gx) * S::BN;
    const int wm = (wave >> 1) * 16 * TM, wn = (wave & 1) * 16 * TN;
    const int lr = tid >> 3, lc = (tid & 7) * 8;
    const int kspan = g.K / g.splits;
    const int nst = kspan / kBK;
    const int64_t a32 = 32 * (int64_t)g.lda, b32 = 32 * (int64_t)g.ldb;
    const uint16_t *pa = g.A + (int64_t)(m0 + lr) * g.lda + lc + (int64_t)bz * kspan;
    const uint16_t *pb = g.B + (int64_t)(n0 + lr) * g.ldb + lc + (int64_t)bz * kspan;
    // the bias gradient: row sums of the A slices (per K split for partials)
    const bool rowsum = (kEpi == kEpiStoreF32 || kEpi == kEpiPartialF32) && g.dbias != nullptr &&
                        n0 == 0;
    float rs[S::RA];
#pragma unroll
    for (int r = 0; r < S::RA; ++r)
        rs[r] = 0.f;

    const int64_t step = g.dstate ? g.dstate[0] : 0;

    f32x4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
            acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // kEpiLossGrad: the lane's targets, loaded before the main loop so that
    // their latency is hidden; the targets are in fragment order (one dword
    // per lane and 16 x 16 block: rows mb .. mb + 3 of the lane's column)
    uint32_t ytile[TM][TN];
    if constexpr (kEpi == kEpiLossGrad) {
        const uint32_t *y32 = reinterpret_cast<const uint32_t *>(g.marks);
        const int64_t nbn = g.N >> 4;
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                ytile[i][j] = y32[(((m0 + wm) / 16 + i) * nbn + (n0 + wn) / 16 + j) * 64 + lane];
    }

    auto load = [&](Stage<kBig> &r, int st) {
        const int64_t o = (int64_t)st * kBK;
#pragma unroll
        for (int q = 0; q < S::RA; ++q)
            r.a[q] = ld16(pa + q * a32 + o);
#pragma unroll
        for (int q = 0; q < S::RB; ++q)
            r.b[q] = ld16(pb + q * b32 + o);
    };
    auto put = [&](const Stage<kBig> &r, int b) {
#pragma unroll
        for (int q = 0; q < S::RA; ++q)
            st16(&sm.a[b][lr + 32 * q][lc], r.a[q]);
#pragma unroll
        for (int q = 0; q < S::RB; ++q)
            st16(&sm.b[b][lr + 32 * q][lc], r.b[q]);
        if (rowsum) {
#pragma unroll
            for (int q = 0; q < S::RA; ++q)
                rs[q] += sum8(r.a[q]);
        }
    };
    auto mma = [&](int b) {
#pragma unroll
        for (int ks = 0; ks < kBK; ks += 32) {
            bf16x8 a[TM], bb[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i)
                a[i] = *reinterpret_cast<const bf16x8 *>(
                    &sm.a[b][wm + i * 16 + (lane & 15)][ks + (lane >> 4) * 8]);
#pragma unroll
            for (int j = 0; j < TN; ++j)
                bb[j] = *reinterpret_cast<const bf16x8 *>(
                    &sm.b[b][wn + j * 16 + (lane & 15)][ks + (lane >> 4) * 8]);
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bb[j], acc[i][j], 0, 0, 0);
        }
    };
    // stage s is computed from LDS buffer s % 2; register set s % 2 carries
    // stage s + 1 (then s + 3) while stage s is computed
    Stage<kBig> r0, r1;
    load(r0, 0);
    if (nst > 1)
        load(r1, 1);
    put(r0, 0);
    if (nst > 2)
        load(r0, 2);
    __syncthreads();
    // pairs of stages, then an odd last one: no MFMA is conditional inside
    // the loop, so the accumulators keep their registers across iterations
    for (int s = 0; s + 1 < nst; s += 2) {
        mma(0);
        put(r1, 1);   // buffer 1 was last read before the previous barrier
        if (s + 3 < nst)
            load(r1, s + 3);
        __syncthreads();
        mma(1);
        if (s + 2 < nst)
            put(r0, 0);
        if (s + 4 < nst)
            load(r0, s + 4);
        __syncthreads();
    }
    if (nst & 1) {
        mma(0);
        __syncthreads();   // buffer 0 is refilled by the workgroup's next tile
    }
    if (rowsum) {
        // the 8 lanes loading one row are consecutive
#pragma unroll
        for (int q = 0; q < S::RA; ++q) {
#pragma unroll
            for (int off = 1; off < 8; off <<= 1)
                rs[q] += __shfl_xor(rs[q], off);
            if ((tid & 7) == 0)
                g.dbias[(kEpi == kEpiPartialF32 ? (int64_t)bz * g.M : 0) + m0 + lr + 32 * q] =
                    rs[q];
        }
    }

    // epilogue: lane holds C[m = mb + 4*(lane/16) + e][n = nb + lane%16], e = 0..3
    if constexpr (kEpi == kEpiAccum) {
        // s = sigmoid(acc + b); per column the sums of s and of s^2 (of the
        // squared pair means when pairs is set: rows mb, mb + 1 and mb + 2,
        // mb + 3 of one lane are pairs) and, with marks, of d = y - s and d^2.
        // Rows at or beyond mvalid count as s = 0 (and d = 0); mvalid is even
        // when pairs is set, so a pair never straddles it.
        const bool full = m0 + S::BM <= g.mvalid;
        const bool with_marks = g.marks != nullptr;
        const int mlane = m0 + wm + (lane >> 4) * 4;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int n = n0 + wn + j * 16 + (lane & 15);
            const float bias = g.bias[n];
            float cs = 0.f, cs2 = 0.f, cd = 0.f, cd2 = 0.f;
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                const int mb = mlane + i * 16;
                float sv[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float sg = sigmoid(acc[i][j][e] + bias);
                    sv[e] = (full || mb + e < g.mvalid) ? sg : 0.f;
                }
                cs += (sv[0] + sv[1]) + (sv[2] + sv[3]);
                if (g.pairs) {
                    const float p0 = 0.5f * (sv[0] + sv[1]), p1 = 0.5f * (sv[2] + sv[3]);
                    cs2 = fmaf(p0, p0, fmaf(p1, p1, cs2));
                } else {
                    cs2 = fmaf(sv[0], sv[0], fmaf(sv[1], sv[1], cs2));
                    cs2 = fmaf(sv[2], sv[2], fmaf(sv[3], sv[3], cs2));
                }
                if (with_marks) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        if (full || mb + e < g.mvalid) {
                            const float y = (float)g.marks[(g.row0 + mb + e) * g.N + n];
                            const float dv = y * g.tscale - sv[e];
                            cd += dv;
                            cd2 = fmaf(dv, dv, cd2);
                        }
                    }
                }
            }
            // lanes l, l^16, l^32, l^48 hold the same column; two waves share
            // each column, so the workgroup's sums meet in LDS
            cs += __shfl_xor(cs, 16);
            cs += __shfl_xor(cs, 32);
            cs2 += __shfl_xor(cs2, 16);
            cs2 += __shfl_xor(cs2, 32);
            if (with_marks) {
                cd += __shfl_xor(cd, 16);
                cd += __shfl_xor(cd, 32);
                cd2 += __shfl_xor(cd2, 16);
                cd2 += __shfl_xor(cd2, 32);
            }
            if (lane < 16) {
                const int cl = n - n0;
                atomicAdd(&sm.col[0][cl], cs);
                atomicAdd(&sm.col[1][cl], cs2);
                if (with_marks) {
                    atomicAdd(&sm.col[2][cl], cd);
                    atomicAdd(&sm.col[3][cl], cd2);
                }
            }
        }
        return;
    }
    // BF16 results go through LDS and leave as 16-byte row segments
    constexpr bool kStaged = kEpi == kEpiGeluFwd || kEpi == kEpiGeluBwd || kEpi == kEpiLossGrad;
    const bool want_loss = kEpi == kEpiLossGrad && g.loss != nullptr;
    float part_l = 0.f;   // the lane's 16 loss terms (monitoring only)
    float *cf = g.Cf;
    if (kEpi == kEpiPartialF32)
        cf += (int64_t)bz * g.M * g.ldc;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int n = n0 + wn + j * 16 + (lane & 15);
        const float bias = (kEpi == kEpiGeluFwd || kEpi == kEpiLogits || kEpi == kEpiLossGrad ||
                            kEpi == kEpiProbs) ? g.bias[n] : 0.f;
#pragma unroll
        for (int i = 0; i < TM; ++i) {
            const int mb = m0 + wm + i * 16 + (lane >> 4) * 4;
            uint16_t hv[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = mb + e;
                const float v = acc[i][j][e];
                if (kEpi == kEpiGeluFwd) {
                    const float z = v + bias;
                    if (cf)   // the pre-activation is kept for the backward pass only
                        cf[(int64_t)m * g.ldc + n] = z;
                    hv[e] = f2bf(gelu(z));
                } else if (kEpi == kEpiLogits) {
                    cf[(int64_t)m * g.ldc + n] = v + bias;
                } else if (kEpi == kEpiGeluBwd) {
                    hv[e] = f2bf(v * gelu_grad(g.Zsaved[(int64_t)m * g.N + n]));
                } else if (kEpi == kEpiStoreF32 || kEpi == kEpiPartialF32) {
                    cf[(int64_t)m * g.ldc + n] = v;
                } else if (kEpi == kEpiLossGrad) {
                    const float z = v + bias;
                    const float y = (float)((ytile[i][j] >> (8 * e)) & 0xffu) * g.tscale;
                    // ez = e^-|z| in (0, 1] serves the logistic function and the
                    // loss (hardware exp2, log2 and reciprocal, about 1 ulp each)
                    const float ez = __builtin_amdgcn_exp2f(-1.4426950408889634f * fabsf(z));
                    const float r = __builtin_amdgcn_rcpf(1.0f + ez);
                    const float sg = z >= 0.f ? r : ez * r;
                    hv[e] = f2bf((sg - y) * g.gscale);
                    if (want_loss)   // BCE with logits: max(z, 0) - z y + log(1 + e^-|z|)
                        part_l += fmaxf(z, 0.f) - z * y +
                                  0.6931471805599453f * __builtin_amdgcn_logf(1.0f + ez);
                } else if (kEpi == kEpiProbs) {
                    if (m < g.mvalid)
                        cf[(int64_t)m * g.ldc + n] = sigmoid(v + bias);
                }
            }
            if constexpr (kStaged) {
                const int ml = mb - m0, nl = n - n0;
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    sm.out[0][ml + e][nl] = hv[e];
                if (g.Cht) {
                    // transposed: 4 consecutive rows of column n are contiguous in C^T
                    uint2 packed;
                    packed.x = (uint32_t)hv[0] | ((uint32_t)hv[1] << 16);
                    packed.y = (uint32_t)hv[2] | ((uint32_t)hv[3] << 16);
                    *reinterpret_cast<uint2 *>(&sm.out[1][nl][ml]) = packed;
                }
            }
        }
    }
    if constexpr (kStaged) {
        // every wave has left the main loop (its last barrier) before writing
        // LDS above; this barrier makes the whole tile visible
        __syncthreads();
        constexpr int RW = S::BN / 8;   // 16-byte words per tile row
        for (int w = tid; w < S::BM * RW; w += 256) {
            const int r = w / RW, c = (w - r * RW) * 8;
            st16(g.Ch + (int64_t)(m0 + r) * g.N + n0 + c, ld16(&sm.out[0][r][c]));
        }
        if (g.Cht) {
            for (int w = tid; w < S::BN * RW; w += 256) {
                const int r = w / RW, c = (w - r * RW) * 8;
                st16(g.Cht + (int64_t)(n0 + r) * g.M + m0 + c, ld16(&sm.out[1][r][c]));
            }
        }
    }
    if (want_loss) {   // one atomic per workgroup
        double pl = (double)part_l;
        for (int off = 32; off > 0; off >>= 1)
            pl += __shfl_xor(pl, off);
        if (lane == 0)
            sm.red[wave] = pl;
        __syncthreads();
        if (tid == 0)
            atomicAdd(g.loss + step, sm.red[0] + sm.red[1] + sm.red[2] + sm.red[3]);
    }
}

// Tiles [0, tiles(g1)) belong to g1, the rest to g2.  Register budgets match
// the workgroups that fit a CU's LDS: 128 x 128 kernels two waves per SIMD
// (256 VGPRs), 64 x 64 kernels four (128 VGPRs); the accumulators then stay
// in architectural registers with no copies between MFMAs.
template <int kEpi1, int kEpi2, int kBig>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kBig ? 2 : 4)))
k_gemm(GemmArgs g1, GemmArgs g2)
{
    __shared__ GemmSmem<kBig> sm;
    if constexpr (kEpi2 == kEpiNone) {
        gemm_tile<kEpi1, kBig>(g1, blockIdx.x, sm);
    } else {
        const int n1 = gemm_tiles(g1, kBig);
        if ((int)blockIdx.x < n1)
            gemm_tile<kEpi1, kBig>(g1, blockIdx.x, sm);
        else
            gemm_tile<kEpi2, kBig>(g2, blockIdx.x - n1, sm);
    }
}

// kEpiAccum: workgroup b sums the tiles of column tile b % (N/BN) in row
// tiles [rt * (b / (N/BN)), + rt) in LDS, then adds them to the FP64 sums.
template <int kBig>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kBig ? 2 : 4)))
k_gemm_accum(GemmArgs g, int rt)
{
    using S = Shape<kBig>;
    __shared__ GemmSmem<kBig> sm;
    const int gx = g.N / S::BN, gy = g.M / S::BM;
    const int ct = blockIdx.x % gx, rg = blockIdx.x / gx;
    for (int i = threadIdx.x; i < 4 * S::BN; i += blockDim.x)
        sm.col[i / S::BN][i % S::BN] = 0.f;
    __syncthreads();
    const int end = min(gy, (rg + 1) * rt);
    for (int r = rg * rt; r < end; ++r)
        gemm_tile<kEpiAccum, kBig>(g, r * gx + ct, sm);
    __syncthreads();
    if (threadIdx.x < S::BN) {
        const int n = ct * S::BN + threadIdx.x;
        // the workgroup's rows (at most 8 x 128) lie in one row of the sums
        const int64_t sr = g.sum_rows ? (g.sum_row0 + (int64_t)rg * rt * S::BM) / g.sum_rows : 0;
        atomicAdd(&g.sums[sr * g.N + n], (double)sm.col[0][threadIdx.x]);
        if (g.sums_sq)
            atomicAdd(&g.sums_sq[sr * g.N + n], (double)sm.col[1][threadIdx.x]);
        if (g.marks && g.diffs)
            atomicAdd(&g.diffs[n], (double)sm.col[2][threadIdx.x]);
        if (g.marks && g.diffs_sq)
            atomicAdd(&g.diffs_sq[n], (double)sm.col[3][threadIdx.x]);
    }
}

// x (FP32, row stride n_in) of rows rowidx[row0 + m] (or row0 + m) -> X0
// (BF16 [M, kin], zero padded) and X0^T ([kin, M]); rows >= mvalid are zero.
// With marks set, the same rows of marks [*, nout] are gathered into ym in
// fragment order: the 16 x 16 block (bm, bn) of the [M, nout] targets is the
// 64 dwords at 64 (bm nout / 16 + bn), dword l holding rows 16 bm + 4 (l / 16)
// + e (byte e) of column 16 bn + l % 16, the lane layout of an MFMA result
// tile.  In a captured training step (dstate set)
// row0 = dstate[0] * M, and this kernel, the step's first, advances the Adam
// step counter dstate[1].
__global__ void __launch_bounds__(256) k_prep_input(const float *x, int n_in, int kin,
                                                    const int32_t *rowidx, int64_t row0, int M,
                                                    int mvalid, uint16_t *x0, uint16_t *x0t,
                                                    const uint8_t *marks, int nout, uint8_t *ym,
                                                    int64_t *dstate)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (dstate) {
        row0 = dstate[0] * M;
        if (t == 0)
            dstate[1] += 1;
    }
    if (t < (int64_t)M * kin) {
        const int m = (int)(t / kin), k = (int)(t % kin);
        float v = 0.f;
        if (m < mvalid && k < n_in) {
            const int64_t row = rowidx ? rowidx[row0 + m] : row0 + m;
            v = x[row * n_in + k];
        }
        const uint16_t h = f2bf(v);
        x0[(int64_t)m * kin + k] = h;
        x0t[(int64_t)k * M + m] = h;
    }
    // lane t: rows 4 mq .. 4 mq + 3, columns 16 q .. 16 q + 15 of the targets,
    // transposed to 16 dwords (byte e of dword c: row 4 mq + e, column 16 q + c);
    // four consecutive lanes fill one block's 256 bytes, consecutive blocks
    // are consecutive in memory
    const int chunks = nout / 16;
    if (marks && t < (int64_t)(M / 4) * chunks) {
        const int64_t blk = t >> 2;
        const int q = (int)(blk % chunks);
        const int mq = 4 * (int)(blk / chunks) + (int)(t & 3);
        uint32_t w[4][4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int m = 4 * mq + e;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (m < mvalid) {
                const int64_t row = rowidx ? rowidx[row0 + m] : row0 + m;
                v = *reinterpret_cast<const uint4 *>(marks + row * nout + q * 16);
            }
            w[e][0] = v.x;
            w[e][1] = v.y;
            w[e][2] = v.z;
            w[e][3] = v.w;
        }
        uint32_t o[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const int sh = 8 * (c & 3);
            o[c] = ((w[0][c >> 2] >> sh) & 0xffu) | (((w[1][c >> 2] >> sh) & 0xffu) << 8) |
                   (((w[2][c >> 2] >> sh) & 0xffu) << 16) | (((w[3][c >> 2] >> sh) & 0xffu) << 24);
        }
        // block (mq / 4, q) starts at dword 64 (mq / 4 * chunks + q); its
        // dwords 16 (mq % 4) .. + 15 are these rows
        uint4 *dst = reinterpret_cast<uint4 *>(ym) + ((int64_t)(mq >> 2) * chunks + q) * 16 +
                     4 * (mq & 3);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            dst[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
    }
}

// Adam over one 32 x 32 tile of one segment per workgroup; weight segments
// also refresh their BF16 copies (the transposed one through LDS so that both
// stores are coalesced).
__global__ void __launch_bounds__(256) k_adam(AdamArgs a)
{
    __shared__ uint16_t tile[32][33];
    __shared__ float corr[2];
    int b = blockIdx.x, s = 0, ct = 1;
    for (; s < a.nseg; ++s) {
        ct = (a.seg[s].cols + 31) / 32;
        const int tiles = ct * ((a.seg[s].rows + 31) / 32);
        if (b < tiles)
            break;
        b -= tiles;
    }
    if (s >= a.nseg)
        return;   // the grid is sized to the segments; never taken
    const AdamSeg sg = a.seg[s];
    const int r0 = (b / ct) * 32, c0 = (b % ct) * 32;
    if (threadIdx.x == 0) {
        const int64_t t = a.dstate ? a.tbase + a.dstate[1] : a.tbase;
        // bias corrections 1/(1 - beta^t)
        corr[0] = (float)(1.0 / (1.0 - exp((double)t * log((double)a.b1))));
        corr[1] = (float)(1.0 / (1.0 - exp((double)t * log((double)a.b2))));
        if (a.dstate && blockIdx.x == 0)
            a.dstate[0] += 1;   // no kernel of this launch reads dstate[0]
    }
    __syncthreads();
    const float c1 = corr[0], c2 = corr[1];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 lanes
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < sg.rows && c < sg.cols) {
            const int64_t i = sg.off + (int64_t)r * sg.cols + c;
            float gi;
            if (sg.nslices > 0) {
                const int64_t e = (int64_t)r * sg.cols + c, sz = (int64_t)sg.rows * sg.cols;
                gi = sg.slices[e];
                for (int k = 1; k < sg.nslices; ++k)
                    gi += sg.slices[k * sz + e];
            } else {
                gi = a.grad[i];
            }
            const float mi = a.b1 * a.m[i] + (1.f - a.b1) * gi;
            const float vi = a.b2 * a.v[i] + (1.f - a.b2) * gi * gi;
            a.m[i] = mi;
            a.v[i] = vi;
            const float pi = a.p[i] - a.lr * (mi * c1) / (sqrtf(vi * c2) + a.eps);
            a.p[i] = pi;
            if (sg.wh) {
                const uint16_t h = f2bf(pi);
                sg.wh[(int64_t)r * sg.cols + c] = h;
                tile[k][tx] = h;
            }
        }
    }
    if (sg.wh) {
        __syncthreads();
        for (int k = ty; k < 32; k += 8) {
            const int c = c0 + k, r = r0 + tx;
            if (r < sg.rows && c < sg.cols)
                sg.wt[(int64_t)c * sg.rows + r] = tile[tx][k];
        }
    }
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

// G = (sum over the split-K slices of acc) * gelu'(Z) as BF16, row-major and
// transposed; the slices are summed in order, so the result is deterministic.
__global__ void __launch_bounds__(256) k_gelu_bwd(const float *acc, int splits, const float *z,
                                                  int M, int N, uint16_t *g, uint16_t *gt)
{
    __shared__ uint16_t tile[32][33];
    const int n0 = blockIdx.x * 32, m0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t slice = (int64_t)M * N;
    // every load of the lane (4 rows x up to kMaxSplitSlices slices) is issued
    // before the first sum
    float part[4][kMaxSplitSlices], zz[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int m = m0 + ty + 8 * r, n = n0 + tx;
        const bool ok = m < M && n < N;
        const int64_t i = (int64_t)m * N + n;
        zz[r] = ok ? z[i] : 0.f;
#pragma unroll
        for (int p = 0; p < kMaxSplitSlices; ++p)
            part[r][p] = ok && p < splits ? acc[i + p * slice] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int k = ty + 8 * r;
        const int m = m0 + k, n = n0 + tx;
        float s = part[r][0];
#pragma unroll
        for (int p = 1; p < kMaxSplitSlices; ++p)
            s += part[r][p];
        if (m < M && n < N) {
            const uint16_t h = f2bf(s * gelu_grad(zz[r]));
            g[(int64_t)m * N + n] = h;
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

// 128 x 128 tiles when every product of the launch divides into them and
// they still give the GPU at least one workgroup per CU; 64 x 64 otherwise.
static bool use_big_tiles(const GemmArgs &g1, const GemmArgs *g2)
{
    // compute units of the current device (the caller has set it)
    int dev = 0, num_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&num_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        num_cu <= 0)
        num_cu = 256;
    auto fits = [](const GemmArgs &g) { return g.M % 128 == 0 && g.N % 128 == 0; };
    if (!fits(g1) || (g2 && !fits(*g2)))
        return false;
    return gemm_tiles(g1, 1) + (g2 ? gemm_tiles(*g2, 1) : 0) >= num_cu;
}

int gemm2_workgroups(const GemmArgs &g1, const GemmArgs *g2)
{
    const bool big = use_big_tiles(g1, g2);
    return gemm_tiles(g1, big) + (g2 ? gemm_tiles(*g2, big) : 0);
}

hipError_t launch_gemm2(int e1, const GemmArgs &g1, int e2, const GemmArgs &g2, hipStream_t s)
{
    if (g1.M % kTile || g1.N % kTile || g1.splits < 1 || g1.K % (g1.splits * kBK))
        return hipErrorInvalidValue;
    if (e2 != kEpiNone && (g2.M % kTile || g2.N % kTile || g2.splits < 1 ||
                           g2.K % (g2.splits * kBK)))
        return hipErrorInvalidValue;
    const bool big = use_big_tiles(g1, e2 == kEpiNone ? nullptr : &g2);
    if (e1 == kEpiAccum && e2 == kEpiNone) {
        constexpr int kRowTiles = 8;
        const int bt = big ? 128 : 64;
        const int gy = g1.M / bt;
        const dim3 grid((g1.N / bt) * ((gy + kRowTiles - 1) / kRowTiles));
        if (big)
            hipLaunchKernelGGL(k_gemm_accum<1>, grid, dim3(256), 0, s, g1, kRowTiles);
        else
            hipLaunchKernelGGL(k_gemm_accum<0>, grid, dim3(256), 0, s, g1, kRowTiles);
        return hipGetLastError();
    }
    const int n = gemm_tiles(g1, big) + (e2 == kEpiNone ? 0 : gemm_tiles(g2, big));
    const dim3 grid(n), block(256);
#define OKA_GEMM_CASE(E1, E2)                                                                 \
    if (e1 == (E1) && e2 == (E2)) {                                                           \
        if (big)                                                                              \
            hipLaunchKernelGGL((k_gemm<(E1), (E2), 1>), grid, block, 0, s, g1, g2);           \
        else                                                                                  \
            hipLaunchKernelGGL((k_gemm<(E1), (E2), 0>), grid, block, 0, s, g1, g2);           \
        return hipGetLastError();                                                             \
    }
    OKA_GEMM_CASE(kEpiGeluFwd, kEpiNone)
    OKA_GEMM_CASE(kEpiLogits, kEpiNone)
    OKA_GEMM_CASE(kEpiGeluBwd, kEpiNone)
    OKA_GEMM_CASE(kEpiStoreF32, kEpiNone)
    OKA_GEMM_CASE(kEpiLossGrad, kEpiNone)
    OKA_GEMM_CASE(kEpiProbs, kEpiNone)
    OKA_GEMM_CASE(kEpiPartialF32, kEpiNone)
    OKA_GEMM_CASE(kEpiStoreF32, kEpiGeluBwd)
    OKA_GEMM_CASE(kEpiStoreF32, kEpiPartialF32)
    OKA_GEMM_CASE(kEpiPartialF32, kEpiGeluBwd)
    OKA_GEMM_CASE(kEpiPartialF32, kEpiPartialF32)
#undef OKA_GEMM_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_gemm(int epi, const GemmArgs &g, hipStream_t s)
{
    GemmArgs none{};
    return launch_gemm2(epi, g, kEpiNone, none, s);
}

hipError_t launch_prep_input(const float *x, int n_in, int kin, const int32_t *rowidx,
                             int64_t row0, int M, int mvalid, uint16_t *x0, uint16_t *x0t,
                             const uint8_t *marks, int nout, uint8_t *ym, hipStream_t s,
                             int64_t *dstate)
{
    if (marks && (nout % 16 || M % 16 || !ym))
        return hipErrorInvalidValue;
    int64_t n = (int64_t)M * kin;
    if (marks)
        n = n > (int64_t)(M / 4) * (nout / 16) ? n : (int64_t)(M / 4) * (nout / 16);
    hipLaunchKernelGGL(k_prep_input, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n_in,
                       kin, rowidx, row0, M, mvalid, x0, x0t, marks, nout, ym, dstate);
    return hipGetLastError();
}

hipError_t launch_adam(const AdamArgs &a, hipStream_t s)
{
    int blocks = 0;
    for (int i = 0; i < a.nseg; ++i) {
        if (a.seg[i].rows < 1 || a.seg[i].cols < 1)
            return hipErrorInvalidValue;
        blocks += ((a.seg[i].cols + 31) / 32) * ((a.seg[i].rows + 31) / 32);
    }
    if (blocks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_adam, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_refresh_bf16(const float *w, int rows, int cols, uint16_t *wh, uint16_t *wt,
                               hipStream_t s)
{
    dim3 grid((cols + 31) / 32, (rows + 31) / 32);
    hipLaunchKernelGGL(k_refresh_bf16, grid, dim3(256), 0, s, w, rows, cols, wh, wt);
    return hipGetLastError();
}

hipError_t launch_gelu_bwd(const float *acc, int splits, const float *z, int M, int N,
                           uint16_t *g, uint16_t *gt, hipStream_t s)
{
    if (splits < 1 || splits > kMaxSplitSlices)
        return hipErrorInvalidValue;
    dim3 grid((N + 31) / 32, (M + 31) / 32);
    hipLaunchKernelGGL(k_gelu_bwd, grid, dim3(256), 0, s, acc, splits, z, M, N, g, gt);
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
