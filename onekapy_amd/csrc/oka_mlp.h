// oka_mlp.h -- argument blocks and launchers of the surrogate-network kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace oka {

enum MlpEpi {
    kEpiNone = -1,     // no second product in a grouped launch
    kEpiGeluFwd = 0,   // Z = acc + b (FP32), H = gelu(Z) (BF16, + transposed)
    kEpiLogits = 1,    // logits = acc + b (FP32)
    kEpiGeluBwd = 2,   // G = acc * gelu'(Zsaved) (BF16, + transposed)
    kEpiStoreF32 = 3,  // C = acc (FP32); with dbias, dbias[m] = sum over K of A[m, :]
    kEpiLossGrad = 4,  // G = (sigmoid(acc + b) - y) * gscale (BF16, + transposed); loss
    kEpiAccum = 5,     // sums[n] += s, sums_sq += s^2, diffs += y - s, diffs_sq += (y - s)^2
    kEpiProbs = 6,     // C = sigmoid(acc + b) (FP32)
    kEpiPartialF32 = 7 // C + z * M * ldc = acc over K range z of `splits` (split-K partial sums)
};

struct GemmArgs {
    const uint16_t *A;       // [M, K] BF16, row stride lda
    const uint16_t *B;       // [N, K] BF16, row stride ldb
    int M, N, K, lda, ldb;
    int mvalid;              // rows >= mvalid are padding (kEpiAccum, kEpiProbs)
    int pairs;               // kEpiAccum: sums_sq takes ((s_2k + s_2k+1) / 2)^2 per row pair
    // kEpiAccum with sum_rows > 0: row m counts into sums row (sum_row0 + m) / sum_rows
    // (sums is then [replicates x N]; sum_row0 and sum_rows multiples of 1024)
    int64_t sum_row0, sum_rows;
    int splits;              // K ranges (kEpiPartialF32), else 1
    const float *bias;       // [N]
    float *Cf;               // FP32 output, row stride ldc
    int ldc;
    uint16_t *Ch;            // BF16 output [M, N]
    uint16_t *Cht;           // BF16 transposed output [N, M]
    const float *Zsaved;     // pre-activation [M, N] for kEpiGeluBwd
    float *dbias;            // kEpiStoreF32: row sums of A (the bias gradient of a dW product)
    const uint8_t *marks;    // targets * (1 / tscale): kEpiLossGrad the batch's rows in the
                             // fragment order of launch_prep_input, kEpiAccum
                             // marks[(row0 + m) * N + n]
    int64_t row0;
    float tscale, gscale;
    double *sums, *sums_sq, *diffs, *diffs_sq;
    double *loss;            // kEpiLossGrad: summed BCE (nullptr: not computed)
    // graph-replayed training: when dstate is set the loss goes to loss[dstate[0]]
    const int64_t *dstate;
};

// One FP32 master segment of the parameter vector updated by Adam: a weight
// matrix [rows, cols] with its BF16 copies (wh row-major, wt transposed), or a
// bias vector laid out as [len / 32, 32] with wh = wt = nullptr.
struct AdamSeg {
    int64_t off;
    int rows, cols;
    uint16_t *wh, *wt;
    // nslices > 0: the segment's gradient is the sum, in slice order, of
    // nslices split-K partial slices [rows x cols] at `slices`, not grad + off
    const float *slices;
    int nslices;
};

struct AdamArgs {
    float *p;
    const float *grad;
    float *m, *v;
    float lr, b1, b2, eps;
    // Adam step t = tbase + dstate[1] when dstate is set (the first kernel of
    // a captured training step advances dstate[1]), else tbase; the update
    // also advances dstate[0], the batch counter that the step's other kernels
    // have finished reading.
    int64_t tbase;
    int64_t *dstate;
    AdamSeg seg[6];
    int nseg;
};

constexpr int kMaxSplitSlices = 8;   // split-K partial slices summed by launch_gelu_bwd

// Block tiles of a product at 64 x 64 (big = 0) or 128 x 128 (big = 1).
__host__ __device__ inline int gemm_tiles(const GemmArgs &g, int big = 0)
{
    const int t = big ? 128 : 64;
    return (g.N / t) * (g.M / t) * g.splits;
}

// Workgroups of a launch of g1 (and g2 when not null): one per block tile.
int gemm2_workgroups(const GemmArgs &g1, const GemmArgs *g2);
// One launch of up to two independent products (the second with e2 != kEpiNone).
hipError_t launch_gemm(int epi, const GemmArgs &g, hipStream_t s);
hipError_t launch_gemm2(int e1, const GemmArgs &g1, int e2, const GemmArgs &g2, hipStream_t s);
// The batch's inputs (and, with marks set, its target rows into ym [M x nout
// bytes] in MFMA fragment order: 16 x 16 blocks of 64 dwords, dword l of block
// (bm, bn) holding rows 16 bm + 4 (l / 16) .. + 3 of column 16 bn + l % 16).
hipError_t launch_prep_input(const float *x, int n_in, int kin, const int32_t *rowidx,
                             int64_t row0, int M, int mvalid, uint16_t *x0, uint16_t *x0t,
                             const uint8_t *marks, int nout, uint8_t *ym, hipStream_t s,
                             int64_t *dstate = nullptr);
hipError_t launch_adam(const AdamArgs &a, hipStream_t s);
// G = (sum of the `splits` partial slices [M, N] of acc) * gelu'(Z) as BF16 (+ transposed).
hipError_t launch_gelu_bwd(const float *acc, int splits, const float *z, int M, int N, uint16_t *g,
                           uint16_t *gt, hipStream_t s);
hipError_t launch_fill_f32(float *p, int64_t n, hipStream_t s);
hipError_t launch_refresh_bf16(const float *w, int rows, int cols, uint16_t *wh, uint16_t *wt,
                               hipStream_t s);
hipError_t launch_zero_f64(double *p, int64_t n, hipStream_t s);

}  // namespace oka
