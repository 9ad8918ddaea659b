// oka_mlp.h -- argument block and launchers of the surrogate-network kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace oka {

enum MlpEpi {
    kEpiGeluFwd = 0,   // Z = acc + b (FP32), H = gelu(Z) (BF16, + transposed)
    kEpiLogits = 1,    // logits = acc + b (FP32)
    kEpiGeluBwd = 2,   // G = acc * gelu'(Zsaved) (BF16, + transposed)
    kEpiStoreF32 = 3,  // C = acc (FP32)
    kEpiLossGrad = 4,  // G = (sigmoid(acc + b) - y) * gscale (BF16, + transposed); loss
    kEpiAccum = 5,     // sums[n] += s, sums_sq += s^2, diffs += y - s, diffs_sq += (y - s)^2
    kEpiProbs = 6,     // C = sigmoid(acc + b) (FP32)
    kEpiAtomicF32 = 7  // C += acc (FP32 atomics; split-K partial sums, grid.z = K splits)
};

struct GemmArgs {
    const uint16_t *A;       // [M, K] BF16, row stride lda
    const uint16_t *B;       // [N, K] BF16, row stride ldb
    int M, N, K, lda, ldb;
    int mvalid;              // rows >= mvalid are padding (kEpiAccum, kEpiProbs)
    const float *bias;       // [N]
    float *Cf;               // FP32 output, row stride ldc
    int ldc;
    uint16_t *Ch;            // BF16 output [M, N]
    uint16_t *Cht;           // BF16 transposed output [N, M]
    const float *Zsaved;     // pre-activation [M, N] for kEpiGeluBwd
    const uint8_t *marks;    // targets: marks[row * N + n] * tscale
    const int32_t *rowidx;   // target/input row of tile row m: rowidx[row0 + m], or row0 + m
    int64_t row0;
    float tscale, gscale;
    double *sums, *sums_sq, *diffs, *diffs_sq, *loss;
    // graph-replayed training: when dstate is set, the tile rows start at
    // dstate[0] * step_rows and the loss goes to loss[dstate[0]]
    const int64_t *dstate;
    int step_rows;
    int ksplit;              // kEpiAtomicF32: K range per grid.z slice
};

hipError_t launch_gemm(int epi, const GemmArgs &g, hipStream_t s, int splits = 1);
hipError_t launch_prep_input(const float *x, int n_in, int kin, const int32_t *rowidx,
                             int64_t row0, int M, int mvalid, uint16_t *x0, uint16_t *x0t,
                             hipStream_t s, const int64_t *dstate = nullptr);
hipError_t launch_bias_grad(const uint16_t *gt, int N, int M, float *db, hipStream_t s);
// Adam; the step number t is dstate[1] + 1 when dstate is set, else `t`.
hipError_t launch_adam(float *p, const float *grad, float *m, float *v, int64_t n, float lr,
                       float b1, float b2, float eps, int64_t t, const int64_t *dstate,
                       hipStream_t s);
hipError_t launch_gelu_bwd(const float *acc, const float *z, int M, int N, uint16_t *g,
                           uint16_t *gt, hipStream_t s);
hipError_t launch_step_advance(int64_t *dstate, hipStream_t s);
hipError_t launch_fill_f32(float *p, int64_t n, hipStream_t s);
hipError_t launch_refresh_bf16(const float *w, int rows, int cols, uint16_t *wh, uint16_t *wt,
                               hipStream_t s);
hipError_t launch_zero_f64(double *p, int64_t n, hipStream_t s);

}  // namespace oka
