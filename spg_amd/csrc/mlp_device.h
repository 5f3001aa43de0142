// Device-side interface of the learned warm start: network shape, GEMM epilogue and kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spgvbd {

constexpr int kMlpFeatures = 24;  // per-vertex input features
constexpr int kMlpIn = 32;        // features zero-padded to one MFMA K step
constexpr int kMlpHidden = 128;
constexpr int kMlpOut = 3;
constexpr int kMlpOutPad = 16;  // output rows of the last weight padded to one MFMA tile
constexpr int kMlpLayers = 4;

// What the GEMM does with an accumulator tile C[row][col] before storing it:
//   v = C; if dactZ: v *= silu'(dactZ[row][col]); if bias: v += bias[col];
//   zOut[row][col] = v (FP32); h = act ? silu(v) : v; hOut[row][col] = bf16(h); hTOut[col][row] = bf16(h).
struct GemmEpilogue {
    const float *bias = nullptr;
    int act = 0;
    const float *dactZ = nullptr;
    int lddz = 0;
    float *zOut = nullptr;
    int ldz = 0;
    __bf16 *hOut = nullptr;
    int ldh = 0;
    __bf16 *hTOut = nullptr;
    int ldht = 0;
};

// C[M x N] = A[M x K] * B[N x K]^T, bf16 operands, FP32 accumulation. M, N multiples of 16, K multiple
// of 32, lda and ldb multiples of 8 (16-byte fragments).
hipError_t launchGemmNT(const __bf16 *A, int lda, const __bf16 *B, int ldb, int M, int N, int K, const GemmEpilogue &ep,
                        hipStream_t stream);
hipError_t launchLossGrad(const float *y, int ldy, const float *t, int outs, int n, int rows, int cols, float *dYf,
                          __bf16 *dY, __bf16 *dYT, double *loss, hipStream_t stream);
hipError_t launchColSum(const float *m, int ld, int rows, int cols, float *out, hipStream_t stream);
hipError_t launchAdam(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps,
                      int64_t t, hipStream_t stream);
hipError_t launchWeightCopies(const float *w, int rows, int cols, __bf16 *wb, __bf16 *wtb, hipStream_t stream);
hipError_t launchPackInput(const float *f, const float *t, const int32_t *index, int n, int rows, __bf16 *X, __bf16 *XT,
                           float *tOut, hipStream_t stream);

}  // namespace spgvbd
