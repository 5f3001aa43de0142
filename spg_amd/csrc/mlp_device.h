// Device-side interface of the learned warm start: network shape, GEMM epilogue and kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "spg_vbd_learned.h"

namespace spgvbd {

constexpr int kMlpFeatures = 24;  // per-vertex input features
constexpr int kMlpIn = 32;        // features zero-padded to one MFMA K step
constexpr int kMlpHidden = 128;
constexpr int kMlpOut = 3;
constexpr int kMlpOutPad = 32;  // output rows of the last weight padded to one MFMA K step
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
hipError_t launchPackInput(const float *f, const float *t, const int32_t *index, const float *mean, const float *invStd,
                           int n, int rows, __bf16 *X, __bf16 *XT, float *tOut, hipStream_t stream);

// Fused inference (all layers in registers, weights staged in LDS): fragments of the weights in the
// kernel's operand order are rebuilt from the FP32 master weights by launchMlpFrag after every change.
size_t mlpFragBytes();
hipError_t launchMlpFrag(const float *w1, const float *w2, const float *w3, const float *w4, void *frag,
                         hipStream_t stream);
// y[row][0..2] = network(f[row]) for rows < n (f: n x 24 FP32, y stride ldy >= 3); at most maxBlocks
// persistent workgroups of 256 threads.
hipError_t launchMlpInfer(const float *f, int n, const float *mean, const float *invStd, const void *frag,
                          const float *b1, const float *b2, const float *b3, const float *b4, float *y, int ldy,
                          int maxBlocks, hipStream_t stream);

// Inputs of the feature kernel (warmstart_kernels.hip), slot numbering of a VBD context.
struct FeatureArgs {
    const float4 *xt = nullptr;    // inertial target (w: mass)
    const float4 *vel = nullptr;   // velocity after the substep's gravity kick
    const float4 *corr = nullptr;  // x* - x~ of the previous substep
    const int32_t *nbPtr = nullptr;
    const int32_t *nbIdx = nullptr;
    const float *nbRest = nullptr;  // rest length of each one-ring edge
    const float4 *stat = nullptr;   // (log10 Y, log10 k_bending, anchored, 0)
    int32_t n = 0;
    float h = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
    float invS = 0.f, invU = 0.f;  // 1 / (h^2 a_ref), 1 / (h a_ref)
    float *out = nullptr;          // n x kMlpFeatures
};

hipError_t launchFeatures(const FeatureArgs &a, hipStream_t stream);
// x = x~ + base + s y (base may be null: x~ + s y)
hipError_t launchApplyWarmStart(float4 *pos, const float4 *xt, const float4 *base, const float *y, int32_t ldy, float s,
                                int32_t n,
                                hipStream_t stream);
// y, out: n rows of 4 floats (the first three are the displacement)
hipError_t launchSmoothPrediction(const float *y, float *out, const int32_t *nbPtr, const int32_t *nbIdx, int32_t n,
                                  hipStream_t stream);
hipError_t launchGate(float4 *pos, const float4 *xt, const int32_t *objOfSlot, const int32_t *reject, int32_t n,
                      hipStream_t stream);

// Columns of the training forward pass's output buffer (the last layer is computed 16 wide).
constexpr int kMlpInferCols = 16;
// Network inference on device-resident features (n x kMlpFeatures FP32) into y (n rows, stride
// ldy >= kMlpOut), ordered on `stream`, with the fused kernel. Throws VbdError. Defined with the
// network object.
void mlpInferDevice(spg_mlp *net, const float *features, int64_t n, float *y, int32_t ldy, hipStream_t stream);
// The same on the network's own stream, after `ready` (recorded by the caller); records `done`, which
// the caller waits on before it reads y.
void mlpInferAfter(spg_mlp *net, const float *features, int64_t n, float *y, int32_t ldy, hipEvent_t ready,
                   hipEvent_t done);
int mlpDevice(const spg_mlp *net);
int mlpHistoryBase(const spg_mlp *net);

}  // namespace spgvbd
