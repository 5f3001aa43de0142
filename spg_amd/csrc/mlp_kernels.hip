// Kernels of the learned warm start: a per-vertex MLP (24 -> 128 -> 128 -> 128 -> 3, SiLU) with BF16
// weights and activations on CDNA4 matrix cores (v_mfma_f32_16x16x32_bf16), FP32 accumulation.
//
// Every GEMM of training and inference is written as C[M x N] = A[M x K] * B[N x K]^T with both operands
// row-major and K contiguous ("NT"), so that each lane's MFMA operand fragment is one 16-byte load:
//   forward      Z = X W      with B = W^T stored as (out x in)
//   input grads  dX = dZ W^T  with B = W stored as (in x out)
//   weight grads dW = X^T dZ  with A = X^T and B = dZ^T (both stored transposed by the epilogues)
// plus the feature kernel, the loss gradient, bias gradients, Adam, and the fused inference kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "mlp_device.h"

namespace spgvbd {

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ float sigmoidf(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoidf(z); }
__device__ __forceinline__ float siluGrad(float z)
{
    const float s = sigmoidf(z);
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ bf16x8 loadFrag(const __bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }
__device__ __forceinline__ bf16x8 zeroFrag()
{
    bf16x8 z;
#pragma unroll
    for (int i = 0; i < 8; ++i) z[i] = static_cast<__bf16>(0.0f);
    return z;
}

// One workgroup of 4 wavefronts computes a 64x64 tile; each wavefront a 32x32 quarter as 2x2 MFMA
// blocks of 16x16. MFMA operand layout (16x16x32): lane l holds row (l % 16) of A and row (l % 16)
// of B at K offsets 8 (l / 16) .. +7 (the same K split for both operands, so the products pair
// up); the accumulator holds C[4 (l / 16) + r][l % 16], r = 0..3.
// Requirements (checked by the launcher): M, N multiples of 16, K multiple of 32, lda/ldb multiples of 8.
__global__ void __launch_bounds__(256) gemmNTKernel(const __bf16 *__restrict__ A, int lda, const __bf16 *__restrict__ B,
                                                    int ldb, int M, int N, int K, GemmEpilogue ep)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const int n0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int r16 = lane & 15;
    const int kq = (lane >> 4) * 8;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bool rowOk[2] = {m0 < M, m0 + 16 < M};
    const bool colOk[2] = {n0 < N, n0 + 16 < N};
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
            a[i] = rowOk[i] ? loadFrag(A + static_cast<int64_t>(m0 + i * 16 + r16) * lda + k0 + kq) : zeroFrag();
#pragma unroll
        for (int j = 0; j < 2; ++j)
            b[j] = colOk[j] ? loadFrag(B + static_cast<int64_t>(n0 + j * 16 + r16) * ldb + k0 + kq) : zeroFrag();
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (!rowOk[i] || !colOk[j]) continue;
            const int col = n0 + j * 16 + r16;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + i * 16 + (lane >> 4) * 4 + r;
                float v = acc[i][j][r];
                if (ep.dactZ) v *= siluGrad(ep.dactZ[static_cast<int64_t>(row) * ep.lddz + col]);
                if (ep.bias) v += ep.bias[col];
                if (ep.zOut) ep.zOut[static_cast<int64_t>(row) * ep.ldz + col] = v;
                const float h = ep.act ? silu(v) : v;
                if (ep.hOut) ep.hOut[static_cast<int64_t>(row) * ep.ldh + col] = static_cast<__bf16>(h);
                if (ep.hTOut) ep.hTOut[static_cast<int64_t>(col) * ep.ldht + row] = static_cast<__bf16>(h);
            }
        }
}

// dY = 2 (y - t) / n on the real rows and the first `outs` columns, zero elsewhere; bf16 row-major and
// transposed copies and an FP32 copy; the squared error summed into *loss.
__global__ void __launch_bounds__(256) lossGradKernel(const float *__restrict__ y, int ldy, const float *__restrict__ t,
                                                      int outs, int n, int rows, int cols, float *__restrict__ dYf,
                                                      __bf16 *__restrict__ dY, __bf16 *__restrict__ dYT,
                                                      double *__restrict__ loss)
{
    const int64_t idx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    const int64_t total = static_cast<int64_t>(rows) * cols;
    float se = 0.f;
    if (idx < total) {
        const int row = static_cast<int>(idx / cols), col = static_cast<int>(idx % cols);
        float g = 0.f;
        if (row < n && col < outs) {
            const float d = y[static_cast<int64_t>(row) * ldy + col] - t[static_cast<int64_t>(row) * outs + col];
            se = d * d;
            g = 2.0f * d / static_cast<float>(n);
        }
        dYf[idx] = g;
        dY[idx] = static_cast<__bf16>(g);
        dYT[static_cast<int64_t>(col) * rows + row] = static_cast<__bf16>(g);
    }
    // wavefront sum, one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) se += __shfl_down(se, off, 64);
    if ((threadIdx.x & 63) == 0 && se != 0.f) atomicAdd(loss, static_cast<double>(se));
}

// Column sums of a row-major FP32 matrix (bias gradients): one block per column, fixed order.
__global__ void __launch_bounds__(256) colSumKernel(const float *__restrict__ m, int ld, int rows, float *__restrict__ out)
{
    __shared__ float part[256];
    const int col = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) s += m[static_cast<int64_t>(r) * ld + col];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[col] = part[0];
}

// Adam (Kingma & Ba 2015) on FP32 master parameters.
__global__ void __launch_bounds__(256) adamKernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                                                  float *__restrict__ v, int64_t n, float lr, float b1, float b2,
                                                  float eps, float c1, float c2)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies of a (rows x cols) FP32 weight: W row-major and W^T row-major.
__global__ void __launch_bounds__(256) weightCopiesKernel(const float *__restrict__ w, int rows, int cols,
                                                          __bf16 *__restrict__ wb, __bf16 *__restrict__ wtb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    const int r = i / cols, c = i % cols;
    const __bf16 b = static_cast<__bf16>(w[i]);
    wb[i] = b;
    wtb[c * rows + r] = b;
}

// Rows of an FP32 (n x 24) feature table, standardised ((f - mean) / std), -> BF16 (rows x 32)
// zero-padded, optional transpose; rows of the table are picked through `index` (a minibatch) or taken
// in order when index is null. The matching targets are gathered into tOut (rows x 3) when given.
__global__ void __launch_bounds__(256) packInputKernel(const float *__restrict__ f, const float *__restrict__ t,
                                                       const int32_t *__restrict__ index,
                                                       const float *__restrict__ mean, const float *__restrict__ invStd,
                                                       int n, int rows, __bf16 *__restrict__ X, __bf16 *__restrict__ XT,
                                                       float *__restrict__ tOut)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= static_cast<int64_t>(rows) * kMlpIn) return;
    const int row = static_cast<int>(i / kMlpIn), col = static_cast<int>(i % kMlpIn);
    float v = 0.f;
    int src = -1;
    if (row < n) src = index ? index[row] : row;
    if (src >= 0 && col < kMlpFeatures) v = (f[static_cast<int64_t>(src) * kMlpFeatures + col] - mean[col]) * invStd[col];
    const __bf16 b = static_cast<__bf16>(v);
    X[i] = b;
    if (XT) XT[static_cast<int64_t>(col) * rows + row] = b;
    if (tOut && col < kMlpOut) tOut[static_cast<int64_t>(row) * kMlpOut + col] =
        src >= 0 ? t[static_cast<int64_t>(src) * kMlpOut + col] : 0.f;
}

// ------------------------------------------------------------------------------------------------
// Fused inference: all four layers in registers, one wavefront per kNT x 16 rows.
//
// Each layer is computed transposed, H^T = W X^T: the weights are the MFMA A operand (M = output
// features) and the rows are the B operand (N = 16 rows per tile), so the accumulator of tile i holds,
// in lane l, output features 16 i + 4 (l / 16) + r (r = 0..3) of row l % 16. The B fragment a layer
// needs for K step kk is, in lane l, 8 features of row l % 16; the lane already holds features
// 32 kk + 4 q + r (tile 2 kk) and 32 kk + 16 + 4 q + r (tile 2 kk + 1), q = l / 16. Using exactly those 8
// (a fixed permutation of K inside the step) needs no data exchange between lanes; the weight
// fragments are stored with the same permutation (mlpFragKernel), so the products pair up. The input
// layer uses the plain order: lane l holds features 8 q .. 8 q + 7.
// Weight fragments (76 KB) and biases are staged in LDS once per persistent workgroup.
constexpr int kFragL1 = 0, kFragL2 = 512, kFragL3 = 2560, kFragL4 = 4608, kFragTotal = 4864;
constexpr int kNT = 4;  // 16-row tiles per wavefront task (each weight fragment read from LDS feeds kNT MFMAs)
constexpr int kTaskRows = 16 * kNT;
constexpr int kBiasTotal = 3 * kMlpHidden + kMlpOutPad;

__device__ __forceinline__ int permutedK(int kk, int q, int p) { return 32 * kk + (p < 4 ? 4 * q + p : 16 + 4 * q + p - 4); }

// frag[f][p] for every fragment of the fused kernel, from the padded FP32 master weights.
__global__ void __launch_bounds__(256) mlpFragKernel(const float *__restrict__ w1, const float *__restrict__ w2,
                                                     const float *__restrict__ w3, const float *__restrict__ w4,
                                                     bf16x8 *__restrict__ frag)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= kFragTotal) return;
    const int lane = f & 63, q = lane >> 4, r16 = lane & 15;
    const int t = f >> 6;  // (tile, kk) index within its layer
    bf16x8 v;
    if (f < kFragL2) {  // layer 1: 8 tiles x 1 step, plain K order, 32 padded inputs
        const int i = t;
#pragma unroll
        for (int p = 0; p < 8; ++p) v[p] = static_cast<__bf16>(w1[(16 * i + r16) * kMlpIn + 8 * q + p]);
    } else if (f < kFragL4) {  // layers 2, 3: 8 tiles x 4 steps
        const float *w = f < kFragL3 ? w2 : w3;
        const int tt = f < kFragL3 ? t - kFragL2 / 64 : t - kFragL3 / 64;
        const int i = tt >> 2, kk = tt & 3;
#pragma unroll
        for (int p = 0; p < 8; ++p) v[p] = static_cast<__bf16>(w[(16 * i + r16) * kMlpHidden + permutedK(kk, q, p)]);
    } else {  // layer 4: 1 tile (outputs 0..15 of the padded 32) x 4 steps
        const int kk = t - kFragL4 / 64;
#pragma unroll
        for (int p = 0; p < 8; ++p) v[p] = static_cast<__bf16>(w4[r16 * kMlpHidden + permutedK(kk, q, p)]);
    }
    frag[f] = v;
}

__device__ __forceinline__ void hiddenEpilogue(const f32x4 (&acc)[8][kNT], const float *bias, int q,
                                               bf16x8 (&hb)[4][kNT])
{
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                hb[kk][nt][r] = static_cast<__bf16>(silu(acc[2 * kk][nt][r] + bias[32 * kk + 4 * q + r]));
                hb[kk][nt][4 + r] = static_cast<__bf16>(silu(acc[2 * kk + 1][nt][r] + bias[32 * kk + 16 + 4 * q + r]));
            }
}

__device__ __forceinline__ void hiddenLayer(const bf16x8 *wl, const bf16x8 (&hb)[4][kNT], int lane,
                                            f32x4 (&acc)[8][kNT])
{
#pragma unroll
    for (int i = 0; i < 8; ++i) {
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt) acc[i][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const bf16x8 a = wl[(i * 4 + kk) * 64 + lane];
#pragma unroll
            for (int nt = 0; nt < kNT; ++nt) acc[i][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, hb[kk][nt], acc[i][nt], 0, 0, 0);
        }
    }
}

__global__ void __launch_bounds__(256) mlpInferKernel(const float *__restrict__ f, int n, const float *__restrict__ mean,
                                                      const float *__restrict__ invStd, const bf16x8 *__restrict__ frag,
                                                      const float *__restrict__ b1, const float *__restrict__ b2,
                                                      const float *__restrict__ b3, const float *__restrict__ b4,
                                                      float *__restrict__ y, int ldy)
{
    __shared__ bf16x8 wl[kFragTotal];
    __shared__ float bl[kBiasTotal];
    __shared__ float nl[2 * kMlpIn];
    for (int i = threadIdx.x; i < kFragTotal; i += blockDim.x) wl[i] = frag[i];
    for (int i = threadIdx.x; i < kMlpHidden; i += blockDim.x) {
        bl[i] = b1[i];
        bl[kMlpHidden + i] = b2[i];
        bl[2 * kMlpHidden + i] = b3[i];
    }
    if (threadIdx.x < kMlpOutPad) bl[3 * kMlpHidden + threadIdx.x] = b4[threadIdx.x];
    if (threadIdx.x < kMlpIn) {
        const bool real = threadIdx.x < kMlpFeatures;
        nl[threadIdx.x] = real ? mean[threadIdx.x] : 0.f;
        nl[kMlpIn + threadIdx.x] = real ? invStd[threadIdx.x] : 0.f;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, q = lane >> 4, r16 = lane & 15;
    const int ntask = (n + kTaskRows - 1) / kTaskRows;
    const int waves = gridDim.x * (blockDim.x >> 6);
    for (int task = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); task < ntask; task += waves) {
        // compiler-level memory barrier: keeps the weight fragments in LDS (re-read per task) instead of
        // hoisting all 76 of them out of the loop into registers
        asm volatile("" ::: "memory");
        const int row0 = task * kTaskRows;
        bf16x8 hb[4][kNT];
        // standardised inputs, plain K order; features 24..31 and rows past n are zero
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt) {
            const int row = row0 + 16 * nt + r16;
            bf16x8 x = zeroFrag();
            if (row < n && q < 3) {
                const float4 *src = reinterpret_cast<const float4 *>(f + static_cast<int64_t>(row) * kMlpFeatures + 8 * q);
                const float4 u0 = src[0], u1 = src[1];
                const float u[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
                for (int p = 0; p < 8; ++p) x[p] = static_cast<__bf16>((u[p] - nl[8 * q + p]) * nl[kMlpIn + 8 * q + p]);
            }
            hb[0][nt] = x;
            hb[1][nt] = zeroFrag();
            hb[2][nt] = zeroFrag();
            hb[3][nt] = zeroFrag();
        }
        f32x4 acc[8][kNT];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const bf16x8 a = wl[kFragL1 + i * 64 + lane];
#pragma unroll
            for (int nt = 0; nt < kNT; ++nt)
                acc[i][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, hb[0][nt], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        }
        hiddenEpilogue(acc, bl, q, hb);
        hiddenLayer(wl + kFragL2, hb, lane, acc);
        hiddenEpilogue(acc, bl + kMlpHidden, q, hb);
        hiddenLayer(wl + kFragL3, hb, lane, acc);
        hiddenEpilogue(acc, bl + 2 * kMlpHidden, q, hb);
        f32x4 o[kNT];
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt) o[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const bf16x8 a = wl[kFragL4 + kk * 64 + lane];
#pragma unroll
            for (int nt = 0; nt < kNT; ++nt) o[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, hb[kk][nt], o[nt], 0, 0, 0);
        }
        if (q == 0) {
#pragma unroll
            for (int nt = 0; nt < kNT; ++nt) {
                const int row = row0 + 16 * nt + r16;
                if (row < n)
#pragma unroll
                    for (int r = 0; r < kMlpOut; ++r) y[static_cast<int64_t>(row) * ldy + r] = o[nt][r] + bl[3 * kMlpHidden + r];
            }
        }
    }
}

inline int blocksFor(int64_t n, int bs) { return static_cast<int>((n + bs - 1) / bs); }

}  // namespace

hipError_t launchGemmNT(const __bf16 *A, int lda, const __bf16 *B, int ldb, int M, int N, int K, const GemmEpilogue &ep,
                        hipStream_t stream)
{
    if (M % 16 || N % 16 || K % 32 || lda % 8 || ldb % 8 || M <= 0 || N <= 0 || K <= 0) return hipErrorInvalidValue;
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    gemmNTKernel<<<grid, 256, 0, stream>>>(A, lda, B, ldb, M, N, K, ep);
    return hipGetLastError();
}

hipError_t launchLossGrad(const float *y, int ldy, const float *t, int outs, int n, int rows, int cols, float *dYf,
                          __bf16 *dY, __bf16 *dYT, double *loss, hipStream_t stream)
{
    lossGradKernel<<<blocksFor(static_cast<int64_t>(rows) * cols, 256), 256, 0, stream>>>(y, ldy, t, outs, n, rows, cols,
                                                                                       dYf, dY, dYT, loss);
    return hipGetLastError();
}

hipError_t launchColSum(const float *m, int ld, int rows, int cols, float *out, hipStream_t stream)
{
    colSumKernel<<<cols, 256, 0, stream>>>(m, ld, rows, out);
    return hipGetLastError();
}

hipError_t launchAdam(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps,
                      int64_t t, hipStream_t stream)
{
    const float c1 = 1.f - powf(b1, static_cast<float>(t)), c2 = 1.f - powf(b2, static_cast<float>(t));
    adamKernel<<<blocksFor(n, 256), 256, 0, stream>>>(p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t launchWeightCopies(const float *w, int rows, int cols, __bf16 *wb, __bf16 *wtb, hipStream_t stream)
{
    weightCopiesKernel<<<blocksFor(rows * cols, 256), 256, 0, stream>>>(w, rows, cols, wb, wtb);
    return hipGetLastError();
}

hipError_t launchPackInput(const float *f, const float *t, const int32_t *index, const float *mean, const float *invStd,
                           int n, int rows, __bf16 *X, __bf16 *XT, float *tOut, hipStream_t stream)
{
    packInputKernel<<<blocksFor(static_cast<int64_t>(rows) * kMlpIn, 256), 256, 0, stream>>>(f, t, index, mean, invStd,
                                                                                           n, rows, X, XT, tOut);
    return hipGetLastError();
}

}  // namespace spgvbd

namespace spgvbd {

hipError_t launchMlpFrag(const float *w1, const float *w2, const float *w3, const float *w4, void *frag,
                         hipStream_t stream)
{
    mlpFragKernel<<<blocksFor(kFragTotal, 256), 256, 0, stream>>>(w1, w2, w3, w4, static_cast<bf16x8 *>(frag));
    return hipGetLastError();
}

size_t mlpFragBytes() { return sizeof(bf16x8) * kFragTotal; }

hipError_t launchMlpInfer(const float *f, int n, const float *mean, const float *invStd, const void *frag,
                          const float *b1, const float *b2, const float *b3, const float *b4, float *y, int ldy,
                          int maxBlocks, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    if (ldy < kMlpOut) return hipErrorInvalidValue;
    const int tasks = (n + kTaskRows - 1) / kTaskRows;
    const int grid = std::max(1, std::min((tasks + 3) / 4, maxBlocks));
    mlpInferKernel<<<grid, 256, 0, stream>>>(f, n, mean, invStd, static_cast<const bf16x8 *>(frag), b1, b2, b3, b4, y,
                                             ldy);
    return hipGetLastError();
}

}  // namespace spgvbd
