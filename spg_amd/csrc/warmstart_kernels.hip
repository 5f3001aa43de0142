// Kernels of the learned warm start that work on the VBD state: per-vertex input features, the
// predicted start x = x~ + s * y, and the per-object gate that falls back to x~.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp_device.h"

namespace spgvbd {

namespace {

__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 xyz(float4 a) { return make_float3(a.x, a.y, a.z); }

// Feature row of one vertex (slot i), all in the slot numbering of the context:
//   0-2   c_i / s           inertial offset: x* - x~ of the previous substep, s = h^2 a_ref
//   3-5   v_i / u           previous velocity (before this substep's gravity kick), u = h a_ref
//   6-8   mean_j c_j / s    one-ring mean of the inertial offset
//   9-11  max_j |c_j - c_i| / s   (per component)
//   12-14 mean_j v_j / u
//   15-17 max_j |v_j - v_i| / u
//   18    log10 Young's modulus of the incident membrane / solid elements (spring stiffness for springs)
//   19    log10 bending stiffness of the incident hinges
//   20    1 if the vertex is anchored
//   21-23 mean, max, min over the one ring of the strain |x~_j - x~_i| / L0_ij - 1 at the inertial target
// j runs over the vertices that share a triangle, tetrahedron or spring with i. A vertex without
// neighbours gets zeros in 6-17 and 21-23.
__global__ void __launch_bounds__(256) featuresKernel(FeatureArgs a)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float3 kick = make_float3(a.h * a.gx, a.h * a.gy, a.h * a.gz);
    const float3 xi = xyz(a.xt[i]);
    const float3 ci = xyz(a.corr[i]);
    const float3 vi = sub3(xyz(a.vel[i]), kick);
    float3 cs = make_float3(0.f, 0.f, 0.f), vs = cs, cm = cs, vm = cs;
    float es = 0.f, emax = 0.f, emin = 0.f;
    const int32_t b = a.nbPtr[i], e = a.nbPtr[i + 1];
    for (int32_t k = b; k < e; ++k) {
        const int32_t j = a.nbIdx[k];
        const float3 cj = xyz(a.corr[j]);
        const float3 vj = sub3(xyz(a.vel[j]), kick);
        cs.x += cj.x; cs.y += cj.y; cs.z += cj.z;
        vs.x += vj.x; vs.y += vj.y; vs.z += vj.z;
        cm.x = fmaxf(cm.x, fabsf(cj.x - ci.x)); cm.y = fmaxf(cm.y, fabsf(cj.y - ci.y)); cm.z = fmaxf(cm.z, fabsf(cj.z - ci.z));
        vm.x = fmaxf(vm.x, fabsf(vj.x - vi.x)); vm.y = fmaxf(vm.y, fabsf(vj.y - vi.y)); vm.z = fmaxf(vm.z, fabsf(vj.z - vi.z));
        const float3 d = sub3(xyz(a.xt[j]), xi);
        const float strain = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z) / a.nbRest[k] - 1.f;
        es += strain;
        emax = k == b ? strain : fmaxf(emax, strain);
        emin = k == b ? strain : fminf(emin, strain);
    }
    const float inv = e > b ? 1.f / static_cast<float>(e - b) : 0.f;
    const float4 st = a.stat[i];
    float *o = a.out + static_cast<int64_t>(i) * kMlpFeatures;
    o[0] = ci.x * a.invS; o[1] = ci.y * a.invS; o[2] = ci.z * a.invS;
    o[3] = vi.x * a.invU; o[4] = vi.y * a.invU; o[5] = vi.z * a.invU;
    o[6] = cs.x * inv * a.invS; o[7] = cs.y * inv * a.invS; o[8] = cs.z * inv * a.invS;
    o[9] = cm.x * a.invS; o[10] = cm.y * a.invS; o[11] = cm.z * a.invS;
    o[12] = vs.x * inv * a.invU; o[13] = vs.y * inv * a.invU; o[14] = vs.z * inv * a.invU;
    o[15] = vm.x * a.invU; o[16] = vm.y * a.invU; o[17] = vm.z * a.invU;
    o[18] = st.x; o[19] = st.y; o[20] = st.z;
    o[21] = es * inv; o[22] = emax; o[23] = emin;
}

// One smoothing pass of the predicted displacement over the one ring: out_i = (y_i + sum_j y_j) / (1 + deg_i)
// (first three columns, row stride 4).
__global__ void __launch_bounds__(256) smoothPredictionKernel(const float *__restrict__ y, float *__restrict__ out,
                                                              const int32_t *__restrict__ nbPtr,
                                                              const int32_t *__restrict__ nbIdx, int32_t n)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 yi = *reinterpret_cast<const float4 *>(y + 4 * static_cast<int64_t>(i));
    float sx = yi.x, sy = yi.y, sz = yi.z;
    const int32_t b = nbPtr[i], e = nbPtr[i + 1];
    for (int32_t k = b; k < e; ++k) {
        const float4 yj = *reinterpret_cast<const float4 *>(y + 4 * static_cast<int64_t>(nbIdx[k]));
        sx += yj.x;
        sy += yj.y;
        sz += yj.z;
    }
    const float inv = 1.f / static_cast<float>(1 + e - b);
    *reinterpret_cast<float4 *>(out + 4 * static_cast<int64_t>(i)) = make_float4(sx * inv, sy * inv, sz * inv, 0.f);
}

// x = x~ + s * y for every vertex (y: the network output, row stride ldy).
__global__ void __launch_bounds__(256) applyWarmStartKernel(float4 *__restrict__ pos, const float4 *__restrict__ xt,
                                                            const float4 *__restrict__ base,
                                                            const float *__restrict__ y, int32_t ldy, float s, int32_t n)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 t = xt[i];
    if (base) {
        const float4 b = base[i];
        t.x += b.x;
        t.y += b.y;
        t.z += b.z;
    }
    const float *yi = y + static_cast<int64_t>(i) * ldy;
    pos[i] = make_float4(t.x + s * yi[0], t.y + s * yi[1], t.z + s * yi[2], 0.f);
}

// x = x~ for the vertices of rejected objects.
__global__ void __launch_bounds__(256) gateKernel(float4 *__restrict__ pos, const float4 *__restrict__ xt,
                                                  const int32_t *__restrict__ objOfSlot,
                                                  const int32_t *__restrict__ reject, int32_t n)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!reject[objOfSlot[i]]) return;
    const float4 t = xt[i];
    pos[i] = make_float4(t.x, t.y, t.z, 0.f);
}

// c = x* - x~ at the end of a substep (the next substep's inertial-offset feature).
inline int blocksOf(int64_t n) { return static_cast<int>((n + 255) / 256); }

}  // namespace

hipError_t launchFeatures(const FeatureArgs &a, hipStream_t stream)
{
    if (a.n <= 0) return hipSuccess;
    featuresKernel<<<blocksOf(a.n), 256, 0, stream>>>(a);
    return hipGetLastError();
}

hipError_t launchApplyWarmStart(float4 *pos, const float4 *xt, const float4 *base, const float *y, int32_t ldy, float s,
                                int32_t n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    applyWarmStartKernel<<<blocksOf(n), 256, 0, stream>>>(pos, xt, base, y, ldy, s, n);
    return hipGetLastError();
}

hipError_t launchSmoothPrediction(const float *y, float *out, const int32_t *nbPtr, const int32_t *nbIdx, int32_t n,
                                  hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    smoothPredictionKernel<<<blocksOf(n), 256, 0, stream>>>(y, out, nbPtr, nbIdx, n);
    return hipGetLastError();
}

hipError_t launchGate(float4 *pos, const float4 *xt, const int32_t *objOfSlot, const int32_t *reject, int32_t n,
                      hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    gateKernel<<<blocksOf(n), 256, 0, stream>>>(pos, xt, objOfSlot, reject, n);
    return hipGetLastError();
}

}  // namespace spgvbd
