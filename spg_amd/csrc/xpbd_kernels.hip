// CDNA4 kernels of the XPBD step (spg/source/sim/solver/xpbd.cpp:16-132): substep boundaries and one
// projection kernel per energy type. A projection launch runs one stencil group (a colour, or a level of
// the serial order) of one energy: a thread per stencil gathers its particles, solves the stencil's
// nC x nC system and writes the displaced particles back. No two stencils of a launch share a
// particle, so the writes need no atomics and the result does not depend on the thread order.
#include <hip/hip_ext.h>

#include "element.h"
#include "vbd_device.h"
#include "xpbd_device.h"
#include "xpbd_element.h"

namespace spgvbd {

namespace {

using V3f = V3<float>;

__device__ __forceinline__ V3f xyz(float4 p) { return {p.x, p.y, p.z}; }
__device__ __forceinline__ float4 moved(float4 p, V3f d) { return make_float4(p.x + d.x, p.y + d.y, p.z + d.z, p.w); }

__device__ __forceinline__ void countFp64(bool fell, unsigned long long *counters)
{
    if (fell && counters) atomicAdd(counters, 1ull);
}

// The projection of stencil record i of type T (xpbd.cpp:84-95 for one stencil).
template <int T>
__device__ __forceinline__ void projectStencil(const XpbdArgs &a, int32_t i)
{
    float4 *pos = a.pos;
    if constexpr (T == kSpring) {
        const int2 id = a.sprI[i];
        const float2 p = a.sprP[i];
        const float4 X0 = pos[id.x], X1 = pos[id.y];
        V3f d0, d1;
        xpbdSpring<float>(xyz(X0), xyz(X1), X0.w, X1.w, p.x, p.y * a.invdt2, d0, d1);
        pos[id.x] = moved(X0, d0);
        pos[id.y] = moved(X1, d1);
    } else if constexpr (T == kAnchor) {
        const int32_t id = a.ancI[i];
        const float4 p = a.ancP[i];
        const float4 X = pos[id];
        pos[id] = moved(X, xpbdAnchor<float>(xyz(X), X.w, V3f{p.x, p.y, p.z}, p.w * a.invdt2));
    } else if constexpr (T == kStvk) {
        const int4 id = a.triI[i];
        const float4 D = a.triD[i], c = a.triC0[i];
        const float2 c2 = a.triC1[i];
        const float s = a.invdt2;
        const float4 X0 = pos[id.x], X1 = pos[id.y], X2 = pos[id.z];
        V3f o0, o1, o2;
        countFp64(xpbdStvkF(xyz(X0), xyz(X1), xyz(X2), X0.w, X1.w, X2.w, D.x, D.y, D.z, D.w, c.x * s, c.y * s, c.z * s,
                            c.w * s, c2.x * s, c2.y * s, o0, o1, o2),
                  a.counters);
        pos[id.x] = moved(X0, o0);
        pos[id.y] = moved(X1, o1);
        pos[id.z] = moved(X2, o2);
    } else if constexpr (T == kBending) {
        const int4 id = a.hinI[i];
        const float2 p = a.hinP[i];
        const float4 X0 = pos[id.x], X1 = pos[id.y], X2 = pos[id.z], X3 = pos[id.w];
        V3f o0, o1, o2, o3;
        xpbdBending<float>(xyz(X0), xyz(X1), xyz(X2), xyz(X3), X0.w, X1.w, X2.w, X3.w, p.x, p.y * a.invdt2, o0, o1,
                           o2, o3);
        pos[id.x] = moved(X0, o0);
        pos[id.y] = moved(X1, o1);
        pos[id.z] = moved(X2, o2);
        pos[id.w] = moved(X3, o3);
    } else if constexpr (T == kSnh) {
        const int4 id = a.tetI[i];
        const float4 d0 = a.tetD0[i], d1 = a.tetD1[i], d2 = a.tetD2[i];
        const float hyd = a.tetH[i];
        const float D[9] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w, d2.x};
        const float s = a.invdt2;
        const float4 X0 = pos[id.x], X1 = pos[id.y], X2 = pos[id.z], X3 = pos[id.w];
        V3f o0, o1, o2, o3;
        countFp64(xpbdSnhF(xyz(X0), xyz(X1), xyz(X2), xyz(X3), X0.w, X1.w, X2.w, X3.w, D, d2.y * s, d2.z * s, d2.w * s,
                           hyd, o0, o1, o2, o3),
                  a.counters);
        pos[id.x] = moved(X0, o0);
        pos[id.y] = moved(X1, o1);
        pos[id.z] = moved(X2, o2);
        pos[id.w] = moved(X3, o3);
    } else if constexpr (T >= 6) {
        constexpr int NS = xpbdFurtherStencil(T);
        const int4 id = a.extI[i];
        const int32_t vv[4] = {id.x, id.y, id.z, id.w};
        float4 X[4];
        double xs[12], ws[4], ps[12], dx[12];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            X[s] = pos[vv[s]];
            xs[3 * s] = X[s].x;
            xs[3 * s + 1] = X[s].y;
            xs[3 * s + 2] = X[s].z;
            ws[s] = X[s].w;
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const float4 c = a.extP[3 * static_cast<int64_t>(i) + q];
            ps[4 * q] = c.x;
            ps[4 * q + 1] = c.y;
            ps[4 * q + 2] = c.z;
            ps[4 * q + 3] = c.w;
        }
        countFp64(xpbdFurtherRecord<true>(T, xs, ws, ps, a.invdt2, dx), a.counters);
#pragma unroll
        for (int s = 0; s < NS; ++s)
            pos[vv[s]] = moved(X[s], V3f{float(dx[3 * s]), float(dx[3 * s + 1]), float(dx[3 * s + 2])});
    }
}

// One stencil group (a colour of one energy) over the whole chip: a thread per record. The grid size and
// threads per block come with the arguments, so the prologue waits for one round of scalar loads.
template <int T>
__global__ void __launch_bounds__(256) xpbdProjectKernel(XpbdArgs a, int32_t begin, int32_t end)
{
    asm volatile("" ::"s"(a.pos), "s"(a.hinI), "s"(a.hinP), "s"(a.triI), "s"(a.triD), "s"(a.triC0), "s"(a.triC1));
    const int32_t blk = xcdContiguousBlock(blockIdx.x, a.gridBlocks);
    const int32_t i = begin + blk * a.blockThreads + static_cast<int32_t>(threadIdx.x);
    if (i >= end) return;
    projectStencil<T>(a, i);
}

// A run of consecutive small stencil groups in one workgroup: the groups in order, each spread over the
// workgroup's threads, a workgroup barrier between two groups (the writes of one group are visible to the
// next: the whole run stays on one CU). Stencils of one group share no particle, so the result equals one
// launch per group.
__global__ void __launch_bounds__(256) xpbdProjectRunKernel(XpbdArgs a, XpbdRun run)
{
    for (int32_t r = 0; r < run.count; ++r) {
        const int32_t t = run.type[r];
        for (int32_t i = run.begin[r] + static_cast<int32_t>(threadIdx.x); i < run.end[r];
             i += static_cast<int32_t>(blockDim.x)) {
            switch (t) {
            case kSpring: projectStencil<kSpring>(a, i); break;
            case kStvk: projectStencil<kStvk>(a, i); break;
            case kBending: projectStencil<kBending>(a, i); break;
            case kSnh: projectStencil<kSnh>(a, i); break;
            case kAnchor: projectStencil<kAnchor>(a, i); break;
            case 6: projectStencil<6>(a, i); break;
            case 7: projectStencil<7>(a, i); break;
            case 8: projectStencil<8>(a, i); break;
            case 9: projectStencil<9>(a, i); break;
            case 10: projectStencil<10>(a, i); break;
            case 11: projectStencil<11>(a, i); break;
            case 12: projectStencil<12>(a, i); break;
            default: projectStencil<13>(a, i);
            }
        }
        __syncthreads();
    }
}

// xpbd.cpp:47-70 with particleGroup.cpp:57-63: x_old = x; v += h g; x += h v (every particle).
__global__ void __launch_bounds__(256) xpbdBeginKernel(float4 *__restrict__ pos, float4 *__restrict__ vel,
                                                       float4 *__restrict__ xold, int32_t n, float h, float hgx,
                                                       float hgy, float hgz)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 x = pos[i], v = vel[i];
    xold[i] = x;
    v.x += hgx;
    v.y += hgy;
    v.z += hgz;
    x.x += v.x * h;
    x.y += v.y * h;
    x.z += v.z * h;
    vel[i] = v;
    pos[i] = x;
}

// End of one substep (particleGroup.cpp:65-72, v = (x - x_old) / h) fused with the start of the next.
// The velocity lives in registers only: inside a step nothing reads vel between two substeps, and the
// end kernel writes the last substep's velocity.
__global__ void __launch_bounds__(256) xpbdBoundaryKernel(float4 *__restrict__ pos, float4 *__restrict__ xold,
                                                          int32_t n, float invhPrev, float h, float hgx, float hgy,
                                                          float hgz)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 x = pos[i];
    const float4 o = xold[i];
    float3 v = make_float3((x.x - o.x) * invhPrev, (x.y - o.y) * invhPrev, (x.z - o.z) * invhPrev);
    xold[i] = x;
    v.x += hgx;
    v.y += hgy;
    v.z += hgz;
    x.x += v.x * h;
    x.y += v.y * h;
    x.z += v.z * h;
    pos[i] = x;
}

__global__ void __launch_bounds__(256) xpbdEndKernel(const float4 *__restrict__ pos, const float4 *__restrict__ xold,
                                                     float4 *__restrict__ vel, int32_t n, float invh)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 x = pos[i], o = xold[i];
    vel[i] = make_float4((x.x - o.x) * invh, (x.y - o.y) * invh, (x.z - o.z) * invh, 0.f);
}

// Check hook: one projection per thread from FP64 inputs in the C-ABI parameter layout, by the device
// functions of the projection kernels (precision 0, with their FP64 fallback) or in double (1).
__global__ void xpbdCheckKernel(int type, int64_t n, const double *__restrict__ xs, const double *__restrict__ w,
                                const double *__restrict__ params, double invdt2, int precision,
                                double *__restrict__ dx)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const int ns = type == kAnchor ? 1 : type == kSpring ? 2 : type == kStvk ? 3 : 4;
    const int np = type == kStvk ? 13 : type == kSnh ? 14 : type == kAnchor ? 4 : 2;
    if (precision == 0)
        xpbdProjectRecord<true>(type, xs + i * ns * 3, w + i * ns, params + i * np, invdt2, dx + i * ns * 3);
    else
        xpbdProjectRecord<false>(type, xs + i * ns * 3, w + i * ns, params + i * np, invdt2, dx + i * ns * 3);
}

inline int blocksFor(int64_t n, int bs) { return static_cast<int>((n + bs - 1) / bs); }

template <int T>
void projectLaunch(const XpbdArgs &args, int32_t begin, int32_t end, int bs, hipStream_t stream, hipEvent_t start,
                   hipEvent_t stop)
{
    const int grid = blocksFor(end - begin, bs);
    XpbdArgs a = args;
    a.gridBlocks = grid;
    a.blockThreads = bs;
    if (start)
        hipExtLaunchKernelGGL(xpbdProjectKernel<T>, dim3(grid), dim3(bs), 0, stream, start, stop, 0, a, begin, end);
    else
        xpbdProjectKernel<T><<<grid, bs, 0, stream>>>(a, begin, end);
}

}  // namespace

hipError_t launchXpbdProject(int type, const XpbdArgs &a, int32_t begin, int32_t end, int blockSize,
                             hipStream_t stream, hipEvent_t start, hipEvent_t stop)
{
    if (end <= begin) return hipSuccess;
    if (blockSize < 64 || blockSize > 256 || blockSize % 64) return hipErrorInvalidValue;
    switch (type) {
    case kSpring: projectLaunch<kSpring>(a, begin, end, blockSize, stream, start, stop); break;
    case kStvk: projectLaunch<kStvk>(a, begin, end, blockSize, stream, start, stop); break;
    case kBending: projectLaunch<kBending>(a, begin, end, blockSize, stream, start, stop); break;
    case kSnh: projectLaunch<kSnh>(a, begin, end, blockSize, stream, start, stop); break;
    case kAnchor: projectLaunch<kAnchor>(a, begin, end, blockSize, stream, start, stop); break;
    case 6: projectLaunch<6>(a, begin, end, blockSize, stream, start, stop); break;
    case 7: projectLaunch<7>(a, begin, end, blockSize, stream, start, stop); break;
    case 8: projectLaunch<8>(a, begin, end, blockSize, stream, start, stop); break;
    case 9: projectLaunch<9>(a, begin, end, blockSize, stream, start, stop); break;
    case 10: projectLaunch<10>(a, begin, end, blockSize, stream, start, stop); break;
    case 11: projectLaunch<11>(a, begin, end, blockSize, stream, start, stop); break;
    case 12: projectLaunch<12>(a, begin, end, blockSize, stream, start, stop); break;
    case 13: projectLaunch<13>(a, begin, end, blockSize, stream, start, stop); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launchXpbdProjectRun(const XpbdArgs &a, const XpbdRun &run, hipStream_t stream, hipEvent_t start,
                                hipEvent_t stop)
{
    if (run.count < 1 || run.count > kXpbdRunMax) return hipErrorInvalidValue;
    for (int32_t r = 0; r < run.count; ++r)
        if (run.type[r] < 1 || run.type[r] > 13 || run.end[r] < run.begin[r]) return hipErrorInvalidValue;
    if (start)
        hipExtLaunchKernelGGL(xpbdProjectRunKernel, dim3(1), dim3(kXpbdRunThreads), 0, stream, start, stop, 0, a, run);
    else
        xpbdProjectRunKernel<<<1, kXpbdRunThreads, 0, stream>>>(a, run);
    return hipGetLastError();
}

hipError_t launchXpbdBegin(float4 *pos, float4 *vel, float4 *xold, int32_t n, float h, float gx, float gy, float gz,
                           hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    xpbdBeginKernel<<<blocksFor(n, 256), 256, 0, stream>>>(pos, vel, xold, n, h, gx, gy, gz);
    return hipGetLastError();
}

hipError_t launchXpbdBoundary(float4 *pos, float4 *xold, int32_t n, float invhPrev, float h, float gx, float gy,
                              float gz, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    xpbdBoundaryKernel<<<blocksFor(n, 256), 256, 0, stream>>>(pos, xold, n, invhPrev, h, gx, gy, gz);
    return hipGetLastError();
}

hipError_t launchXpbdEnd(const float4 *pos, const float4 *xold, float4 *vel, int32_t n, float invh,
                         hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    xpbdEndKernel<<<blocksFor(n, 256), 256, 0, stream>>>(pos, xold, vel, n, invh);
    return hipGetLastError();
}

hipError_t launchXpbdProjectCheck(int type, int64_t n, const double *xs, const double *w, const double *params,
                                  double invdt2, int precision, double *dx, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    xpbdCheckKernel<<<blocksFor(n, 64), 64, 0, stream>>>(type, n, xs, w, params, invdt2, precision, dx);
    return hipGetLastError();
}

}  // namespace spgvbd
