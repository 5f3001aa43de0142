// CDNA4 kernels of the VBD step: inertial predict (K1), per-colour vertex block solve (K2),
// velocity update (K3), FP64 residual (K4) and the per-element derivative hook.
#include "element.h"
#include "vbd_device.h"

namespace spgvbd {

namespace {

using V3f = V3<float>;
using S3f = Sym3<float>;

__device__ __forceinline__ V3f xyz(float4 p) { return {p.x, p.y, p.z}; }

__device__ __forceinline__ void acc(V3f &g, S3f &H, const V3f &ge, const S3f &He)
{
    g.x += ge.x;
    g.y += ge.y;
    g.z += ge.z;
    H.xx += He.xx;
    H.xy += He.xy;
    H.xz += He.xz;
    H.yy += He.yy;
    H.yz += He.yz;
    H.zz += He.zz;
}

template <int G>
__device__ __forceinline__ float groupSum(float v)
{
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
}

// K1 (vbd.cpp:48-83, particleGroup.cpp:57-63): x_old = x; v += h g; x += h v; x_tilde = x.
__global__ void __launch_bounds__(256) beginSubstepKernel(float4 *__restrict__ pos, float4 *__restrict__ xold,
                                                          float4 *__restrict__ vel, float4 *__restrict__ xt,
                                                          const float *__restrict__ mass, int32_t n, float h,
                                                          float gx, float gy, float gz)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 x = pos[i];
    float4 v = vel[i];
    xold[i] = x;
    v.x += h * gx;
    v.y += h * gy;
    v.z += h * gz;
    x.x += v.x * h;
    x.y += v.y * h;
    x.z += v.z * h;
    vel[i] = v;
    pos[i] = x;
    xt[i] = make_float4(x.x, x.y, x.z, mass[i]);
}

// K3 (vbd.cpp:173-185, particleGroup.cpp:65-72): v = (x - x_old) / h.
__global__ void __launch_bounds__(256) endSubstepKernel(const float4 *__restrict__ pos,
                                                        const float4 *__restrict__ xold, float4 *__restrict__ vel,
                                                        int32_t n, float invh)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 x = pos[i], xo = xold[i];
    vel[i] = make_float4((x.x - xo.x) * invh, (x.y - xo.y) * invh, (x.z - xo.z) * invh, 0.f);
}

__global__ void __launch_bounds__(256) copyXtKernel(float4 *__restrict__ pos, const float4 *__restrict__ xt, int32_t n)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 t = xt[i];
    pos[i] = make_float4(t.x, t.y, t.z, 0.f);
}

// 3x3 block solve of l_vertexDescent (vbd.cpp:139-146): if det(H) > 1e-10 then x += H^-1 f.
// FP32 adjugate solve; when det is near the 1e-10 threshold or small relative to |H|^3 (condition
// worse than ~1e4) the decision and the solve are redone in FP64 from the same accumulated H, f.
__device__ __forceinline__ bool blockSolve(const S3f &H, const V3f &f, V3f &dx, unsigned long long *counters)
{
    const float c00 = H.yy * H.zz - H.yz * H.yz;
    const float c01 = H.xz * H.yz - H.xy * H.zz;
    const float c02 = H.xy * H.yz - H.xz * H.yy;
    const float det = H.xx * c00 + H.xy * c01 + H.xz * c02;
    const float scale = fmaxf(fmaxf(fmaxf(fabsf(H.xx), fabsf(H.yy)), fmaxf(fabsf(H.zz), fabsf(H.xy))),
                              fmaxf(fabsf(H.xz), fabsf(H.yz)));
    const bool wellPosed = det > 2e-10f && det > 1e-4f * scale * scale * scale;
    if (wellPosed) {
        const float c11 = H.xx * H.zz - H.xz * H.xz;
        const float c12 = H.xy * H.xz - H.xx * H.yz;
        const float c22 = H.xx * H.yy - H.xy * H.xy;
        const float id = 1.0f / det;
        dx.x = (c00 * f.x + c01 * f.y + c02 * f.z) * id;
        dx.y = (c01 * f.x + c11 * f.y + c12 * f.z) * id;
        dx.z = (c02 * f.x + c12 * f.y + c22 * f.z) * id;
        return true;
    }
    const double xx = H.xx, xy = H.xy, xz = H.xz, yy = H.yy, yz = H.yz, zz = H.zz;
    const double d00 = yy * zz - yz * yz, d01 = xz * yz - xy * zz, d02 = xy * yz - xz * yy;
    const double det64 = xx * d00 + xy * d01 + xz * d02;
    if (det64 > 1e-10) {
        const double d11 = xx * zz - xz * xz, d12 = xy * xz - xx * yz, d22 = xx * yy - xy * xy;
        const double fx = f.x, fy = f.y, fz = f.z;
        dx.x = static_cast<float>((d00 * fx + d01 * fy + d02 * fz) / det64);
        dx.y = static_cast<float>((d01 * fx + d11 * fy + d12 * fz) / det64);
        dx.z = static_cast<float>((d02 * fx + d12 * fy + d22 * fz) / det64);
        atomicAdd(&counters[0], 1ull);
        return true;
    }
    atomicAdd(&counters[1], 1ull);
    return false;
}

// K2: one colour sweep. G lanes of a wavefront cooperate on one vertex: lane l of the group takes
// incidences l, l+G, ... of each energy type, then the 3-vector force and the 6 Hessian entries
// are summed over the group with cross-lane shuffles and lane 0 solves and writes.
template <int G>
__global__ void __launch_bounds__(256) solveColorKernel(SolveArgs a)
{
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t vid = a.begin + t / G;
    const int32_t lane = t & (G - 1);
    const bool active = vid < a.end;
    V3f g = {0.f, 0.f, 0.f};
    S3f H = zeroSym<float>();
    V3f xo = {0.f, 0.f, 0.f};
    if (active) {
        xo = xyz(a.pos[vid]);
        if (a.triPtr) {
            const int32_t e = a.triPtr[vid + 1];
            for (int32_t k = a.triPtr[vid] + lane; k < e; k += G) {
                const int2 o = a.triO[k];
                const float4 gg = a.triG[k];
                const float4 kk = a.triK[k];
                V3f ge;
                S3f He;
                stvkVertex<float>(xo, xyz(a.pos[o.x]), xyz(a.pos[o.y]), gg.x, gg.y, gg.z, gg.w, kk.x, kk.y, kk.z, ge,
                                  He);
                acc(g, H, ge, He);
            }
        }
        if (a.hinPtr) {
            const int32_t e = a.hinPtr[vid + 1];
            for (int32_t k = a.hinPtr[vid] + lane; k < e; k += G) {
                const int4 v = a.hinV[k];
                const float2 p = a.hinP[k];
                const int slot = v.x == vid ? 0 : (v.y == vid ? 1 : (v.z == vid ? 2 : 3));
                V3f ge;
                S3f He;
                bendingVertex<float>(xyz(a.pos[v.x]), xyz(a.pos[v.y]), xyz(a.pos[v.z]), xyz(a.pos[v.w]), slot, p.x,
                                     p.y, ge, He);
                acc(g, H, ge, He);
            }
        }
        if (a.tetPtr) {
            const int32_t e = a.tetPtr[vid + 1];
            for (int32_t k = a.tetPtr[vid] + lane; k < e; k += G) {
                const int4 o = a.tetO[k];
                const float4 ga = a.tetGa[k], gb = a.tetGb[k], gc = a.tetGc[k];
                V3f ge;
                S3f He;
                snhVertex<float>(xo, xyz(a.pos[o.x]), xyz(a.pos[o.y]), xyz(a.pos[o.z]), xyz(ga), xyz(gb), xyz(gc),
                                 ga.w, gb.w, gc.w, ge, He);
                acc(g, H, ge, He);
            }
        }
        if (a.sprPtr) {
            const int32_t e = a.sprPtr[vid + 1];
            for (int32_t k = a.sprPtr[vid] + lane; k < e; k += G) {
                const float4 s = a.spr[k];
                V3f ge;
                S3f He;
                springVertex<float>(xo, xyz(a.pos[__float_as_int(s.x)]), s.y, s.z, ge, He);
                acc(g, H, ge, He);
            }
        }
        if (a.ancPtr) {
            const int32_t e = a.ancPtr[vid + 1];
            for (int32_t k = a.ancPtr[vid] + lane; k < e; k += G) {
                const float4 s = a.anc[k];
                V3f ge;
                S3f He;
                anchorVertex<float>(xo, xyz(s), s.w, ge, He);
                acc(g, H, ge, He);
            }
        }
    }
    if (G > 1) {
        g.x = groupSum<G>(g.x);
        g.y = groupSum<G>(g.y);
        g.z = groupSum<G>(g.z);
        H.xx = groupSum<G>(H.xx);
        H.xy = groupSum<G>(H.xy);
        H.xz = groupSum<G>(H.xz);
        H.yy = groupSum<G>(H.yy);
        H.yz = groupSum<G>(H.yz);
        H.zz = groupSum<G>(H.zz);
    }
    if (active && lane == 0) {
        const float4 xt = a.xt[vid];
        const float mdt2 = xt.w * a.invdt2;
        // f = -(m/h^2)(x - x_tilde) - sum grad; H = (m/h^2) I + sum H_e (vbd.cpp:102-138)
        const V3f f = {-mdt2 * (xo.x - xt.x) - g.x, -mdt2 * (xo.y - xt.y) - g.y, -mdt2 * (xo.z - xt.z) - g.z};
        H.xx += mdt2;
        H.yy += mdt2;
        H.zz += mdt2;
        V3f dx;
        if (blockSolve(H, f, dx, a.counters)) a.pos[vid] = make_float4(xo.x + dx.x, xo.y + dx.y, xo.z + dx.z, 0.f);
    }
}

using V3d = V3<double>;
using S3d = Sym3<double>;

__device__ __forceinline__ V3d xyzd(float4 p) { return {double(p.x), double(p.y), double(p.z)}; }

// K4: r_v = -(m/h^2)(x_v - x_tilde_v) - sum_e grad_v E_e(x), all in FP64 from the FP32 state and the
// exact prepared parameters (implicitEulerBase.cpp:102-143 force, x-form of
// implicitEulerNewtonRobust.cpp:65-66). Writes |r_v|^2 per vertex slot.
__global__ void __launch_bounds__(256) residualKernel(ResidualArgs a)
{
    const int32_t vid = blockIdx.x * blockDim.x + threadIdx.x;
    if (vid >= a.n) return;
    const float4 xf = a.pos[vid], tf = a.xt[vid];
    const V3d xo = xyzd(xf);
    const double mdt2 = a.mass[vid] * a.invdt2;
    V3d r = {-mdt2 * (xo.x - double(tf.x)), -mdt2 * (xo.y - double(tf.y)), -mdt2 * (xo.z - double(tf.z))};
    V3d ge;
    S3d He;
    if (a.triPtr) {
        for (int32_t k = a.triPtr[vid]; k < a.triPtr[vid + 1]; ++k) {
            const int32_t el = a.triElem[k];
            const int4 v = a.triV[el];
            const double *p = a.triP + 8 * static_cast<int64_t>(el);
            // rows of Dinv: g1 = (p0, p1), g2 = (p2, p3), g0 = -(g1 + g2)
            const double g[3][2] = {{-(p[0] + p[2]), -(p[1] + p[3])}, {p[0], p[1]}, {p[2], p[3]}};
            const int s = v.x == vid ? 0 : (v.y == vid ? 1 : 2);
            const int ia = s == 0 ? 1 : 0, ib = s == 2 ? 1 : 2;
            const int vv[3] = {v.x, v.y, v.z};
            stvkVertex<double>(xo, xyzd(a.pos[vv[ia]]), xyzd(a.pos[vv[ib]]), g[ia][0], g[ia][1], g[ib][0], g[ib][1],
                               p[4], p[5], p[6], ge, He);
            r = r - ge;
        }
    }
    if (a.hinPtr) {
        for (int32_t k = a.hinPtr[vid]; k < a.hinPtr[vid + 1]; ++k) {
            const int32_t el = a.hinElem[k];
            const int4 v = a.hinV[el];
            const double2 p = a.hinP[el];
            const int slot = v.x == vid ? 0 : (v.y == vid ? 1 : (v.z == vid ? 2 : 3));
            bendingVertex<double>(xyzd(a.pos[v.x]), xyzd(a.pos[v.y]), xyzd(a.pos[v.z]), xyzd(a.pos[v.w]), slot, p.x,
                                  p.y, ge, He);
            r = r - ge;
        }
    }
    if (a.tetPtr) {
        for (int32_t k = a.tetPtr[vid]; k < a.tetPtr[vid + 1]; ++k) {
            const int32_t el = a.tetElem[k];
            const int4 v = a.tetV[el];
            const double *p = a.tetP + 12 * static_cast<int64_t>(el);
            V3d gr[4];
            gr[1] = {p[0], p[1], p[2]};
            gr[2] = {p[3], p[4], p[5]};
            gr[3] = {p[6], p[7], p[8]};
            gr[0] = -(gr[1] + gr[2] + gr[3]);
            const int vv[4] = {v.x, v.y, v.z, v.w};
            const int s = v.x == vid ? 0 : (v.y == vid ? 1 : (v.z == vid ? 2 : 3));
            int o[3], m = 0;
            for (int q = 0; q < 4; ++q)
                if (q != s) o[m++] = q;
            snhVertex<double>(xo, xyzd(a.pos[vv[o[0]]]), xyzd(a.pos[vv[o[1]]]), xyzd(a.pos[vv[o[2]]]), gr[o[0]],
                              gr[o[1]], gr[o[2]], p[9], p[10], p[11], ge, He);
            r = r - ge;
        }
    }
    if (a.sprPtr) {
        for (int32_t k = a.sprPtr[vid]; k < a.sprPtr[vid + 1]; ++k) {
            const int32_t el = a.sprElem[k];
            const int2 v = a.sprV[el];
            const double2 p = a.sprP[el];
            const int32_t other = v.x == vid ? v.y : v.x;
            springVertex<double>(xo, xyzd(a.pos[other]), p.x, p.y, ge, He);
            r = r - ge;
        }
    }
    if (a.ancPtr) {
        for (int32_t k = a.ancPtr[vid]; k < a.ancPtr[vid + 1]; ++k) {
            const int32_t el = a.ancElem[k];
            const double4 p = a.ancP[el];
            anchorVertex<double>(xo, V3d{p.x, p.y, p.z}, p.w, ge, He);
            r = r - ge;
        }
    }
    a.r2[vid] = dot(r, r);
}

// Sum of r2 over each object's slots, in a fixed order (one block per object).
__global__ void __launch_bounds__(256) reduceObjectsKernel(const double *__restrict__ r2,
                                                           const int32_t *__restrict__ objSlots,
                                                           const int64_t *__restrict__ objOffsets,
                                                           double *__restrict__ out)
{
    __shared__ double part[256];
    const int32_t obj = blockIdx.x;
    double s = 0.0;
    for (int64_t j = objOffsets[obj] + threadIdx.x; j < objOffsets[obj + 1]; j += blockDim.x) s += r2[objSlots[j]];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[obj] = sqrt(part[0]);
}

// Per-element derivative hook: evaluates the same device functions the solve uses, for every
// vertex of each element, from the prepared FP64 parameters (params layout of spg_vbd.h).
template <class R>
__global__ void __launch_bounds__(64) elementKernel(int type, int64_t n, const double *xs, const double *params,
                                                    double *grad, double *hess)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    int ns = 0, np = 0;
    switch (type) {
    case kSpring: ns = 2; np = 2; break;
    case kStvk: ns = 3; np = 13; break;
    case kBending: ns = 4; np = 2; break;
    case kSnh: ns = 4; np = 12; break;
    case kAnchor: ns = 1; np = 4; break;
    default: return;
    }
    const double *xe = xs + i * ns * 3;
    const double *p = params + i * np;
    V3<R> x[4];
    for (int s = 0; s < ns; ++s) x[s] = {R(xe[3 * s]), R(xe[3 * s + 1]), R(xe[3 * s + 2])};
    for (int s = 0; s < ns; ++s) {
        V3<R> ge = {R(0), R(0), R(0)};
        Sym3<R> He = zeroSym<R>();
        if (type == kSpring) {
            springVertex<R>(x[s], x[1 - s], R(p[0]), R(p[1]), ge, He);
        } else if (type == kAnchor) {
            anchorVertex<R>(x[0], V3<R>{R(p[0]), R(p[1]), R(p[2])}, R(p[3]), ge, He);
        } else if (type == kStvk) {
            const double g[3][2] = {{-(p[0] + p[2]), -(p[1] + p[3])}, {p[0], p[1]}, {p[2], p[3]}};
            const int ia = s == 0 ? 1 : 0, ib = s == 2 ? 1 : 2;
            stvkVertex<R>(x[s], x[ia], x[ib], R(g[ia][0]), R(g[ia][1]), R(g[ib][0]), R(g[ib][1]), R(p[4]), R(p[5]),
                          R(p[12]), ge, He);
        } else if (type == kBending) {
            bendingVertex<R>(x[0], x[1], x[2], x[3], s, R(p[0]), R(p[1]), ge, He);
        } else {
            double gr[4][3];
            for (int c = 0; c < 3; ++c) {
                gr[1][c] = p[c];
                gr[2][c] = p[3 + c];
                gr[3][c] = p[6 + c];
                gr[0][c] = -(p[c] + p[3 + c] + p[6 + c]);
            }
            int o[3], m = 0;
            for (int q = 0; q < 4; ++q)
                if (q != s) o[m++] = q;
            auto G = [&](int q) { return V3<R>{R(gr[q][0]), R(gr[q][1]), R(gr[q][2])}; };
            snhVertex<R>(x[s], x[o[0]], x[o[1]], x[o[2]], G(o[0]), G(o[1]), G(o[2]), R(p[9]), R(p[10]), R(p[11]), ge,
                         He);
        }
        double *gd = grad + (i * ns + s) * 3;
        gd[0] = ge.x;
        gd[1] = ge.y;
        gd[2] = ge.z;
        double *hd = hess + (i * ns + s) * 9;
        hd[0] = He.xx;
        hd[1] = He.xy;
        hd[2] = He.xz;
        hd[3] = He.xy;
        hd[4] = He.yy;
        hd[5] = He.yz;
        hd[6] = He.xz;
        hd[7] = He.yz;
        hd[8] = He.zz;
    }
}

inline int blocksFor(int64_t n, int bs) { return static_cast<int>((n + bs - 1) / bs); }

}  // namespace

hipError_t launchBeginSubstep(float4 *pos, float4 *xold, float4 *vel, float4 *xt, const float *mass, int32_t n,
                              float h, float gx, float gy, float gz, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    beginSubstepKernel<<<blocksFor(n, 256), 256, 0, stream>>>(pos, xold, vel, xt, mass, n, h, gx, gy, gz);
    return hipGetLastError();
}

hipError_t launchEndSubstep(const float4 *pos, const float4 *xold, float4 *vel, int32_t n, float invh,
                            hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    endSubstepKernel<<<blocksFor(n, 256), 256, 0, stream>>>(pos, xold, vel, n, invh);
    return hipGetLastError();
}

hipError_t launchCopyXt(float4 *pos, const float4 *xt, int32_t n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    copyXtKernel<<<blocksFor(n, 256), 256, 0, stream>>>(pos, xt, n);
    return hipGetLastError();
}

hipError_t launchSolveColor(const SolveArgs &a, int lanesPerVertex, int blockSize, int /*ldsStaging*/,
                            hipStream_t stream)
{
    const int64_t nv = a.end - a.begin;
    if (nv <= 0) return hipSuccess;
    if (blockSize > 256 || blockSize < 64 || blockSize % 64) return hipErrorInvalidValue;
    const int grid = blocksFor(nv * lanesPerVertex, blockSize);
    switch (lanesPerVertex) {
    case 1: solveColorKernel<1><<<grid, blockSize, 0, stream>>>(a); break;
    case 2: solveColorKernel<2><<<grid, blockSize, 0, stream>>>(a); break;
    case 4: solveColorKernel<4><<<grid, blockSize, 0, stream>>>(a); break;
    case 8: solveColorKernel<8><<<grid, blockSize, 0, stream>>>(a); break;
    case 16: solveColorKernel<16><<<grid, blockSize, 0, stream>>>(a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launchResidual(const ResidualArgs &a, hipStream_t stream)
{
    if (a.n <= 0) return hipSuccess;
    residualKernel<<<blocksFor(a.n, 256), 256, 0, stream>>>(a);
    return hipGetLastError();
}

hipError_t launchReduceObjects(const double *r2, const int32_t *objSlots, const int64_t *objOffsets, int32_t nObjects,
                               double *out, hipStream_t stream)
{
    if (nObjects <= 0) return hipSuccess;
    reduceObjectsKernel<<<nObjects, 256, 0, stream>>>(r2, objSlots, objOffsets, out);
    return hipGetLastError();
}

hipError_t launchElementDerivatives(int type, int64_t n, const double *xs, const double *params, int precision,
                                    double *grad, double *hess, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    if (precision == 0)
        elementKernel<float><<<blocksFor(n, 64), 64, 0, stream>>>(type, n, xs, params, grad, hess);
    else
        elementKernel<double><<<blocksFor(n, 64), 64, 0, stream>>>(type, n, xs, params, grad, hess);
    return hipGetLastError();
}

}  // namespace spgvbd
