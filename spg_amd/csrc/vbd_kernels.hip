// CDNA4 kernels of the VBD step: inertial predict (K1), per-colour vertex block solve (K2),
// velocity update (K3), FP64 residual (K4) and the per-element derivative hook.
#include <hip/hip_ext.h>

#include <type_traits>

#include "element.h"
#include "element_eval.h"
#include "implicit_element.h"
#include "vbd_device.h"

namespace spgvbd {

namespace {

using V3f = V3<float>;
using S3f = Sym3<float>;

__device__ __forceinline__ V3f xyz(float4 p) { return {p.x, p.y, p.z}; }

// Adds an element's gradient and block only when `take`, by selects: a dropped contribution (a padding record,
// evaluated on stand-in positions) reaches neither g nor H. The kernel instantiation with the further energies
// uses it for every class; the five-energy instantiations branch around the evaluation instead (acc), which
// measured faster there (about 1.5% of the 512x512 cloth's step) and slower with the further energies (2 to 8% of
// that cloth's colour launch).
__device__ __forceinline__ void accIf(bool take, V3f &g, S3f &H, const V3f &ge, const S3f &He)
{
    g.x += take ? ge.x : 0.f;
    g.y += take ? ge.y : 0.f;
    g.z += take ? ge.z : 0.f;
    H.xx += take ? He.xx : 0.f;
    H.xy += take ? He.xy : 0.f;
    H.xz += take ? He.xz : 0.f;
    H.yy += take ? He.yy : 0.f;
    H.yz += take ? He.yz : 0.f;
    H.zz += take ? He.zz : 0.f;
}

// Adds an element's gradient and block.
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

// Evaluates one incidence (eval writes its gradient and block) and adds it when `ok`: BRANCH = true skips the
// evaluation of padding records by a branch, false evaluates every record (padding on stand-in positions) and
// drops it by selects (accIf). Both add the same values; the choice is by measured speed per instantiation.
template <bool BRANCH, class F>
__device__ __forceinline__ void accElement(bool ok, V3f &g, S3f &H, F &&eval)
{
    V3f ge;
    S3f He;
    if constexpr (BRANCH) {
        if (ok) {
            eval(ge, He);
            acc(g, H, ge, He);
        }
    } else {
        eval(ge, He);
        accIf(ok, g, H, ge, He);
    }
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

// K3 (vbd.cpp:173-185, particleGroup.cpp:65-72): v = (x - x_old) / h; with it the substep's correction
// x - x~ (the offset the history and learned starts of the next substep use), from the same read of x.
__global__ void __launch_bounds__(256) endSubstepKernel(const float4 *__restrict__ pos,
                                                        const float4 *__restrict__ xold,
                                                        const float4 *__restrict__ xt, float4 *__restrict__ vel,
                                                        float4 *__restrict__ corr, int32_t n, float invh)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 x = pos[i], xo = xold[i], t = xt[i];
    vel[i] = make_float4((x.x - xo.x) * invh, (x.y - xo.y) * invh, (x.z - xo.z) * invh, 0.f);
    corr[i] = make_float4(x.x - t.x, x.y - t.y, x.z - t.z, 0.f);
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
// Returns kSolvedFp32, kSolvedFp64 (dx set) or kSkipped (det(H) <= 1e-10: x stays).
constexpr int kSolvedFp32 = 0, kSolvedFp64 = 1, kSkipped = 2;
__device__ __forceinline__ int blockSolve(const S3f &H, const V3f &f, V3f &dx)
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
        return kSolvedFp32;
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
        return kSolvedFp64;
    }
    return kSkipped;
}

template <int G, class T>
__device__ __forceinline__ T groupSum(T v, int32_t lane64);

// The end of a vertex's update, shared by the solve kernels: the group's partial forces and Hessians are
// summed in a fixed shuffle order, the group's first lane adds the inertia term, solves the 3x3 block and
// writes the position; the FP64-fallback and skip counts are summed per wavefront (one atomic each at most).
template <int G>
__device__ __forceinline__ void finishVertex(const SolveArgs &a, bool active, int32_t vid, int32_t lane64,
                                             int32_t lig, const V3f &xo, const float4 &xt, V3f g, S3f H)
{
    if (G > 1) {
        g.x = groupSum<G>(g.x, lane64);
        g.y = groupSum<G>(g.y, lane64);
        g.z = groupSum<G>(g.z, lane64);
        H.xx = groupSum<G>(H.xx, lane64);
        H.xy = groupSum<G>(H.xy, lane64);
        H.xz = groupSum<G>(H.xz, lane64);
        H.yy = groupSum<G>(H.yy, lane64);
        H.yz = groupSum<G>(H.yz, lane64);
        H.zz = groupSum<G>(H.zz, lane64);
    }
    int path = -1;
    if (active && lig == 0) {
        const float mdt2 = xt.w * a.invdt2;
        // f = -(m/h^2)(x - x_tilde) - sum grad; H = (m/h^2) I + sum H_e (vbd.cpp:102-138)
        const V3f f = {-mdt2 * (xo.x - xt.x) - g.x, -mdt2 * (xo.y - xt.y) - g.y, -mdt2 * (xo.z - xt.z) - g.z};
        H.xx += mdt2;
        H.yy += mdt2;
        H.zz += mdt2;
        V3f dx;
        path = blockSolve(H, f, dx);
        if (path != kSkipped) a.pos[vid] = make_float4(xo.x + dx.x, xo.y + dx.y, xo.z + dx.z, 0.f);
    }
    const unsigned long long fb = __ballot(path == kSolvedFp64), sk = __ballot(path == kSkipped);
    if (lane64 == 0) {
        if (fb) atomicAdd(&a.counters[0], static_cast<unsigned long long>(__popcll(fb)));
        if (sk) atomicAdd(&a.counters[1], static_cast<unsigned long long>(__popcll(sk)));
    }
}

// K2: one colour sweep. A group of G lanes of one wavefront cooperates on one vertex; each
// wavefront holds floor(64 / G) groups (the remaining lanes idle when G does not divide 64). Lane l
// of a group takes incidences l, l+G, ... of each incidence class (StVK triangles, hinge-edge and
// hinge-wing slots of bending elements, SNH tetrahedra, springs, anchors and, with EXT, the further
// energies' classes), then the force and the six Hessian entries are summed over the group in a fixed
// order with cross-lane shuffles and the group's first lane solves the 3x3 block and writes the position.
// LDS = true: the workgroup first copies the element-record rows of its contiguous vertex range (one
// contiguous span per record array) into LDS with coalesced loads, then every lane reads its records
// from LDS (the further energies' records excepted); positions are gathered from global memory as in the
// direct variant.
// Blocks are renumbered so that each XCD works on one contiguous range of the colour
// (xcdContiguousBlock), keeping neighbouring positions in that XCD's L2.

// The sum of v over the G lanes of a group, valid in the group's first lane (the other lanes get partial sums
// that reach into the next group; nobody reads them). acc_k at a lane is the sum of the 2^k lanes from it, and
// step `off` adds to the first lane the sum of its lanes off .. min(2 off, G) - 1, which is acc_j at lane
// off with 2^j = min(off, G - off) (a power of two for every G here: 1, 2, 3, 4, 6, 8, 16). These are the
// additions, in the same order, of a butterfly whose steps are masked to the group's lanes, without the masks.
template <int G, class T>
__device__ __forceinline__ T groupSum(T v, int32_t lane64)
{
    static_assert(G == 1 || G == 2 || G == 3 || G == 4 || G == 6 || G == 8 || G == 16, "lanes per vertex");
    T acc[5];
    acc[0] = v;
    int k = 0;
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
        const int len = off < G - off ? off : G - off;
        const int j = len == 1 ? 0 : len == 2 ? 1 : len == 4 ? 2 : 3;
        v = v + __shfl(acc[j], lane64 + off, 64);
        acc[++k] = v;
    }
    return v;
}

// Copies count records starting at src[first] into LDS at *cursor (16-byte aligned), with the whole
// workgroup, and advances the cursor. T is 8 or 16 bytes.
template <class T>
__device__ __forceinline__ const T *stageRows(const T *src, int64_t first, int32_t count, int32_t threads,
                                              unsigned char *&cursor)
{
    T *dst = reinterpret_cast<T *>(cursor);
    for (int32_t i = threadIdx.x; i < count; i += threads) dst[i] = src[first + i];
    cursor += (static_cast<size_t>(count) * sizeof(T) + 15) & ~size_t(15);
    return dst;
}

// LDS bytes the staged variant needs for one workgroup (host side mirrors the kernel's layout).
__host__ __device__ inline size_t solveLdsBytes(const SolveArgs &a, int G, int blockSize)
{
    const int64_t nvb = static_cast<int64_t>(blockSize / 64) * (64 / G);
    auto seg = [&](int32_t w, size_t bytes) { return (static_cast<size_t>(nvb * w) * bytes + 15) & ~size_t(15); };
    return seg(a.triW, 8) + 2 * seg(a.triW, 16) + seg(a.hinEW, 16) + seg(a.hinEW, 8) + seg(a.hinWW, 16) +
           seg(a.hinWW, 8) + 4 * seg(a.tetW, 16) + seg(a.sprW, 16);
}

// LDS bytes of the neighbourhood-staged variant: per vertex group of the workgroup (and one spare per
// wavefront for its idle lanes when G does not divide 64) a slab of ceil(16 / G) G float4 (the row's at most
// 16 entries, rounded up to whole lane strides).
__host__ __device__ inline size_t nbSlabBytes(int G, int blockSize)
{
    const size_t slab = static_cast<size_t>((16 + G - 1) / G) * G;
    const size_t slabs = static_cast<size_t>(64 / G + (64 % G ? 1 : 0));
    return static_cast<size_t>(blockSize / 64) * slabs * slab * 16;
}
__host__ __device__ inline size_t solveNbLdsBytes(const SolveArgs &, int G, int blockSize)
{
    return nbSlabBytes(G, blockSize);
}

// The further energies' closed-form slot blocks. Each takes the element's positions in stencil order and FP32
// coefficients that already hold what depends on the slot (the gamma row of a membrane or tetrahedron slot, the
// spring's sign, the quadratic bending's own coefficient, which end or wing of a hinge), so the lanes of a
// wavefront run one instruction stream whatever slots they hold. The solve reads these coefficients from its
// per-incidence records; the per-element hook (extSlot) derives them from the FP64 parameters and the slot.

// Baraff-Witkin bending (baraffWitkinBendingEnergy.cpp:9-36), the gradient and 3x3 block of one hinge slot in
// closed form. With e0 = x1 - x0, e3 = x2 - x1, e4 = x3 - x1 and s0 = 1/|e0| held constant:
//   X = n1 . n2 = -[(e0.e0)(e3.e4) - (e0.e4)(e0.e3)]          (Binet-Cauchy identity)
//   Y = s0 (n1 x n2) . e0 = -s0 (e0.e0) D,  D = e0 . (e3 x e4)  (n1 x n2 = -D e0)
// both polynomials in (e0, e3, e4); theta = atan2(Y, X), E = 1/2 k (theta - theta0)^2. A slot's block sums the
// second derivatives of X and Y over the edge vectors it moves (x0: -e0; x1: e0 - e3 - e4; x2: e3; x3: e4); in the
// symmetric sum the skew parts cancel and the wing slots (x2, x3), on which X and Y are linear, get none.
// WING false: slot 0 (second false) or 1 (second true); WING true: slot 2 or 3. Slot 1's terms are slot 0's with
// the opposite sign plus the terms of the edges e3, e4 that x1 also moves (weight m = 1, 0 for slot 0).
// T: float (the solve, with the block) or double (the residual, HESS false: the gradient only).
template <bool WING, class T = float, bool HESS = true>
__device__ __forceinline__ void bendingBwRec(const V3<T> *x, T theta0, T k, bool second, V3<T> &g, Sym3<T> &H)
{
    const V3<T> e0 = x[1] - x[0], e3 = x[2] - x[1], e4 = x[3] - x[1];
    const T q = dot(e0, e0), s0 = T(1) / sqrt(q);
    const T d34 = dot(e3, e4), d03 = dot(e0, e3), d04 = dot(e0, e4);
    const V3<T> g0 = cross(e3, e4), g3 = cross(e4, e0), g4 = cross(e0, e3);
    const T D = dot(e0, g0);
    const T X = -(q * d34 - d04 * d03), Y = -s0 * q * D;
    V3<T> gX, gY;
    Sym3<T> HX = zeroSym<T>(), HY = zeroSym<T>();
    if constexpr (!WING) {
        const T m = second ? T(1) : T(0), sg = second ? T(1) : T(-1);
        const V3<T> X0 = (T(-2) * d34) * e0 + d03 * e4 + d04 * e3;
        const V3<T> Y0 = (-s0) * ((T(2) * D) * e0 + q * g0);
        const V3<T> g34 = g3 + g4;
        gX = sg * X0 - m * ((d04 * e0 - q * e4) + (d03 * e0 - q * e3));
        gY = sg * Y0 + (m * s0 * q) * g34;
        if constexpr (HESS) {
            addSymOuter(HX, T(1), e4, e3);
            addDiag(HX, T(-2) * (d34 + m * (d04 + d03 + q)));
            addSymOuter(HX, m, e0, e3 + e4);
            addOuter(HX, T(2) * m, e0);
            addDiag(HY, T(-2) * s0 * D);
            addSymOuter(HY, T(-2) * s0, e0, g0 - m * g34);
        }
    } else {
        const V3<T> ew = second ? e3 : e4, gw = second ? g4 : g3;
        gX = (second ? d03 : d04) * e0 - q * ew;
        gY = (-s0 * q) * gw;
    }
    const T ir2 = T(1) / (X * X + Y * Y);
    const V3<T> gt = ir2 * (X * gY - Y * gX);  // grad theta
    const T kd = k * (hingeAtan2(Y, X) - theta0);
    g = kd * gt;
    if constexpr (HESS) {
        // k grad theta grad theta^T + k (theta - theta0) [(X HY - Y HX) / r2 - sym(grad theta, u) / r2]
        const V3<T> u = X * gX + Y * gY;
        const T a = kd * ir2;
        H = {a * (X * HY.xx - Y * HX.xx), a * (X * HY.xy - Y * HX.xy), a * (X * HY.xz - Y * HX.xz),
             a * (X * HY.yy - Y * HX.yy), a * (X * HY.yz - Y * HX.yz), a * (X * HY.zz - Y * HX.zz)};
        addSymOuter(H, -a, gt, u);
        addOuter(H, k, gt);
    }
}

// four coefficients of a record: float4 in FP32, double4 in FP64
template <class T>
using Coef4 = std::conditional_t<std::is_same_v<T, float>, float4, double4>;

// Baraff-Witkin membrane (membraneBaraffWitkinEnergy.cpp:9-68), one triangle slot in closed form. F = [f0 f1] =
// [x1 - x0, x2 - x0] Dinv, C = (|f0| - 1, |f1| - 1, f0 . f1), E = 1/2 sum k_i C_i^2 (k_i the effective stiffness
// diagonal, area included); slot a moves f0 by gamma_a0 (ga) and f1 by gamma_a1 (gb) (the rows of Dinv, gamma_0
// their negated sum), so its block is gamma_a0^2 (A0 + k2 f1 f1^T) + gamma_a1^2 (A1 + k2 f0 f0^T)
// + gamma_a0 gamma_a1 k2 (2 C2 I + f0 f1^T + f1 f0^T) with Ai = k_i (fi fi^T / |fi|^3 + C_i / |fi| I).
template <class T = float, bool HESS = true>
__device__ __forceinline__ void membraneBwRec(const V3<T> *x, Coef4<T> d, T k0, T k1, T k2, T ga, T gb, V3<T> &g,
                                              Sym3<T> &H)
{
    const V3<T> u = x[1] - x[0], v = x[2] - x[0];
    const V3<T> f0 = d.x * u + d.z * v, f1 = d.y * u + d.w * v;
    const T n0 = sqrt(dot(f0, f0)), n1 = sqrt(dot(f1, f1));
    const T c0 = n0 - T(1), c1 = n1 - T(1), c2 = dot(f0, f1);
    const T i0 = T(1) / n0, i1 = T(1) / n1;
    const V3<T> dEdf0 = (k0 * c0 * i0) * f0 + (k2 * c2) * f1, dEdf1 = (k1 * c1 * i1) * f1 + (k2 * c2) * f0;
    g = ga * dEdf0 + gb * dEdf1;
    if constexpr (HESS) {
        H = zeroSym<T>();
        addDiag(H, ga * ga * k0 * c0 * i0 + gb * gb * k1 * c1 * i1 + T(2) * ga * gb * k2 * c2);
        addOuter(H, ga * ga * k0 * i0 * i0 * i0 + gb * gb * k2, f0);
        addOuter(H, gb * gb * k1 * i1 * i1 * i1 + ga * ga * k2, f1);
        addSymOuter(H, ga * gb * k2, f0, f1);
    }
}

// k (f f^T / |f|^3 + (|f| - 1) / |f| I) and k (|f| - 1) f / |f|, scaled by c^2 and c: one unit-stretch term
// 1/2 k (|f| - 1)^2 of a membrane whose f moves by c times the slot's displacement.
template <class T = float, bool HESS = true>
__device__ __forceinline__ void unitStretchSlot(V3<T> f, T k, T c, V3<T> &g, Sym3<T> &H)
{
    const T n = sqrt(dot(f, f)), in = T(1) / n;
    g = g + (c * k * (n - T(1)) * in) * f;
    if constexpr (HESS) {
        addDiag(H, c * c * k * (n - T(1)) * in);
        addOuter(H, c * c * k * in * in * in, f);
    }
}

// Choi membrane (membraneChoiEnergy.cpp:9-28): four unit stretches, of f0, f1, gp = (f0 + f1)/sqrt 2 and
// gm = (f0 - f1)/sqrt 2; the slot moves them by ga, gb, (ga +- gb)/sqrt 2.
template <class T = float, bool HESS = true>
__device__ __forceinline__ void membraneChoiRec(const V3<T> *x, Coef4<T> d, Coef4<T> k, T ga, T gb, V3<T> &g,
                                                Sym3<T> &H)
{
    const V3<T> u = x[1] - x[0], v = x[2] - x[0];
    const V3<T> f0 = d.x * u + d.z * v, f1 = d.y * u + d.w * v;
    const T r = T(0.70710678118654752440);
    g = {T(0), T(0), T(0)};
    H = zeroSym<T>();
    unitStretchSlot<T, HESS>(f0, k.x, ga, g, H);
    unitStretchSlot<T, HESS>(f1, k.y, gb, g, H);
    unitStretchSlot<T, HESS>(r * (f0 + f1), k.z, r * (ga + gb), g, H);
    unitStretchSlot<T, HESS>(r * (f0 - f1), k.w, r * (ga - gb), g, H);
}

// Quadratic bending (quadraticBendingEnergy.cpp:9-30): E = 1/2 keff |C|^2 with C = sum c_s x_s, so slot a has
// gradient keff c_a C and block keff c_a^2 I (cs = c_a).
template <class T = float, bool HESS = true>
__device__ __forceinline__ void bendingQuadraticRec(const V3<T> *x, Coef4<T> c, T keff, T cs, V3<T> &g, Sym3<T> &H)
{
    const V3<T> C = c.x * x[0] + c.y * x[1] + c.z * x[2] + c.w * x[3];
    const T kc = keff * cs;
    g = kc * C;
    if constexpr (HESS) {
        H = zeroSym<T>();
        addDiag(H, kc * cs);
    }
}

// Squared continuum spring (springSquaredContinuumEnergy.cpp:9-19): C = |d|^2 / L0^2 - 1, E = 1/2 Keff C^2,
// d = x0 - x1: both slots have the block Keff ((2/L0^2)^2 d d^T + C (2/L0^2) I); the gradient has sign sg
// (+1 for slot 0, -1 for slot 1).
template <class T = float, bool HESS = true>
__device__ __forceinline__ void springSquaredContinuumRec(const V3<T> *x, T L0, T k, T sg, V3<T> &g, Sym3<T> &H)
{
    const V3<T> d = x[0] - x[1];
    const T L2 = L0 * L0;
    const T C = dot(d, d) / L2 - T(1), s = T(2) / L2;
    g = (sg * k * C * s) * d;
    if constexpr (HESS) {
        H = zeroSym<T>();
        addDiag(H, k * C * s);
        addOuter(H, k * s * s, d);
    }
}

// Tetrahedra: F = [f0 f1 f2] = [x1 - x0, x2 - x0, x3 - x0] Dinv (Dinv row-major in dv[0..8]) and the slot moves
// column c by gm[c] (row a - 1 of Dinv, gamma_0 the negated column sums), so it moves F by u gamma^T and
// w = F gamma is F's rate.
template <class T = float>
__device__ __forceinline__ void tetFrameRec(const V3<T> *x, const T *dv, V3<T> *f)
{
    const V3<T> u = x[1] - x[0], v = x[2] - x[0], w = x[3] - x[0];
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] = dv[c] * u + dv[3 + c] * v + dv[6 + c] * w;
}

// StVK tetrahedron (stvkEnergy.cpp:41-63, energy form): P = F S, S = lambda tr(E) I + 2 mu E; the slot block is
// (gamma^T S gamma) I + (lambda + mu) w w^T + mu |gamma|^2 F F^T.
template <class T = float, bool HESS = true>
__device__ __forceinline__ void tetStvkRec(const V3<T> *x, const T *dv, const T *gm, T lam, T mu, V3<T> &g, Sym3<T> &H)
{
    V3<T> f[3];
    tetFrameRec<T>(x, dv, f);
    T S[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) S[i][j] = mu * (dot(f[i], f[j]) - (i == j ? T(1) : T(0)));
    const T lt = T(0.5) * lam * (dot(f[0], f[0]) + dot(f[1], f[1]) + dot(f[2], f[2]) - T(3));
    S[0][0] += lt;
    S[1][1] += lt;
    S[2][2] += lt;
    T Sg[3], gSg = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        Sg[i] = S[i][0] * gm[0] + S[i][1] * gm[1] + S[i][2] * gm[2];
        gSg += gm[i] * Sg[i];
    }
    g = Sg[0] * f[0] + Sg[1] * f[1] + Sg[2] * f[2];
    if constexpr (HESS) {
        const V3<T> w = gm[0] * f[0] + gm[1] * f[1] + gm[2] * f[2];
        const T g2 = gm[0] * gm[0] + gm[1] * gm[1] + gm[2] * gm[2];
        H = zeroSym<T>();
        addDiag(H, gSg);
        addOuter(H, lam + mu, w);
#pragma unroll
        for (int c = 0; c < 3; ++c) addOuter(H, mu * g2, f[c]);
    }
}

// Stable squared Neo-Hookean (stableSquaredNeoHookeanEnergy.cpp:9-31): E = 1/2 lambda (J - 1)^2
// + 1/2 mu (|F|^2 - 3)^2; J is affine along the slot's rank-one update, so the block is lambda c c^T
// + mu (4 w w^T + 2 (|F|^2 - 3) |gamma|^2 I) with c = cof(F) gamma.
template <class T = float, bool HESS = true>
__device__ __forceinline__ void tetSsnhRec(const V3<T> *x, const T *dv, const T *gm, T lam, T mu, V3<T> &g, Sym3<T> &H)
{
    V3<T> f[3];
    tetFrameRec<T>(x, dv, f);
    const V3<T> k0 = cross(f[1], f[2]), k1 = cross(f[2], f[0]), k2 = cross(f[0], f[1]);
    const T J = dot(f[0], k0), s = dot(f[0], f[0]) + dot(f[1], f[1]) + dot(f[2], f[2]) - T(3);
    const V3<T> c = gm[0] * k0 + gm[1] * k1 + gm[2] * k2;
    const V3<T> w = gm[0] * f[0] + gm[1] * f[1] + gm[2] * f[2];
    g = (lam * (J - T(1))) * c + (T(2) * mu * s) * w;
    if constexpr (HESS) {
        const T g2 = gm[0] * gm[0] + gm[1] * gm[1] + gm[2] * gm[2];
        H = zeroSym<T>();
        addOuter(H, lam, c);
        addOuter(H, T(4) * mu, w);
        addDiag(H, T(2) * mu * s * g2);
    }
}

// The slot-dependent coefficients from an element's FP64 parameters (the same values the host writes into the
// solve's records): membrane gamma rows, tetrahedron gamma rows (the negated column sums summed in FP64).
template <int LOCAL>
__device__ __forceinline__ void triGammaOf(float4 d, float &ga, float &gb)
{
    ga = LOCAL == 0 ? -(d.x + d.z) : LOCAL == 1 ? d.x : d.z;
    gb = LOCAL == 0 ? -(d.y + d.w) : LOCAL == 1 ? d.y : d.w;
}
template <int LOCAL>
__device__ __forceinline__ void tetRecOf(const double *p, float *dv, float *gm)
{
#pragma unroll
    for (int i = 0; i < 9; ++i) dv[i] = float(p[i]);
#pragma unroll
    for (int c = 0; c < 3; ++c)
        gm[c] = LOCAL == 0 ? -float(p[c] + p[3 + c] + p[6 + c]) : float(p[3 * (LOCAL - 1) + c]);
}

// The closed-form slot block of further type TYPE (7 ... 13) from the element's FP64 parameters.
template <int TYPE, int LOCAL>
__device__ __forceinline__ void extRecOf(const V3f *xs, const double *p, V3f &g, S3f &H)
{
    if constexpr (TYPE == 11) {
        bendingBwRec<(LOCAL >= 2)>(xs, float(p[0]), float(p[1]), LOCAL == 1 || LOCAL == 3, g, H);
    } else if constexpr (TYPE == 8 || TYPE == 9) {
        const float4 d = make_float4(float(p[0]), float(p[1]), float(p[2]), float(p[3]));
        float ga, gb;
        triGammaOf<LOCAL>(d, ga, gb);
        if constexpr (TYPE == 8)
            membraneBwRec(xs, d, float(p[4]), float(p[5]), float(p[6]), ga, gb, g, H);
        else
            membraneChoiRec(xs, d, make_float4(float(p[4]), float(p[5]), float(p[6]), float(p[7])), ga, gb, g, H);
    } else if constexpr (TYPE == 10) {
        bendingQuadraticRec(xs, make_float4(float(p[0]), float(p[1]), float(p[2]), float(p[3])), float(p[4]),
                            float(p[LOCAL]), g, H);
    } else if constexpr (TYPE == 7) {
        springSquaredContinuumRec(xs, float(p[0]), float(p[1]), LOCAL == 0 ? 1.f : -1.f, g, H);
    } else {
        float dv[9], gm[3];
        tetRecOf<LOCAL>(p, dv, gm);
        if constexpr (TYPE == 12)
            tetStvkRec(xs, dv, gm, float(p[9]), float(p[10]), g, H);
        else
            tetSsnhRec(xs, dv, gm, float(p[9]), float(p[10]), g, H);
    }
}

// A further energy's gradient w.r.t. its stencil slot LOCAL and, with HESS, the slot's 3x3 Hessian block. In
// FP32 with the type known at compile time (7 ... 13) the block is the closed form above; otherwise the
// closed-form full gradient (implicit_element.h) at the element's positions is taken in one forward-mode pass
// whose three tangents are the slot's coordinates (the block's columns, symmetrised). The slot is a template
// parameter so the slot's gradient is picked at compile time (a run-time pick puts the per-slot gradients in
// scratch memory) and the other slots' outputs are dropped. The per-element derivative hook (extElementKernel)
// instantiates it in FP32 with the block: the closed forms for types 7 ... 13 and the forward-mode pass for the
// continuum spring (type 6, whose solve path is the spring class). TYPE = 0: the energy type is read at run time
// (`type`); otherwise it is TYPE.
template <int NS, int LOCAL, class T, bool HESS, int TYPE = 0>
__device__ __forceinline__ void extSlot(int type, const float4 *__restrict__ pos, int4 v,
                                        const double *__restrict__ p, V3<T> &g, Sym3<T> &H)
{
    const int32_t vv[4] = {v.x, v.y, v.z, v.w};
    V3<T> xs[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const float4 q = pos[vv[s]];
        xs[s] = {T(q.x), T(q.y), T(q.z)};
    }
    if constexpr (!HESS) {
        V3<T> gd[NS];
        if constexpr (TYPE == 0)
            fullGrad<NS, T>(type, xs, p, gd);
        else
            fullGradOf<TYPE, T>(xs, p, gd);
        g = gd[LOCAL];
    } else if constexpr (TYPE >= 7 && TYPE <= 13 && std::is_same_v<T, float>) {
        extRecOf<TYPE, LOCAL>(xs, p, g, H);
    } else {
        using D3 = ad::Dual3T<T>;
        V3<D3> x[NS], gd[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const T o = s == LOCAL ? T(1) : T(0);
            x[s] = {D3(xs[s].x, o, T(0), T(0)), D3(xs[s].y, T(0), o, T(0)), D3(xs[s].z, T(0), T(0), o)};
        }
        if constexpr (TYPE == 0)
            fullGrad<NS, D3>(type, x, p, gd);
        else
            fullGradOf<TYPE, D3>(x, p, gd);
        const V3<D3> &gl = gd[LOCAL];
        g = {gl.x.v, gl.y.v, gl.z.v};
        const T h = T(0.5);
        H = {gl.x.d[0], h * (gl.x.d[1] + gl.y.d[0]), h * (gl.x.d[2] + gl.z.d[0]), gl.y.d[1],
             h * (gl.y.d[2] + gl.z.d[1]), gl.z.d[2]};
    }
}
template <int NS, class T, bool HESS, int TYPE = 0>
__device__ __forceinline__ void extSlotAny(int type, int local, const float4 *__restrict__ pos, int4 v,
                                           const double *__restrict__ p, V3<T> &g, Sym3<T> &H)
{
    if constexpr (NS == 2) {
        if (local == 0)
            extSlot<2, 0, T, HESS, TYPE>(type, pos, v, p, g, H);
        else
            extSlot<2, 1, T, HESS, TYPE>(type, pos, v, p, g, H);
    } else if constexpr (NS == 3) {
        if (local == 0)
            extSlot<3, 0, T, HESS, TYPE>(type, pos, v, p, g, H);
        else if (local == 1)
            extSlot<3, 1, T, HESS, TYPE>(type, pos, v, p, g, H);
        else
            extSlot<3, 2, T, HESS, TYPE>(type, pos, v, p, g, H);
    } else {
        switch (local) {
        case 0: extSlot<4, 0, T, HESS, TYPE>(type, pos, v, p, g, H); break;
        case 1: extSlot<4, 1, T, HESS, TYPE>(type, pos, v, p, g, H); break;
        case 2: extSlot<4, 2, T, HESS, TYPE>(type, pos, v, p, g, H); break;
        default: extSlot<4, 3, T, HESS, TYPE>(type, pos, v, p, g, H);
        }
    }
}

// One further-energy incidence class of the vertex's records (fixed-width row, G lanes striding over it): the
// record's coefficients already hold what depends on the vertex's own slot, so every lane runs the class's one
// closed form.
template <int C, int G>
__device__ __forceinline__ void extClassRun(const SolveArgs &a, int32_t vid, int32_t lig, V3f &g, S3f &H)
{
    constexpr int NS = xclsStencil(C), NP = xclsPlanes(C);
    const int32_t w = a.xrW[C];
    for (int32_t j = lig; j < w; j += G) {
        const int64_t r = static_cast<int64_t>(vid) * w + j;
        const int4 v = a.xrV[C][r];
        float4 q[NP];
#pragma unroll
        for (int i = 0; i < NP; ++i) q[i] = a.xrP[C][i * a.xrPlane[C] + r];
        const bool ok = v.x >= 0;
        const int32_t vv[4] = {v.x, v.y, v.z, v.w};
        V3f xs[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) xs[s] = xyz(a.pos[ok ? vv[s] : vid]);
        V3f ge;
        S3f He;
        if constexpr (C == 0) {
            springSquaredContinuumRec(xs, q[0].x, q[0].y, q[0].z, ge, He);
        } else if constexpr (C == 1) {
            const float4 d = q[0];
            const float sl = q[1].w;
            const float ga = sl == 0.f ? -(d.x + d.z) : sl == 1.f ? d.x : d.z;
            const float gb = sl == 0.f ? -(d.y + d.w) : sl == 1.f ? d.y : d.w;
            membraneBwRec(xs, d, q[1].x, q[1].y, q[1].z, ga, gb, ge, He);
        } else if constexpr (C == 2) {
            membraneChoiRec(xs, q[0], q[1], q[2].x, q[2].y, ge, He);
        } else if constexpr (C == 3) {
            bendingQuadraticRec(xs, q[0], q[1].x, q[1].y, ge, He);
        } else if constexpr (C == 4 || C == 5) {
            bendingBwRec<C == 5>(xs, q[0].x, q[0].y, q[0].z != 0.f, ge, He);
        } else {
            const float dv[9] = {q[0].x, q[0].y, q[0].z, q[0].w, q[1].x, q[1].y, q[1].z, q[1].w, q[2].x};
            const float gm[3] = {q[2].y, q[2].z, q[2].w};
            if constexpr (C == 6)
                tetStvkRec(xs, dv, gm, q[3].x, q[3].y, ge, He);
            else
                tetSsnhRec(xs, dv, gm, q[3].x, q[3].y, ge, He);
        }
        accIf(ok, g, H, ge, He);
    }
}

template <int G, bool LDS, bool EXT>
__global__ void __launch_bounds__(256) solveColorKernel(SolveArgs a)
{
    constexpr int GPW = 64 / G;  // groups per wavefront
    extern __shared__ float4 ldsRows[];
    const int32_t lane64 = threadIdx.x & 63;
    const int32_t blk = xcdContiguousBlock(blockIdx.x, a.gridBlocks);
    const int32_t wave = blk * a.wavesPerBlock + (threadIdx.x >> 6);
    const int32_t grp = lane64 / G;
    const int32_t lig = lane64 - grp * G;
    const int32_t vid = a.begin + wave * GPW + grp;
    const bool active = grp < GPW && vid < a.end;
    int64_t rowBase = 0;  // record row of vertex vid is vid - rowBase
    if constexpr (LDS) {
        const int32_t v0 = a.begin + blk * a.wavesPerBlock * GPW;
        const int32_t nvb = max(0, min(a.wavesPerBlock * GPW, a.end - v0));
        const int32_t nt = a.wavesPerBlock * 64;
        unsigned char *cur = reinterpret_cast<unsigned char *>(ldsRows);
        a.triO = stageRows(a.triO, int64_t(v0) * a.triW, nvb * a.triW, nt, cur);
        a.triG = stageRows(a.triG, int64_t(v0) * a.triW, nvb * a.triW, nt, cur);
        a.triM = stageRows(a.triM, int64_t(v0) * a.triW, nvb * a.triW, nt, cur);
        a.hinEV = stageRows(a.hinEV, int64_t(v0) * a.hinEW, nvb * a.hinEW, nt, cur);
        a.hinEP = stageRows(a.hinEP, int64_t(v0) * a.hinEW, nvb * a.hinEW, nt, cur);
        a.hinWV = stageRows(a.hinWV, int64_t(v0) * a.hinWW, nvb * a.hinWW, nt, cur);
        a.hinWP = stageRows(a.hinWP, int64_t(v0) * a.hinWW, nvb * a.hinWW, nt, cur);
        a.tetO = stageRows(a.tetO, int64_t(v0) * a.tetW, nvb * a.tetW, nt, cur);
        a.tetGa = stageRows(a.tetGa, int64_t(v0) * a.tetW, nvb * a.tetW, nt, cur);
        a.tetGb = stageRows(a.tetGb, int64_t(v0) * a.tetW, nvb * a.tetW, nt, cur);
        a.tetGc = stageRows(a.tetGc, int64_t(v0) * a.tetW, nvb * a.tetW, nt, cur);
        a.spr = stageRows(a.spr, int64_t(v0) * a.sprW, nvb * a.sprW, nt, cur);
        rowBase = v0;
        __syncthreads();
    }
    // the arguments of the first loads, taken here (an empty asm that reads them), so the compiler issues all
    // their scalar loads in the prologue's one round instead of a second round inside the branch below
    asm volatile("" ::"s"(a.xt), "s"(a.ancPtr), "s"(a.triO), "s"(a.triG), "s"(a.triM), "s"(a.triW));
    V3f g = {0.f, 0.f, 0.f};
    S3f H = zeroSym<float>();
    V3f xo = {0.f, 0.f, 0.f};
    float4 xt4 = make_float4(0.f, 0.f, 0.f, 0.f);
    if (active) {
        const float4 xo4 = a.pos[vid];
        xt4 = a.xt[vid];
        int32_t ab = 0, ae = 0;
        if (a.ancPtr) {
            ab = a.ancPtr[vid];
            ae = a.ancPtr[vid + 1];
        }
        xo = xyz(xo4);
        const int64_t row = vid - rowBase;
        for (int32_t j = lig; j < a.triW; j += G) {
            const int64_t r = row * a.triW + j;
            const int2 o = a.triO[r];
            const float4 gg = a.triG[r];
            const float4 kk = a.triM[r];
            const bool ok = o.x >= 0;
            accElement<!EXT>(ok, g, H, [&](V3f &ge, S3f &He) {
                stvkVertex<float>(xo, xyz(a.pos[ok ? o.x : vid]), xyz(a.pos[ok ? o.y : vid]), gg.x, gg.y, gg.z, gg.w,
                                  kk.x, kk.y, kk.z, ge, He);
            });
        }
        for (int32_t j = lig; j < a.hinEW; j += G) {
            const int64_t r = row * a.hinEW + j;
            const int4 v = a.hinEV[r];
            const float2 p = a.hinEP[r];
            const bool ok = v.x >= 0;
            accElement<!EXT>(ok, g, H, [&](V3f &ge, S3f &He) {
                // the vertex is edge end x0 or x1: its own position is xo, only the other end is gathered
                const bool second = v.y == vid;
                const V3f pe = xyz(a.pos[ok ? (second ? v.x : v.y) : vid]);
                bendingEdge<float>(second ? pe : xo, second ? xo : pe, xyz(a.pos[ok ? v.z : vid]),
                                   xyz(a.pos[ok ? v.w : vid]), second, p.x, p.y, ge, He);
            });
        }
        for (int32_t j = lig; j < a.hinWW; j += G) {
            const int64_t r = row * a.hinWW + j;
            const int4 v = a.hinWV[r];
            const float2 p = a.hinWP[r];
            const bool ok = v.x >= 0;
            accElement<!EXT>(ok, g, H, [&](V3f &ge, S3f &He) {
                // the vertex is wing x2 or x3: its own position is xo, only the other wing is gathered
                const bool second = v.w == vid;
                const V3f pw = xyz(a.pos[ok ? (second ? v.z : v.w) : vid]);
                bendingWing<float>(xyz(a.pos[ok ? v.x : vid]), xyz(a.pos[ok ? v.y : vid]), second ? pw : xo,
                                   second ? xo : pw, second, p.x, p.y, ge, He);
            });
        }
        for (int32_t j = lig; j < a.tetW; j += G) {
            const int64_t r = row * a.tetW + j;
            const int4 o = a.tetO[r];
            const float4 ga = a.tetGa[r], gb = a.tetGb[r], gc = a.tetGc[r];
            const bool ok = o.x >= 0;
            accElement<!EXT>(ok, g, H, [&](V3f &ge, S3f &He) {
                snhVertex<float>(xo, xyz(a.pos[ok ? o.x : vid]), xyz(a.pos[ok ? o.y : vid]),
                                 xyz(a.pos[ok ? o.z : vid]), xyz(ga), xyz(gb), xyz(gc), ga.w, gb.w, gc.w, ge, He);
            });
        }
        for (int32_t j = lig; j < a.sprW; j += G) {
            const float4 s = a.spr[row * a.sprW + j];
            const int32_t other = __float_as_int(s.x);
            const bool ok = other >= 0;
            accElement<!EXT>(ok, g, H, [&](V3f &ge, S3f &He) {
                springVertex<float>(xo, xyz(a.pos[ok ? other : vid]), s.y, s.z, ge, He);
            });
        }
        for (int32_t k = ab + lig; k < ae; k += G) {
            const float4 s = a.anc[k];
            V3f ge;
            S3f He;
            anchorVertex<float>(xo, xyz(s), s.w, ge, He);
            acc(g, H, ge, He);
        }
        if constexpr (EXT) {
            extClassRun<0, G>(a, vid, lig, g, H);
            extClassRun<1, G>(a, vid, lig, g, H);
            extClassRun<2, G>(a, vid, lig, g, H);
            extClassRun<3, G>(a, vid, lig, g, H);
            extClassRun<4, G>(a, vid, lig, g, H);
            extClassRun<5, G>(a, vid, lig, g, H);
            extClassRun<6, G>(a, vid, lig, g, H);
            extClassRun<7, G>(a, vid, lig, g, H);
        }
    }
    finishVertex<G>(a, active, vid, lane64, lig, xo, xt4, g, H);
}

// K2, neighbourhood-staged variant (lds_staging 2). A vertex needs the positions of few distinct vertices (on
// the cloth grid 12: its one-ring and the far wing of each hinge it is a wing of) but reads them once per
// incident element (54 times). Here the vertex's group first gathers its own position and those of its
// distinct neighbours (entry e of its neighbour row, lanes striding by G) into an LDS slab, and the element
// records name their vertices by 5-bit index into the slab (0: the vertex itself), so each position is read
// from memory once per vertex and the texture data path, busy most of the launch in solveColorKernel, carries
// a quarter of the position loads. The first record of each class does not depend on positions and is loaded
// before the gather, so a vertex waits for two rounds of memory loads. The element functions see the same
// values in the same order as in solveColorKernel: the result is bitwise the same.
constexpr int kNbMax = 16;  // neighbour-row entries per vertex at most
// incidence classes compiled into a neighbourhood-staged kernel (the launcher picks the instantiation that
// holds the context's classes, so a cloth kernel keeps no registers for tetrahedra)
constexpr int kClsTri = 1, kClsHinge = 2, kClsTet = 4, kClsSpring = 8, kClsAll = 15;

template <int G, int CLS>
__global__ void __launch_bounds__(256) solveColorNbKernel(SolveArgs a)
{
    constexpr int GPW = 64 / G;
    constexpr int R = (kNbMax + G - 1) / G;  // slab entries per lane at most
    extern __shared__ float4 nbSlab[];
    const int32_t lane64 = threadIdx.x & 63;
    const int32_t blk = xcdContiguousBlock(blockIdx.x, a.gridBlocks);
    const int32_t wib = threadIdx.x >> 6;
    const int32_t wave = blk * a.wavesPerBlock + wib;
    const int32_t grp = lane64 / G;
    const int32_t lig = lane64 - grp * G;
    const int32_t vid = a.begin + wave * GPW + grp;
    const bool active = grp < GPW && vid < a.end;
    const int32_t vsafe = active ? vid : a.begin;  // a slot inside the colour, for the loads of idle lanes
    // R G entries per group: every lane writes all R of its entries (those past the row are never read), so
    // the position loads are issued together instead of one per branch. When G does not divide 64 the idle
    // lanes past the last group (grp == GPW) write into one spare slab per wavefront.
    constexpr int SLABS = GPW + (64 % G ? 1 : 0);
    float4 *slab = nbSlab + static_cast<int64_t>(wib * SLABS + grp) * (R * G);
    asm volatile("" ::"s"(a.xt), "s"(a.nbRow), "s"(a.triG), "s"(a.triL), "s"(a.hinEL), "s"(a.hinWL), "s"(a.triW),
                 "s"(a.hinEW), "s"(a.hinWW), "s"(a.nbW));
    const int64_t row = vid;
    // records that do not depend on positions: the first of each class, and the neighbour row
    const bool t0 = (CLS & kClsTri) && active && lig < a.triW;
    const bool e0 = (CLS & kClsHinge) && active && lig < a.hinEW;
    const bool w0 = (CLS & kClsHinge) && active && lig < a.hinWW;
    const bool q0 = (CLS & kClsTet) && active && lig < a.tetW;
    const bool s0 = (CLS & kClsSpring) && active && lig < a.sprW;
    float4 tg = {}, tk = {}, ga = {}, gb = {}, gc = {}, sp = {};
    NbHinge he = {}, hw = {};
    int32_t tl = -1;
    if constexpr ((CLS & kClsTri) != 0) {
        if (t0) {
            tg = a.triG[row * a.triW + lig];
            tk = a.triL[row * a.triW + lig];
        }
    }
    if constexpr ((CLS & kClsHinge) != 0) {
        if (e0) he = a.hinEL[row * a.hinEW + lig];
        if (w0) hw = a.hinWL[row * a.hinWW + lig];
    }
    if constexpr ((CLS & kClsTet) != 0) {
        if (q0) {
            tl = a.tetL[row * a.tetW + lig];
            ga = a.tetGa[row * a.tetW + lig];
            gb = a.tetGb[row * a.tetW + lig];
            gc = a.tetGc[row * a.tetW + lig];
        }
    }
    if constexpr ((CLS & kClsSpring) != 0) {
        if (s0) sp = a.sprL[row * a.sprW + lig];
    }
    int32_t ids[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int32_t e = lig + k * G;
        ids[k] = (active && e < a.nbW) ? a.nbRow[row * a.nbW + e] : -1;
    }
    const float4 xt4 = a.xt[vsafe];
    int32_t ab = 0, ae = 0;
    if (active && a.ancPtr) {
        ab = a.ancPtr[vid];
        ae = a.ancPtr[vid + 1];
    }
    // the neighbourhood into LDS: every lane issues its loads (padding entries read the vertex's own
    // position), then writes them
    float px[R], py[R], pz[R];  // (an array of float4 here is not kept in registers)
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const float4 p = a.pos[ids[k] >= 0 ? ids[k] : vsafe];
        px[k] = p.x;
        py[k] = p.y;
        pz[k] = p.z;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) slab[lig + k * G] = make_float4(px[k], py[k], pz[k], 0.f);
    __syncthreads();
    V3f g = {0.f, 0.f, 0.f};
    S3f H = zeroSym<float>();
    V3f xo = {0.f, 0.f, 0.f};
    if (active) {
        xo = xyz(slab[0]);
        if constexpr ((CLS & kClsTri) != 0) {
            auto tri = [&](const float4 &gg, const float4 &kk) {
                const int32_t q = __float_as_int(kk.w);
                if (q >= 0) {
                    V3f ge;
                    S3f He;
                    stvkVertex<float>(xo, xyz(slab[q & 31]), xyz(slab[(q >> 5) & 31]), gg.x, gg.y, gg.z, gg.w, kk.x,
                                      kk.y, kk.z, ge, He);
                    acc(g, H, ge, He);
                }
            };
            if (t0) tri(tg, tk);
            for (int32_t j = lig + G; j < a.triW; j += G) tri(a.triG[row * a.triW + j], a.triL[row * a.triW + j]);
        }
        if constexpr ((CLS & kClsHinge) != 0) {
            auto hinE = [&](const NbHinge &r) {
                const int32_t q = r.slots;
                if (q >= 0) {
                    const int32_t l1 = (q >> 5) & 31;
                    V3f ge;
                    S3f He;
                    bendingEdge<float>(xyz(slab[q & 31]), xyz(slab[l1]), xyz(slab[(q >> 10) & 31]),
                                       xyz(slab[(q >> 15) & 31]), l1 == 0, r.theta0, r.keff, ge, He);
                    acc(g, H, ge, He);
                }
            };
            if (e0) hinE(he);
            for (int32_t j = lig + G; j < a.hinEW; j += G) hinE(a.hinEL[row * a.hinEW + j]);
            auto hinW = [&](const NbHinge &r) {
                const int32_t q = r.slots;
                if (q >= 0) {
                    const int32_t l3 = (q >> 15) & 31;
                    V3f ge;
                    S3f He;
                    bendingWing<float>(xyz(slab[q & 31]), xyz(slab[(q >> 5) & 31]), xyz(slab[(q >> 10) & 31]),
                                       xyz(slab[l3]), l3 == 0, r.theta0, r.keff, ge, He);
                    acc(g, H, ge, He);
                }
            };
            if (w0) hinW(hw);
            for (int32_t j = lig + G; j < a.hinWW; j += G) hinW(a.hinWL[row * a.hinWW + j]);
        }
        if constexpr ((CLS & kClsTet) != 0) {
            auto tet = [&](int32_t q, const float4 &ra, const float4 &rb, const float4 &rc) {
                if (q >= 0) {
                    V3f ge;
                    S3f He;
                    snhVertex<float>(xo, xyz(slab[q & 31]), xyz(slab[(q >> 5) & 31]), xyz(slab[(q >> 10) & 31]),
                                     xyz(ra), xyz(rb), xyz(rc), ra.w, rb.w, rc.w, ge, He);
                    acc(g, H, ge, He);
                }
            };
            if (q0) tet(tl, ga, gb, gc);
            for (int32_t j = lig + G; j < a.tetW; j += G) {
                const int64_t r = row * a.tetW + j;
                tet(a.tetL[r], a.tetGa[r], a.tetGb[r], a.tetGc[r]);
            }
        }
        if constexpr ((CLS & kClsSpring) != 0) {
            auto spring = [&](const float4 &r) {
                const int32_t q = __float_as_int(r.x);
                if (q >= 0) {
                    V3f ge;
                    S3f He;
                    springVertex<float>(xo, xyz(slab[q]), r.y, r.z, ge, He);
                    acc(g, H, ge, He);
                }
            };
            if (s0) spring(sp);
            for (int32_t j = lig + G; j < a.sprW; j += G) spring(a.sprL[row * a.sprW + j]);
        }
        for (int32_t k = ab + lig; k < ae; k += G) {
            const float4 r = a.anc[k];
            V3f ge;
            S3f He;
            anchorVertex<float>(xo, xyz(r), r.w, ge, He);
            acc(g, H, ge, He);
        }
    }
    finishVertex<G>(a, active, vid, lane64, lig, xo, xt4, g, H);
}

using V3d = V3<double>;
using S3d = Sym3<double>;

__device__ __forceinline__ V3d xyzd(float4 p) { return {double(p.x), double(p.y), double(p.z)}; }

// The FP64 StVK gradient at a triangle slot: p01 .. p67 the element's prepared parameters (rows of Dinv
// g1 = (p0, p1), g2 = (p2, p3), g0 = -(g1 + g2); kd, ko, ks), s0 / s2 whether the vertex is slot 0 / slot 2, xa, xb
// the other two slots' positions in stencil order.
__device__ __forceinline__ V3d triSlotGradD(double2 p01, double2 p23, double2 p45, double2 p67, bool s0, bool s2,
                                            const V3d &xo, const V3d &xa, const V3d &xb)
{
    const double g00 = -(p01.x + p23.x), g01 = -(p01.y + p23.y);
    const double ga0 = s0 ? p01.x : g00, ga1 = s0 ? p01.y : g01;
    const double gb0 = s2 ? p01.x : p23.x, gb1 = s2 ? p01.y : p23.y;
    V3d ge;
    S3d He;
    stvkVertex<double>(xo, xa, xb, ga0, ga1, gb0, gb1, p45.x, p45.y, p67.x, ge, He);
    return ge;
}

// The FP64 Stable Neo-Hookean gradient at a tetrahedron slot: p the element's twelve prepared parameters (rows
// of Dinv g1 = p0..2, g2 = p3..5, g3 = p6..8, g0 = -(g1 + g2 + g3); lambda, mu, alpha), s0 / s01 / s012 whether
// the vertex's slot is 0 / at most 1 / at most 2, x0 .. x2 the other three slots' positions in stencil order.
__device__ __forceinline__ V3d tetSlotGradD(const double2 *p, bool s0, bool s01, bool s012, const V3d &xo,
                                            const V3d &x0, const V3d &x1, const V3d &x2)
{
    const double2 p01 = p[0], p23 = p[1], p45 = p[2], p67 = p[3], p89 = p[4], pab = p[5];
    const V3d g1 = {p01.x, p01.y, p23.x}, g2 = {p23.y, p45.x, p45.y}, g3 = {p67.x, p67.y, p89.x};
    const V3d g0 = -(g1 + g2 + g3);
    const V3d ga = s0 ? g1 : g0, gb = s01 ? g2 : g1, gc = s012 ? g3 : g2;
    V3d ge;
    S3d He;
    snhVertex<double>(xo, x0, x1, x2, ga, gb, gc, p89.y, pab.x, pab.y, ge, He);
    return ge;
}

// The FP64 gradient at slot `slot` of a further-energy element of class C from its exact parameters p: the
// closed forms of the solve (gradient only), with the slot-dependent coefficients picked by selects.
template <int C>
__device__ __forceinline__ V3d extSlotGradD(const V3d *xs, const double *p, int slot)
{
    V3d g;
    S3d H;
    if constexpr (C == 0) {
        springSquaredContinuumRec<double, false>(xs, p[0], p[1], slot == 0 ? 1.0 : -1.0, g, H);
    } else if constexpr (C == 1 || C == 2) {
        const double4 d = make_double4(p[0], p[1], p[2], p[3]);
        const double ga = slot == 0 ? -(d.x + d.z) : slot == 1 ? d.x : d.z;
        const double gb = slot == 0 ? -(d.y + d.w) : slot == 1 ? d.y : d.w;
        if constexpr (C == 1)
            membraneBwRec<double, false>(xs, d, p[4], p[5], p[6], ga, gb, g, H);
        else
            membraneChoiRec<double, false>(xs, d, make_double4(p[4], p[5], p[6], p[7]), ga, gb, g, H);
    } else if constexpr (C == 3) {
        const double cs = slot == 0 ? p[0] : slot == 1 ? p[1] : slot == 2 ? p[2] : p[3];
        bendingQuadraticRec<double, false>(xs, make_double4(p[0], p[1], p[2], p[3]), p[4], cs, g, H);
    } else if constexpr (C == 4 || C == 5) {
        bendingBwRec<C == 5, double, false>(xs, p[0], p[1], slot == 1 || slot == 3, g, H);
    } else {
        double dv[9], gm[3];
#pragma unroll
        for (int i = 0; i < 9; ++i) dv[i] = p[i];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            gm[c] = slot == 0 ? -(p[c] + p[3 + c] + p[6 + c]) : slot == 1 ? p[c] : slot == 2 ? p[3 + c] : p[6 + c];
        if constexpr (C == 6)
            tetStvkRec<double, false>(xs, dv, gm, p[9], p[10], g, H);
        else
            tetSsnhRec<double, false>(xs, dv, gm, p[9], p[10], g, H);
    }
    return g;
}

// One further-energy class of the residual: the vertex's row of (element, its slot in the element) in the solve's
// class layout, each entry's FP64 slot gradient from the element's exact parameters and slots.
template <int C, int G>
__device__ __forceinline__ void extResidualClass(const ResidualArgs &a, int64_t row, int32_t lig, V3d &r)
{
    constexpr int NS = xclsStencil(C);
    const int32_t w = a.xrW[C];
    for (int32_t j = lig; j < w; j += G) {
        const int2 e = a.xrR[C][row * w + j];
        if (e.x < 0) continue;
        const int4 v = a.extV[e.x];
        const int32_t vv[4] = {v.x, v.y, v.z, v.w};
        V3d xs[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) xs[s] = xyzd(a.pos[vv[s]]);
        r = r - extSlotGradD<C>(xs, a.extP + static_cast<int64_t>(kExtParams) * e.x, e.y);
    }
}

// K4: r_v = -(m/h^2)(x_v - x_tilde_v) - sum_e grad_v E_e(x), all in FP64 from the FP32 state and the
// exact prepared parameters (implicitEulerBase.cpp:102-143 force, x-form of
// implicitEulerNewtonRobust.cpp:65-66). Writes |r_v|^2 per particle (particle order).
// A group of G lanes per vertex, vertices in particle order (neighbours and the elements they share are
// processed close together in time, and with XCD-contiguous block numbering on one XCD, whose L2 then serves
// the elements and positions they share). Lane l takes entries l, l + G, ... of the vertex's fixed-width row of
// element ids in each class, reads the element's slots and exact parameters from its table and the stencil's
// positions, and the group sums its partial forces in a fixed shuffle order. The vertex's slot in an element
// is found by comparing ids, and the other slots and their shape vectors are picked by selects, so nothing is
// indexed at run time in private memory.
template <int G, bool EXT>
__global__ void __launch_bounds__(256) residualKernel(ResidualArgs a)
{
    constexpr int GPW = 64 / G;
    const int32_t lane64 = threadIdx.x & 63;
    const int32_t wave = xcdContiguousBlock(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int32_t grp = lane64 / G, lig = lane64 - grp * G;
    const int32_t gi = wave * GPW + grp;
    const bool active = grp < GPW && gi < a.n;
    const int32_t vid = active ? a.order[gi] : 0;
    const int64_t row = vid;
    V3d r = {0.0, 0.0, 0.0};
    if (active) {
        const float4 xf = a.pos[vid];
        const V3d xo = xyzd(xf);
        if (lig == 0) {
            const float4 tf = a.xt[vid];
            const double mdt2 = a.mass[vid] * a.invdt2;
            r = V3d{-mdt2 * (xo.x - double(tf.x)), -mdt2 * (xo.y - double(tf.y)), -mdt2 * (xo.z - double(tf.z))};
        }
        V3d ge;
        S3d He;
        for (int32_t j = lig; j < a.triW; j += G) {
            const int32_t el = a.triRow[row * a.triW + j];
            if (el < 0) continue;
            const int4 v = a.triV[el];
            const double2 *p = reinterpret_cast<const double2 *>(a.triP + 8 * static_cast<int64_t>(el));
            // the other two slots in stencil order
            const bool s0 = v.x == vid, s2 = v.z == vid;
            const int32_t ia = s0 ? v.y : v.x, ib = s2 ? v.y : v.z;
            r = r - triSlotGradD(p[0], p[1], p[2], p[3], s0, s2, xo, xyzd(a.pos[ia]), xyzd(a.pos[ib]));
        }
        // hinges: k (theta - theta0) once per hinge (hingeKcKernel), the slot's gradient of theta here
        for (int32_t j = lig; j < a.hinEW; j += G) {  // hinge-edge incidences (slots 0, 1)
            const int32_t el = a.hinERow[row * a.hinEW + j];
            if (el < 0) continue;
            const int4 v = a.hinV[el];
            r = r - bendingEdgeGrad<double>(xyzd(a.pos[v.x]), xyzd(a.pos[v.y]), xyzd(a.pos[v.z]), xyzd(a.pos[v.w]),
                                            v.y == vid, a.hinKc[el]);
        }
        for (int32_t j = lig; j < a.hinWW; j += G) {  // hinge-wing incidences (slots 2, 3)
            const int32_t el = a.hinWRow[row * a.hinWW + j];
            if (el < 0) continue;
            const int4 v = a.hinV[el];
            r = r - bendingWingGrad<double>(xyzd(a.pos[v.x]), xyzd(a.pos[v.y]), xyzd(a.pos[v.z]), xyzd(a.pos[v.w]),
                                            v.w == vid, a.hinKc[el]);
        }
        for (int32_t j = lig; j < a.tetW; j += G) {
            const int32_t el = a.tetRow[row * a.tetW + j];
            if (el < 0) continue;
            const int4 v = a.tetV[el];
            const double2 *p = reinterpret_cast<const double2 *>(a.tetP + 12 * static_cast<int64_t>(el));
            // the other three slots in stencil order: s = 0: (1, 2, 3), 1: (0, 2, 3), 2: (0, 1, 3), 3: (0, 1, 2)
            const bool s0 = v.x == vid, s01 = s0 || v.y == vid, s012 = s01 || v.z == vid;
            const int32_t i0 = s0 ? v.y : v.x, i1 = s01 ? v.z : v.y, i2 = s012 ? v.w : v.z;
            r = r - tetSlotGradD(p, s0, s01, s012, xo, xyzd(a.pos[i0]), xyzd(a.pos[i1]), xyzd(a.pos[i2]));
        }
        for (int32_t j = lig; j < a.sprW; j += G) {
            const int32_t el = a.sprRow[row * a.sprW + j];
            if (el < 0) continue;
            const int2 v = a.sprV[el];
            const double2 p = a.sprP[el];
            springVertex<double>(xo, xyzd(a.pos[v.x == vid ? v.y : v.x]), p.x, p.y, ge, He);
            r = r - ge;
        }
        if (a.ancPtr) {
            for (int32_t k = a.ancPtr[vid] + lig; k < a.ancPtr[vid + 1]; k += G) {
                const int32_t el = a.ancElem[k];
                const double4 p = a.ancP[el];
                anchorVertex<double>(xo, V3d{p.x, p.y, p.z}, p.w, ge, He);
                r = r - ge;
            }
        }
        if constexpr (EXT) {
            extResidualClass<0, G>(a, row, lig, r);
            extResidualClass<1, G>(a, row, lig, r);
            extResidualClass<2, G>(a, row, lig, r);
            extResidualClass<3, G>(a, row, lig, r);
            extResidualClass<4, G>(a, row, lig, r);
            extResidualClass<5, G>(a, row, lig, r);
            extResidualClass<6, G>(a, row, lig, r);
            extResidualClass<7, G>(a, row, lig, r);
        }
    }
    r.x = groupSum<G>(r.x, lane64);
    r.y = groupSum<G>(r.y, lane64);
    r.z = groupSum<G>(r.z, lane64);
    if (active && lig == 0) a.r2[gi] = dot(r, r);
}

// Per-object |r| in two deterministic stages. Stage 1: block b sums the squared residuals of segment b (at most
// kReduceSegment consecutive particles of one object): thread t adds entries t, t + 256, ... in that order, all its
// loads issued at once, then the block's 256 sums are added in a fixed tree order. Stage 2: one wavefront per
// object: lane l adds the object's segment partials l, l + 64, ... in that order, the lanes' sums are added in a
// fixed shuffle tree, and lane 0 takes the square root.
__global__ void __launch_bounds__(256) reduceSegmentsKernel(const double *__restrict__ r2,
                                                            const int64_t *__restrict__ segBegin,
                                                            double *__restrict__ partial)
{
    constexpr int K = kReduceSegment / 256;
    __shared__ double part[256];
    const int64_t b = segBegin[blockIdx.x], e = segBegin[blockIdx.x + 1];
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int64_t j = b + threadIdx.x + 256 * k;
        v[k] = j < e ? r2[j] : 0.0;
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) s += v[k];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = part[0];
}
__global__ void __launch_bounds__(256) reduceObjectsKernel(const double *__restrict__ partial,
                                                           const int32_t *__restrict__ objSeg, int32_t nObjects,
                                                           double *__restrict__ out)
{
    const int32_t o = blockIdx.x * 4 + (threadIdx.x >> 6);  // uniform over the wavefront
    const int32_t lane = threadIdx.x & 63;
    if (o >= nObjects) return;
    double s = 0.0;
    for (int32_t k = objSeg[o] + lane; k < objSeg[o + 1]; k += 64) s += partial[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) out[o] = sqrt(s);
}

// Per-element derivative hook: evaluates the same device functions the solve uses, for every
// vertex of each element, from the prepared FP64 parameters (params layout of spg_vbd.h).
template <class R>
__global__ void __launch_bounds__(64) elementKernel(int type, int64_t n, const double *xs, const double *params,
                                                    int form, double *grad, double *hess)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const int ns = elemStencilSize(type), np = elemParams(type);
    if (!ns) return;
    elementAll<R>(type, xs + i * ns * 3, params + i * np, form, grad + i * ns * 3, hess + i * ns * 9);
}

// Calibration copy: one float4 read and one float4 write per thread.
__global__ void __launch_bounds__(256) copyKernel(const float4 *__restrict__ src, float4 *__restrict__ dst, int64_t n)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

inline int blocksFor(int64_t n, int bs) { return static_cast<int>((n + bs - 1) / bs); }

}  // namespace

hipError_t launchCopy(const float4 *src, float4 *dst, int64_t n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    copyKernel<<<blocksFor(n, 256), 256, 0, stream>>>(src, dst, n);
    return hipGetLastError();
}

hipError_t launchBeginSubstep(float4 *pos, float4 *xold, float4 *vel, float4 *xt, const float *mass, int32_t n,
                              float h, float gx, float gy, float gz, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    beginSubstepKernel<<<blocksFor(n, 256), 256, 0, stream>>>(pos, xold, vel, xt, mass, n, h, gx, gy, gz);
    return hipGetLastError();
}

hipError_t launchEndSubstep(const float4 *pos, const float4 *xold, const float4 *xt, float4 *vel, float4 *corr,
                            int32_t n, float invh, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    endSubstepKernel<<<blocksFor(n, 256), 256, 0, stream>>>(pos, xold, xt, vel, corr, n, invh);
    return hipGetLastError();
}

hipError_t launchCopyXt(float4 *pos, const float4 *xt, int32_t n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    copyXtKernel<<<blocksFor(n, 256), 256, 0, stream>>>(pos, xt, n);
    return hipGetLastError();
}

template <int G, bool LDS, bool EXT>
void solveColorLaunchE(const SolveArgs &a, int grid, int blockSize, size_t lds, hipStream_t stream, hipEvent_t start,
                       hipEvent_t stop)
{
    if (start)  // timestamps of the dispatch itself (no extra packets between the events and the kernel)
        hipExtLaunchKernelGGL(solveColorKernel<G, LDS, EXT>, dim3(grid), dim3(blockSize), lds, stream, start, stop, 0,
                              a);
    else
        solveColorKernel<G, LDS, EXT><<<grid, blockSize, lds, stream>>>(a);
}
// the kernel with the further energies' incidence class is used only when such energies exist, so the
// five-energy path keeps its registers
template <int G, bool LDS>
void solveColorLaunch(const SolveArgs &a, int grid, int blockSize, size_t lds, hipStream_t stream, hipEvent_t start,
                      hipEvent_t stop)
{
    if (a.extPtr)
        solveColorLaunchE<G, LDS, true>(a, grid, blockSize, lds, stream, start, stop);
    else
        solveColorLaunchE<G, LDS, false>(a, grid, blockSize, lds, stream, start, stop);
}

template <int G>
void solveColorLaunchG(const SolveArgs &a, int grid, int blockSize, bool staged, hipStream_t stream, hipEvent_t start,
                       hipEvent_t stop)
{
    if (staged)
        solveColorLaunch<G, true>(a, grid, blockSize, solveLdsBytes(a, G, blockSize), stream, start, stop);
    else
        solveColorLaunch<G, false>(a, grid, blockSize, 0, stream, start, stop);
}

template <int G, int CLS>
void solveColorNbLaunchC(const SolveArgs &a, int grid, int blockSize, hipStream_t stream, hipEvent_t start,
                         hipEvent_t stop)
{
    const size_t lds = solveNbLdsBytes(a, G, blockSize);
    if (start)
        hipExtLaunchKernelGGL(solveColorNbKernel<G, CLS>, dim3(grid), dim3(blockSize), lds, stream, start, stop, 0,
                              a);
    else
        solveColorNbKernel<G, CLS><<<grid, blockSize, lds, stream>>>(a);
}

// the instantiation for the classes present: surface (StVK + hinges), solid (SNH), or all
template <int G>
void solveColorNbLaunch(const SolveArgs &a, int grid, int blockSize, hipStream_t stream, hipEvent_t start,
                        hipEvent_t stop)
{
    const int cls = (a.triW ? kClsTri : 0) | (a.hinEW || a.hinWW ? kClsHinge : 0) | (a.tetW ? kClsTet : 0) |
                    (a.sprW ? kClsSpring : 0);
    if ((cls & ~(kClsTri | kClsHinge)) == 0)
        solveColorNbLaunchC<G, kClsTri | kClsHinge>(a, grid, blockSize, stream, start, stop);
    else if (cls == kClsTet)
        solveColorNbLaunchC<G, kClsTet>(a, grid, blockSize, stream, start, stop);
    else
        solveColorNbLaunchC<G, kClsAll>(a, grid, blockSize, stream, start, stop);
}

hipError_t launchSolveColor(const SolveArgs &args, int lanesPerVertex, int blockSize, int ldsStaging,
                            hipStream_t stream, hipEvent_t start, hipEvent_t stop)
{
    const int64_t nv = args.end - args.begin;
    if (nv <= 0) return hipSuccess;
    if (blockSize > 256 || blockSize < 64 || blockSize % 64) return hipErrorInvalidValue;
    if (lanesPerVertex < 1 || lanesPerVertex > 16) return hipErrorInvalidValue;
    if (ldsStaging < 0 || ldsStaging > 2) return hipErrorInvalidValue;
    if (ldsStaging == 1 && solveLdsBytes(args, lanesPerVertex, blockSize) > kSolveMaxLds) return hipErrorInvalidValue;
    if (ldsStaging == 2 && (args.nbW < 1 || args.nbW > kNbMax || args.extPtr ||
                            solveNbLdsBytes(args, lanesPerVertex, blockSize) > kSolveMaxLds))
        return hipErrorInvalidValue;
    // wavefronts of 64 / G vertex groups, blockSize / 64 wavefronts per block
    const int64_t waves = (nv + 64 / lanesPerVertex - 1) / (64 / lanesPerVertex);
    const int grid = blocksFor(waves, blockSize / 64);
    const bool st = ldsStaging != 0;
    SolveArgs a = args;
    a.gridBlocks = grid;
    a.wavesPerBlock = blockSize / 64;
    if (ldsStaging == 2) {
        switch (lanesPerVertex) {
        case 1: solveColorNbLaunch<1>(a, grid, blockSize, stream, start, stop); break;
        case 2: solveColorNbLaunch<2>(a, grid, blockSize, stream, start, stop); break;
        case 3: solveColorNbLaunch<3>(a, grid, blockSize, stream, start, stop); break;
        case 4: solveColorNbLaunch<4>(a, grid, blockSize, stream, start, stop); break;
        case 6: solveColorNbLaunch<6>(a, grid, blockSize, stream, start, stop); break;
        case 8: solveColorNbLaunch<8>(a, grid, blockSize, stream, start, stop); break;
        case 16: solveColorNbLaunch<16>(a, grid, blockSize, stream, start, stop); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (lanesPerVertex) {
    case 1: solveColorLaunchG<1>(a, grid, blockSize, st, stream, start, stop); break;
    case 2: solveColorLaunchG<2>(a, grid, blockSize, st, stream, start, stop); break;
    case 3: solveColorLaunchG<3>(a, grid, blockSize, st, stream, start, stop); break;
    case 4: solveColorLaunchG<4>(a, grid, blockSize, st, stream, start, stop); break;
    case 6: solveColorLaunchG<6>(a, grid, blockSize, st, stream, start, stop); break;
    case 8: solveColorLaunchG<8>(a, grid, blockSize, st, stream, start, stop); break;
    case 16: solveColorLaunchG<16>(a, grid, blockSize, st, stream, start, stop); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

size_t solveStagedLdsBytes(const SolveArgs &a, int lanesPerVertex, int blockSize, int ldsStaging)
{
    if (ldsStaging == 2) return solveNbLdsBytes(a, lanesPerVertex, blockSize);
    return ldsStaging == 1 ? solveLdsBytes(a, lanesPerVertex, blockSize) : 0;
}

// K4, neighbourhood-staged variant: the same residual in the same order of operations (bitwise the same result),
// with the vertex's group first gathering its own and its distinct neighbours' positions into an LDS slab (the
// neighbour rows of the staged solve), so each position is read from memory once per vertex, and row entries
// that carry the element id and the 5-bit slab indices of its slots in stencil order (the vertex itself is
// index 0). Vertices run in particle order, as in residualKernel (neighbouring vertices share elements, so a
// wavefront's element loads touch fewer cache lines than in colour-major slot order), and the first entry of every
// class and its element's parameters are loaded before the gather.
template <int G, int CLS>
__global__ void __launch_bounds__(256) residualNbKernel(ResidualArgs a)
{
    constexpr int GPW = 64 / G;
    constexpr int R = (kNbMax + G - 1) / G;
    constexpr int SLABS = GPW + (64 % G ? 1 : 0);
    extern __shared__ float4 nbSlab[];
    const int32_t lane64 = threadIdx.x & 63;
    const int32_t wib = threadIdx.x >> 6;
    const int32_t wave = xcdContiguousBlock(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + wib;
    const int32_t grp = lane64 / G, lig = lane64 - grp * G;
    const int32_t gi = wave * GPW + grp;
    const bool active = grp < GPW && gi < a.n;
    const int32_t vid = active ? a.order[gi] : 0;
    const int64_t row = vid;
    float4 *slab = nbSlab + static_cast<int64_t>(wib * SLABS + grp) * (R * G);
    // first round: row entries, neighbour ids, the vertex's own data
    const int2 none = make_int2(-1, -1);
    int2 et = none, ee = none, ew = none, eq = none, es = none;
    if constexpr ((CLS & kClsTri) != 0)
        if (active && lig < a.triW) et = a.triRL[row * a.triW + lig];
    if constexpr ((CLS & kClsHinge) != 0) {
        if (active && lig < a.hinEW) ee = a.hinERL[row * a.hinEW + lig];
        if (active && lig < a.hinWW) ew = a.hinWRL[row * a.hinWW + lig];
    }
    if constexpr ((CLS & kClsTet) != 0)
        if (active && lig < a.tetW) eq = a.tetRL[row * a.tetW + lig];
    if constexpr ((CLS & kClsSpring) != 0)
        if (active && lig < a.sprW) es = a.sprRL[row * a.sprW + lig];
    int32_t ids[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int32_t e = lig + k * G;
        ids[k] = (active && e < a.nbW) ? a.nbRow[row * a.nbW + e] : -1;
    }
    const float4 tf = a.xt[vid];
    const double mass = a.mass[vid];
    int32_t ab = 0, ae = 0;
    if (active && a.ancPtr) {
        ab = a.ancPtr[vid];
        ae = a.ancPtr[vid + 1];
    }
    // second round: the first entries' element parameters and hinge angles, and the neighbourhood's positions
    double2 t01 = {}, t23 = {}, t45 = {}, t67 = {}, sp = {};
    double ke = 0.0, kw = 0.0;
    if constexpr ((CLS & kClsTri) != 0) {
        if (et.x >= 0) {
            const double2 *p = reinterpret_cast<const double2 *>(a.triP + 8 * static_cast<int64_t>(et.x));
            t01 = p[0];
            t23 = p[1];
            t45 = p[2];
            t67 = p[3];
        }
    }
    if constexpr ((CLS & kClsHinge) != 0) {
        if (ee.x >= 0) ke = a.hinKc[ee.x];
        if (ew.x >= 0) kw = a.hinKc[ew.x];
    }
    if constexpr ((CLS & kClsSpring) != 0)
        if (es.x >= 0) sp = a.sprP[es.x];
    float px[R], py[R], pz[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const float4 p = a.pos[ids[k] >= 0 ? ids[k] : vid];
        px[k] = p.x;
        py[k] = p.y;
        pz[k] = p.z;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) slab[lig + k * G] = make_float4(px[k], py[k], pz[k], 0.f);
    __syncthreads();
    auto at = [&](int32_t pk, int s) { return xyzd(slab[(pk >> (5 * s)) & 31]); };
    V3d r = {0.0, 0.0, 0.0};
    if (active) {
        const V3d xo = at(0, 0);
        if (lig == 0) {
            const double mdt2 = mass * a.invdt2;
            r = V3d{-mdt2 * (xo.x - double(tf.x)), -mdt2 * (xo.y - double(tf.y)), -mdt2 * (xo.z - double(tf.z))};
        }
        if constexpr ((CLS & kClsTri) != 0) {
            // the slots' indices: the vertex is slot 0 / slot 2 when that index is 0
            auto tri = [&](int32_t pk, double2 p01, double2 p23, double2 p45, double2 p67) {
                const int32_t l0 = pk & 31, l1 = (pk >> 5) & 31, l2 = (pk >> 10) & 31;
                const bool s0 = l0 == 0, s2 = l2 == 0;
                const int32_t ia = s0 ? l1 : l0, ib = s2 ? l1 : l2;
                r = r - triSlotGradD(p01, p23, p45, p67, s0, s2, xo, xyzd(slab[ia]), xyzd(slab[ib]));
            };
            if (et.x >= 0) tri(et.y, t01, t23, t45, t67);
            for (int32_t j = lig + G; j < a.triW; j += G) {
                const int2 e = a.triRL[row * a.triW + j];
                if (e.x < 0) continue;
                const double2 *p = reinterpret_cast<const double2 *>(a.triP + 8 * static_cast<int64_t>(e.x));
                tri(e.y, p[0], p[1], p[2], p[3]);
            }
        }
        if constexpr ((CLS & kClsHinge) != 0) {
            auto edge = [&](int32_t pk, double kc) {
                r = r - bendingEdgeGrad<double>(at(pk, 0), at(pk, 1), at(pk, 2), at(pk, 3), ((pk >> 5) & 31) == 0, kc);
            };
            if (ee.x >= 0) edge(ee.y, ke);
            for (int32_t j = lig + G; j < a.hinEW; j += G) {
                const int2 e = a.hinERL[row * a.hinEW + j];
                if (e.x >= 0) edge(e.y, a.hinKc[e.x]);
            }
            auto wing = [&](int32_t pk, double kc) {
                r = r - bendingWingGrad<double>(at(pk, 0), at(pk, 1), at(pk, 2), at(pk, 3), ((pk >> 15) & 31) == 0, kc);
            };
            if (ew.x >= 0) wing(ew.y, kw);
            for (int32_t j = lig + G; j < a.hinWW; j += G) {
                const int2 e = a.hinWRL[row * a.hinWW + j];
                if (e.x >= 0) wing(e.y, a.hinKc[e.x]);
            }
        }
        if constexpr ((CLS & kClsTet) != 0) {
            auto tet = [&](int2 e) {
                const int32_t pk = e.y;
                const int32_t l0 = pk & 31, l1 = (pk >> 5) & 31, l2 = (pk >> 10) & 31, l3 = (pk >> 15) & 31;
                const bool s0 = l0 == 0, s01 = s0 || l1 == 0, s012 = s01 || l2 == 0;
                const int32_t i0 = s0 ? l1 : l0, i1 = s01 ? l2 : l1, i2 = s012 ? l3 : l2;
                const double2 *p = reinterpret_cast<const double2 *>(a.tetP + 12 * static_cast<int64_t>(e.x));
                r = r - tetSlotGradD(p, s0, s01, s012, xo, xyzd(slab[i0]), xyzd(slab[i1]), xyzd(slab[i2]));
            };
            if (eq.x >= 0) tet(eq);
            for (int32_t j = lig + G; j < a.tetW; j += G) {
                const int2 e = a.tetRL[row * a.tetW + j];
                if (e.x >= 0) tet(e);
            }
        }
        if constexpr ((CLS & kClsSpring) != 0) {
            auto spring = [&](int32_t pk, double2 p) {
                V3d ge;
                S3d He;
                springVertex<double>(xo, at(pk, 0), p.x, p.y, ge, He);
                r = r - ge;
            };
            if (es.x >= 0) spring(es.y, sp);
            for (int32_t j = lig + G; j < a.sprW; j += G) {
                const int2 e = a.sprRL[row * a.sprW + j];
                if (e.x >= 0) spring(e.y, a.sprP[e.x]);
            }
        }
        for (int32_t k = ab + lig; k < ae; k += G) {
            const double4 p = a.ancP[a.ancElem[k]];
            V3d ge;
            S3d He;
            anchorVertex<double>(xo, V3d{p.x, p.y, p.z}, p.w, ge, He);
            r = r - ge;
        }
    }
    r.x = groupSum<G>(r.x, lane64);
    r.y = groupSum<G>(r.y, lane64);
    r.z = groupSum<G>(r.z, lane64);
    if (active && lig == 0) a.r2[gi] = dot(r, r);
}

// The residual's first pass over the hinges: kC = k (theta - theta0) per hinge in FP64 (discreteBendingEnergy.cpp
// :20-27 angle), so that the four slots of a hinge share one angle evaluation. Blocks are numbered XCD-contiguous
// (xcdContiguousBlock), so consecutive hinges, which share particles, gather them through one XCD's L2.
__global__ void __launch_bounds__(256) hingeKcKernel(const float4 *__restrict__ pos, const int4 *__restrict__ hinV,
                                                    const double2 *__restrict__ hinP, int32_t n,
                                                    double *__restrict__ kc)
{
    const int32_t i = xcdContiguousBlock(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int4 v = hinV[i];
    const double2 p = hinP[i];
    const double th = bendingTheta<double>(xyzd(pos[v.x]), xyzd(pos[v.y]), xyzd(pos[v.z]), xyzd(pos[v.w]));
    kc[i] = p.y * (th - p.x);
}

template <int G>
void residualLaunchG(const ResidualArgs &a, hipStream_t stream)
{
    const int blocks = blocksFor(a.n, 256 / 64 * (64 / G));
    if (a.extPtr) {
        residualKernel<G, true><<<blocks, 256, 0, stream>>>(a);
    } else if (a.nbW > 0 && a.nbW <= kNbMax) {
        const size_t lds = nbSlabBytes(G, 256);
        const int cls = (a.triW ? kClsTri : 0) | (a.hinEW || a.hinWW ? kClsHinge : 0) | (a.tetW ? kClsTet : 0) |
                        (a.sprW ? kClsSpring : 0);
        if ((cls & ~(kClsTri | kClsHinge)) == 0)
            residualNbKernel<G, kClsTri | kClsHinge><<<blocks, 256, lds, stream>>>(a);
        else if (cls == kClsTet)
            residualNbKernel<G, kClsTet><<<blocks, 256, lds, stream>>>(a);
        else
            residualNbKernel<G, kClsAll><<<blocks, 256, lds, stream>>>(a);
    } else {
        residualKernel<G, false><<<blocks, 256, 0, stream>>>(a);
    }
}

hipError_t launchResidual(const ResidualArgs &a, hipStream_t stream)
{
    if (a.n <= 0) return hipSuccess;
    if (a.nHinges > 0) {
        if (!a.hinKc) return hipErrorInvalidValue;
        hingeKcKernel<<<blocksFor(a.nHinges, 256), 256, 0, stream>>>(a.pos, a.hinV, a.hinP, a.nHinges, a.hinKc);
    }
    // lanes per vertex: the widest row's entries in one pass where they fit 6 lanes (the cloth's 6 / 6 / 6),
    // otherwise 8 lanes striding over longer rows (the cube's 24 tetrahedra)
    const int32_t w = max(max(max(a.triW, a.hinEW), max(a.hinWW, a.tetW)), a.sprW);
    if (w <= 4)
        residualLaunchG<4>(a, stream);
    else if (w <= 6)
        residualLaunchG<6>(a, stream);
    else
        residualLaunchG<8>(a, stream);
    return hipGetLastError();
}

hipError_t launchReduceObjects(const double *r2, const int64_t *segBegin, int32_t nSegments, const int32_t *objSeg,
                               int32_t nObjects, double *partial, double *out, hipStream_t stream)
{
    if (nObjects <= 0 || nSegments <= 0) return hipSuccess;
    reduceSegmentsKernel<<<nSegments, 256, 0, stream>>>(r2, segBegin, partial);
    reduceObjectsKernel<<<blocksFor(nObjects, 4), 256, 0, stream>>>(partial, objSeg, nObjects, out);
    return hipGetLastError();
}

// The solve's slot gradients and blocks of further-energy elements (types 6 ... 13), for derivative checks:
// element i has its stencil's positions at pos[NS i ... NS i + NS - 1] and its prepared parameters at prm + np i.
template <int TYPE>
__global__ void __launch_bounds__(64) extElementKernel(int64_t n, const float4 *__restrict__ pos,
                                                       const double *__restrict__ prm, int np, double *grad,
                                                       double *hess)
{
    constexpr int NS = fullStencilSize(TYPE);
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = static_cast<int32_t>(i * NS);
    const int4 v = make_int4(b, b + 1, NS > 2 ? b + 2 : -1, NS > 3 ? b + 3 : -1);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        V3f g;
        S3f H;
        extSlotAny<NS, float, true, TYPE>(TYPE, s, pos, v, prm + i * np, g, H);
        double *gd = grad + (i * NS + s) * 3, *hd = hess + (i * NS + s) * 9;
        gd[0] = g.x;
        gd[1] = g.y;
        gd[2] = g.z;
        const float hv[9] = {H.xx, H.xy, H.xz, H.xy, H.yy, H.yz, H.xz, H.yz, H.zz};
#pragma unroll
        for (int k = 0; k < 9; ++k) hd[k] = hv[k];
    }
}

hipError_t launchExtElementDerivatives(int type, int64_t n, const float4 *pos, const double *prm, int np,
                                       double *grad, double *hess, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    const unsigned blocks = blocksFor(n, 64);
    switch (type) {
    case 6: extElementKernel<6><<<blocks, 64, 0, stream>>>(n, pos, prm, np, grad, hess); break;
    case 7: extElementKernel<7><<<blocks, 64, 0, stream>>>(n, pos, prm, np, grad, hess); break;
    case 8: extElementKernel<8><<<blocks, 64, 0, stream>>>(n, pos, prm, np, grad, hess); break;
    case 9: extElementKernel<9><<<blocks, 64, 0, stream>>>(n, pos, prm, np, grad, hess); break;
    case 10: extElementKernel<10><<<blocks, 64, 0, stream>>>(n, pos, prm, np, grad, hess); break;
    case 11: extElementKernel<11><<<blocks, 64, 0, stream>>>(n, pos, prm, np, grad, hess); break;
    case 12: extElementKernel<12><<<blocks, 64, 0, stream>>>(n, pos, prm, np, grad, hess); break;
    case 13: extElementKernel<13><<<blocks, 64, 0, stream>>>(n, pos, prm, np, grad, hess); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launchElementDerivatives(int type, int64_t n, const double *xs, const double *params, int precision,
                                    int form, double *grad, double *hess, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    if (precision == 0)
        elementKernel<float><<<blocksFor(n, 64), 64, 0, stream>>>(type, n, xs, params, form, grad, hess);
    else
        elementKernel<double><<<blocksFor(n, 64), 64, 0, stream>>>(type, n, xs, params, form, grad, hess);
    return hipGetLastError();
}

}  // namespace spgvbd
