// VBD on rigid bodies (include/spg_rigid.h): FP64 kernels for the three phases of a substep, the device layout
// built from spg RigidBodyGroups, and the C-ABI.
//
// Layout in HBM. Bodies are renumbered colour-major ("slots"): the bodies of colour 0 of every group first, then
// colour 1, ..., each colour a contiguous slot range that one launch covers (kRbLanes lanes per body). A slot's
// pose (centre and rotation matrix, 12 doubles = 96 bytes, read with 16-byte loads) is the record its neighbours
// gather; the rest of its state is stored as arrays of 3 or 9 doubles per body (v, omega, theta, xt (inertial
// prediction), xold, thetaold in 3; local inertia, its inverse and the rotation at the substep start in 9). World
// inertias R Il R^T are recomputed where used instead of stored. Element parameters are packed 16-byte aligned
// (8 doubles for springs and anchors, 2 for the alignment); a slot's incidence is a CSR row of (other body's slot,
// parameter offset << 3 | type << 1 | stencil slot) entries in energy-major,
// ascending-stencil order, the order in which the reference sums them (vbd.cpp:105-135 over
// computeStencilInfoPerVertex, vbd.cpp:231-268).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "coloring.h"
#include "host_util.h"
#include "rigid_element.h"
#include "spg_rigid.h"
#include "vbd_device.h"

using namespace spgvbd;
using namespace spgrigid;

namespace {

constexpr int kRbBlock = 256;
// Lanes cooperating on one body's incidence row in the colour solve. One lane per body measured fastest on a
// 512 x 512 lattice (8 incidence entries per body): 78.7 us per sweep, against 82 with 2 lanes and 115 with 4; the
// shuffle reduction stays for rows long enough to want more.
constexpr int kRbLanes = 1;
constexpr int kPose = 12;    // per-slot pose record: centre (3) then rotation matrix (9, row-major), 96 bytes

struct RbState {
    double *pose;                                // kPose per slot
    double *v, *w, *th, *xt, *xold, *thold;      // 3 per slot
    const double *Il, *IlInv;                    // 9 per slot: local inertia and its inverse
    double *Rold;                                // 9 per slot: rotation at the substep start
    const double *m;
};

__device__ __forceinline__ void loadPose(const double *pose, int64_t b, double *x, double *R)
{
    const double2 *q = reinterpret_cast<const double2 *>(pose + kPose * b);
    double v[12];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double2 t = q[k];
        v[2 * k] = t.x;
        v[2 * k + 1] = t.y;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = v[d];
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k] = v[3 + k];
}

__device__ __forceinline__ void storePose(double *pose, int64_t b, const double *x, const double *R)
{
    double2 *q = reinterpret_cast<double2 *>(pose + kPose * b);
    const double v[12] = {x[0], x[1], x[2], R[0], R[1], R[2], R[3], R[4], R[5], R[6], R[7], R[8]};
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = make_double2(v[2 * k], v[2 * k + 1]);
}

// world inertia R Il R^T (RigidBodyGroup::updateInertia, rigidBodyGroup.cpp:195-199)
__device__ __forceinline__ void worldInertia(const double *R, const double *Il, double *I)
{
    double T[9];
    matMul3(R, Il, T);
    matMulT3(T, R, I);
}

// Gyroscopic prediction (vbd.cpp:62-72), RigidBodyGroup::integrateVelocities with updateThetas/updateInertias
// (rigidBodyGroup.cpp:97-113), and the substep-start copies (getPositions into xOld, vbd.cpp:47-55).
__global__ void __launch_bounds__(kRbBlock) rbBeginKernel(RbState s, int64_t nb, double h, double gx, double gy,
                                                          double gz)
{
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double x[3], v[3], w[3], R[9], I[9], Iinv[9], th[3], Il[9];
    loadPose(s.pose, b, x, R);
    for (int d = 0; d < 3; ++d) {
        v[d] = s.v[3 * b + d];
        w[d] = s.w[3 * b + d];
        th[d] = s.th[3 * b + d];
        s.xold[3 * b + d] = x[d];
        s.thold[3 * b + d] = th[d];
    }
    for (int k = 0; k < 9; ++k) Il[k] = s.Il[9 * b + k];
    worldInertia(R, Il, I);
    {
        double T[9];
        matMul3(R, s.IlInv + 9 * b, T);
        matMulT3(T, R, Iinv);
    }
    v[0] += h * gx;
    v[1] += h * gy;
    v[2] += h * gz;
    double Iw[3], c[3], acc[3];
    matVec3(I, w, Iw);
    cross3(w, Iw, c);
    c[0] = -c[0];
    c[1] = -c[1];
    c[2] = -c[2];
    matVec3(Iinv, c, acc);
    for (int d = 0; d < 3; ++d) w[d] += h * acc[d];
    for (int d = 0; d < 3; ++d) x[d] += v[d] * h;
    const double wn = sqrt(dot3(w, w));
    if (wn != 0) {
        const double a[3] = {w[0] / wn, w[1] / wn, w[2] / wn};
        double dR[9], Rn[9];
        angleAxisToR(wn * h, a, dR);
        matMul3(dR, R, Rn);
        for (int k = 0; k < 9; ++k) R[k] = Rn[k];
    }
    double thn[3], Rold[9];
    rToAa(R, thn);
    aaToR(th, Rold);
    storePose(s.pose, b, x, R);
    for (int d = 0; d < 3; ++d) {
        s.xt[3 * b + d] = x[d];
        s.v[3 * b + d] = v[d];
        s.w[3 * b + d] = w[d];
        s.th[3 * b + d] = thn[d];
    }
    for (int k = 0; k < 9; ++k) s.Rold[9 * b + k] = Rold[k];
}

// omega from the rotation increment since the substep start: log(R Rold^T) / h
// (RigidBodyGroup::computeIntegratedOmega, rigidBodyGroup.cpp:130-134)
__device__ __forceinline__ void integratedOmega(const double *R, const double *Rold, double invh, double *w)
{
    double D[9];
    matMulT3(R, Rold, D);
    rToAa(D, w);
    w[0] *= invh;
    w[1] *= invh;
    w[2] *= invh;
}

struct RbSolveArgs {
    RbState s;
    const int64_t *incPtr;
    const int2 *inc;          // (other body's slot or -1, parameter offset / 2 << 3 | (type - 21) << 1 | stencil slot)
    const double *elemP;      // per element its parameters, 16-byte aligned: 8 doubles (springs, anchor), 2 (alignment)
    unsigned long long *skipped;
    double h;
};

// One colour: the 6x6 block descent of every body in [begin, end) (vbd.cpp:101-145 with the RigidBodyGroup
// branches) and updateElementPositionFromDx (rigidBodyGroup.cpp:64-77). kRbLanes consecutive lanes share a body:
// each sums every kRbLanes-th entry of its incidence row (lane 0 also the inertia terms) into a force and the
// upper triangle of the block, the partial sums are combined by lane shuffles, and every lane of the group then
// solves the same block; lane 0 writes the body.
__global__ void __launch_bounds__(kRbBlock) rbSolveKernel(RbSolveArgs a, int64_t begin, int64_t end)
{
    const int32_t blk = xcdContiguousBlock(int32_t(blockIdx.x), int32_t(gridDim.x));
    const int64_t t = int64_t(blk) * blockDim.x + threadIdx.x;
    const int64_t b = begin + t / kRbLanes;
    const int lane = int(t % kRbLanes);
    if (b >= end) return;  // whole groups: kRbLanes divides the wavefront and the block
    const RbState &s = a.s;
    const double invh = 1.0 / a.h, invh2 = invh * invh;
    double x[3], R[9], f[6], Hu[21];
#pragma unroll
    for (int k = 0; k < 6; ++k) f[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 21; ++k) Hu[k] = 0.0;
    loadPose(s.pose, b, x, R);
    if (lane == 0) {
        const double mh2 = s.m[b] * invh2;
        for (int d = 0; d < 3; ++d) f[d] = -mh2 * (x[d] - s.xt[3 * b + d]);
        double wi[3], I[9], Il[9], Ro[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            Il[k] = s.Il[9 * b + k];
            Ro[k] = s.Rold[9 * b + k];
        }
        worldInertia(R, Il, I);
        integratedOmega(R, Ro, invh, wi);
        const double dw[3] = {wi[0] - s.w[3 * b], wi[1] - s.w[3 * b + 1], wi[2] - s.w[3 * b + 2]};
        double Idw[3];
        matVec3(I, dw, Idw);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            f[3 + d] = -invh * Idw[d];
            addB<true>(Hu, d, d, mh2);
#pragma unroll
            for (int q = 0; q < 3; ++q) addB<true>(Hu, 3 + d, 3 + q, invh2 * I[3 * d + q]);
        }
    }
    const int64_t e0 = a.incPtr[b], e1 = a.incPtr[b + 1];
    for (int64_t j = e0 + lane; j < e1; j += kRbLanes) {
        const int2 en = a.inc[j];
        const int type = kRbSpring + ((en.y >> 1) & 3);
        const int slot = en.y & 1;
        const double *pp = a.elemP + 2 * int64_t(uint32_t(en.y) >> 3);
        double p[kRbParamStride];
        if (type == kRbAlignment) {
            p[0] = pp[0];
        } else {
            const double2 *p2 = reinterpret_cast<const double2 *>(pp);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double2 v = p2[q];
                p[2 * q] = v.x;
                p[2 * q + 1] = v.y;
            }
        }
        if (type == kRbAnchor) {
            rbAccumulate<true>(type, 0, x, R, x, R, p, f, Hu);
            continue;
        }
        double xo[3], Ro[9];
        loadPose(s.pose, en.x, xo, Ro);
        if (slot == 0)
            rbAccumulate<true>(type, 0, x, R, xo, Ro, p, f, Hu);
        else
            rbAccumulate<true>(type, 1, xo, Ro, x, R, p, f, Hu);
    }
#pragma unroll
    for (int m = 1; m < kRbLanes; m <<= 1) {
#pragma unroll
        for (int k = 0; k < 6; ++k) f[k] += __shfl_xor(f[k], m, kRbLanes);
#pragma unroll
        for (int k = 0; k < 21; ++k) Hu[k] += __shfl_xor(Hu[k], m, kRbLanes);
    }
    double H[36];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 6; ++j) {
            const double v = Hu[i * 6 - i * (i - 1) / 2 + (j - i)];
            H[6 * i + j] = v;
            H[6 * j + i] = v;
        }
    double dx[6];
    const double det = solve6(H, f, dx);
    if (lane != 0) return;
    if (det > 1e-10) {
        double xn[3];
        for (int d = 0; d < 3; ++d) xn[d] = x[d] + dx[d];
        const double dn = sqrt(dx[3] * dx[3] + dx[4] * dx[4] + dx[5] * dx[5]);
        if (dn != 0) {
            const double ax[3] = {dx[3] / dn, dx[4] / dn, dx[5] / dn};
            double dR[9], Rn[9];
            angleAxisToR(dn, ax, dR);
            matMul3(dR, R, Rn);
#pragma unroll
            for (int k = 0; k < 9; ++k) R[k] = Rn[k];
        }
        double th[3];
        rToAa(R, th);
        storePose(s.pose, b, xn, R);
        for (int d = 0; d < 3; ++d) s.th[3 * b + d] = th[d];
    } else {
        atomicAdd(a.skipped, 1ull);
    }
}

// computeIntegratedVelocities (rigidBodyGroup.cpp:115-128)
__global__ void __launch_bounds__(kRbBlock) rbEndKernel(RbState s, int64_t nb, double invh)
{
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double x[3], R[9];
    loadPose(s.pose, b, x, R);
    for (int d = 0; d < 3; ++d) s.v[3 * b + d] = (x[d] - s.xold[3 * b + d]) * invh;
    double w[3];
    integratedOmega(R, s.Rold + 9 * b, invh, w);
    for (int d = 0; d < 3; ++d) s.w[3 * b + d] = w[d];
}

__global__ void rbElementKernel(int type, int64_t n, const double *com, const double *rot, const double *prm,
                                double *grad, double *hess)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ns = rbStencil(type), np = rbParams(type);
    double p[kRbParamStride] = {};
    for (int q = 0; q < np; ++q) p[q] = prm[np * i + q];
    if (type == kRbSpringSquared) p[6] = p[6] * p[6];
    for (int sl = 0; sl < ns; ++sl) {
        double f[6] = {}, B[36] = {};
        const double *c0 = com + 3 * ns * i, *r0 = rot + 9 * ns * i;
        rbAccumulate(type, sl, c0, r0, ns > 1 ? c0 + 3 : c0, ns > 1 ? r0 + 9 : r0, p, f, B);
        for (int d = 0; d < 6; ++d) grad[6 * ns * i + 6 * sl + d] = -f[d];
        for (int k = 0; k < 36; ++k) hess[36 * (ns * i + sl) + k] = B[k];
    }
}

thread_local std::string g_lastError;

}  // namespace

struct spg_rigid_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int64_t nb = 0, nElem = 0;
    int32_t nColors = 0;
    std::vector<int64_t> groupOffset;        // bodies of group g: [groupOffset[g], groupOffset[g + 1]) in input order
    std::vector<int64_t> slotOf;             // input body index (over all groups) -> slot
    std::vector<std::vector<int32_t>> colors;  // per group, per body
    std::vector<int64_t> colorStart;         // slot range of colour c: [colorStart[c], colorStart[c + 1])
    int64_t incEntries = 0, solveBytes = 0;
    DevBuf<double> pose, v, w, th, xt, xold, thold, Il, IlInv, Rold, m, elemP;
    DevBuf<int64_t> incPtr;
    DevBuf<int2> inc;
    DevBuf<unsigned long long> skipped;
    bool useGraph = true;
    hipGraphExec_t graphExec = nullptr;
    double gDt = 0, gG[3] = {0, 0, 0};
    int32_t gSub = 0, gIt = 0;

    ~spg_rigid_ctx()
    {
        if (graphExec) (void)hipGraphExecDestroy(graphExec);
        if (stream) (void)hipStreamDestroy(stream);
    }
    RbState state() const
    {
        return RbState{pose.p, v.p, w.p, th.p, xt.p, xold.p, thold.p, Il.p, IlInv.p, Rold.p, m.p};
    }
    RbSolveArgs solveArgs(double h) const
    {
        return RbSolveArgs{state(), incPtr.p, inc.p, elemP.p, skipped.p, h};
    }
};

namespace {

template <class F>
int guarded(spg_rigid_ctx *ctx, F &&f)
{
    try {
        f();
        if (ctx) ctx->err.clear();
        return SPG_VBD_OK;
    } catch (const VbdError &e) {
        (ctx ? ctx->err : g_lastError) = e.what();
        return e.code;
    } catch (const std::exception &e) {
        (ctx ? ctx->err : g_lastError) = e.what();
        return SPG_VBD_ERR_INVALID;
    }
}

bool finiteN(const double *p, int64_t n)
{
    for (int64_t k = 0; k < n; ++k)
        if (!std::isfinite(p[k])) return false;
    return true;
}

void inverse3(const double *m, double *r)
{
    const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    if (det == 0 || !std::isfinite(det)) throw VbdError(SPG_VBD_ERR_INVALID, "singular local inertia");
    const double id = 1.0 / det;
    r[0] = c00 * id;
    r[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    r[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    r[3] = c01 * id;
    r[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    r[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    r[6] = c02 * id;
    r[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    r[8] = (m[0] * m[4] - m[1] * m[3]) * id;
}

void buildContext(spg_rigid_ctx &c, const spg_rigid_group *groups, int32_t nGroups)
{
    if (nGroups < 0 || (nGroups > 0 && !groups)) throw VbdError(SPG_VBD_ERR_INVALID, "invalid group list");
    c.groupOffset.assign(1, 0);
    for (int32_t g = 0; g < nGroups; ++g) {
        const auto &G = groups[g];
        if (G.n_bodies < 0) throw VbdError(SPG_VBD_ERR_INVALID, "negative body count");
        if (G.n_bodies > 0 && (!G.x || !G.theta || !G.mass || !G.inertia))
            throw VbdError(SPG_VBD_ERR_INVALID, "null body array");
        if (G.n_energies < 0 || (G.n_energies > 0 && !G.energies))
            throw VbdError(SPG_VBD_ERR_INVALID, "invalid energy list");
        c.groupOffset.push_back(c.groupOffset.back() + G.n_bodies);
    }
    c.nb = c.groupOffset.back();
    if (c.nb >= (int64_t(1) << 31)) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "too many bodies");
    // energies: validation as in addStencil, colouring per group (coloring::colorVertices over the stencils)
    struct Elem {
        int64_t b0, b1;
        int32_t type;
        double p[kRbParamStride];
    };
    std::vector<Elem> elems;
    std::vector<std::vector<std::pair<int64_t, int>>> bodyInc(c.nb);  // (element, slot) per input body
    c.colors.resize(nGroups);
    for (int32_t g = 0; g < nGroups; ++g) {
        const auto &G = groups[g];
        const int64_t off = c.groupOffset[g];
        std::vector<FlatStencilSet> sets;
        for (int32_t ei = 0; ei < G.n_energies; ++ei) {
            const auto &E = G.energies[ei];
            if (E.type < kRbSpring || E.type > kRbAlignment)
                throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "unknown rigid-body energy type");
            if (E.n_stencils < 0 || (E.n_stencils > 0 && (!E.stencils || !E.params)))
                throw VbdError(SPG_VBD_ERR_INVALID, "invalid energy arrays");
            const int ss = rbStencil(E.type), np = rbParams(E.type);
            for (int64_t i = 0; i < E.n_stencils; ++i) {
                const double *q = E.params + np * i;
                if (!finiteN(q, np)) throw VbdError(SPG_VBD_ERR_NONFINITE, "non-finite energy parameter");
                if (E.type == kRbSpring && q[6] <= 0)
                    throw VbdError(SPG_VBD_ERR_INVALID, "Zero/negative rest length invalid");
                if (E.type == kRbSpringSquared && q[6] < 0)
                    throw VbdError(SPG_VBD_ERR_INVALID, "Negative rest length invalid");
            }
            for (int64_t k = 0; k < E.n_stencils * ss; ++k)  // anchors add no graph edge, so check here
                if (E.stencils[k] < 0 || E.stencils[k] >= G.n_bodies)
                    throw VbdError(SPG_VBD_ERR_INVALID, "stencil index out of range");
            sets.push_back(FlatStencilSet{E.stencils, E.n_stencils * ss, ss});
        }
        c.colors[g] = colorVertices(G.n_bodies, sets);  // throws on out-of-range indices like Graph::addEdge
        for (int32_t ei = 0; ei < G.n_energies; ++ei) {
            const auto &E = G.energies[ei];
            const int ss = rbStencil(E.type), np = rbParams(E.type);
            for (int64_t i = 0; i < E.n_stencils; ++i) {
                Elem el{};
                el.type = E.type;
                el.b0 = off + E.stencils[ss * i];
                el.b1 = ss > 1 ? off + E.stencils[ss * i + 1] : -1;
                for (int q = 0; q < np; ++q) el.p[q] = E.params[np * i + q];
                if (E.type == kRbSpringSquared) el.p[6] = el.p[6] * el.p[6];  // m_squaredRestLength
                const int64_t e = static_cast<int64_t>(elems.size());
                elems.push_back(el);
                bodyInc[el.b0].push_back({e, 0});
                if (ss > 1) bodyInc[el.b1].push_back({e, 1});
            }
        }
    }
    c.nElem = static_cast<int64_t>(elems.size());
    // colour-major slots
    int32_t nc = 0;
    for (const auto &cg : c.colors)
        for (auto col : cg) nc = std::max(nc, col + 1);
    c.nColors = nc;
    std::vector<int64_t> count(nc + 1, 0);
    for (const auto &cg : c.colors)
        for (auto col : cg) ++count[col + 1];
    c.colorStart.assign(nc + 1, 0);
    for (int32_t k = 0; k < nc; ++k) c.colorStart[k + 1] = c.colorStart[k] + count[k + 1];
    c.slotOf.assign(c.nb, -1);
    {
        std::vector<int64_t> next(c.colorStart.begin(), c.colorStart.end() - 1);
        for (int32_t g = 0; g < nGroups; ++g)
            for (int64_t i = 0; i < groups[g].n_bodies; ++i) c.slotOf[c.groupOffset[g] + i] = next[c.colors[g][i]]++;
    }
    // per-slot host arrays
    const int64_t nb = c.nb;
    std::vector<double> hpose(kPose * nb), hv(3 * nb, 0.0), hw(3 * nb, 0.0), hth(3 * nb), hm(nb), hIl(9 * nb),
        hIli(9 * nb);
    for (int32_t g = 0; g < nGroups; ++g) {
        const auto &G = groups[g];
        for (int64_t i = 0; i < G.n_bodies; ++i) {
            const int64_t s = c.slotOf[c.groupOffset[g] + i];
            if (!finiteN(G.x + 3 * i, 3) || !finiteN(G.theta + 3 * i, 3) || !finiteN(G.mass + i, 1) ||
                !finiteN(G.inertia + 9 * i, 9) || (G.v && !finiteN(G.v + 3 * i, 3)) ||
                (G.omega && !finiteN(G.omega + 3 * i, 3)))
                throw VbdError(SPG_VBD_ERR_NONFINITE, "non-finite body data");
            for (int d = 0; d < 3; ++d) {
                hpose[kPose * s + d] = G.x[3 * i + d];
                hth[3 * s + d] = G.theta[3 * i + d];
                if (G.v) hv[3 * s + d] = G.v[3 * i + d];
                if (G.omega) hw[3 * s + d] = G.omega[3 * i + d];
            }
            hm[s] = G.mass[i];
            for (int k = 0; k < 9; ++k) hIl[9 * s + k] = G.inertia[9 * i + k];
            inverse3(G.inertia + 9 * i, hIli.data() + 9 * s);
            aaToR(hth.data() + 3 * s, hpose.data() + kPose * s + 3);  // setPositions: R from theta
        }
    }
    // element parameters, packed with 16-byte alignment (the alignment energy's single stiffness takes 16 bytes
    // instead of a full 64-byte record), and incidence rows in slot order
    std::vector<int64_t> poff(c.nElem);
    std::vector<double> ep;
    for (int64_t e = 0; e < c.nElem; ++e) {
        poff[e] = static_cast<int64_t>(ep.size());
        const int width = elems[e].type == kRbAlignment ? 2 : kRbParamStride;
        ep.insert(ep.end(), elems[e].p, elems[e].p + width);
    }
    if (static_cast<int64_t>(ep.size()) / 2 >= (int64_t(1) << 28))
        throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "too many rigid-body elements");
    std::vector<int64_t> ptr(nb + 1, 0);
    std::vector<int2> inc;
    std::vector<int64_t> bodyOfSlot(nb);
    for (int64_t b = 0; b < nb; ++b) bodyOfSlot[c.slotOf[b]] = b;
    int64_t bytes = 0;
    for (int64_t s = 0; s < nb; ++s) {
        const auto &row = bodyInc[bodyOfSlot[s]];
        for (const auto &pr : row) {
            const auto &el = elems[pr.first];
            const int64_t other = el.b1 < 0 ? -1 : c.slotOf[pr.second == 0 ? el.b1 : el.b0];
            inc.push_back(make_int2(int32_t(other), int32_t(((poff[pr.first] / 2) << 3) | ((el.type - kRbSpring) << 1) |
                                                            pr.second)));
            // the entry, the parameters it uses and the other body's pose
            bytes += 8 + 8 * rbParams(el.type) + (el.type == kRbAnchor ? 0 : 8 * kPose);
        }
        ptr[s + 1] = static_cast<int64_t>(inc.size());
        // read: pose, row pointers, mass, xt, omega, local inertia, Rold; written: pose, theta
        bytes += 8 * kPose + 16 + 8 + 24 + 24 + 72 + 72 + 8 * kPose + 24;
    }
    c.incEntries = static_cast<int64_t>(inc.size());
    c.solveBytes = bytes;
    auto up = [&](DevBuf<double> &d, const std::vector<double> &h) { d.upload(h, c.stream); };
    up(c.pose, hpose);
    up(c.v, hv);
    up(c.w, hw);
    up(c.th, hth);
    up(c.m, hm);
    up(c.Il, hIl);
    up(c.IlInv, hIli);
    c.xt.alloc(3 * nb);
    c.xold.alloc(3 * nb);
    c.thold.alloc(3 * nb);
    c.Rold.alloc(9 * nb);
    c.elemP.upload(ep, c.stream);
    c.incPtr.upload(ptr, c.stream);
    c.inc.upload(inc, c.stream);
    c.skipped.alloc(1);
    hipCheck(hipMemsetAsync(c.skipped.p, 0, sizeof(unsigned long long), c.stream), "memset");
    hipCheck(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
}

unsigned blocksFor(int64_t n) { return static_cast<unsigned>((n + kRbBlock - 1) / kRbBlock); }

// One substep: begin, `iterations` sweeps over the colours, end. `events` (optional) brackets each colour launch.
void enqueueSubstep(spg_rigid_ctx &c, double h, int32_t iterations, const double *g,
                    std::vector<std::pair<hipEvent_t, hipEvent_t>> *events)
{
    if (c.nb == 0) return;
    const RbState s = c.state();
    hipLaunchKernelGGL(rbBeginKernel, dim3(blocksFor(c.nb)), dim3(kRbBlock), 0, c.stream, s, c.nb, h, g[0], g[1],
                       g[2]);
    hipCheck(hipGetLastError(), "begin launch");
    const RbSolveArgs a = c.solveArgs(h);
    size_t ev = 0;
    for (int32_t it = 0; it < iterations; ++it)
        for (int32_t col = 0; col < c.nColors; ++col) {
            const int64_t b0 = c.colorStart[col], b1 = c.colorStart[col + 1];
            if (b1 == b0) continue;
            if (events) hipCheck(hipEventRecord((*events)[ev].first, c.stream), "hipEventRecord");
            hipLaunchKernelGGL(rbSolveKernel, dim3(blocksFor(kRbLanes * (b1 - b0))), dim3(kRbBlock), 0, c.stream, a, b0,
                               b1);
            hipCheck(hipGetLastError(), "solve launch");
            if (events) hipCheck(hipEventRecord((*events)[ev++].second, c.stream), "hipEventRecord");
        }
    hipLaunchKernelGGL(rbEndKernel, dim3(blocksFor(c.nb)), dim3(kRbBlock), 0, c.stream, s, c.nb, 1.0 / h);
    hipCheck(hipGetLastError(), "end launch");
}

void checkGroup(const spg_rigid_ctx *ctx, int32_t group)
{
    if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
    if (group < 0 || group + 1 >= static_cast<int32_t>(ctx->groupOffset.size()))
        throw VbdError(SPG_VBD_ERR_INVALID, "group index out of range");
}

}  // namespace

extern "C" {

int spg_rigid_create(spg_rigid_ctx **out, const spg_rigid_group *groups, int32_t n_groups, int32_t device)
{
    if (!out) return SPG_VBD_ERR_INVALID;
    *out = nullptr;
    auto c = std::make_unique<spg_rigid_ctx>();
    const int rc = guarded(nullptr, [&] {
        int ndev = 0;
        hipCheck(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        if (device < 0 || device >= ndev) throw VbdError(SPG_VBD_ERR_INVALID, "invalid HIP device");
        c->device = device;
        hipCheck(hipSetDevice(device), "hipSetDevice");
        hipCheck(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        buildContext(*c, groups, n_groups);
    });
    if (rc == SPG_VBD_OK) *out = c.release();
    return rc;
}

void spg_rigid_destroy(spg_rigid_ctx *ctx)
{
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        delete ctx;
    }
}

const char *spg_rigid_last_error(const spg_rigid_ctx *ctx) { return ctx ? ctx->err.c_str() : g_lastError.c_str(); }

int spg_rigid_step(spg_rigid_ctx *ctx, double dt, int32_t nsubsteps, int32_t iterations, const double *gravity)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        if (!(dt > 0) || !std::isfinite(dt) || nsubsteps < 1 || iterations < 0 || !gravity || !finiteN(gravity, 3))
            throw VbdError(SPG_VBD_ERR_INVALID, "invalid step arguments");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        if (ctx->nb == 0) return;
        const double h = dt / nsubsteps;
        if (!ctx->useGraph) {
            for (int32_t sub = 0; sub < nsubsteps; ++sub) enqueueSubstep(*ctx, h, iterations, gravity, nullptr);
        } else {
            if (!ctx->graphExec || ctx->gDt != dt || ctx->gSub != nsubsteps || ctx->gIt != iterations ||
                std::memcmp(ctx->gG, gravity, sizeof(ctx->gG)) != 0) {
                if (ctx->graphExec) {
                    hipCheck(hipGraphExecDestroy(ctx->graphExec), "hipGraphExecDestroy");
                    ctx->graphExec = nullptr;
                }
                hipGraph_t graph = nullptr;
                hipCheck(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
                try {
                    for (int32_t sub = 0; sub < nsubsteps; ++sub) enqueueSubstep(*ctx, h, iterations, gravity, nullptr);
                } catch (...) {
                    (void)hipStreamEndCapture(ctx->stream, &graph);
                    if (graph) (void)hipGraphDestroy(graph);
                    throw;
                }
                hipCheck(hipStreamEndCapture(ctx->stream, &graph), "hipStreamEndCapture");
                const hipError_t ie = hipGraphInstantiate(&ctx->graphExec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                hipCheck(ie, "hipGraphInstantiate");
                ctx->gDt = dt;
                ctx->gSub = nsubsteps;
                ctx->gIt = iterations;
                std::memcpy(ctx->gG, gravity, sizeof(ctx->gG));
            }
            hipCheck(hipGraphLaunch(ctx->graphExec, ctx->stream), "hipGraphLaunch");
        }
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int spg_rigid_set_graph(spg_rigid_ctx *ctx, int32_t enabled)
{
    if (!ctx) return SPG_VBD_ERR_INVALID;
    ctx->useGraph = enabled != 0;
    return SPG_VBD_OK;
}

int spg_rigid_profile_substep(spg_rigid_ctx *ctx, double h, int32_t iterations, const double *gravity,
                              double *solve_ms, int32_t *launches)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        if (!(h > 0) || iterations < 0 || !gravity || !finiteN(gravity, 3))
            throw VbdError(SPG_VBD_ERR_INVALID, "invalid step arguments");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        int32_t n = 0;
        for (int32_t col = 0; col < ctx->nColors; ++col)
            if (ctx->colorStart[col + 1] > ctx->colorStart[col]) ++n;
        n *= iterations;
        std::vector<std::pair<hipEvent_t, hipEvent_t>> ev(n);
        for (auto &e : ev) {
            hipCheck(hipEventCreate(&e.first), "hipEventCreate");
            hipCheck(hipEventCreate(&e.second), "hipEventCreate");
        }
        double total = 0;
        try {
            enqueueSubstep(*ctx, h, iterations, gravity, &ev);
            hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
            for (auto &e : ev) {
                float ms = 0;
                hipCheck(hipEventElapsedTime(&ms, e.first, e.second), "hipEventElapsedTime");
                total += ms;
            }
        } catch (...) {
            for (auto &e : ev) {
                (void)hipEventDestroy(e.first);
                (void)hipEventDestroy(e.second);
            }
            throw;
        }
        for (auto &e : ev) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        if (solve_ms) *solve_ms = total;
        if (launches) *launches = n;
    });
}

int spg_rigid_get_state(spg_rigid_ctx *ctx, int32_t group, double *x, double *theta, double *v, double *omega)
{
    return guarded(ctx, [&] {
        checkGroup(ctx, group);
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        const int64_t b0 = ctx->groupOffset[group], n = ctx->groupOffset[group + 1] - b0;
        if (n == 0) return;
        // fetch(src, width, offset, dst): 3 doubles at `offset` of each width-double slot record
        auto fetch = [&](const DevBuf<double> &src, int width, int offset, double *dst) {
            if (!dst) return;
            std::vector<double> buf(width * ctx->nb);
            hipCheck(hipMemcpyAsync(buf.data(), src.p, buf.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
                     "download");
            hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
            for (int64_t i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) dst[3 * i + d] = buf[width * ctx->slotOf[b0 + i] + offset + d];
        };
        fetch(ctx->pose, kPose, 0, x);
        fetch(ctx->th, 3, 0, theta);
        fetch(ctx->v, 3, 0, v);
        fetch(ctx->w, 3, 0, omega);
    });
}

int spg_rigid_set_state(spg_rigid_ctx *ctx, int32_t group, const double *x, const double *theta, const double *v,
                        const double *omega)
{
    return guarded(ctx, [&] {
        checkGroup(ctx, group);
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        const int64_t b0 = ctx->groupOffset[group], n = ctx->groupOffset[group + 1] - b0;
        for (const double *p : {x, theta, v, omega})
            if (p && !finiteN(p, 3 * n)) throw VbdError(SPG_VBD_ERR_NONFINITE, "non-finite state");
        if (n == 0) return;
        const size_t nb = static_cast<size_t>(ctx->nb);
        auto sync = [&] { hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize"); };
        // edit(dst, width, fill): read-modify-write of the slot records of this group
        auto edit = [&](DevBuf<double> &dst, int width, auto &&fill) {
            std::vector<double> buf(width * nb);
            hipCheck(hipMemcpyAsync(buf.data(), dst.p, buf.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
                     "download");
            sync();
            for (int64_t i = 0; i < n; ++i) fill(i, buf.data() + width * ctx->slotOf[b0 + i]);
            hipCheck(hipMemcpyAsync(dst.p, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream),
                     "upload");
            sync();
        };
        auto copy3 = [](const double *src) {
            return [src](int64_t i, double *rec) { std::copy(src + 3 * i, src + 3 * i + 3, rec); };
        };
        if (v) edit(ctx->v, 3, copy3(v));
        if (omega) edit(ctx->w, 3, copy3(omega));
        if (theta) edit(ctx->th, 3, copy3(theta));
        if (x || theta)  // setPositions: the rotation matrix follows theta (world inertias derive from it)
            edit(ctx->pose, kPose, [&](int64_t i, double *rec) {
                if (x) std::copy(x + 3 * i, x + 3 * i + 3, rec);
                if (theta) aaToR(theta + 3 * i, rec + 3);
            });
    });
}

int spg_rigid_get_rotations(spg_rigid_ctx *ctx, int32_t group, double *R)
{
    return guarded(ctx, [&] {
        checkGroup(ctx, group);
        const int64_t b0 = ctx->groupOffset[group], n = ctx->groupOffset[group + 1] - b0;
        if (n == 0) return;
        if (!R) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        std::vector<double> buf(kPose * ctx->nb);
        hipCheck(hipMemcpyAsync(buf.data(), ctx->pose.p, buf.size() * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream),
                 "download");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        for (int64_t i = 0; i < n; ++i) {
            const double *rec = buf.data() + kPose * ctx->slotOf[b0 + i];
            std::copy(rec + 3, rec + kPose, R + 9 * i);
        }
    });
}

int spg_rigid_set_rotations(spg_rigid_ctx *ctx, int32_t group, const double *R)
{
    return guarded(ctx, [&] {
        checkGroup(ctx, group);
        const int64_t b0 = ctx->groupOffset[group], n = ctx->groupOffset[group + 1] - b0;
        if (n == 0) return;
        if (!R) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        if (!finiteN(R, 9 * n)) throw VbdError(SPG_VBD_ERR_NONFINITE, "non-finite state");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        std::vector<double> buf(kPose * ctx->nb);
        hipCheck(hipMemcpyAsync(buf.data(), ctx->pose.p, buf.size() * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream),
                 "download");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        for (int64_t i = 0; i < n; ++i) std::copy(R + 9 * i, R + 9 * i + 9, buf.data() + kPose * ctx->slotOf[b0 + i] + 3);
        hipCheck(hipMemcpyAsync(ctx->pose.p, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice,
                                ctx->stream),
                 "upload");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int spg_rigid_get_colors(const spg_rigid_ctx *ctx, int32_t group, int32_t *colors, int32_t *n_colors)
{
    return guarded(const_cast<spg_rigid_ctx *>(ctx), [&] {
        checkGroup(ctx, group);
        const auto &cg = ctx->colors[group];
        if (colors) std::copy(cg.begin(), cg.end(), colors);
        if (n_colors) {
            int32_t nc = 0;
            for (auto c : cg) nc = std::max(nc, c + 1);
            *n_colors = nc;
        }
    });
}

int spg_rigid_get_stats(spg_rigid_ctx *ctx, spg_rigid_stats *st)
{
    return guarded(ctx, [&] {
        if (!ctx || !st) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        unsigned long long sk = 0;
        hipCheck(hipMemcpyAsync(&sk, ctx->skipped.p, sizeof(sk), hipMemcpyDeviceToHost, ctx->stream), "download");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        st->n_colors = ctx->nColors;
        st->n_bodies = ctx->nb;
        st->n_elements = ctx->nElem;
        st->incidence_entries = ctx->incEntries;
        st->skipped_blocks = static_cast<int64_t>(sk);
        st->solve_bytes = ctx->solveBytes;
    });
}

int spg_rigid_element_derivatives(int32_t type, int64_t n, const double *com, const double *rot,
                                  const double *params, double *grad, double *hess_blocks, int32_t device)
{
    return guarded(nullptr, [&] {
        if (type < kRbSpring || type > kRbAlignment) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "unknown type");
        if (n < 0 || (n > 0 && (!com || !rot || !params || !grad || !hess_blocks)))
            throw VbdError(SPG_VBD_ERR_INVALID, "invalid arguments");
        if (n == 0) return;
        const int ns = rbStencil(type), np = rbParams(type);
        if (device < 0) {
            for (int64_t i = 0; i < n; ++i) {
                double p[kRbParamStride] = {};
                for (int q = 0; q < np; ++q) p[q] = params[np * i + q];
                if (type == kRbSpringSquared) p[6] = p[6] * p[6];
                for (int sl = 0; sl < ns; ++sl) {
                    double f[6] = {}, B[36] = {};
                    const double *c0 = com + 3 * ns * i, *r0 = rot + 9 * ns * i;
                    rbAccumulate(type, sl, c0, r0, ns > 1 ? c0 + 3 : c0, ns > 1 ? r0 + 9 : r0, p, f, B);
                    for (int d = 0; d < 6; ++d) grad[6 * ns * i + 6 * sl + d] = -f[d];
                    std::copy(B, B + 36, hess_blocks + 36 * (ns * i + sl));
                }
            }
            return;
        }
        int ndev = 0;
        hipCheck(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        if (device >= ndev) throw VbdError(SPG_VBD_ERR_INVALID, "invalid HIP device");
        hipCheck(hipSetDevice(device), "hipSetDevice");
        DevBuf<double> dc, dr, dp, dg, dh;
        dc.alloc(3 * ns * n);
        dr.alloc(9 * ns * n);
        dp.alloc(np * n);
        dg.alloc(6 * ns * n);
        dh.alloc(36 * ns * n);
        hipCheck(hipMemcpy(dc.p, com, 3 * ns * n * sizeof(double), hipMemcpyHostToDevice), "upload");
        hipCheck(hipMemcpy(dr.p, rot, 9 * ns * n * sizeof(double), hipMemcpyHostToDevice), "upload");
        hipCheck(hipMemcpy(dp.p, params, np * n * sizeof(double), hipMemcpyHostToDevice), "upload");
        hipLaunchKernelGGL(rbElementKernel, dim3(blocksFor(n)), dim3(kRbBlock), 0, nullptr, type, n, dc.p, dr.p, dp.p,
                           dg.p, dh.p);
        hipCheck(hipGetLastError(), "element launch");
        hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hipCheck(hipMemcpy(grad, dg.p, 6 * ns * n * sizeof(double), hipMemcpyDeviceToHost), "download");
        hipCheck(hipMemcpy(hess_blocks, dh.p, 36 * ns * n * sizeof(double), hipMemcpyDeviceToHost), "download");
    });
}

}  // extern "C"
