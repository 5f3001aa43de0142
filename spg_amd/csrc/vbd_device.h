// Device-side data layout of a VBD context and the kernel launch interface.
//
// Layout in HBM (one context, all objects concatenated):
//   * Vertex slots are numbered colour-major: colour 0 of object 0, colour 0 of object 1, ...,
//     colour 1 of object 0, ... with each (object, colour) run in ascending particle index, the order
//     of the reference's colour groups (spg/source/sim/solver/vbd.cpp:220-223). A colour sweep is one
//     contiguous slot range, so the vertex solve writes positions with unit stride.
//   * Per-vertex state: float4 streams pos (x), xt (inertial target, w = mass), xold, vel.
//   * Per energy type, an incidence stream in vertex-slot order: each incidence (vertex, element)
//     carries its own copy of the element record, rewritten relative to the owning vertex, so the
//     solve reads its element data as contiguous SoA streams; row pointers give each vertex's range.
//     Within a vertex the order is the reference's (energy in addEnergy order, ascending stencil,
//     vbd.cpp:250-259).
//   * FP64 element tables (vertex slots + exact prepared parameters) and per-incidence element ids,
//     read only by the FP64 residual kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spgvbd {

// Workgroups are dealt round-robin over the 8 XCDs of the chip. This renumbering gives XCD x the
// x-th contiguous share of the logical block range, so blocks that touch neighbouring data share
// that XCD's L2.
__device__ __forceinline__ int32_t xcdContiguousBlock(int32_t b, int32_t nb)
{
    // XCDs below r = nb mod 8 hold q + 1 blocks, the others q: XCD x starts at x q + min(x, r) (branch-free,
    // so the kernel's argument loads are not held behind a branch)
    const int32_t xcd = b & 7, i = b >> 3, q = nb >> 3, r = nb & 7;
    return xcd * q + min(xcd, r) + i;
}

enum EnergyType { kSpring = 1, kStvk = 2, kBending = 3, kSnh = 4, kAnchor = 5 };

// Incidence rows of the solve. Dense classes use fixed-width rows: the j-th incidence of vertex slot v
// is record v * width + j, and rows shorter than the width are padded with records whose first
// index is -1 (their contribution is dropped). Anchors, which few vertices have, use row offsets.
// Incidence classes of the further energies in the solve: 0 squared continuum spring (type 7), 1 Baraff-Witkin
// membrane (8), 2 Choi membrane (9), 3 quadratic bending (10), 4 Baraff-Witkin bending at a hinge end (type 11,
// slot 0 or 1), 5 the same at a wing (slot 2 or 3), 6 StVK tetrahedron (12), 7 stable squared Neo-Hookean (13).
// Coefficient planes per record (float4 each):
//   0: (L0, k, +1 for slot 0 / -1 for slot 1, 0)
//   1: (Dinv 00, 01, 10, 11), (k0, k1, k2, slot)
//   2: (Dinv 00, 01, 10, 11), (k0, k1, k2, k3), (gamma_a0, gamma_a1, 0, 0)
//   3: (c0, c1, c2, c3), (keff, c of the slot, 0, 0)
//   4, 5: (theta0, keff, 1 for slot 1 or 3 else 0, 0)
//   6, 7: (Dinv 0 ... 3), (Dinv 4 ... 7), (Dinv 8, gamma of the slot), (lambda, mu, 0, 0)
constexpr int kXCls = 8;
__host__ __device__ constexpr int xclsOf(int type, int slot)
{
    switch (type) {
    case 7: return 0;
    case 8: return 1;
    case 9: return 2;
    case 10: return 3;
    case 11: return slot < 2 ? 4 : 5;
    case 12: return 6;
    default: return 7;
    }
}
__host__ __device__ constexpr int xclsPlanes(int c)
{
    return (c == 0 || c == 4 || c == 5) ? 1 : (c == 1 || c == 3) ? 2 : c == 2 ? 3 : 4;
}
__host__ __device__ constexpr int xclsStencil(int c) { return c == 0 ? 2 : c <= 2 ? 3 : 4; }
__host__ __device__ constexpr int xclsType(int c) { return c <= 4 ? c + 7 : c + 6; }  // the energy type of class c

// A hinge record of the neighbourhood-staged solve: the hinge's (theta0, keff) and its four slots as 5-bit indices
// into the vertex's neighbour row (i0 | i1 << 5 | i2 << 10 | i3 << 15; -1: padding). 12 bytes, read with one
// three-dword load.
struct NbHinge {
    float theta0, keff;
    int32_t slots;
};
static_assert(sizeof(NbHinge) == 12, "hinge records are 12 bytes");

struct SolveArgs {
    float4 *pos;
    const float4 *xt;
    // spring: (other slot bits, L0, k, 0)
    int32_t sprW;
    const float4 *spr;
    // anchor: row offsets, (ax, ay, az, k)
    const int32_t *ancPtr;
    const float4 *anc;
    // StVK: other slots, (g_a, g_b), (kd, ko, ks, 0)
    int32_t triW;
    const int2 *triO;
    const float4 *triG;
    const float4 *triM;
    // bending, split by the owner's slot: hinge-edge (slots 0, 1) and hinge-wing (slots 2, 3)
    // incidences; canonical slots v0..v3, (theta0, keff)
    int32_t hinEW;
    const int4 *hinEV;
    const float2 *hinEP;
    int32_t hinWW;
    const int4 *hinWV;
    const float2 *hinWP;
    // SNH: other slots (3), (g_a, lambda), (g_b, mu), (g_c, alpha)
    int32_t tetW;
    const int4 *tetO;
    const float4 *tetGa;
    const float4 *tetGb;
    const float4 *tetGc;
    // the further energies (types 7 ... 13; continuum springs join the spring class): per-slot incidence counts
    // as CSR offsets (extPtr; null: no further energies, and the kernels' instantiations without these classes
    // run), and element tables (up to four slots, unused -1, and kExtParams FP64 parameters) read by the FP64
    // residual
    const int32_t *extPtr;
    const int4 *extV;
    const double *extP;
    // the solve's records of the further energies, one fixed-width row per vertex slot and incidence class
    // (xcls*): record j of slot v at v * xrW[c] + j holds the element's slots in stencil order (xrV; padding
    // records -1) and xclsPlanes(c) float4 of FP32 coefficients (plane i at xrP[c] + i * xrPlane[c]) that include
    // what depends on the incidence's own slot in the element
    int32_t xrW[kXCls];  // 0: class absent
    const int4 *xrV[kXCls];
    const float4 *xrP[kXCls];
    int64_t xrPlane[kXCls];
    unsigned long long *counters;  // [0] FP64 fallbacks, [1] skipped blocks
    int32_t begin, end;            // slot range of the colour
    float invdt2;
    // neighbourhood layout (lds_staging 2): per vertex slot a row of nbW slots (entry 0 the vertex itself, then
    // its distinct neighbours in order of first use, -1 padding), and records that name their vertices by 5-bit
    // index into that row: StVK (kd, ko, ks, a | b << 5), hinges NbHinge, SNH a | b << 5 | c << 10 (with
    // tetGa..c), spring (index, L0, k, 0); padding records -1.
    // nbW is 0 when the layout does not exist (a vertex with more than 31 neighbours, or further energies).
    int32_t nbW;
    const int32_t *nbRow;
    const float4 *triL;
    const NbHinge *hinEL;
    const NbHinge *hinWL;
    const int32_t *tetL;
    const float4 *sprL;
    // launch shape, set by the launcher: the kernel reads it with its other arguments instead of from the
    // dispatch packet, so its prologue waits for one round of scalar loads
    int32_t gridBlocks, wavesPerBlock;
};
constexpr int kExtParams = 12;  // parameters stored per further-energy element (the largest needs 11)

struct ResidualArgs {
    const float4 *pos;
    const float4 *xt;
    const double *mass;  // exact FP64 masses per slot
    double invdt2;
    int32_t n;
    // per type: incidence row pointers (shared with the solve) and element ids; element tables
    const int32_t *sprPtr;
    const int32_t *sprElem;
    const int2 *sprV;
    const double2 *sprP;  // (L0, k)
    const int32_t *ancPtr;
    const int32_t *ancElem;
    const double4 *ancP;  // (a, k)
    const int32_t *triPtr;
    const int32_t *triElem;
    const int4 *triV;
    const double *triP;   // 8 per element: Dinv (4), kd, ko, ks, 0
    const int32_t *hinEPtr;
    const int32_t *hinEElem;
    const int32_t *hinWPtr;
    const int32_t *hinWElem;
    const int4 *hinV;
    const double2 *hinP;  // (theta0, keff)
    const int32_t *tetPtr;
    const int32_t *tetElem;
    const int4 *tetV;
    const double *tetP;   // 12 per element: Dinv (9), lambda, mu, alpha
    const int32_t *extPtr;  // the further energies, as in SolveArgs (null: none)
    const int4 *extV;
    const double *extP;
    // their rows in the solve's class layout: per class, vertex slot and entry (element id, the vertex's slot in
    // the element); -1 padding
    int32_t xrW[kXCls];
    const int2 *xrR[kXCls];
    const int32_t *order; // slot of the i-th particle (objects in order, particles ascending)
    double *r2;           // out: |r_v|^2 of the i-th particle (particle order)
    // fixed-width rows of element ids per vertex slot and class, in the solve's widths (-1: padding)
    int32_t triW, hinEW, hinWW, tetW, sprW;
    const int32_t *triRow;
    const int32_t *hinERow;
    const int32_t *hinWRow;
    const int32_t *tetRow;
    const int32_t *sprRow;
    double *hinKc;  // work: per hinge k (theta - theta0) at `pos` (hingeKcKernel, launched first)
    int32_t nHinges;
    // neighbourhood-staged variant (nbW > 0, no further energies): the solve's neighbour rows, and per class
    // row entries (element id, 5-bit indices into the vertex's neighbour row of the element's slots in stencil
    // order; springs: of the other end) in the same widths (padding -1)
    int32_t nbW;
    const int32_t *nbRow;
    const int2 *triRL;
    const int2 *hinERL;
    const int2 *hinWRL;
    const int2 *tetRL;
    const int2 *sprRL;
};

// kernel launchers (vbd_kernels.hip); all asynchronous on `stream`
hipError_t launchBeginSubstep(float4 *pos, float4 *xold, float4 *vel, float4 *xt, const float *mass, int32_t n,
                              float h, float gx, float gy, float gz, hipStream_t stream);
// v = (x - x_old) / h and corr = x - x~
hipError_t launchEndSubstep(const float4 *pos, const float4 *xold, const float4 *xt, float4 *vel, float4 *corr,
                            int32_t n, float invh, hipStream_t stream);
// LDS staging of the element-record rows is available while one workgroup's rows fit this many bytes
constexpr size_t kSolveMaxLds = 65536;
// LDS bytes per workgroup of the staged variants (lds_staging 1: element rows, 2: neighbourhood slabs; 0: none)
size_t solveStagedLdsBytes(const SolveArgs &a, int lanesPerVertex, int blockSize, int ldsStaging);
// start/stop: optional events stamped with the dispatch's own begin/end (nullptr: untimed launch)
hipError_t launchSolveColor(const SolveArgs &a, int lanesPerVertex, int blockSize, int ldsStaging,
                            hipStream_t stream, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
hipError_t launchResidual(const ResidualArgs &a, hipStream_t stream);
// entries of one object's slot list summed by one block of the per-object residual reduction
constexpr int64_t kReduceSegment = 4096;
// r2 in particle order; segment k sums r2[segBegin[k] .. segBegin[k + 1])
hipError_t launchReduceObjects(const double *r2, const int64_t *segBegin, int32_t nSegments, const int32_t *objSeg,
                               int32_t nObjects, double *partial, double *out, hipStream_t stream);
hipError_t launchCopyXt(float4 *pos, const float4 *xt, int32_t n, hipStream_t stream);
hipError_t launchCopy(const float4 *src, float4 *dst, int64_t n, hipStream_t stream);
hipError_t launchExtElementDerivatives(int type, int64_t n, const float4 *pos, const double *prm, int np,
                                       double *grad, double *hess, hipStream_t stream);
hipError_t launchElementDerivatives(int type, int64_t n, const double *xs, const double *params, int precision,
                                    int form, double *grad, double *hess, hipStream_t stream);

}  // namespace spgvbd
