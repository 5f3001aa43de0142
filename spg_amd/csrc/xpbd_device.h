// Device-side data layout of an XPBD context and the kernel launch interface.
//
// Layout in HBM (one context, all objects concatenated, particles in object order, ascending index):
//   * Per particle: float4 pos (xyz, w = inverse mass), vel and xold (substep start).
//   * Per energy type, the stencil records in launch order: sorted by (energy position in its object,
//     group, object, stencil index), where a group is a stencil colour (parallel Gauss-Seidel,
//     xpbd.cpp:134-160) or a serial level (see coloring.h). One projection launch covers one
//     contiguous record range, so the records stream with unit stride; within a launch no two
//     stencils share a particle, so every particle of the range is read and written once.
//     Records (SoA, FP32, compliance as stored by effectiveCompliance(), divided by dt^2 in the kernel):
//       spring   int2 particles, float2 (L0, compliance)                              16 B
//       anchor   int particle, float4 (anchor xyz, compliance)                         20 B
//       membrane int4 particles, float4 Dinv, float4 (c00 c01 c02 c11), float2 (c12 c22)  56 B
//       bending  int4 particles, float2 (theta0, compliance)                          24 B
//       SNH      int4 particles, float4 x3 (Dinv 9, c00 c01 c11), float hydrostatic term  68 B
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spgvbd {

struct XpbdArgs {
    float4 *pos;
    const int2 *sprI;
    const float2 *sprP;
    const int32_t *ancI;
    const float4 *ancP;
    const int4 *triI;
    const float4 *triD;
    const float4 *triC0;
    const float2 *triC1;
    const int4 *hinI;
    const float2 *hinP;
    const int4 *tetI;
    const float4 *tetD0;  // D00 D01 D02 D10
    const float4 *tetD1;  // D11 D12 D20 D21
    const float4 *tetD2;  // D22 c00 c01 c11
    const float *tetH;    // hydrostatic correction term
    // further energies (types 6 ... 13): slots (unused -1) and 12 parameters in the spg_xpbd_energy layout
    const int4 *extI;
    const float4 *extP;  // 3 per record
    unsigned long long *counters;  // [0] projections solved in FP64 (ill-conditioned A)
    float invdt2;
    int32_t gridBlocks, blockThreads;  // launch shape, set by the launcher
};

// Consecutive small stencil groups run by one workgroup (xpbdProjectRunKernel), in order.
constexpr int kXpbdRunMax = 8;          // groups per run
constexpr int kXpbdRunThreads = 256;    // threads of the run's workgroup
constexpr int32_t kXpbdSmallGroup = 1024;  // groups of at most this many stencils join runs
struct XpbdRun {
    int32_t count;
    int32_t type[kXpbdRunMax];
    int32_t begin[kXpbdRunMax];
    int32_t end[kXpbdRunMax];
};

constexpr int kXpbdNumTypes = 14;  // type ids 1 ... 13
// compulsory bytes of one stencil record (see the layout above)
constexpr double kXpbdRecordBytes[kXpbdNumTypes] = {0, 16, 56, 24, 68, 20, 64, 64, 64, 64, 64, 64, 64, 64};

// One projection launch over records [begin, end) of `type` (1 spring, 2 membrane, 3 bending, 4 SNH,
// 5 anchor, 6 ... 13 the further energies). start/stop: optional events stamped with the dispatch's own begin/end.
hipError_t launchXpbdProject(int type, const XpbdArgs &a, int32_t begin, int32_t end, int blockSize,
                             hipStream_t stream, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
// One workgroup running the groups of `run` in order with a workgroup barrier between them.
hipError_t launchXpbdProjectRun(const XpbdArgs &a, const XpbdRun &run, hipStream_t stream, hipEvent_t start = nullptr,
                                hipEvent_t stop = nullptr);
// Substep boundaries (xpbd.cpp:47-70,110-121): begin: xold = x, v += h g, x += h v;
// between two substeps: v = (x - xold) / h_prev, then begin, with v kept in registers (vel is not written);
// end: v = (x - xold) / h. g arguments: h g.
hipError_t launchXpbdBegin(float4 *pos, float4 *vel, float4 *xold, int32_t n, float h, float gx, float gy, float gz,
                           hipStream_t stream);
hipError_t launchXpbdBoundary(float4 *pos, float4 *xold, int32_t n, float invhPrev, float h, float gx, float gy,
                              float gz, hipStream_t stream);
hipError_t launchXpbdEnd(const float4 *pos, const float4 *xold, float4 *vel, int32_t n, float invh,
                         hipStream_t stream);
// Check hook: n independent projections of one type from FP64 inputs (xs: n * size * 3, w: n * size,
// params in the C-ABI layout), computed in float (precision 0) or double (1); dx: n * size * 3.
hipError_t launchXpbdProjectCheck(int type, int64_t n, const double *xs, const double *w, const double *params,
                                  double invdt2, int precision, double *dx, hipStream_t stream);

}  // namespace spgvbd
