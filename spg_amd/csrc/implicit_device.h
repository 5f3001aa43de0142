// Device-side layout and launch interface of the implicit Euler family (global force, stiffness matrix,
// conjugate gradient).
//
// Layout in HBM (one context, all objects concatenated, particle-major DOFs as
// ParticleGroup::getPositions, spg/source/sim/simObject/particleGroup.cpp:7-14):
//   * x, v: FP64, 3 per particle.
//   * Per energy type an element table: global particle ids (size per stencil) and the prepared FP64
//     parameters of spg_vbd_energy.
//   * Element derivatives: per element its gradient (3N doubles) and Hessian (3N x 3N, column-major:
//     entry (r, c) at c * 3N + r), element-major.
//   * Global matrices are symmetric 3x3 block-sparse matrices stored by their upper blocks only: row i
//     holds the blocks (i, j), j >= i, of the particles sharing a stencil with i, columns ascending, the
//     diagonal block first; a block is 9 doubles, row-major. A lower block (i, j), j < i, is the
//     transpose of the upper block (j, i): each row also lists those (upper index and j, ascending j), so
//     a product reads every stored block twice, the second time from L2 (the rows that hold it were
//     streamed just before by the same XCD). The Hessian H is assembled per upper block from a
//     contribution list (element Hessian sub-blocks, in energy order and ascending stencil), the diagonal
//     blocks symmetrised, so H is exactly symmetric. K = -H is the reference's stiffness matrix
//     (implicitEulerBase.cpp:180-222); system matrices are a M + c H.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spgvbd {

struct ImplicitElems {
    int type;              // SPG_VBD_* energy type (1 ... 13)
    int64_t n;
    const int32_t *idx;    // n * size global particle ids
    const double *prm;     // n * params
    double *grad;          // n * 3 size
    double *hess;          // n * (3 size)^2, column-major per element
};

struct SymBsr {
    int64_t rows;          // particles
    const int64_t *upPtr;  // rows + 1: upper blocks of row i are [upPtr[i], upPtr[i + 1]), the first is (i, i)
    const int32_t *upCol;  // column of every upper block
    const int64_t *loPtr;  // rows + 1: lower blocks of row i are [loPtr[i], loPtr[i + 1])
    const int32_t *loIdx;  // upper block (j, i) holding lower block (i, j), j ascending
    const int32_t *loCol;  // that j
    double *val;           // 9 per upper block
};

// gradients and Hessians of all elements of one type at x
hipError_t launchElementDerivatives(const ImplicitElems &e, const double *x, hipStream_t stream);
// gradients only (e.hess unused)
hipError_t launchElementGradients(const ImplicitElems &e, const double *x, hipStream_t stream);
// f_i = m_i g - sum of the incident element gradient slots (incidence: offsets in units of 3 doubles into
// the gradient buffer, CSR over particles, energy order, ascending stencil)
hipError_t launchGatherForce(int64_t nParticles, const double *mass, double gx, double gy, double gz,
                             const int64_t *incPtr, const int64_t *incOff, const double *gradBase, double *f,
                             hipStream_t stream);
// H's upper blocks from contribution lists (offsets in doubles of element Hessian sub-block (a, b)
// origins, and the stride 3N of that element's columns packed as (offset << 4 | N3)); diagonal blocks
// (upRow[u] == upCol[u]) are symmetrised
hipError_t launchAssembleHessian(int64_t nUpper, const int32_t *upRow, const int32_t *upCol, const int64_t *contribPtr,
                                 const int64_t *contrib, const double *hessBase, double *H, hipStream_t stream);
// A = c H + a M over the stored blocks (M = diag(mass) on the diagonal blocks, diagBlock[i] = block index
// of (i, i))
hipError_t launchCombine(int64_t nBlocks, const double *H, double c, int64_t nParticles, const int64_t *diagBlock,
                         const double *mass, double a, double *A, hipStream_t stream);
// y = s * (A x) for the symmetric block matrix; partial (may be null): per-workgroup partial sums of
// x . y, spmvPartials(rows) of them
hipError_t launchSymSpmv(const SymBsr &A, const double *x, double s, double *y, double *partial, hipStream_t stream);
int spmvPartials(int64_t rows);
// the full block rows of s * A (rowPtr/col: the full pattern, diagBlock: index of (i, i) in it) into out
hipError_t launchExpandSym(const SymBsr &A, double s, const int64_t *rowPtr, const int32_t *col,
                           const int64_t *diagBlock, double *out, hipStream_t stream);

// Conjugate gradient (Eigen's algorithm with the diagonal preconditioner). Scalars live on the device:
struct CgScalars {
    double rhsNorm2, threshold, residualNorm2;
    double absNew[2];  // r.D^-1 r, by iteration parity
    int32_t done;   // 1 once converged (or the right-hand side is 0)
    int32_t iters;  // completed iterations (Eigen's convention)
};
constexpr int kCgBlocks = 512;  // workgroups of the vector kernels (fixed: deterministic reductions)
// a CG needs two partial-sum buffers: pPartial of max(2 kCgBlocks, spmvPartials(rows)) doubles (also used
// by launchCgInit) and uPartial of 2 kCgBlocks
hipError_t launchCgInit(int64_t n, const double *b, const double *invdiag, double *x, double *r, double *p,
                        double *partial, CgScalars *s, double tol, hipStream_t stream);
// iteration `iteration` (0-based, launch order): tmp = A p; alpha; x += alpha p, r -= alpha tmp;
// convergence test; beta; p = D^-1 r + beta p. Three kernels. Non-null events are recorded right before and
// after the product (to time it inside a solve).
hipError_t launchCgIteration(const SymBsr &A, int64_t n, const double *invdiag, double *x, double *r, double *p,
                             double *tmp, double *pPartial, double *uPartial, CgScalars *s, int32_t maxIters,
                             int32_t iteration, hipStream_t stream, hipEvent_t spmvStart = nullptr,
                             hipEvent_t spmvEnd = nullptr);
// invdiag = 1 / diagonal of A (1 where it is 0), from the diagonal blocks
hipError_t launchInvDiag(int64_t nParticles, const int64_t *diagBlock, const double *A, double *invdiag,
                         hipStream_t stream);

// out = a x + b y + c z (y, z may be null)
hipError_t launchLinComb(int64_t n, double *out, double a, const double *x, double b, const double *y, double c,
                         const double *z, hipStream_t stream);
// out_d = a * m_{d/3} * x_d + b * y_d (mass per particle; y may be null)
hipError_t launchMassLinComb(int64_t n, double *out, double a, const double *mass, const double *x, double b,
                             const double *y, hipStream_t stream);
// deterministic |x|^2 into *out (device)
hipError_t launchNorm2(int64_t n, const double *x, double *partial, double *out, hipStream_t stream);

}  // namespace spgvbd
