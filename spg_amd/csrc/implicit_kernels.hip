// CDNA4 kernels of the implicit Euler family: FP64 element derivatives, force gather, block-sparse
// Hessian assembly, BSR SpMV and the conjugate gradient, all with fixed summation orders (the results
// do not depend on scheduling).
#include <hip/hip_runtime.h>

#include <cfloat>

#include "implicit_device.h"
#include "implicit_element.h"

namespace spgvbd {

namespace {

inline int blocksFor(int64_t n, int bs) { return static_cast<int>((n + bs - 1) / bs); }

// One element per thread: gradient (tangent-free values of the first pass) and the Hessian one column
// per forward-mode pass, written straight to the element's column-major slot. One instantiation per
// energy type keeps each kernel's registers to that energy's closed form.
template <int TYPE>
__global__ void __launch_bounds__(128) elementDerivKernel(ImplicitElems e, const double *__restrict__ x)
{
    constexpr int NS = fullStencilSize(TYPE);
    constexpr int N = 3 * NS;
    constexpr int NP = fullParams(TYPE);
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= e.n) return;
    double p[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) p[k] = e.prm[i * NP + k];
    double xs[N];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int64_t v = e.idx[i * NS + s];
        xs[3 * s] = x[3 * v];
        xs[3 * s + 1] = x[3 * v + 1];
        xs[3 * s + 2] = x[3 * v + 2];
    }
    double *H = e.hess + i * N * N;
    double *g = e.grad + i * N;
#pragma unroll 1
    for (int j = 0; j < N; ++j) {
        V3<Dual1> xd[NS], gd[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s)
            xd[s] = {Dual1(xs[3 * s], 3 * s == j ? 1.0 : 0.0), Dual1(xs[3 * s + 1], 3 * s + 1 == j ? 1.0 : 0.0),
                     Dual1(xs[3 * s + 2], 3 * s + 2 == j ? 1.0 : 0.0)};
        fullGrad<NS, Dual1>(TYPE, xd, p, gd);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            H[j * N + 3 * s] = gd[s].x.d;
            H[j * N + 3 * s + 1] = gd[s].y.d;
            H[j * N + 3 * s + 2] = gd[s].z.d;
        }
        if (j == 0)
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                g[3 * s] = gd[s].x.v;
                g[3 * s + 1] = gd[s].y.v;
                g[3 * s + 2] = gd[s].z.v;
            }
    }
}

// Gradient only (no Hessian), in plain FP64: the force of the line-search trials.
template <int TYPE>
__global__ void __launch_bounds__(256) elementGradKernel(ImplicitElems e, const double *__restrict__ x)
{
    constexpr int NS = fullStencilSize(TYPE);
    constexpr int NP = fullParams(TYPE);
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= e.n) return;
    double p[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) p[k] = e.prm[i * NP + k];
    V3<double> xd[NS], gd[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int64_t v = e.idx[i * NS + s];
        xd[s] = {x[3 * v], x[3 * v + 1], x[3 * v + 2]};
    }
    fullGrad<NS, double>(TYPE, xd, p, gd);
    double *g = e.grad + i * 3 * NS;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        g[3 * s] = gd[s].x;
        g[3 * s + 1] = gd[s].y;
        g[3 * s + 2] = gd[s].z;
    }
}

// Workgroups are dealt round-robin over the 8 XCDs; this renumbering gives each XCD a contiguous band of
// the logical range, so neighbouring particles (which share elements and blocks) share that XCD's L2.
__device__ __forceinline__ int32_t xcdContiguous(int32_t b, int32_t nb)
{
    const int32_t xcd = b & 7, i = b >> 3, q = nb >> 3, r = nb & 7;
    return xcd < r ? xcd * (q + 1) + i : r * (q + 1) + (xcd - r) * q + i;
}

__global__ void __launch_bounds__(256) gatherForceKernel(int64_t n, const double *__restrict__ mass, double gx,
                                                         double gy, double gz, const int64_t *__restrict__ incPtr,
                                                         const int64_t *__restrict__ incOff,
                                                         const double *__restrict__ grad, double *__restrict__ f)
{
    const int64_t i = xcdContiguous(blockIdx.x, gridDim.x) * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const double m = mass[i];
    double fx = gx * m, fy = gy * m, fz = gz * m;
    for (int64_t k = incPtr[i]; k < incPtr[i + 1]; ++k) {
        const double *g = grad + 3 * incOff[k];
        fx -= g[0];
        fy -= g[1];
        fz -= g[2];
    }
    f[3 * i] = fx;
    f[3 * i + 1] = fy;
    f[3 * i + 2] = fz;
}

__global__ void __launch_bounds__(256) assembleHessianKernel(int64_t nUpper, const int32_t *__restrict__ upRow,
                                                             const int32_t *__restrict__ upCol,
                                                             const int64_t *__restrict__ contribPtr,
                                                             const int64_t *__restrict__ contrib,
                                                             const double *__restrict__ hess, double *__restrict__ H)
{
    const int64_t u = xcdContiguous(blockIdx.x, gridDim.x) * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (u >= nUpper) return;
    double b[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t k = contribPtr[u]; k < contribPtr[u + 1]; ++k) {
        const int64_t c = contrib[k];
        const double *h = hess + (c >> 4);
        const int N3 = static_cast<int>(c & 15);
        for (int cc = 0; cc < 3; ++cc)
            for (int r = 0; r < 3; ++r) b[r * 3 + cc] += h[cc * N3 + r];
    }
    double *out = H + 9 * u;
    if (upRow[u] == upCol[u]) {  // diagonal block: symmetric part
        for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) out[r * 3 + cc] = 0.5 * (b[r * 3 + cc] + b[cc * 3 + r]);
        return;
    }
    for (int k = 0; k < 9; ++k) out[k] = b[k];
}

__global__ void __launch_bounds__(256) combineKernel(int64_t nBlocks, const double *__restrict__ H, double c,
                                                     double *__restrict__ A)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= 9 * nBlocks) return;
    A[k] = c * H[k];
}

__global__ void __launch_bounds__(256) addMassKernel(int64_t n, const int64_t *__restrict__ diagBlock,
                                                     const double *__restrict__ mass, double a, double *__restrict__ A)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    double *d = A + 9 * diagBlock[i];
    const double am = a * mass[i];
    d[0] += am;
    d[4] += am;
    d[8] += am;
}

__device__ __forceinline__ double blockSumFixed(double v, double *sh)
{
    // wave-level fixed butterfly, then the waves' sums in order
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) s += sh[w];
    __syncthreads();
    return s;
}

// Symmetric block-row SpMV: kSpmvG lanes per block row; lane l takes the row's upper blocks l, l + G, ...
// whole (9 values, the column, the 3 x entries), then its lower blocks l, l + G, ... as transposes of the
// upper blocks that hold them, then a fixed butterfly over the G lanes. Workgroups are renumbered so
// each XCD streams a contiguous band of rows: the upper blocks a row reads transposed belong to rows a
// band streamed just before. Eight lanes per row (more than a cloth row's four upper blocks) keep fewer rows
// in flight per XCD, so those blocks are more often still in its L2 when they are read again: 40.5 us per
// product at 512x512 against 46.3 us with four lanes (two: 61.2, sixteen: 54.8).
constexpr int kSpmvG = 8;
constexpr int kSpmvBlock = 1024;

__global__ void __launch_bounds__(kSpmvBlock) symSpmvKernel(SymBsr A, const double *__restrict__ x, double s,
                                                           double *__restrict__ y, double *__restrict__ partial,
                                                           const int32_t *__restrict__ done)
{
    __shared__ double sh[kSpmvBlock / 64];
    if (done && *done) return;  // no other kernel runs meanwhile: the same value in every thread
    const int32_t blk = xcdContiguous(static_cast<int32_t>(blockIdx.x), static_cast<int32_t>(gridDim.x));
    const int64_t t = blk * static_cast<int64_t>(kSpmvBlock) + threadIdx.x;
    const int64_t row = t / kSpmvG;
    const int lig = static_cast<int>(t % kSpmvG);
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    if (row < A.rows) {
        const int64_t u1 = A.upPtr[row + 1];
        for (int64_t k = A.upPtr[row] + lig; k < u1; k += kSpmvG) {
            const double *v = A.val + 9 * k;
            const int64_t c = 3 * static_cast<int64_t>(A.upCol[k]);
            const double x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
            y0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
            y1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
            y2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
        }
        const int64_t l1 = A.loPtr[row + 1];
        for (int64_t k = A.loPtr[row] + lig; k < l1; k += kSpmvG) {
            const double *v = A.val + 9 * static_cast<int64_t>(A.loIdx[k]);
            const int64_t c = 3 * static_cast<int64_t>(A.loCol[k]);
            const double x0 = x[c], x1 = x[c + 1], x2 = x[c + 2];
            y0 += v[0] * x0 + v[3] * x1 + v[6] * x2;
            y1 += v[1] * x0 + v[4] * x1 + v[7] * x2;
            y2 += v[2] * x0 + v[5] * x1 + v[8] * x2;
        }
    }
#pragma unroll
    for (int o = kSpmvG / 2; o > 0; o >>= 1) {
        y0 += __shfl_xor(y0, o, kSpmvG);
        y1 += __shfl_xor(y1, o, kSpmvG);
        y2 += __shfl_xor(y2, o, kSpmvG);
    }
    double dot = 0.0;
    if (row < A.rows && lig == 0) {
        const double z0 = s * y0, z1 = s * y1, z2 = s * y2;
        y[3 * row] = z0;
        y[3 * row + 1] = z1;
        y[3 * row + 2] = z2;
        dot = x[3 * row] * z0 + x[3 * row + 1] * z1 + x[3 * row + 2] * z2;
    }
    if (partial) {  // per-workgroup partial sums, in row order
        const double b = blockSumFixed(dot, sh);
        if (threadIdx.x == 0) partial[blk] = b;
    }
}

// Full block rows of s * A from the upper storage: one thread per block row.
__global__ void __launch_bounds__(256) expandSymKernel(SymBsr A, double s, const int64_t *__restrict__ rowPtr,
                                                       const int32_t *__restrict__ col,
                                                       const int64_t *__restrict__ diagBlock, double *__restrict__ out)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= A.rows) return;
    for (int64_t k = rowPtr[i]; k < rowPtr[i + 1]; ++k) {
        const int32_t j = col[k];
        double *o = out + 9 * k;
        if (j >= i) {
            const double *v = A.val + 9 * (A.upPtr[i] + (k - diagBlock[i]));
            for (int e = 0; e < 9; ++e) o[e] = s * v[e];
            continue;
        }
        // upper block (j, i): binary search of column i in row j's upper blocks
        int64_t lo = A.upPtr[j], hi = A.upPtr[j + 1];
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) / 2;
            if (A.upCol[mid] <= i) lo = mid;
            else hi = mid;
        }
        const double *v = A.val + 9 * lo;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) o[r * 3 + c] = s * v[c * 3 + r];
    }
}

// Fixed-order partial sums: block b owns a contiguous chunk; its threads sum strided, then a fixed
// block reduction.
template <class F>
__device__ __forceinline__ double chunkSum(int64_t n, F term, double *sh)
{
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b = blockIdx.x * chunk, e = b + chunk < n ? b + chunk : n;
    double acc = 0.0;
    for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) acc += term(i);
    return blockSumFixed(acc, sh);
}

__device__ __forceinline__ double finalSum(const double *__restrict__ partial, int n, double *sh)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
    return blockSumFixed(acc, sh);
}

__global__ void __launch_bounds__(256) cgInitKernel(int64_t n, const double *__restrict__ b,
                                                    const double *__restrict__ invdiag, double *__restrict__ x,
                                                    double *__restrict__ r, double *__restrict__ p,
                                                    double *__restrict__ partial)
{
    __shared__ double sh[4];
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = blockIdx.x * chunk, e0 = b0 + chunk < n ? b0 + chunk : n;
    for (int64_t i = b0 + threadIdx.x; i < e0; i += blockDim.x) {
        const double bi = b[i];
        x[i] = 0.0;
        r[i] = bi;
        p[i] = invdiag[i] * bi;
    }
    const double bb = chunkSum(n, [&](int64_t i) { return b[i] * b[i]; }, sh);
    const double rp = chunkSum(n, [&](int64_t i) { return b[i] * (invdiag[i] * b[i]); }, sh);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = bb;
        partial[gridDim.x + blockIdx.x] = rp;
    }
}

__global__ void __launch_bounds__(1024) cgInitFinalKernel(const double *__restrict__ partial, int nb, CgScalars *s,
                                                          double tol)
{
    __shared__ double sh[16];
    const double bb = finalSum(partial, nb, sh);
    const double rp = finalSum(partial + nb, nb, sh);
    if (threadIdx.x == 0) {
        s->rhsNorm2 = bb;
        s->absNew[0] = rp;
        s->absNew[1] = 0.0;
        s->iters = 0;
        s->residualNorm2 = bb;  // r = b - A 0
        const double th = tol * tol * bb;
        s->threshold = th > DBL_MIN ? th : DBL_MIN;
        s->done = (bb == 0.0 || bb < s->threshold) ? 1 : 0;
    }
}

// The sum of partial[0, n) in a fixed order, the same value in every thread (and every workgroup).
__device__ __forceinline__ double blockTotal(const double *__restrict__ partial, int n, double *sh)
{
    // thread t adds entries t, t + B, t + 2B, ... in that order; eight of its loads are issued together
    double acc = 0.0;
    const int B = static_cast<int>(blockDim.x);
    int i = threadIdx.x;
    for (; i + 7 * B < n; i += 8 * B) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = partial[i + k * B];
#pragma unroll
        for (int k = 0; k < 8; ++k) acc += v[k];
    }
    for (; i < n; i += B) acc += partial[i];
    const double t = blockSumFixed(acc, sh);
    if (threadIdx.x == 0) sh[0] = t;
    __syncthreads();
    const double v = sh[0];
    __syncthreads();
    return v;
}

// The flag read once per workgroup, so every thread of the workgroup takes the same branch.
__device__ __forceinline__ bool blockDone(const CgScalars *s, int32_t *flag)
{
    if (threadIdx.x == 0) *flag = s->done;
    __syncthreads();
    return *flag != 0;
}

// alpha = absNew / p.Ap (every workgroup reduces the product's partial sums itself, in the same order);
// x += alpha p, r -= alpha A p; per-workgroup partial sums of r.r and r.D^-1 r.
__global__ void __launch_bounds__(256) cgUpdateKernel(int64_t n, const double *__restrict__ invdiag,
                                                      double *__restrict__ x, double *__restrict__ r,
                                                      const double *__restrict__ p, const double *__restrict__ tmp,
                                                      const double *__restrict__ pPartial, int nbP,
                                                      double *__restrict__ partial, const CgScalars *s, int parity)
{
    __shared__ double sh[4];
    __shared__ int32_t flag;
    if (blockDone(s, &flag)) return;
    const double alpha = s->absNew[parity] / blockTotal(pPartial, nbP, sh);
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = blockIdx.x * chunk, e0 = b0 + chunk < n ? b0 + chunk : n;
    double rr = 0.0, rz = 0.0;
    for (int64_t i = b0 + threadIdx.x; i < e0; i += blockDim.x) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * tmp[i];
        r[i] = ri;
        rr += ri * ri;
        rz += ri * (invdiag[i] * ri);
    }
    const double a = blockSumFixed(rr, sh);
    const double c = blockSumFixed(rz, sh);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = a;
        partial[gridDim.x + blockIdx.x] = c;
    }
}

// Convergence test and beta (every workgroup reduces the update's partial sums itself; workgroup 0
// stores the scalars), then p = D^-1 r + beta p. absNew is double-buffered by iteration parity because
// the other workgroups still read the old value while workgroup 0 stores the new one. A workgroup that
// already sees the flag set by workgroup 0 skips p, which no later iteration reads.
__global__ void __launch_bounds__(256) cgDirectionKernel(int64_t n, const double *__restrict__ invdiag,
                                                         const double *__restrict__ r, double *__restrict__ p,
                                                         const double *__restrict__ uPartial, int nbU, CgScalars *s,
                                                         int32_t maxIters, int parity)
{
    __shared__ double sh[4];
    __shared__ int32_t flag;
    if (blockDone(s, &flag)) return;
    const double rr = blockTotal(uPartial, nbU, sh);
    const double rz = blockTotal(uPartial + nbU, nbU, sh);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (rr < s->threshold) {
        if (lead) {
            s->residualNorm2 = rr;
            s->done = 1;
        }
        return;
    }
    const double beta = rz / s->absNew[parity];
    if (lead) {
        s->residualNorm2 = rr;
        s->absNew[parity ^ 1] = rz;
        s->iters += 1;
        if (s->iters >= maxIters) s->done = 2;  // iteration budget spent (x is the last iterate)
    }
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = blockIdx.x * chunk, e0 = b0 + chunk < n ? b0 + chunk : n;
    for (int64_t i = b0 + threadIdx.x; i < e0; i += blockDim.x) p[i] = invdiag[i] * r[i] + beta * p[i];
}

__global__ void __launch_bounds__(256) invDiagKernel(int64_t n, const int64_t *__restrict__ diagBlock,
                                                     const double *__restrict__ A, double *__restrict__ invdiag)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const double *d = A + 9 * diagBlock[i];
    for (int k = 0; k < 3; ++k) {
        const double a = d[4 * k];
        invdiag[3 * i + k] = a != 0.0 ? 1.0 / a : 1.0;
    }
}

__global__ void __launch_bounds__(256) linCombKernel(int64_t n, double *__restrict__ out, double a,
                                                     const double *__restrict__ x, double b,
                                                     const double *__restrict__ y, double c,
                                                     const double *__restrict__ z)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    double v = a * x[i];
    if (y) v += b * y[i];
    if (z) v += c * z[i];
    out[i] = v;
}

__global__ void __launch_bounds__(256) massLinCombKernel(int64_t n, double *__restrict__ out, double a,
                                                         const double *__restrict__ mass,
                                                         const double *__restrict__ x, double b,
                                                         const double *__restrict__ y)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    double v = a * mass[i / 3] * x[i];
    if (y) v += b * y[i];
    out[i] = v;
}

__global__ void __launch_bounds__(256) norm2PartialKernel(int64_t n, const double *__restrict__ x,
                                                          double *__restrict__ partial)
{
    __shared__ double sh[4];
    const double t = chunkSum(n, [&](int64_t i) { return x[i] * x[i]; }, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void __launch_bounds__(1024) sumFinalKernel(const double *__restrict__ partial, int nb, double *out)
{
    __shared__ double sh[16];
    const double t = finalSum(partial, nb, sh);
    if (threadIdx.x == 0) *out = t;
}

}  // namespace

#define SPG_ELEMENT_CASES(KERNEL, BS)                                                                              \
    switch (e.type) {                                                                                          \
    case 1: KERNEL<1><<<grid, BS, 0, stream>>>(e, x); break;                                                   \
    case 2: KERNEL<2><<<grid, BS, 0, stream>>>(e, x); break;                                                   \
    case 3: KERNEL<3><<<grid, BS, 0, stream>>>(e, x); break;                                                   \
    case 4: KERNEL<4><<<grid, BS, 0, stream>>>(e, x); break;                                                   \
    case 5: KERNEL<5><<<grid, BS, 0, stream>>>(e, x); break;                                                   \
    case 6: KERNEL<6><<<grid, BS, 0, stream>>>(e, x); break;                                                   \
    case 7: KERNEL<7><<<grid, BS, 0, stream>>>(e, x); break;                                                   \
    case 8: KERNEL<8><<<grid, BS, 0, stream>>>(e, x); break;                                                   \
    case 9: KERNEL<9><<<grid, BS, 0, stream>>>(e, x); break;                                                   \
    case 10: KERNEL<10><<<grid, BS, 0, stream>>>(e, x); break;                                                 \
    case 11: KERNEL<11><<<grid, BS, 0, stream>>>(e, x); break;                                                 \
    case 12: KERNEL<12><<<grid, BS, 0, stream>>>(e, x); break;                                                 \
    case 13: KERNEL<13><<<grid, BS, 0, stream>>>(e, x); break;                                                 \
    default: return hipErrorInvalidValue;                                                                      \
    }

hipError_t launchElementDerivatives(const ImplicitElems &e, const double *x, hipStream_t stream)
{
    if (e.n <= 0) return hipSuccess;
    const int grid = blocksFor(e.n, 128);
    SPG_ELEMENT_CASES(elementDerivKernel, 128)
    return hipGetLastError();
}

hipError_t launchElementGradients(const ImplicitElems &e, const double *x, hipStream_t stream)
{
    if (e.n <= 0) return hipSuccess;
    const int grid = blocksFor(e.n, 256);
    SPG_ELEMENT_CASES(elementGradKernel, 256)
    return hipGetLastError();
}

#undef SPG_ELEMENT_CASES

hipError_t launchGatherForce(int64_t nParticles, const double *mass, double gx, double gy, double gz,
                             const int64_t *incPtr, const int64_t *incOff, const double *gradBase, double *f,
                             hipStream_t stream)
{
    if (nParticles <= 0) return hipSuccess;
    gatherForceKernel<<<blocksFor(nParticles, 256), 256, 0, stream>>>(nParticles, mass, gx, gy, gz, incPtr, incOff,
                                                                       gradBase, f);
    return hipGetLastError();
}

hipError_t launchAssembleHessian(int64_t nUpper, const int32_t *upRow, const int32_t *upCol, const int64_t *contribPtr,
                                 const int64_t *contrib, const double *hessBase, double *H, hipStream_t stream)
{
    if (nUpper <= 0) return hipSuccess;
    assembleHessianKernel<<<blocksFor(nUpper, 256), 256, 0, stream>>>(nUpper, upRow, upCol, contribPtr, contrib,
                                                                       hessBase, H);
    return hipGetLastError();
}

hipError_t launchCombine(int64_t nBlocks, const double *H, double c, int64_t nParticles, const int64_t *diagBlock,
                         const double *mass, double a, double *A, hipStream_t stream)
{
    if (nBlocks <= 0) return hipSuccess;
    combineKernel<<<blocksFor(9 * nBlocks, 256), 256, 0, stream>>>(nBlocks, H, c, A);
    if (nParticles > 0 && a != 0.0)
        addMassKernel<<<blocksFor(nParticles, 256), 256, 0, stream>>>(nParticles, diagBlock, mass, a, A);
    return hipGetLastError();
}

namespace {
int spmvGrid(int64_t rows) { return blocksFor(rows * kSpmvG, kSpmvBlock); }
}  // namespace

int spmvPartials(int64_t rows) { return spmvGrid(rows); }

hipError_t launchSymSpmv(const SymBsr &A, const double *x, double s, double *y, double *partial, hipStream_t stream)
{
    if (A.rows <= 0) return hipSuccess;
    symSpmvKernel<<<spmvGrid(A.rows), kSpmvBlock, 0, stream>>>(A, x, s, y, partial, nullptr);
    return hipGetLastError();
}

hipError_t launchExpandSym(const SymBsr &A, double s, const int64_t *rowPtr, const int32_t *col,
                           const int64_t *diagBlock, double *out, hipStream_t stream)
{
    if (A.rows <= 0) return hipSuccess;
    expandSymKernel<<<blocksFor(A.rows, 256), 256, 0, stream>>>(A, s, rowPtr, col, diagBlock, out);
    return hipGetLastError();
}

hipError_t launchCgInit(int64_t n, const double *b, const double *invdiag, double *x, double *r, double *p,
                        double *partial, CgScalars *s, double tol, hipStream_t stream)
{
    cgInitKernel<<<kCgBlocks, 256, 0, stream>>>(n, b, invdiag, x, r, p, partial);
    cgInitFinalKernel<<<1, 1024, 0, stream>>>(partial, kCgBlocks, s, tol);
    return hipGetLastError();
}

hipError_t launchCgIteration(const SymBsr &A, int64_t n, const double *invdiag, double *x, double *r, double *p,
                             double *tmp, double *pPartial, double *uPartial, CgScalars *s, int32_t maxIters,
                             int32_t iteration, hipStream_t stream, hipEvent_t spmvStart, hipEvent_t spmvEnd)
{
    if (spmvStart) {
        const hipError_t e = hipEventRecord(spmvStart, stream);
        if (e != hipSuccess) return e;
    }
    const int grid = spmvGrid(A.rows);
    symSpmvKernel<<<grid, kSpmvBlock, 0, stream>>>(A, p, 1.0, tmp, pPartial, &s->done);
    if (spmvEnd) {
        const hipError_t e = hipEventRecord(spmvEnd, stream);
        if (e != hipSuccess) return e;
    }
    const int parity = iteration & 1;
    cgUpdateKernel<<<kCgBlocks, 256, 0, stream>>>(n, invdiag, x, r, p, tmp, pPartial, grid, uPartial, s, parity);
    cgDirectionKernel<<<kCgBlocks, 256, 0, stream>>>(n, invdiag, r, p, uPartial, kCgBlocks, s, maxIters, parity);
    return hipGetLastError();
}

hipError_t launchInvDiag(int64_t nParticles, const int64_t *diagBlock, const double *A, double *invdiag,
                         hipStream_t stream)
{
    if (nParticles <= 0) return hipSuccess;
    invDiagKernel<<<blocksFor(nParticles, 256), 256, 0, stream>>>(nParticles, diagBlock, A, invdiag);
    return hipGetLastError();
}

hipError_t launchLinComb(int64_t n, double *out, double a, const double *x, double b, const double *y, double c,
                         const double *z, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    linCombKernel<<<blocksFor(n, 256), 256, 0, stream>>>(n, out, a, x, b, y, c, z);
    return hipGetLastError();
}

hipError_t launchMassLinComb(int64_t n, double *out, double a, const double *mass, const double *x, double b,
                             const double *y, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    massLinCombKernel<<<blocksFor(n, 256), 256, 0, stream>>>(n, out, a, mass, x, b, y);
    return hipGetLastError();
}

hipError_t launchNorm2(int64_t n, const double *x, double *partial, double *out, hipStream_t stream)
{
    norm2PartialKernel<<<kCgBlocks, 256, 0, stream>>>(n, x, partial);
    sumFinalKernel<<<1, 1024, 0, stream>>>(partial, kCgBlocks, out);
    return hipGetLastError();
}

}  // namespace spgvbd
