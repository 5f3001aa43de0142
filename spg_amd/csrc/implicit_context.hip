// Host side of the implicit Euler family in libspg_vbd.so: element tables, the block-sparse pattern of
// the system matrices with its contribution lists, the solver loops of Baraff-Witkin, NewtonDv,
// NewtonDx, NewtonRobust and BDF2 over the GPU kernels, and the C-ABI of include/spg_implicit.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "host_util.h"
#include "implicit_device.h"
#include "implicit_element.h"
#include "spg_implicit.h"

using namespace spgvbd;

namespace {

thread_local std::string g_lastError;

constexpr int kNumTypes = SPG_VBD_STABLE_SQUARED_NEO_HOOKEAN;
int stencilSizeOf(int type) { return type >= 1 && type <= kNumTypes ? fullStencilSize(type) : 0; }
int paramsOf(int type) { return type >= 1 && type <= kNumTypes ? fullParams(type) : 0; }

bool finite3(const double *p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

}  // namespace

struct spg_implicit_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int32_t solver = 0;
    int32_t nObjects = 0;
    std::vector<int64_t> objOffset;
    int64_t np = 0;  // particles
    int64_t ndof = 0;

    DevBuf<double> x, v, mass, f, grad, hess;
    struct Table {
        int type = 0;
        int64_t n = 0;
        DevBuf<int32_t> idx;
        DevBuf<double> prm;
        int64_t gradOff = 0, hessOff = 0;  // in doubles
    };
    Table tables[kNumTypes + 1];
    DevBuf<int64_t> incPtr, incOff;
    // block-sparse pattern: the full one (exports) and the upper storage of the symmetric matrices
    int64_t nBlocks = 0, nUpper = 0;
    DevBuf<int64_t> rowPtr, diagBlock, contribPtr, contrib, upPtr, loPtr;
    DevBuf<int32_t> col, upRow, upCol, loIdx, loCol;
    std::vector<int64_t> hRowPtr;
    std::vector<int32_t> hCol;
    DevBuf<double> H, A, invdiag;  // upper blocks only
    // conjugate gradient
    DevBuf<double> cgr, cgp, cgtmp, partial, partial2;
    DevBuf<CgScalars> cgs;
    CgScalars *hostCg = nullptr;  // pinned
    double *hostNorm = nullptr;   // pinned
    DevBuf<double> normOut;
    // solver vectors and state
    DevBuf<double> x0, v0, vi, dv, rhs, tmp, xm1, vm1;
    bool haveHistory = false;
    double currentDt = -1.0;
    int consecutiveSuccessful = 0;
    // statistics
    spg_implicit_stats stats{};
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evCopy = nullptr;
    // one (start, end) pair per iteration chunk of a timed solve: the chunk's first product is sampled
    std::vector<std::pair<hipEvent_t, hipEvent_t>> spmvEvents;

    ~spg_implicit_ctx()
    {
        if (hostCg) (void)hipHostFree(hostCg);
        if (hostNorm) (void)hipHostFree(hostNorm);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (evCopy) (void)hipEventDestroy(evCopy);
        for (auto &e : spmvEvents) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    SymBsr sym(double *vals) const { return SymBsr{np, upPtr.p, upCol.p, loPtr.p, loIdx.p, loCol.p, vals}; }

    void timedBegin()
    {
        if (timing) hipCheck(hipEventRecord(ev0, stream), "hipEventRecord");
    }
    double timedEnd()
    {
        if (!timing) return 0.0;
        hipCheck(hipEventRecord(ev1, stream), "hipEventRecord");
        hipCheck(hipEventSynchronize(ev1), "hipEventSynchronize");
        float ms = 0.f;
        hipCheck(hipEventElapsedTime(&ms, ev0, ev1), "hipEventElapsedTime");
        return ms;
    }

    // element gradients (and Hessians when `hessians`) at positions xp; then the force f at xp
    void derivatives(const double *xp, bool hessians, const double *g)
    {
        for (int t = 1; t <= kNumTypes; ++t) {
            Table &tb = tables[t];
            if (!tb.n) continue;
            ImplicitElems e{t, tb.n, tb.idx.p, tb.prm.p, grad.p + tb.gradOff, hessians ? hess.p + tb.hessOff : nullptr};
            hipCheck(hessians ? launchElementDerivatives(e, xp, stream) : launchElementGradients(e, xp, stream),
                     "element derivative launch");
        }
        hipCheck(launchGatherForce(np, mass.p, g[0], g[1], g[2], incPtr.p, incOff.p, grad.p, f.p, stream),
                 "force launch");
        if (hessians)
            hipCheck(launchAssembleHessian(nUpper, upRow.p, upCol.p, contribPtr.p, contrib.p, hess.p, H.p, stream),
                     "assembly launch");
    }

    // A = a M + c H, its inverse diagonal; CG on A dv = b (zero initial guess) into out
    void solve(double a, double c, const double *b, double *out, double tol, int32_t maxIters)
    {
        hipCheck(launchCombine(nUpper, H.p, c, np, upPtr.p, mass.p, a, A.p, stream), "combine launch");
        hipCheck(launchInvDiag(np, upPtr.p, A.p, invdiag.p, stream), "invdiag launch");
        timedBegin();
        hipCheck(launchCgInit(ndof, b, invdiag.p, out, cgr.p, cgp.p, partial.p, cgs.p, tol, stream), "cg launch");
        // Iterations go in chunks; the scalars are read back after each chunk while the next one is
        // already queued, and every kernel of an iteration returns at once after convergence.
        constexpr int kChunk = 8;
        int launched = 0;
        std::vector<int> sampled;  // iteration whose product each event pair timed
        auto chunk = [&] {
            for (int k = 0; k < kChunk && launched < maxIters; ++k, ++launched) {
                hipEvent_t e0 = nullptr, e1 = nullptr;
                if (timing && k == 0) {
                    if (sampled.size() == spmvEvents.size()) {
                        std::pair<hipEvent_t, hipEvent_t> e{nullptr, nullptr};
                        hipCheck(hipEventCreate(&e.first), "hipEventCreate");
                        const hipError_t err = hipEventCreate(&e.second);
                        if (err != hipSuccess) (void)hipEventDestroy(e.first);
                        hipCheck(err, "hipEventCreate");
                        spmvEvents.push_back(e);
                    }
                    e0 = spmvEvents[sampled.size()].first;
                    e1 = spmvEvents[sampled.size()].second;
                    sampled.push_back(launched);
                }
                hipCheck(launchCgIteration(sym(A.p), ndof, invdiag.p, out, cgr.p, cgp.p, cgtmp.p, partial.p,
                                           partial2.p, cgs.p, maxIters, launched, stream, e0, e1),
                         "cg launch");
            }
        };
        auto readBack = [&] {
            hipCheck(hipMemcpyAsync(hostCg, cgs.p, sizeof(CgScalars), hipMemcpyDeviceToHost, stream), "download");
            hipCheck(hipEventRecord(evCopy, stream), "hipEventRecord");
        };
        chunk();
        readBack();
        for (;;) {
            const bool more = launched < maxIters;
            if (more) chunk();
            hipCheck(hipEventSynchronize(evCopy), "hipEventSynchronize");
            if (hostCg->done || !more) break;
            readBack();
        }
        hipCheck(hipMemcpyAsync(hostCg, cgs.p, sizeof(CgScalars), hipMemcpyDeviceToHost, stream), "download");
        hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
        const double ms = timedEnd();
        stats.cg_ms += ms;
        stats.linear_solves += 1;
        stats.cg_iterations += hostCg->iters;
        stats.last_cg_error = hostCg->rhsNorm2 > 0 ? std::sqrt(hostCg->residualNorm2 / hostCg->rhsNorm2) : 0.0;
        if (timing) {
            const int64_t spmvs = hostCg->rhsNorm2 == 0 ? 0 : hostCg->iters + (hostCg->done == 1 ? 1 : 0);
            stats.spmv_launches += spmvs;
            stats.spmv_bytes += spmvs * spmvBytes();
            // sampled products that ran (the later ones return at once after convergence)
            for (size_t k = 0; k < sampled.size(); ++k) {
                if (sampled[k] >= spmvs) break;
                float t = 0.f;
                hipCheck(hipEventElapsedTime(&t, spmvEvents[k].first, spmvEvents[k].second), "hipEventElapsedTime");
                stats.spmv_sampled_ms += t;
                stats.spmv_sampled += 1;
            }
        }
    }

    // compulsory bytes of one product: 72 B + 4 B column per upper block, 4 B index + 4 B column per
    // lower block, x (24 B) and y (24 B) per block row and the two row-pointer arrays
    double spmvBytes() const
    {
        return 76.0 * nUpper + 8.0 * (nUpper - np) + 48.0 * np + 16.0 * (np + 1);
    }

    double norm(const double *vec)
    {
        hipCheck(launchNorm2(ndof, vec, partial.p, normOut.p, stream), "norm launch");
        hipCheck(hipMemcpyAsync(hostNorm, normOut.p, sizeof(double), hipMemcpyDeviceToHost, stream), "download");
        hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
        return std::sqrt(*hostNorm);
    }

    void copy(double *dst, const double *src)
    {
        hipCheck(hipMemcpyAsync(dst, src, sizeof(double) * ndof, hipMemcpyDeviceToDevice, stream), "copy");
    }
    void lin(double *out, double a, const double *p, double b = 0, const double *q = nullptr, double c = 0,
             const double *r = nullptr)
    {
        hipCheck(launchLinComb(ndof, out, a, p, b, q, c, r, stream), "lincomb launch");
    }

    void substep(double dtStep, int32_t nsub, const double *g)
    {
        const double dt = dtStep / nsub;
        const double tol = SPG_IMPLICIT_CG_TOLERANCE;
        const int32_t maxIt = SPG_IMPLICIT_CG_MAX_ITERATIONS;
        auto assemble = [&](const double *xp, bool hessians) {
            timedBegin();
            derivatives(xp, hessians, g);
            stats.assembly_ms += timedEnd();
        };
        switch (solver) {
        case SPG_IMPLICIT_BARAFF_WITKIN: {  // implicitEulerBaraffWitkin.cpp:31-53
            copy(v0.p, v.p);
            assemble(x.p, true);
            // rhs = dt (f + dt K v0) with K = -H
            hipCheck(launchSymSpmv(sym(H.p), v0.p, -1.0, tmp.p, nullptr, stream), "spmv launch");
            lin(rhs.p, 1.0, f.p, dt, tmp.p);
            lin(rhs.p, dt, rhs.p);
            solve(1.0, dt * dt, rhs.p, dv.p, tol, maxIt);
            lin(v.p, 1.0, v0.p, 1.0, dv.p);
            lin(x.p, 1.0, x.p, dt, v.p);
            break;
        }
        case SPG_IMPLICIT_NEWTON_DV: {  // implicitEulerNewtonDv.cpp:30-61
            copy(x0.p, x.p);
            copy(v0.p, v.p);
            lin(x.p, 1.0, x0.p, dt, v0.p);
            assemble(x.p, true);
            lin(rhs.p, dt, f.p);
            solve(1.0, dt * dt, rhs.p, dv.p, tol, maxIt);
            lin(v.p, 1.0, v0.p, 1.0, dv.p);
            lin(x.p, 1.0, x0.p, dt, v.p);
            break;
        }
        case SPG_IMPLICIT_NEWTON_DX: {  // implicitEulerNewtonDx.cpp:30-58
            copy(x0.p, x.p);
            lin(x.p, 1.0, x.p, dt, v.p);
            assemble(x.p, true);
            const double invdt = 1. / dt;
            solve(invdt * invdt, 1.0, f.p, dv.p, tol, maxIt);
            lin(x.p, 1.0, x.p, 1.0, dv.p);
            lin(v.p, invdt, x.p, -invdt, x0.p);
            break;
        }
        case SPG_IMPLICIT_BDF2: {  // bdf2.cpp:31-73
            copy(x0.p, x.p);
            copy(v0.p, v.p);
            if (!haveHistory) {
                copy(xm1.p, x0.p);
                copy(vm1.p, v0.p);
                haveHistory = true;
            }
            assemble(x.p, true);
            // w = x0 - x_{-1} + 2 dt v0; tmp = K w = -H w
            lin(rhs.p, 1.0, x0.p, -1.0, xm1.p, 2 * dt, v0.p);
            hipCheck(launchSymSpmv(sym(H.p), rhs.p, -1.0, tmp.p, nullptr, stream), "spmv launch");
            // tmp = 2/3 dt (f + 1/3 K w); rhs = 1/3 M (v0 - v_{-1}) + tmp
            lin(tmp.p, 2.0 / 3.0 * dt, f.p, 2.0 / 3.0 * dt / 3.0, tmp.p);
            lin(dv.p, 1.0, v0.p, -1.0, vm1.p);
            hipCheck(launchMassLinComb(ndof, rhs.p, 1.0 / 3.0, mass.p, dv.p, 1.0, tmp.p, stream), "lincomb launch");
            solve(1.0, 4.0 / 9.0 * (dt * dt), rhs.p, dv.p, tol, maxIt);
            lin(v.p, 1.0, v0.p, 1.0, dv.p);
            lin(x.p, 4.0 / 3.0, x0.p, -1.0 / 3.0, xm1.p, 2.0 / 3.0 * dt, v.p);
            copy(xm1.p, x0.p);
            copy(vm1.p, v0.p);
            break;
        }
        default:
            robustSubstep(dtStep, nsub, g);
        }
    }

    // implicitEulerNewtonRobust.cpp:31-122
    void robustSubstep(double dtStep, int32_t nsub, const double *g)
    {
        const double newtonThreshold = 1e-5, splitting = 0.75;
        const int maxNewton = 5, maxLineSearch = 5, increaseAfter = 5;
        const double maxDt = dtStep / nsub;
        if (currentDt == -1. || currentDt > maxDt) currentDt = maxDt;
        const double h = currentDt;
        bool ok = true;
        copy(x0.p, x.p);
        copy(v0.p, v.p);
        lin(x.p, 1.0, x0.p, h, v0.p);
        copy(vi.p, v0.p);
        int newton = 0;
        double stepResidual = 0.;
        do {
            timedBegin();
            derivatives(x.p, true, g);
            stats.assembly_ms += timedEnd();
            // rhs = h f - M (vi - v0)
            lin(tmp.p, 1.0, vi.p, -1.0, v0.p);
            hipCheck(launchMassLinComb(ndof, rhs.p, -1.0, mass.p, tmp.p, h, f.p, stream), "lincomb launch");
            const double initialResidual = norm(rhs.p);
            solve(1.0, h * h, rhs.p, dv.p, SPG_IMPLICIT_CG_TOLERANCE, SPG_IMPLICIT_CG_MAX_ITERATIONS);
            stepResidual = initialResidual;
            int ls = 0;
            double alpha = 1.;
            while (stepResidual >= initialResidual && ls < maxLineSearch) {
                lin(v.p, 1.0, vi.p, alpha, dv.p);
                lin(x.p, 1.0, x0.p, h, v.p);
                derivatives(x.p, false, g);
                lin(tmp.p, 1.0, v.p, -1.0, v0.p);
                hipCheck(launchMassLinComb(ndof, rhs.p, -1.0, mass.p, tmp.p, h, f.p, stream), "lincomb launch");
                stepResidual = norm(rhs.p);
                if (stepResidual < initialResidual) copy(vi.p, v.p);
                alpha *= 0.5;
                ++ls;
            }
            if (stepResidual >= initialResidual && stepResidual > newtonThreshold) ok = false;
            ++newton;
        } while (stepResidual > newtonThreshold && ok && newton < maxNewton);
        stats.newton_iterations += newton;
        if (stepResidual > newtonThreshold) ok = false;
        if (!ok) {
            copy(x.p, x0.p);
            copy(v.p, v0.p);
            currentDt *= splitting;
            consecutiveSuccessful = 0;
            ++stats.failed_substeps;
        } else {
            ++consecutiveSuccessful;
            if (consecutiveSuccessful > increaseAfter && currentDt < maxDt) currentDt /= splitting;
        }
        stats.current_dt = currentDt;
    }
};

namespace {

template <class F>
int guarded(spg_implicit_ctx *ctx, F &&f)
{
    try {
        f();
        if (ctx) ctx->err.clear();
        return SPG_VBD_OK;
    } catch (const VbdError &e) {
        if (ctx) ctx->err = e.what();
        g_lastError = e.what();
        return e.code;
    } catch (const std::exception &e) {
        if (ctx) ctx->err = e.what();
        g_lastError = e.what();
        return SPG_VBD_ERR_INVALID;
    }
}

void validateEnergy(const spg_vbd_energy &en, int64_t np)
{
    const int ss = stencilSizeOf(en.type), pp = paramsOf(en.type);
    if (!ss) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "unsupported energy type " + std::to_string(en.type));
    if (en.n_stencils < 0 || (en.n_stencils > 0 && (!en.stencils || !en.params)))
        throw VbdError(SPG_VBD_ERR_INVALID, "invalid energy arrays");
    for (int64_t i = 0; i < en.n_stencils * ss; ++i)
        if (en.stencils[i] < 0 || en.stencils[i] >= np) throw VbdError(SPG_VBD_ERR_INVALID, "stencil index out of range");
    for (int64_t i = 0; i < en.n_stencils * pp; ++i)
        if (!std::isfinite(en.params[i])) throw VbdError(SPG_VBD_ERR_INVALID, "non-finite energy parameter");
    if (en.type == SPG_VBD_SPRING || en.type == SPG_VBD_SPRING_CONTINUUM || en.type == SPG_VBD_SPRING_SQUARED_CONTINUUM)
        for (int64_t i = 0; i < en.n_stencils; ++i)
            if (en.params[i * pp] <= 0) throw VbdError(SPG_VBD_ERR_INVALID, "Zero/negative rest length invalid");
    if (en.type == SPG_VBD_DISCRETE_BENDING || en.type == SPG_VBD_BARAFF_WITKIN_BENDING)
        for (int64_t i = 0; i < en.n_stencils; ++i)
            if (en.params[i * pp] < -3.12 || en.params[i * pp] > 3.12)
                throw VbdError(SPG_VBD_ERR_INVALID, "Invalid restTheta");
}

void buildContext(spg_implicit_ctx &c, const spg_vbd_object *objs, int32_t nObjects)
{
    if (nObjects < 1 || !objs) throw VbdError(SPG_VBD_ERR_INVALID, "at least one object required");
    c.nObjects = nObjects;
    c.objOffset.assign(nObjects + 1, 0);
    for (int32_t o = 0; o < nObjects; ++o) {
        const auto &ob = objs[o];
        if (ob.n_particles < 1 || !ob.x || !ob.mass || ob.n_energies < 0 || (ob.n_energies > 0 && !ob.energies))
            throw VbdError(SPG_VBD_ERR_INVALID, "invalid object arrays");
        c.objOffset[o + 1] = c.objOffset[o] + ob.n_particles;
    }
    c.np = c.objOffset[nObjects];
    c.ndof = 3 * c.np;
    if (c.np >= (int64_t(1) << 31) - 1) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "too many particles");
    std::vector<double> hx(c.ndof), hv(c.ndof, 0.0), hm(c.np);
    for (int32_t o = 0; o < nObjects; ++o) {
        const auto &ob = objs[o];
        const int64_t b = c.objOffset[o];
        for (int64_t i = 0; i < ob.n_particles; ++i) {
            if (!finite3(ob.x + 3 * i) || (ob.v && !finite3(ob.v + 3 * i)))
                throw VbdError(SPG_VBD_ERR_NONFINITE, "non-finite state");
            if (!std::isfinite(ob.mass[i]) || ob.mass[i] < 0) throw VbdError(SPG_VBD_ERR_INVALID, "invalid mass");
            hm[b + i] = ob.mass[i];
            for (int d = 0; d < 3; ++d) {
                hx[3 * (b + i) + d] = ob.x[3 * i + d];
                if (ob.v) hv[3 * (b + i) + d] = ob.v[3 * i + d];
            }
        }
        for (int32_t e = 0; e < ob.n_energies; ++e) validateEnergy(ob.energies[e], ob.n_particles);
    }

    // element tables per type in (object, energy, stencil) order
    struct Src {
        int32_t object, energy;
    };
    std::vector<Src> srcs[kNumTypes + 1];
    int64_t count[kNumTypes + 1] = {};
    for (int32_t o = 0; o < nObjects; ++o)
        for (int32_t e = 0; e < objs[o].n_energies; ++e) {
            const auto &en = objs[o].energies[e];
            srcs[en.type].push_back({o, e});
            count[en.type] += en.n_stencils;
        }
    int64_t gradTotal = 0, hessTotal = 0;
    std::vector<std::vector<int64_t>> elemBase(nObjects);
    for (int32_t o = 0; o < nObjects; ++o) elemBase[o].assign(objs[o].n_energies, 0);
    for (int t = 1; t <= kNumTypes; ++t) {
        auto &tb = c.tables[t];
        tb.type = t;
        tb.n = count[t];
        const int ss = stencilSizeOf(t), pp = paramsOf(t), N = 3 * ss;
        tb.gradOff = gradTotal;
        tb.hessOff = hessTotal;
        gradTotal += tb.n * N;
        hessTotal += tb.n * N * N;
        std::vector<int32_t> idx;
        std::vector<double> prm;
        idx.reserve(tb.n * ss);
        prm.reserve(tb.n * pp);
        int64_t k = 0;
        for (const auto &s : srcs[t]) {
            const auto &en = objs[s.object].energies[s.energy];
            elemBase[s.object][s.energy] = k;
            for (int64_t i = 0; i < en.n_stencils * ss; ++i)
                idx.push_back(static_cast<int32_t>(c.objOffset[s.object] + en.stencils[i]));
            prm.insert(prm.end(), en.params, en.params + en.n_stencils * pp);
            k += en.n_stencils;
        }
        tb.idx.upload(idx, c.stream);
        tb.prm.upload(prm, c.stream);
    }
    c.grad.alloc(std::max<int64_t>(gradTotal, 1));
    c.hess.alloc(std::max<int64_t>(hessTotal, 1));

    // force incidence: per particle, its (energy order, ascending stencil, slot) gradient slots
    std::vector<std::vector<int64_t>> inc(c.np);
    // block pattern: per particle the particles sharing a stencil (and itself)
    std::vector<std::vector<int32_t>> nb(c.np);
    for (int64_t i = 0; i < c.np; ++i) nb[i].push_back(static_cast<int32_t>(i));
    for (int32_t o = 0; o < nObjects; ++o)
        for (int32_t e = 0; e < objs[o].n_energies; ++e) {
            const auto &en = objs[o].energies[e];
            const int ss = stencilSizeOf(en.type);
            const auto &tb = c.tables[en.type];
            for (int64_t s = 0; s < en.n_stencils; ++s) {
                const int64_t el = elemBase[o][e] + s;
                for (int a = 0; a < ss; ++a) {
                    const int64_t va = c.objOffset[o] + en.stencils[s * ss + a];
                    inc[va].push_back(tb.gradOff / 3 + el * ss + a);
                    for (int b = 0; b < ss; ++b)
                        nb[va].push_back(static_cast<int32_t>(c.objOffset[o] + en.stencils[s * ss + b]));
                }
            }
        }
    std::vector<int64_t> incPtr(c.np + 1, 0), incOff;
    for (int64_t i = 0; i < c.np; ++i) {
        incPtr[i + 1] = incPtr[i] + static_cast<int64_t>(inc[i].size());
        incOff.insert(incOff.end(), inc[i].begin(), inc[i].end());
    }
    inc.clear();
    inc.shrink_to_fit();
    c.hRowPtr.assign(c.np + 1, 0);
    c.hCol.clear();
    std::vector<int64_t> diag(c.np);
    for (int64_t i = 0; i < c.np; ++i) {
        auto &r = nb[i];
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        c.hRowPtr[i + 1] = c.hRowPtr[i] + static_cast<int64_t>(r.size());
        for (int32_t j : r) {
            if (j == i) diag[i] = static_cast<int64_t>(c.hCol.size());
            c.hCol.push_back(j);
        }
    }
    nb.clear();
    nb.shrink_to_fit();
    c.nBlocks = c.hRowPtr[c.np];
    auto blockOf = [&](int64_t i, int32_t j) {
        const auto b = c.hCol.begin() + c.hRowPtr[i], e = c.hCol.begin() + c.hRowPtr[i + 1];
        return static_cast<int64_t>(std::lower_bound(b, e, j) - c.hCol.begin());
    };
    // upper blocks (i <= j) in block order; per row the upper blocks (j, i), j < i, read transposed
    std::vector<int64_t> upperOfBlock(c.nBlocks, -1), upPtr(c.np + 1, 0), loPtr(c.np + 1, 0);
    std::vector<int32_t> upRow, upCol;
    for (int64_t i = 0; i < c.np; ++i) {
        for (int64_t k = c.hRowPtr[i]; k < c.hRowPtr[i + 1]; ++k) {
            const int32_t j = c.hCol[k];
            if (j < i) {
                ++loPtr[i + 1];
                continue;
            }
            upperOfBlock[k] = static_cast<int64_t>(upRow.size());
            upRow.push_back(static_cast<int32_t>(i));
            upCol.push_back(j);
        }
        upPtr[i + 1] = static_cast<int64_t>(upRow.size());
    }
    c.nUpper = static_cast<int64_t>(upRow.size());
    if (c.nUpper > std::numeric_limits<int32_t>::max())
        throw VbdError(SPG_VBD_ERR_INVALID, "system matrix has more than 2^31 - 1 stored blocks");
    for (int64_t i = 0; i < c.np; ++i) loPtr[i + 1] += loPtr[i];
    std::vector<int32_t> loIdx(loPtr[c.np]), loCol(loPtr[c.np]);
    {
        std::vector<int64_t> fillLo(loPtr.begin(), loPtr.end() - 1);
        for (int64_t u = 0; u < c.nUpper; ++u)  // rows ascending, so every lower list is in ascending j
            if (upCol[u] != upRow[u]) {
                const int64_t at = fillLo[upCol[u]]++;
                loIdx[at] = static_cast<int32_t>(u);
                loCol[at] = upRow[u];
            }
    }
    // contributions: element Hessian sub-block (a, b) with v_a <= v_b to upper block (v_a, v_b), kept in
    // (object, energy, stencil, a, b) order within each block
    std::vector<int64_t> cnt(c.nUpper + 1, 0);
    std::vector<std::pair<int64_t, int64_t>> items;  // (upper block, packed origin)
    for (int32_t o = 0; o < nObjects; ++o)
        for (int32_t e = 0; e < objs[o].n_energies; ++e) {
            const auto &en = objs[o].energies[e];
            const int ss = stencilSizeOf(en.type), N = 3 * ss;
            const auto &tb = c.tables[en.type];
            for (int64_t s = 0; s < en.n_stencils; ++s) {
                const int64_t el = elemBase[o][e] + s;
                for (int a = 0; a < ss; ++a)
                    for (int b = 0; b < ss; ++b) {
                        const int64_t va = c.objOffset[o] + en.stencils[s * ss + a];
                        const int64_t vb = c.objOffset[o] + en.stencils[s * ss + b];
                        if (va > vb) continue;
                        const int64_t u = upperOfBlock[blockOf(va, static_cast<int32_t>(vb))];
                        const int64_t origin = tb.hessOff + el * N * N + int64_t(3 * b) * N + 3 * a;
                        items.push_back({u, (origin << 4) | N});
                        ++cnt[u + 1];
                    }
            }
        }
    for (int64_t u = 0; u < c.nUpper; ++u) cnt[u + 1] += cnt[u];
    std::vector<int64_t> contrib(items.size()), fill(cnt.begin(), cnt.end() - 1);
    for (const auto &it : items) contrib[fill[it.first]++] = it.second;
    items.clear();
    items.shrink_to_fit();

    c.x.upload(hx, c.stream);
    c.v.upload(hv, c.stream);
    c.mass.upload(hm, c.stream);
    c.f.alloc(c.ndof);
    c.incPtr.upload(incPtr, c.stream);
    c.incOff.upload(incOff.empty() ? std::vector<int64_t>{0} : incOff, c.stream);
    c.rowPtr.upload(c.hRowPtr, c.stream);
    c.col.upload(c.hCol, c.stream);
    c.diagBlock.upload(diag, c.stream);
    c.upPtr.upload(upPtr, c.stream);
    c.loPtr.upload(loPtr, c.stream);
    c.upRow.upload(upRow, c.stream);
    c.upCol.upload(upCol, c.stream);
    c.loIdx.upload(loIdx, c.stream);
    c.loCol.upload(loCol, c.stream);
    c.contribPtr.upload(cnt, c.stream);
    c.contrib.upload(contrib.empty() ? std::vector<int64_t>{0} : contrib, c.stream);
    c.H.alloc(9 * c.nUpper);
    c.A.alloc(9 * c.nUpper);
    c.invdiag.alloc(c.ndof);
    for (auto *b : {&c.cgr, &c.cgp, &c.cgtmp, &c.x0, &c.v0, &c.vi, &c.dv, &c.rhs, &c.tmp, &c.xm1, &c.vm1})
        b->alloc(c.ndof);
    c.partial.alloc(std::max<int64_t>(2 * kCgBlocks, spmvPartials(c.np)));
    c.partial2.alloc(2 * kCgBlocks);
    c.cgs.alloc(1);
    c.normOut.alloc(1);
    hipCheck(hipHostMalloc(reinterpret_cast<void **>(&c.hostCg), sizeof(CgScalars)), "hipHostMalloc");
    hipCheck(hipHostMalloc(reinterpret_cast<void **>(&c.hostNorm), sizeof(double)), "hipHostMalloc");
    hipCheck(hipEventCreate(&c.ev0), "hipEventCreate");
    hipCheck(hipEventCreate(&c.ev1), "hipEventCreate");
    hipCheck(hipEventCreateWithFlags(&c.evCopy, hipEventDisableTiming), "hipEventCreate");
    hipCheck(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
    c.stats.current_dt = -1.0;
}

void checkObject(const spg_implicit_ctx *ctx, int32_t object)
{
    if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
    if (object < 0 || object >= ctx->nObjects) throw VbdError(SPG_VBD_ERR_INVALID, "object index out of range");
}

}  // namespace

extern "C" {

int spg_implicit_create(spg_implicit_ctx **out, const spg_vbd_object *objects, int32_t n_objects, int32_t solver,
                        int32_t device)
{
    if (!out) return SPG_VBD_ERR_INVALID;
    *out = nullptr;
    auto c = std::make_unique<spg_implicit_ctx>();
    const int rc = guarded(nullptr, [&] {
        if (solver < SPG_IMPLICIT_BARAFF_WITKIN || solver > SPG_IMPLICIT_BDF2)
            throw VbdError(SPG_VBD_ERR_INVALID, "unknown implicit solver");
        int ndev = 0;
        hipCheck(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        if (device < 0 || device >= ndev) throw VbdError(SPG_VBD_ERR_INVALID, "invalid HIP device");
        c->device = device;
        c->solver = solver;
        hipCheck(hipSetDevice(device), "hipSetDevice");
        hipCheck(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        buildContext(*c, objects, n_objects);
    });
    if (rc == SPG_VBD_OK) *out = c.release();
    return rc;
}

void spg_implicit_destroy(spg_implicit_ctx *ctx)
{
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        delete ctx;
    }
}

const char *spg_implicit_last_error(const spg_implicit_ctx *ctx) { return ctx ? ctx->err.c_str() : g_lastError.c_str(); }

int spg_implicit_num_dofs(const spg_implicit_ctx *ctx, int64_t *n)
{
    if (!ctx || !n) return SPG_VBD_ERR_INVALID;
    *n = ctx->ndof;
    return SPG_VBD_OK;
}

int spg_implicit_num_blocks(const spg_implicit_ctx *ctx, int64_t *n)
{
    if (!ctx || !n) return SPG_VBD_ERR_INVALID;
    *n = ctx->nBlocks;
    return SPG_VBD_OK;
}

int spg_implicit_set_state(spg_implicit_ctx *ctx, int32_t object, const double *x, const double *v)
{
    return guarded(ctx, [&] {
        checkObject(ctx, object);
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        const int64_t b = ctx->objOffset[object], n = ctx->objOffset[object + 1] - b;
        for (int64_t i = 0; i < n; ++i)
            if ((x && !finite3(x + 3 * i)) || (v && !finite3(v + 3 * i)))
                throw VbdError(SPG_VBD_ERR_NONFINITE, "non-finite state");
        if (x)
            hipCheck(hipMemcpyAsync(ctx->x.p + 3 * b, x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream),
                     "upload");
        if (v)
            hipCheck(hipMemcpyAsync(ctx->v.p + 3 * b, v, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream),
                     "upload");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int spg_implicit_get_state(spg_implicit_ctx *ctx, int32_t object, double *x, double *v)
{
    return guarded(ctx, [&] {
        checkObject(ctx, object);
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        const int64_t b = ctx->objOffset[object], n = ctx->objOffset[object + 1] - b;
        if (x)
            hipCheck(hipMemcpyAsync(x, ctx->x.p + 3 * b, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream),
                     "download");
        if (v)
            hipCheck(hipMemcpyAsync(v, ctx->v.p + 3 * b, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream),
                     "download");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int spg_implicit_step(spg_implicit_ctx *ctx, double dt, int32_t nsubsteps, const double gravity[3])
{
    return guarded(ctx, [&] {
        if (!ctx || !gravity) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        if (!(dt > 0) || nsubsteps < 1) throw VbdError(SPG_VBD_ERR_INVALID, "invalid dt/substeps");
        if (!finite3(gravity)) throw VbdError(SPG_VBD_ERR_INVALID, "non-finite gravity");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        for (int32_t s = 0; s < nsubsteps; ++s) ctx->substep(dt, nsubsteps, gravity);
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        ctx->stats.steps += 1;
    });
}

int spg_implicit_reset(spg_implicit_ctx *ctx)
{
    if (!ctx) return SPG_VBD_ERR_INVALID;
    ctx->haveHistory = false;
    ctx->currentDt = -1.0;
    ctx->consecutiveSuccessful = 0;
    ctx->stats.current_dt = -1.0;
    return SPG_VBD_OK;
}

int spg_implicit_system_force(spg_implicit_ctx *ctx, const double gravity[3], double *f)
{
    return guarded(ctx, [&] {
        if (!ctx || !gravity || !f) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->derivatives(ctx->x.p, false, gravity);
        hipCheck(hipMemcpyAsync(f, ctx->f.p, sizeof(double) * ctx->ndof, hipMemcpyDeviceToHost, ctx->stream),
                 "download");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int spg_implicit_stiffness(spg_implicit_ctx *ctx, int64_t *row_ptr, int32_t *cols, double *blocks)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        if (row_ptr) std::copy(ctx->hRowPtr.begin(), ctx->hRowPtr.end(), row_ptr);
        if (cols) std::copy(ctx->hCol.begin(), ctx->hCol.end(), cols);
        if (!blocks) return;
        const double g[3] = {0, 0, 0};
        ctx->derivatives(ctx->x.p, true, g);
        // K = -H, expanded to the full block rows
        DevBuf<double> full;
        full.alloc(9 * ctx->nBlocks);
        hipCheck(launchExpandSym(ctx->sym(ctx->H.p), -1.0, ctx->rowPtr.p, ctx->col.p, ctx->diagBlock.p, full.p,
                                 ctx->stream),
                 "expand launch");
        hipCheck(hipMemcpyAsync(blocks, full.p, sizeof(double) * 9 * ctx->nBlocks, hipMemcpyDeviceToHost, ctx->stream),
                 "download");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int spg_implicit_cg_solve(spg_implicit_ctx *ctx, double alpha, double beta, const double *b, double tolerance,
                          int32_t max_iterations, double *x, int32_t *iterations, double *error)
{
    return guarded(ctx, [&] {
        if (!ctx || !b || !x) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        if (!(tolerance >= 0) || max_iterations < 0) throw VbdError(SPG_VBD_ERR_INVALID, "invalid tolerance/iterations");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        const double g[3] = {0, 0, 0};
        ctx->derivatives(ctx->x.p, true, g);
        hipCheck(hipMemcpyAsync(ctx->rhs.p, b, sizeof(double) * ctx->ndof, hipMemcpyHostToDevice, ctx->stream),
                 "upload");
        // alpha M + beta K = alpha M - beta H
        ctx->solve(alpha, -beta, ctx->rhs.p, ctx->dv.p, tolerance, max_iterations);
        hipCheck(hipMemcpyAsync(x, ctx->dv.p, sizeof(double) * ctx->ndof, hipMemcpyDeviceToHost, ctx->stream),
                 "download");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        if (iterations) *iterations = ctx->hostCg->iters;
        if (error) *error = ctx->stats.last_cg_error;
    });
}

int spg_implicit_element_derivatives(int32_t type, int64_t n, const double *xs, const double *params, double *grad,
                                     double *hess, int32_t device)
{
    return guarded(nullptr, [&] {
        const int ns = stencilSizeOf(type), np = paramsOf(type);
        if (!ns) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "unknown energy type");
        if (n < 0 || (n > 0 && (!xs || !params || !grad || !hess)))
            throw VbdError(SPG_VBD_ERR_INVALID, "invalid arguments");
        const int N = 3 * ns;
        if (device < 0) {  // host instantiation of the same closed forms
            for (int64_t i = 0; i < n; ++i) {
                switch (ns) {
                case 1: elementGradHessFull<1>(type, xs + i * N, params + i * np, grad + i * N, hess + i * N * N); break;
                case 2: elementGradHessFull<2>(type, xs + i * N, params + i * np, grad + i * N, hess + i * N * N); break;
                case 3: elementGradHessFull<3>(type, xs + i * N, params + i * np, grad + i * N, hess + i * N * N); break;
                default: elementGradHessFull<4>(type, xs + i * N, params + i * np, grad + i * N, hess + i * N * N);
                }
            }
            return;
        }
        if (n == 0) return;
        hipCheck(hipSetDevice(device), "hipSetDevice");
        hipStream_t stream = nullptr;
        hipCheck(hipStreamCreate(&stream), "hipStreamCreate");
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } sg{stream};
        std::vector<int32_t> idx(n * ns);
        for (int64_t k = 0; k < n * ns; ++k) idx[k] = static_cast<int32_t>(k);
        DevBuf<int32_t> dIdx;
        DevBuf<double> dX, dP, dG, dH;
        dIdx.upload(idx, stream);
        dX.upload(std::vector<double>(xs, xs + n * N), stream);
        dP.upload(std::vector<double>(params, params + n * np), stream);
        dG.alloc(n * N);
        dH.alloc(n * N * N);
        const ImplicitElems e{type, n, dIdx.p, dP.p, dG.p, dH.p};
        hipCheck(launchElementDerivatives(e, dX.p, stream), "element derivative launch");
        std::vector<double> hH(n * N * N);
        hipCheck(hipMemcpyAsync(grad, dG.p, sizeof(double) * n * N, hipMemcpyDeviceToHost, stream), "download");
        hipCheck(hipMemcpyAsync(hH.data(), dH.p, sizeof(double) * n * N * N, hipMemcpyDeviceToHost, stream), "download");
        hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
        for (int64_t i = 0; i < n; ++i)  // column-major per element on the device
            for (int r = 0; r < N; ++r)
                for (int c = 0; c < N; ++c) hess[i * N * N + r * N + c] = hH[i * N * N + c * N + r];
    });
}

int spg_implicit_bench_spmv(spg_implicit_ctx *ctx, double alpha, double beta, int32_t repeats, double *ms)
{
    return guarded(ctx, [&] {
        if (!ctx || !ms || repeats < 1) throw VbdError(SPG_VBD_ERR_INVALID, "invalid arguments");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        const double g[3] = {0, 0, 0};
        ctx->derivatives(ctx->x.p, true, g);
        hipCheck(launchCombine(ctx->nUpper, ctx->H.p, -beta, ctx->np, ctx->upPtr.p, ctx->mass.p, alpha, ctx->A.p,
                               ctx->stream),
                 "combine launch");
        hipCheck(launchSymSpmv(ctx->sym(ctx->A.p), ctx->v.p, 1.0, ctx->tmp.p, nullptr, ctx->stream), "spmv launch");
        hipCheck(hipEventRecord(ctx->ev0, ctx->stream), "hipEventRecord");
        for (int32_t r = 0; r < repeats; ++r)
            hipCheck(launchSymSpmv(ctx->sym(ctx->A.p), ctx->v.p, 1.0, ctx->tmp.p, nullptr, ctx->stream),
                     "spmv launch");
        hipCheck(hipEventRecord(ctx->ev1, ctx->stream), "hipEventRecord");
        hipCheck(hipEventSynchronize(ctx->ev1), "hipEventSynchronize");
        float t = 0.f;
        hipCheck(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1), "hipEventElapsedTime");
        *ms = t / repeats;
    });
}

int spg_implicit_get_stats(const spg_implicit_ctx *ctx, spg_implicit_stats *stats)
{
    if (!ctx || !stats) return SPG_VBD_ERR_INVALID;
    *stats = ctx->stats;
    return SPG_VBD_OK;
}

int spg_implicit_reset_stats(spg_implicit_ctx *ctx)
{
    if (!ctx) return SPG_VBD_ERR_INVALID;
    const double cd = ctx->stats.current_dt;
    ctx->stats = spg_implicit_stats{};
    ctx->stats.current_dt = cd;
    return SPG_VBD_OK;
}

int spg_implicit_set_timing(spg_implicit_ctx *ctx, int32_t enabled)
{
    if (!ctx) return SPG_VBD_ERR_INVALID;
    ctx->timing = enabled != 0;
    return SPG_VBD_OK;
}

int spg_implicit_synchronize(spg_implicit_ctx *ctx)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

}  // extern "C"
