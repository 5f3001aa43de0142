// Host side of libspg_vbd.so: builds the device layout of a VBD context from the spg particle
// groups and energies, orchestrates the step on one HIP stream, and implements the C-ABI of
// include/spg_vbd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "coloring.h"
#include "spg_vbd.h"
#include "spg_vbd_learned.h"
#include "element_eval.h"
#include "vbd_device.h"
#include "implicit_element.h"
#include "host_util.h"
#include "mlp_device.h"

using namespace spgvbd;

namespace {

thread_local std::string g_lastError;

// energy types 1 ... 5 have dedicated incidence classes; 6 ... 13 (the further energies) share one
bool isFurther(int type) { return type >= SPG_VBD_SPRING_CONTINUUM && type <= SPG_VBD_STABLE_SQUARED_NEO_HOOKEAN; }
int stencilSizeOf(int type)
{
    return type >= SPG_VBD_SPRING && type <= SPG_VBD_STABLE_SQUARED_NEO_HOOKEAN ? fullStencilSize(type) : 0;
}
int paramsOf(int type)
{
    return type >= SPG_VBD_SPRING && type <= SPG_VBD_STABLE_SQUARED_NEO_HOOKEAN ? fullParams(type) : 0;
}
// compulsory bytes of one element record read per incidence (roofline model)
double recordBytes(int type)
{
    switch (type) {
    case SPG_VBD_SPRING: return 16;
    case SPG_VBD_MEMBRANE_STVK: return 40;
    case SPG_VBD_DISCRETE_BENDING: return 24;
    case SPG_VBD_STABLE_NEO_HOOKEAN: return 64;
    case SPG_VBD_SPRING_ANCHOR: return 20;
    }
    return 0.0;
}
// the same for a further-energy incidence of solve class cls: the element's four slots and the FP32
// coefficients the class reads (not the padding of its planes)
double furtherRecordBytes(int cls)
{
    static constexpr int kFloats[kXCls] = {3, 8, 10, 6, 3, 3, 14, 14};
    return 16.0 + 4.0 * kFloats[cls];
}

bool finite3(const double *p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

// Work memory of spg_vbd_residual_noalloc: the per-particle squared residuals, the per-hinge k (theta - theta0),
// the per-segment partial sums and the per-object norms of the residual's two-stage reduction. One copy per GPU,
// held until the process ends and shared by every call on that GPU. In host code `scratch` is only a stand-in
// for the symbol; its device address on each GPU comes from scratchOnDevice.
constexpr int64_t kScratchParticles = int64_t(1) << 20;  // 1,048,576 particles
constexpr int64_t kScratchHinges = int64_t(1) << 21;     // 2,097,152 bending hinges
constexpr int64_t kScratchSegments = 4096;               // residual segments (at most 4096 particles of one object)
constexpr int64_t kScratchObjects = 4096;
constexpr int64_t kScratchDoubles = kScratchParticles + kScratchHinges + kScratchSegments + kScratchObjects;  // 24.06 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kScratchMaxDevices = 64;
std::mutex g_scratchMutex;
double *g_scratchPtr[kScratchMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current: resolved once per device and kept.
double *scratchOnDevice(int device)
{
    if (device < 0 || device >= kScratchMaxDevices)
        throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "device index beyond the allocation-free residual's device table");
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[device]) {
        void *p = nullptr;
        hipCheck(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw VbdError(SPG_VBD_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratchPtr[device] = static_cast<double *>(p);
    }
    return g_scratchPtr[device];
}

float intBitsAsFloat(int32_t i)
{
    float f;
    std::memcpy(&f, &i, sizeof(f));
    return f;
}

}  // namespace

struct spg_vbd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    double *scratchDev = nullptr;  // this device's `scratch` (spg_vbd_residual_noalloc)

    // objects and vertex numbering
    int32_t nObjects = 0;
    std::vector<int64_t> objOffset;     // global particle id of object o's particle 0
    int64_t nv = 0;
    std::vector<int32_t> slotOfGid, gidOfSlot;
    std::vector<int32_t> localColors;   // per global particle
    std::vector<int32_t> numColors;     // per object
    int32_t maxColors = 0;
    std::vector<int32_t> colorBegin;    // slot range of each colour, size maxColors + 1
    std::vector<int64_t> colorActiveV;  // vertices of objects that have this colour (roofline model)
    std::vector<double> colorBytes;     // compulsory bytes of one sweep of this colour
    double residualBytes = 0.0;         // compulsory bytes of one FP64 residual evaluation
    std::vector<int64_t> objSlotsOffsets;

    // device state
    DevBuf<float4> pos, xt, xold, vel;
    DevBuf<float4> savedPos, savedVel;  // device checkpoint
    DevBuf<float> mass;
    DevBuf<double> massd;  // FP64 masses for the residual
    DevBuf<unsigned long long> counters;
    DevBuf<double> r2, norms, normsAux;
    DevBuf<double> kc, kcCheck;  // per hinge k (theta - theta0) of the residual (solver's stream, check stream)
    // pipelined convergence checks of iterate_to_tolerance: a second stream computes the residual of
    // position snapshots while the solver's stream keeps sweeping
    // Two units in flight: the next sweeps are queued while the host reads a check (hiding the host
    // round trip). Deeper speculation does not pay: the residual and the sweeps each fill the GPU, so
    // sweeps ahead of a passing check only delay it.
    static constexpr int kPipe = 2;
    hipStream_t checkStream = nullptr;
    DevBuf<float4> snap;                      // kPipe x nv positions
    DevBuf<double> r2Check, partialCheck, normsCheck;  // normsCheck: kPipe x nObjects
    double *hostNormsCheck = nullptr;         // pinned, (kPipe + 1) x nObjects; the last row is |r(x~)|
    hipEvent_t evSnap[kPipe] = {}, evCheck[kPipe] = {}, evInertialCheck = nullptr;
    hipEvent_t evFeatures = nullptr, evPrediction = nullptr;  // learned start: features ready, prediction ready
    DevBuf<int32_t> objSlots;
    DevBuf<int64_t> objOffsets;
    // segments of the per-object residual reduction over the particle-ordered squared residuals:
    // particles segBegin[k] .. segBegin[k + 1] - 1 belong to one object; object o owns segments
    // objSeg[o] .. objSeg[o + 1] - 1
    DevBuf<int64_t> segBegin;
    DevBuf<int32_t> objSeg;
    DevBuf<double> segPartial;
    int32_t nSegments = 0;
    std::vector<double> r0Cache;  // |r(x~)| per object of the open substep
    bool r0Valid = false;
    std::vector<double> rCurCache;  // |r(x)| per object at the current positions
    bool rCurValid = false;
    bool posIsInertial = false;  // x == x~ (right after begin_substep)
    bool r0OnCheckStream = false;  // |r(x~)| enqueued on the check stream, not yet read back

    // incidence streams
    // row offsets (CSR, used by the residual; anchors also by the solve)
    DevBuf<int32_t> sprPtr, ancPtr, triPtr, hinEPtr, hinWPtr, tetPtr;
    // fixed-width record rows of the solve (anchors: CSR records)
    int32_t sprW = 0, triW = 0, hinEW = 0, hinWW = 0, tetW = 0;
    DevBuf<float4> spr, anc;
    DevBuf<int2> triO;
    DevBuf<float4> triG, triM;
    DevBuf<int4> hinEV, hinWV;
    DevBuf<float2> hinEP, hinWP;
    DevBuf<int4> tetO;
    DevBuf<float4> tetGa, tetGb, tetGc;
    // neighbourhood layout of the LDS-staged solve (lds_staging 2; nbW 0: not built)
    int32_t nbW = 0;
    DevBuf<int32_t> nbRow, tetL;
    DevBuf<float4> triL, sprL;
    DevBuf<NbHinge> hinEL, hinWL;
    // further energies: CSR rows per slot, elements (slots, padded FP64 parameters) for the residual, and the
    // solve's per-class fixed-width record rows
    DevBuf<int32_t> extPtr;
    DevBuf<int4> extV;
    DevBuf<double> extP;
    int32_t xrW[kXCls] = {};
    DevBuf<int4> xrV[kXCls];
    DevBuf<float4> xrP[kXCls];
    DevBuf<int2> xrR[kXCls];  // the residual's rows in the same layout: (element, slot in the element)
    // FP64 tables for the residual
    DevBuf<int32_t> sprElem, ancElem, triElem, hinEElem, hinWElem, tetElem;
    DevBuf<int32_t> triRowR, hinERowR, hinWRowR, tetRowR, sprRowR;  // element ids in the solve's row widths
    // the neighbourhood-staged residual's row entries (element id, slot indices into the neighbour row)
    DevBuf<int2> triRL, hinERL, hinWRL, tetRL, sprRL;
    DevBuf<int2> sprV;
    DevBuf<double2> sprP;
    DevBuf<double4> ancP;
    DevBuf<int4> triV;
    DevBuf<double> triP;
    DevBuf<int4> hinVe;
    DevBuf<double2> hinPe;
    DevBuf<int4> tetV;
    DevBuf<double> tetP;

    // learned warm start: x* - x~ of the last substep, one-ring tables, static features, workspaces
    DevBuf<float4> corr, savedCorr;
    DevBuf<int32_t> nbPtr, nbIdx, objOfSlot, reject;
    DevBuf<float> nbRest, feat, pred;
    DevBuf<float4> stat;
    double g[3] = {0.0, 0.0, 0.0};  // gravity of the open substep
    std::vector<int32_t> predicted;  // per object: the open substep started from a network prediction
    bool anyPredicted = false;
    int64_t predictionFallbacks = 0;
    // per object |r(start)| / |r(x~)| of the open substep's start as its gate measured it (NaN: not measured)
    std::vector<double> startRho;

    spg_vbd_launch launch{4, 256, 0, 0};
    bool timing = false;
    spg_vbd_stats stats{};
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;  // recorded, not yet accumulated
    std::vector<int32_t> eventColors;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> eventPool;

    bool inSubstep = false;
    double h = 0.0;

    hipGraphExec_t graphExec = nullptr;
    hipGraph_t graph = nullptr;
    std::array<double, 6> graphKey{};  // h, gx, gy, gz, iterations, nsubsteps
    int graphLaunchHash = -1;

    ~spg_vbd_ctx()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        dropGraph();
        if (evFeatures) (void)hipEventDestroy(evFeatures);
        if (evPrediction) (void)hipEventDestroy(evPrediction);
        if (checkStream) (void)hipStreamSynchronize(checkStream);
        for (int k = 0; k < kPipe; ++k) {
            if (evSnap[k]) (void)hipEventDestroy(evSnap[k]);
            if (evCheck[k]) (void)hipEventDestroy(evCheck[k]);
        }
        if (evInertialCheck) (void)hipEventDestroy(evInertialCheck);
        if (hostNormsCheck) (void)hipHostFree(hostNormsCheck);
        if (checkStream) (void)hipStreamDestroy(checkStream);
        for (auto &e : events) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        for (auto &e : eventPool) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }

    void dropGraph()
    {
        if (graphExec) (void)hipGraphExecDestroy(graphExec);
        if (graph) (void)hipGraphDestroy(graph);
        graphExec = nullptr;
        graph = nullptr;
    }

    SolveArgs solveArgs(int32_t color, float invdt2)
    {
        SolveArgs a{};
        a.pos = pos.p;
        a.xt = xt.p;
        a.sprW = sprW;
        a.spr = spr.p;
        a.ancPtr = ancPtr.p;
        a.anc = anc.p;
        a.triW = triW;
        a.triO = triO.p;
        a.triG = triG.p;
        a.triM = triM.p;
        a.hinEW = hinEW;
        a.hinEV = hinEV.p;
        a.hinEP = hinEP.p;
        a.hinWW = hinWW;
        a.hinWV = hinWV.p;
        a.hinWP = hinWP.p;
        a.tetW = tetW;
        a.tetO = tetO.p;
        a.tetGa = tetGa.p;
        a.tetGb = tetGb.p;
        a.tetGc = tetGc.p;
        a.nbW = nbW;
        a.nbRow = nbRow.p;
        a.triL = triL.p;
        a.hinEL = hinEL.p;
        a.hinWL = hinWL.p;
        a.tetL = tetL.p;
        a.sprL = sprL.p;
        a.extPtr = extPtr.p;
        a.extV = extV.p;
        a.extP = extP.p;
        for (int k = 0; k < kXCls; ++k) {
            a.xrW[k] = xrW[k];
            a.xrV[k] = xrV[k].p;
            a.xrP[k] = xrP[k].p;
            a.xrPlane[k] = static_cast<int64_t>(nv) * xrW[k];
        }
        a.counters = counters.p;
        a.begin = colorBegin[color];
        a.end = colorBegin[color + 1];
        a.invdt2 = invdt2;
        return a;
    }

    std::pair<hipEvent_t, hipEvent_t> takeEvents()
    {
        if (eventPool.empty()) {
            hipEvent_t a, b;
            hipCheck(hipEventCreate(&a), "hipEventCreate");
            hipCheck(hipEventCreate(&b), "hipEventCreate");
            return {a, b};
        }
        auto e = eventPool.back();
        eventPool.pop_back();
        return e;
    }

    void drainEvents()
    {
        if (events.empty()) return;
        hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
        for (size_t i = 0; i < events.size(); ++i) {
            float ms = 0.f;
            hipCheck(hipEventElapsedTime(&ms, events[i].first, events[i].second), "hipEventElapsedTime");
            stats.solve_ms += ms;
            stats.solve_launches += 1;
            const int32_t c = eventColors[i];
            stats.vertex_updates += colorBegin[c + 1] - colorBegin[c];
            stats.algorithmic_bytes += colorBytes[c];
            eventPool.push_back(events[i]);
        }
        events.clear();
        eventColors.clear();
    }

    // positions changed: the cached residual at the current positions is stale
    void positionsChanged()
    {
        rCurValid = false;
        posIsInertial = false;
    }

    void sweepColors(int32_t iterations, bool timed)
    {
        if (iterations > 0) positionsChanged();
        const float invdt2 = static_cast<float>(1.0 / (h * h));
        for (int32_t it = 0; it < iterations; ++it)
            for (int32_t c = 0; c < maxColors; ++c) {
                const SolveArgs a = solveArgs(c, invdt2);
                std::pair<hipEvent_t, hipEvent_t> ev{};
                if (timed) ev = takeEvents();
                hipCheck(launchSolveColor(a, launch.lanes_per_vertex, launch.block_size, launch.lds_staging, stream,
                                          ev.first, ev.second),
                         "vertex solve launch");
                if (timed) {
                    events.push_back(ev);
                    eventColors.push_back(c);
                    if (events.size() >= 8192) drainEvents();
                }
            }
    }

    // A convergence check still in flight from an earlier iterate_to_tolerance reads x~: the next
    // substep's writes to x~ are ordered after it (on the device, no host wait). Called outside graph
    // capture, before a new substep is enqueued.
    void orderAfterChecks()
    {
        for (int k = 0; checkStream && k < kPipe; ++k)
            hipCheck(hipStreamWaitEvent(stream, evCheck[k], 0), "hipStreamWaitEvent");
        if (evInertialCheck) hipCheck(hipStreamWaitEvent(stream, evInertialCheck, 0), "hipStreamWaitEvent");
    }

    void beginSubstep(double hh, const double *g)
    {
        h = hh;
        r0Valid = false;
        rCurValid = false;
        posIsInertial = true;
        r0OnCheckStream = false;
        anyPredicted = false;
        predicted.assign(nObjects, 0);
        startRho.assign(nObjects, std::numeric_limits<double>::quiet_NaN());
        for (int k = 0; k < 3; ++k) this->g[k] = g[k];
        hipCheck(launchBeginSubstep(pos.p, xold.p, vel.p, xt.p, mass.p, static_cast<int32_t>(nv), static_cast<float>(hh),
                                    static_cast<float>(g[0]), static_cast<float>(g[1]), static_cast<float>(g[2]),
                                    stream),
                 "begin substep");
        inSubstep = true;
    }

    void endSubstep()
    {
        positionsChanged();
        hipCheck(launchEndSubstep(pos.p, xold.p, xt.p, vel.p, corr.p, static_cast<int32_t>(nv),
                                  static_cast<float>(1.0 / h), stream),
                 "end substep");
        inSubstep = false;
    }

    // Launches the FP64 residual at x~ or at x and its per-object reduction into normsDev (device, one
    // double per object), on the context's stream, without waiting.
    void enqueueResidual(bool atInertial, double *normsDev)
    {
        enqueueResidualOf(atInertial ? xt.p : pos.p, r2.p, kc.p, segPartial.p, normsDev, stream);
    }

    // The same for arbitrary positions, per-vertex, per-hinge and per-segment work buffers and stream.
    void enqueueResidualOf(const float4 *positions, double *r2Buf, double *kcBuf, double *partialBuf, double *normsDev,
                           hipStream_t s)
    {
        ResidualArgs a{};
        a.pos = positions;
        a.xt = xt.p;
        a.mass = massd.p;
        a.invdt2 = 1.0 / (h * h);
        a.n = static_cast<int32_t>(nv);
        a.sprPtr = sprPtr.p;
        a.sprElem = sprElem.p;
        a.sprV = sprV.p;
        a.sprP = sprP.p;
        a.ancPtr = ancPtr.p;
        a.ancElem = ancElem.p;
        a.ancP = ancP.p;
        a.triPtr = triPtr.p;
        a.triElem = triElem.p;
        a.triV = triV.p;
        a.triP = triP.p;
        a.hinEPtr = hinEPtr.p;
        a.hinEElem = hinEElem.p;
        a.hinWPtr = hinWPtr.p;
        a.hinWElem = hinWElem.p;
        a.hinV = hinVe.p;
        a.hinP = hinPe.p;
        a.tetPtr = tetPtr.p;
        a.tetElem = tetElem.p;
        a.tetV = tetV.p;
        a.tetP = tetP.p;
        a.extPtr = extPtr.p;
        a.extV = extV.p;
        a.extP = extP.p;
        for (int k = 0; k < kXCls; ++k) {
            a.xrW[k] = xrW[k];
            a.xrR[k] = xrR[k].p;
        }
        a.order = objSlots.p;
        a.r2 = r2Buf;
        a.triW = triW;
        a.hinEW = hinEW;
        a.hinWW = hinWW;
        a.tetW = tetW;
        a.sprW = sprW;
        a.triRow = triRowR.p;
        a.hinERow = hinERowR.p;
        a.hinWRow = hinWRowR.p;
        a.tetRow = tetRowR.p;
        a.sprRow = sprRowR.p;
        a.nbW = nbW;
        a.nbRow = nbRow.p;
        a.triRL = triRL.p;
        a.hinERL = hinERL.p;
        a.hinWRL = hinWRL.p;
        a.tetRL = tetRL.p;
        a.sprRL = sprRL.p;
        a.hinKc = kcBuf;
        a.nHinges = static_cast<int32_t>(hinVe.n);
        hipCheck(launchResidual(a, s), "residual launch");
        hipCheck(launchReduceObjects(r2Buf, segBegin.p, nSegments, objSeg.p, nObjects, partialBuf, normsDev, s),
                 "reduce launch");
    }

    // Per-object FP64 |r| at the current positions with the device's `scratch` as the work memory of the residual
    // and its reduction (every entry a call reads is written first by that call: no clearing). The same kernels
    // in the same order as residual(), so the norms are bitwise those of spg_vbd_residual; nothing is cached.
    void residualNoAlloc(double *out)
    {
        if (!inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "residual needs an open substep (begin_substep)");
        if (nv > kScratchParticles)
            throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "allocation-free residual: more than 1048576 particles (" +
                                                        std::to_string(nv) + ")");
        if (static_cast<int64_t>(hinVe.n) > kScratchHinges)
            throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "allocation-free residual: more than 2097152 hinges (" +
                                                        std::to_string(hinVe.n) + ")");
        if (nObjects > kScratchObjects)
            throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "allocation-free residual: more than 4096 objects (" +
                                                        std::to_string(nObjects) + ")");
        if (nSegments > kScratchSegments)
            throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "allocation-free residual: more than 4096 reduction segments (" +
                                                        std::to_string(nSegments) + ")");
        if (!scratchDev) throw VbdError(SPG_VBD_ERR_STATE, "allocation-free residual: scratch address not resolved");
        if (nSegments == 0) {  // every object is empty
            std::fill(out, out + nObjects, 0.0);
            return;
        }
        double *r2s = scratchDev, *kcs = scratchDev + kScratchParticles, *partial = kcs + kScratchHinges,
               *nrm = partial + kScratchSegments;
        enqueueResidualOf(posIsInertial ? xt.p : pos.p, r2s, kcs, partial, nrm, stream);
        hipCheck(hipMemcpyAsync(out, nrm, sizeof(double) * nObjects, hipMemcpyDeviceToHost, stream), "download");
        hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    // Per-object FP64 |r| at x~ or at the current positions. Both are cached until they can change:
    // x~ is fixed for the whole substep, x until the next sweep or position write.
    void residual(bool atInertial, double *out)
    {
        if (!inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "residual needs an open substep (begin_substep)");
        if (!atInertial && posIsInertial) atInertial = true;  // x == x~: the same positions
        if (atInertial && !r0Valid && r0OnCheckStream) {
            hipCheck(hipStreamSynchronize(checkStream), "hipStreamSynchronize");
            r0Cache.assign(hostNormsCheck + kPipe * nObjects, hostNormsCheck + (kPipe + 1) * nObjects);
            r0Valid = true;
            r0OnCheckStream = false;
        }
        if (atInertial && r0Valid) {
            std::copy(r0Cache.begin(), r0Cache.end(), out);
            return;
        }
        if (!atInertial && rCurValid) {
            std::copy(rCurCache.begin(), rCurCache.end(), out);
            return;
        }
        enqueueResidual(atInertial, norms.p);
        hipCheck(hipMemcpyAsync(out, norms.p, sizeof(double) * nObjects, hipMemcpyDeviceToHost, stream), "download");
        hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
        if (atInertial) {
            r0Cache.assign(out, out + nObjects);
            r0Valid = true;
        } else {
            rCurCache.assign(out, out + nObjects);
            rCurValid = true;
        }
    }

    // acceleration scale of the learned features: |g|, or 1 without gravity
    double refAccel() const
    {
        const double a = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        return a > 0.0 ? a : 1.0;
    }

    void computeFeatures()
    {
        if (!inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "features need an open substep (begin_substep)");
        feat.reserve(static_cast<size_t>(nv) * kMlpFeatures);
        FeatureArgs a;
        a.xt = xt.p;
        a.vel = vel.p;
        a.corr = corr.p;
        a.nbPtr = nbPtr.p;
        a.nbIdx = nbIdx.p;
        a.nbRest = nbRest.p;
        a.stat = stat.p;
        a.n = static_cast<int32_t>(nv);
        a.h = static_cast<float>(h);
        a.gx = static_cast<float>(g[0]);
        a.gy = static_cast<float>(g[1]);
        a.gz = static_cast<float>(g[2]);
        a.invS = static_cast<float>(1.0 / (h * h * refAccel()));
        a.invU = static_cast<float>(1.0 / (h * refAccel()));
        a.out = feat.p;
        hipCheck(launchFeatures(a, stream), "features launch");
    }

    // Objects whose substep started from a network prediction: while one of them has a residual larger
    // than at x~ (or a non-finite one) its start is dropped and it continues from x~, as plain VBD.
    // Returns whether any object fell back (their positions are x~ again).
    bool fallBackFromPredictions(const std::vector<double> &r0, const std::vector<double> &r)
    {
        if (!anyPredicted) return false;
        std::vector<int32_t> rej(nObjects, 0);
        bool any = false;
        for (int32_t o = 0; o < nObjects; ++o)
            if (predicted[o] && !(r[o] <= r0[o])) {
                rej[o] = 1;
                predicted[o] = 0;
                ++predictionFallbacks;
                any = true;
            }
        anyPredicted = std::find(predicted.begin(), predicted.end(), 1) != predicted.end();
        if (any) {
            hipCheck(hipMemcpyAsync(reject.p, rej.data(), sizeof(int32_t) * nObjects, hipMemcpyHostToDevice, stream),
                     "upload");
            hipCheck(launchGate(pos.p, xt.p, objOfSlot.p, reject.p, static_cast<int32_t>(nv), stream), "gate launch");
            hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
            positionsChanged();
        }
        return any;
    }

    // gateMode 0: keep the prediction; 1: gate now (two FP64 residuals); 2: keep it and let the first
    // check of iterate_to_tolerance apply the same gate (no residual here; rho is reported as NaN).
    void warmStart(spg_mlp *net, int32_t smoothing, int32_t gateMode, int32_t *accepted, double *rho)
    {
        if (!inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "warm_start needs an open substep (begin_substep)");
        if (mlpDevice(net) != device) throw VbdError(SPG_VBD_ERR_INVALID, "network and context on different devices");
        if (smoothing < 0 || smoothing > 64) throw VbdError(SPG_VBD_ERR_INVALID, "smoothing passes must be in 0..64");
        // the deferred gate needs |r(x~)| in its first check: start it first, beside features and inference
        if (gateMode == 2 && !r0Valid && !r0OnCheckStream) enqueueInertialResidualOnCheckStream();
        if (!evFeatures) {
            hipCheck(hipEventCreateWithFlags(&evFeatures, hipEventDisableTiming), "hipEventCreate");
            hipCheck(hipEventCreateWithFlags(&evPrediction, hipEventDisableTiming), "hipEventCreate");
        }
        pred.reserve(static_cast<size_t>(nv) * 8);  // two ping-pong rows of 4 floats per vertex
        float *y = pred.p, *ySpare = pred.p + static_cast<size_t>(nv) * 4;
        hipCheck(hipMemsetAsync(y, 0, sizeof(float) * 4 * nv, stream), "memset");
        computeFeatures();
        hipCheck(hipEventRecord(evFeatures, stream), "hipEventRecord");
        // the inference runs on the network's stream while this stream computes |r(x~)| for the gate
        mlpInferAfter(net, feat.p, nv, y, 4, evFeatures, evPrediction);
        hipCheck(hipSetDevice(device), "hipSetDevice");
        const bool needResiduals = gateMode == 1 || (gateMode == 0 && rho);
        const bool needR0 = needResiduals && !r0Valid && !r0OnCheckStream;
        if (needR0) enqueueResidual(true, normsAux.p);
        hipCheck(hipStreamWaitEvent(stream, evPrediction, 0), "hipStreamWaitEvent");
        for (int32_t k = 0; k < smoothing; ++k) {
            hipCheck(launchSmoothPrediction(y, ySpare, nbPtr.p, nbIdx.p, static_cast<int32_t>(nv), stream), "smooth");
            std::swap(y, ySpare);
        }
        // a network trained on the change of the previous substep's offset predicts on top of that offset
        hipCheck(launchApplyWarmStart(pos.p, xt.p, mlpHistoryBase(net) ? corr.p : nullptr, y, 4,
                                      static_cast<float>(h * h * refAccel()), static_cast<int32_t>(nv), stream),
                 "warm start launch");
        positionsChanged();
        finishStart(gateMode, needResiduals, needR0, accepted, rho);
    }

    // The end of a start (network prediction or history extrapolation) already written to the positions:
    // gate 1 compares |r(start)| with |r(x~)| per object now and sends rejected objects back to x~; gate 2
    // leaves that test to the first check of iterate_to_tolerance; gate 0 keeps the start. Marks the kept
    // objects as started from a prediction and reports accepted / rho.
    void finishStart(int32_t gateMode, bool needResiduals, bool needR0, int32_t *accepted, double *rho)
    {
        const bool gate = gateMode == 1, deferGate = gateMode == 2;
        std::vector<int32_t> rej(nObjects, 0);
        std::vector<double> r0(nObjects, 0.0), rp(nObjects, 0.0);
        if (needResiduals) {  // |r(pred)| after the inference, one synchronisation for both norms
            enqueueResidual(false, norms.p);
            if (needR0)
                hipCheck(hipMemcpyAsync(r0.data(), normsAux.p, sizeof(double) * nObjects, hipMemcpyDeviceToHost, stream),
                         "download");
            hipCheck(hipMemcpyAsync(rp.data(), norms.p, sizeof(double) * nObjects, hipMemcpyDeviceToHost, stream),
                     "download");
            hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
            if (needR0) {
                r0Cache = r0;
                r0Valid = true;
            } else {
                residual(true, r0.data());  // cached, or waiting for the check stream's copy
            }
            rCurCache = rp;
            rCurValid = true;
        }
        if (gate) {
            bool any = false;
            for (int32_t o = 0; o < nObjects; ++o) {
                rej[o] = !(rp[o] <= r0[o]);
                any = any || rej[o];
            }
            if (any) {
                hipCheck(hipMemcpyAsync(reject.p, rej.data(), sizeof(int32_t) * nObjects, hipMemcpyHostToDevice, stream),
                         "upload");
                hipCheck(launchGate(pos.p, xt.p, objOfSlot.p, reject.p, static_cast<int32_t>(nv), stream), "gate launch");
                positionsChanged();
            }
        }
        predicted.assign(nObjects, 0);
        for (int32_t o = 0; o < nObjects; ++o) {
            predicted[o] = !rej[o];
            if (accepted) accepted[o] = !rej[o];
            if (rho)
                rho[o] = needResiduals ? (r0[o] > 0.0 ? rp[o] / r0[o] : rp[o]) : std::numeric_limits<double>::quiet_NaN();
        }
        anyPredicted = std::find(predicted.begin(), predicted.end(), 1) != predicted.end();
        startRho.assign(nObjects, std::numeric_limits<double>::quiet_NaN());
        if (needResiduals)
            for (int32_t o = 0; o < nObjects; ++o) startRho[o] = r0[o] > 0.0 ? rp[o] / r0[o] : rp[o];
        // a deferred gate leaves the prediction queued on the stream for the iterations that follow
        if (!deferGate) hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    // Start from the history: x = x~ + (x* - x~) of the previous substep (the solution's offset from its own
    // inertial guess carried over, the quantity the network's first three inputs hold), with the gate of
    // warmStart.
    void historyStart(int32_t gateMode, int32_t *accepted, double *rho)
    {
        if (!inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "history_start needs an open substep (begin_substep)");
        if (gateMode == 2 && !r0Valid && !r0OnCheckStream) enqueueInertialResidualOnCheckStream();
        const bool needResiduals = gateMode == 1 || (gateMode == 0 && rho);
        const bool needR0 = needResiduals && !r0Valid && !r0OnCheckStream;
        if (needR0) enqueueResidual(true, normsAux.p);
        hipCheck(launchApplyWarmStart(pos.p, xt.p, nullptr, reinterpret_cast<const float *>(corr.p), 4, 1.0f,
                                      static_cast<int32_t>(nv), stream),
                 "history start launch");
        positionsChanged();
        finishStart(gateMode, needResiduals, needR0, accepted, rho);
    }

    // largest per-object |r| / |r(x~)|; NaN when any object's residual is not finite
    double relMax(const std::vector<double> &r0, const std::vector<double> &r) const
    {
        double m = 0.0;
        for (int32_t o = 0; o < nObjects; ++o) {
            const double q = r0[o] > 0 ? r[o] / r0[o] : r[o];
            if (!std::isfinite(q)) return std::numeric_limits<double>::quiet_NaN();
            m = std::max(m, q);
        }
        return m;
    }

    void ensureCheckPipeline()
    {
        if (checkStream) return;
        hipCheck(hipStreamCreateWithFlags(&checkStream, hipStreamNonBlocking), "hipStreamCreate");
        for (int k = 0; k < kPipe; ++k) {
            hipCheck(hipEventCreateWithFlags(&evSnap[k], hipEventDisableTiming), "hipEventCreate");
            hipCheck(hipEventCreateWithFlags(&evCheck[k], hipEventDisableTiming), "hipEventCreate");
        }
        hipCheck(hipEventCreateWithFlags(&evInertialCheck, hipEventDisableTiming), "hipEventCreate");
        snap.alloc(static_cast<size_t>(kPipe) * nv);
        r2Check.alloc(nv);
        kcCheck.alloc(hinVe.n);
        partialCheck.alloc(nSegments);
        normsCheck.alloc(static_cast<size_t>(kPipe) * nObjects);
        hipCheck(hipHostMalloc(reinterpret_cast<void **>(&hostNormsCheck), sizeof(double) * (kPipe + 1) * nObjects),
                 "hipHostMalloc");
    }

    // |r(x~)| on the check stream into the pinned row kPipe of hostNormsCheck (read after the next check).
    void enqueueInertialResidualOnCheckStream()
    {
        hipCheck(hipEventRecord(evSnap[0], stream), "hipEventRecord");  // x~ is written by begin_substep
        hipCheck(hipStreamWaitEvent(checkStream, evSnap[0], 0), "hipStreamWaitEvent");
        enqueueResidualOf(xt.p, r2Check.p, kcCheck.p, partialCheck.p, normsAux.p, checkStream);
        hipCheck(hipMemcpyAsync(hostNormsCheck + kPipe * nObjects, normsAux.p, sizeof(double) * nObjects,
                                hipMemcpyDeviceToHost, checkStream),
                 "download");
        hipCheck(hipEventRecord(evInertialCheck, checkStream), "hipEventRecord");
        r0OnCheckStream = true;
    }

    // Sweeps in units of `every` until max_j |r_j(x)| / |r_j(x~)| <= tol after a unit, or max sweeps.
    // After each unit the positions are copied to a snapshot whose FP64 residual is computed on the
    // check stream while the solver's stream goes on with the next units (at most kPipe in flight).
    // When a check passes, the positions are restored from its snapshot, so the result (state, sweep
    // count, rho) is that of checking synchronously after every unit. Unless the residual at the
    // start is already known, the first unit has no sweeps: its check is that of the start (for a
    // network prediction whose gate was deferred, this check is the gate), and |r(x~)| is computed on
    // the check stream as well.
    void iterateToTolerance(double tol, int32_t maxSweeps, int32_t every, int32_t *sweepsOut, double *rhoOut)
    {
        std::vector<double> r0(nObjects), r(nObjects);
        const bool startKnown = rCurValid || (posIsInertial && r0Valid);
        if (startKnown || maxSweeps == 0) {
            residual(true, r0.data());
            residual(false, r.data());
            if (fallBackFromPredictions(r0, r)) residual(false, r.data());
            const double cur = relMax(r0, r);
            if (cur <= tol || maxSweeps == 0) {
                *sweepsOut = 0;
                *rhoOut = cur;
                return;
            }
        }
        ensureCheckPipeline();
        bool firstUnitEmpty = !startKnown;
        // |r(x~)|: known, equal to the first check (x == x~), or computed first on the check stream
        const bool r0FromFirstCheck = !r0Valid && !r0OnCheckStream && firstUnitEmpty && posIsInertial;
        if (r0Valid) r0 = r0Cache;
        if (!r0Valid && !r0FromFirstCheck && !r0OnCheckStream) enqueueInertialResidualOnCheckStream();
        bool r0Pending = !r0Valid;
        std::vector<int32_t> sweepsAt{0};  // sweeps done after unit u: sweepsAt[u + 1]
        int32_t issued = 0, checked = 0, done = 0;
        const size_t nvs = static_cast<size_t>(nv);
        for (;;) {
            while (sweepsAt[issued] < maxSweeps && issued - checked < kPipe) {
                const int32_t n = (issued == 0 && firstUnitEmpty) ? 0 : std::min(every, maxSweeps - sweepsAt[issued]);
                sweepColors(n, timing);
                sweepsAt.push_back(sweepsAt[issued] + n);
                const int k = issued % kPipe;
                // the last check that read this snapshot slot (this call or an earlier one) must be done
                hipCheck(hipStreamWaitEvent(stream, evCheck[k], 0), "hipStreamWaitEvent");
                hipCheck(launchCopy(pos.p, snap.p + k * nvs, nv, stream), "snapshot");
                hipCheck(hipEventRecord(evSnap[k], stream), "hipEventRecord");
                hipCheck(hipStreamWaitEvent(checkStream, evSnap[k], 0), "hipStreamWaitEvent");
                enqueueResidualOf(snap.p + k * nvs, r2Check.p, kcCheck.p, partialCheck.p, normsCheck.p + k * nObjects,
                                  checkStream);
                hipCheck(hipMemcpyAsync(hostNormsCheck + k * nObjects, normsCheck.p + k * nObjects,
                                        sizeof(double) * nObjects, hipMemcpyDeviceToHost, checkStream),
                         "download");
                hipCheck(hipEventRecord(evCheck[k], checkStream), "hipEventRecord");
                ++issued;
            }
            const int k = checked % kPipe;
            hipCheck(hipEventSynchronize(evCheck[k]), "hipEventSynchronize");
            std::copy(hostNormsCheck + k * nObjects, hostNormsCheck + (k + 1) * nObjects, r.begin());
            if (r0Pending) {  // the check stream computed |r(x~)| before this first check
                if (r0FromFirstCheck)
                    r0 = r;  // the first check is at x == x~
                else
                    std::copy(hostNormsCheck + kPipe * nObjects, hostNormsCheck + (kPipe + 1) * nObjects, r0.begin());
                r0Cache = r0;
                r0Valid = true;
                r0Pending = false;
                r0OnCheckStream = false;
            }
            ++checked;
            if (checked == 1 && firstUnitEmpty && anyPredicted) {  // the deferred gate's check
                startRho.assign(nObjects, std::numeric_limits<double>::quiet_NaN());
                for (int32_t o = 0; o < nObjects; ++o)
                    if (predicted[o]) startRho[o] = r0[o] > 0.0 ? r[o] / r0[o] : r[o];
            }
            bool fallBack = false;
            for (int32_t o = 0; anyPredicted && o < nObjects; ++o) fallBack = fallBack || (predicted[o] && !(r[o] <= r0[o]));
            const bool stop = fallBack || relMax(r0, r) <= tol || sweepsAt[checked] >= maxSweeps;
            if (!stop) continue;
            // the solver's stream may be ahead of this unit: go back to its snapshot (a check still in
            // flight for a later unit finishes on its own; its slot is guarded by evCheck)
            if (issued > checked) hipCheck(launchCopy(snap.p + k * nvs, pos.p, nv, stream), "rollback");
            hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
            positionsChanged();
            rCurCache = r;
            rCurValid = true;
            done = sweepsAt[checked];
            issued = checked;
            sweepsAt.resize(checked + 1);
            firstUnitEmpty = false;
            if (fallBack) {  // predicted objects above x~ go back to x~; the pipeline refills from there
                fallBackFromPredictions(r0, r);
                residual(false, r.data());
                if (!(relMax(r0, r) <= tol) && done < maxSweeps) continue;
            }
            break;
        }
        *sweepsOut = done;
        *rhoOut = relMax(r0, r);
    }

    void enqueueStep(double dt, int32_t nsub, int32_t iterations, const double *g, bool timed)
    {
        const double hh = dt / nsub;
        for (int32_t s = 0; s < nsub; ++s) {
            beginSubstep(hh, g);
            sweepColors(iterations, timed);
            endSubstep();
        }
    }

    void step(double dt, int32_t nsub, int32_t iterations, const double *g)
    {
        if (inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "step called inside an open substep");
        if (!(dt > 0) || nsub < 1 || iterations < 0) throw VbdError(SPG_VBD_ERR_INVALID, "invalid dt/substeps/iterations");
        orderAfterChecks();
        if (launch.use_graph && !timing) {
            const std::array<double, 6> key{dt / nsub, g[0], g[1], g[2], double(iterations), double(nsub)};
            const int lh = launch.lanes_per_vertex * 100000 + launch.block_size * 10 + launch.lds_staging;
            if (!graphExec || key != graphKey || lh != graphLaunchHash) {
                dropGraph();
                hipCheck(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
                try {
                    enqueueStep(dt, nsub, iterations, g, false);
                } catch (...) {
                    hipGraph_t tmp = nullptr;
                    (void)hipStreamEndCapture(stream, &tmp);
                    if (tmp) (void)hipGraphDestroy(tmp);
                    throw;
                }
                hipCheck(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
                hipCheck(hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
                graphKey = key;
                graphLaunchHash = lh;
            }
            hipCheck(hipGraphLaunch(graphExec, stream), "hipGraphLaunch");
            h = dt / nsub;
            return;
        }
        enqueueStep(dt, nsub, iterations, g, timing);
    }
};

namespace {

template <class F>
int guarded(spg_vbd_ctx *ctx, F &&f)
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

struct Incidence {
    int32_t slot;    // owning vertex slot
    int32_t elem;    // element id within its type table
    int32_t local;   // position of the owning vertex in the stencil
};

// Counting sort of incidences by owning slot, stable: keeps energy order then ascending stencil.
std::vector<int32_t> sortBySlot(const std::vector<Incidence> &inc, int64_t nv, std::vector<int32_t> &ptr)
{
    ptr.assign(nv + 1, 0);
    for (const auto &i : inc) ++ptr[i.slot + 1];
    for (int64_t v = 0; v < nv; ++v) ptr[v + 1] += ptr[v];
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1), order(inc.size());
    for (int32_t k = 0; k < static_cast<int32_t>(inc.size()); ++k) order[fill[inc[k].slot]++] = k;
    return order;
}

int32_t rowWidth(const std::vector<int32_t> &ptr)
{
    int32_t w = 0;
    for (size_t v = 0; v + 1 < ptr.size(); ++v) w = std::max(w, ptr[v + 1] - ptr[v]);
    return w;
}

// Fixed-width rows of the solve: record j of vertex slot v at v * width + j, padded with `pad`;
// make(k) builds the record of the k-th incidence in slot order.
template <class Rec, class Make>
std::vector<Rec> toRows(const std::vector<int32_t> &ptr, int32_t width, Rec pad, Make make)
{
    const size_t nv = ptr.size() - 1;
    std::vector<Rec> rows(nv * static_cast<size_t>(width), pad);
    for (size_t v = 0; v < nv; ++v)
        for (int32_t k = ptr[v]; k < ptr[v + 1]; ++k) rows[v * width + (k - ptr[v])] = make(k);
    return rows;
}

void buildFeatureTables(spg_vbd_ctx &c, const spg_vbd_object *objs, int32_t nObjects);

// The neighbourhood layout of the LDS-staged solve (lds_staging 2), from the fixed-width rows of the direct
// solve: per vertex slot v a row [v, then each distinct vertex its records name, in order of first use over
// the classes in solve order], and per record the 5-bit indices of its vertices in that row. Not built
// (c.nbW = 0) when some vertex names more than kNbLayoutMax - 1 other vertices.
constexpr int32_t kNbLayoutMax = 16;

// element-id rows of the residual in the solve's widths (host copies)
struct ResidualRowsH {
    std::vector<int32_t> tri, hin[2], tet, spr;
};

void buildNeighbourhoodLayout(spg_vbd_ctx &c, const std::vector<float4> &sprRows, const std::vector<int2> &triO,
                              const std::vector<float4> &triM, const std::vector<int4> (&hinV)[2],
                              const std::vector<float2> (&hinP)[2], const std::vector<int4> &tetO,
                              const ResidualRowsH &res, const std::vector<int4> &triE, const std::vector<int4> &hinE,
                              const std::vector<int4> &tetE, const std::vector<int2> &sprE)
{
    const int64_t nv = c.nv;
    std::vector<int32_t> ent;  // row of the vertex being built
    auto localOf = [&](int32_t s) -> int32_t {
        for (size_t k = 0; k < ent.size(); ++k)
            if (ent[k] == s) return static_cast<int32_t>(k);
        ent.push_back(s);
        return static_cast<int32_t>(ent.size() - 1);
    };
    // pass 1: row widths
    int32_t width = 1;
    std::vector<std::vector<int32_t>> rows(nv);
    for (int64_t v = 0; v < nv; ++v) {
        ent.assign(1, static_cast<int32_t>(v));
        for (int32_t j = 0; j < c.triW; ++j) {
            const int2 o = triO[v * c.triW + j];
            if (o.x >= 0) {
                localOf(o.x);
                localOf(o.y);
            }
        }
        for (int cls = 0; cls < 2; ++cls) {
            const int32_t w = cls == 0 ? c.hinEW : c.hinWW;
            for (int32_t j = 0; j < w; ++j) {
                const int4 h = hinV[cls][v * w + j];
                if (h.x >= 0) {
                    localOf(h.x);
                    localOf(h.y);
                    localOf(h.z);
                    localOf(h.w);
                }
            }
        }
        for (int32_t j = 0; j < c.tetW; ++j) {
            const int4 o = tetO[v * c.tetW + j];
            if (o.x >= 0) {
                localOf(o.x);
                localOf(o.y);
                localOf(o.z);
            }
        }
        for (int32_t j = 0; j < c.sprW; ++j) {
            int32_t other;
            std::memcpy(&other, &sprRows[v * c.sprW + j].x, sizeof(other));
            if (other >= 0) localOf(other);
        }
        if (static_cast<int32_t>(ent.size()) > kNbLayoutMax) {
            c.nbW = 0;
            return;
        }
        width = std::max(width, static_cast<int32_t>(ent.size()));
        rows[v] = ent;
    }
    // pass 2: the row table and the records with local indices
    std::vector<int32_t> nb(static_cast<size_t>(nv) * width, -1);
    for (int64_t v = 0; v < nv; ++v) std::copy(rows[v].begin(), rows[v].end(), nb.begin() + v * width);
    auto loc = [&](int64_t v, int32_t s) {
        const auto &r = rows[v];
        for (size_t k = 0; k < r.size(); ++k)
            if (r[k] == s) return static_cast<int32_t>(k);
        throw VbdError(SPG_VBD_ERR_INVALID, "neighbourhood layout: slot missing from its row");
    };
    std::vector<float4> triL(triM.size()), sprL(sprRows.size());
    std::vector<NbHinge> hinL[2];
    std::vector<int32_t> tetL(tetO.size(), -1);
    for (int64_t v = 0; v < nv; ++v) {
        for (int32_t j = 0; j < c.triW; ++j) {
            const size_t r = v * c.triW + j;
            const int2 o = triO[r];
            const int32_t pk = o.x >= 0 ? loc(v, o.x) | loc(v, o.y) << 5 : -1;
            triL[r] = make_float4(triM[r].x, triM[r].y, triM[r].z, intBitsAsFloat(pk));
        }
        for (int32_t j = 0; j < c.tetW; ++j) {
            const size_t r = v * c.tetW + j;
            const int4 o = tetO[r];
            tetL[r] = o.x >= 0 ? loc(v, o.x) | loc(v, o.y) << 5 | loc(v, o.z) << 10 : -1;
        }
        for (int32_t j = 0; j < c.sprW; ++j) {
            const size_t r = v * c.sprW + j;
            int32_t other;
            std::memcpy(&other, &sprRows[r].x, sizeof(other));
            sprL[r] = make_float4(intBitsAsFloat(other >= 0 ? loc(v, other) : -1), sprRows[r].y, sprRows[r].z, 0.f);
        }
    }
    for (int cls = 0; cls < 2; ++cls) {
        const int32_t w = cls == 0 ? c.hinEW : c.hinWW;
        hinL[cls].resize(hinV[cls].size());
        for (int64_t v = 0; v < nv; ++v)
            for (int32_t j = 0; j < w; ++j) {
                const size_t r = v * w + j;
                const int4 h = hinV[cls][r];
                const int32_t pk =
                    h.x >= 0 ? loc(v, h.x) | loc(v, h.y) << 5 | loc(v, h.z) << 10 | loc(v, h.w) << 15 : -1;
                hinL[cls][r] = NbHinge{hinP[cls][r].x, hinP[cls][r].y, pk};
            }
    }
    // the residual's row entries: element id and the slab indices of its slots in stencil order
    auto resEntries = [&](const std::vector<int32_t> &ids, int32_t w, auto slotsOf) {
        std::vector<int2> out(ids.size(), make_int2(-1, -1));
        for (int64_t v = 0; v < nv; ++v)
            for (int32_t j = 0; j < w; ++j) {
                const size_t r = v * w + j;
                if (ids[r] >= 0) out[r] = make_int2(ids[r], slotsOf(v, ids[r]));
            }
        return out;
    };
    auto pack4 = [&](int64_t v, const int4 &e, int n) {
        const int32_t s[4] = {e.x, e.y, e.z, e.w};
        int32_t pk = 0;
        for (int k = 0; k < n; ++k) pk |= loc(v, s[k]) << (5 * k);
        return pk;
    };
    const auto triRL = resEntries(res.tri, c.triW, [&](int64_t v, int32_t e) { return pack4(v, triE[e], 3); });
    const auto hinERL = resEntries(res.hin[0], c.hinEW, [&](int64_t v, int32_t e) { return pack4(v, hinE[e], 4); });
    const auto hinWRL = resEntries(res.hin[1], c.hinWW, [&](int64_t v, int32_t e) { return pack4(v, hinE[e], 4); });
    const auto tetRL = resEntries(res.tet, c.tetW, [&](int64_t v, int32_t e) { return pack4(v, tetE[e], 4); });
    const auto sprRL = resEntries(res.spr, c.sprW, [&](int64_t v, int32_t e) {
        return loc(v, sprE[e].x == v ? sprE[e].y : sprE[e].x);
    });
    if (!triRL.empty()) c.triRL.upload(triRL, c.stream);
    if (!hinERL.empty()) c.hinERL.upload(hinERL, c.stream);
    if (!hinWRL.empty()) c.hinWRL.upload(hinWRL, c.stream);
    if (!tetRL.empty()) c.tetRL.upload(tetRL, c.stream);
    if (!sprRL.empty()) c.sprRL.upload(sprRL, c.stream);
    c.nbW = width;
    c.nbRow.upload(nb, c.stream);
    if (!triL.empty()) c.triL.upload(triL, c.stream);
    if (!hinL[0].empty()) c.hinEL.upload(hinL[0], c.stream);
    if (!hinL[1].empty()) c.hinWL.upload(hinL[1], c.stream);
    if (!tetL.empty()) c.tetL.upload(tetL, c.stream);
    if (!sprL.empty()) c.sprL.upload(sprL, c.stream);
    // the host vectors are released when this returns; the uploads must have read them
    hipCheck(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
}

// FP32 coefficients of one further-energy incidence (element parameters p, the incidence's slot in the stencil),
// in the planes of its solve class (the table beside xclsOf).
void furtherCoefficients(int cls, int slot, const double *p, float4 *q)
{
    auto f = [](double x) { return static_cast<float>(x); };
    switch (cls) {
    case 0: q[0] = make_float4(f(p[0]), f(p[1]), slot == 0 ? 1.f : -1.f, 0.f); break;
    case 1:
        q[0] = make_float4(f(p[0]), f(p[1]), f(p[2]), f(p[3]));
        q[1] = make_float4(f(p[4]), f(p[5]), f(p[6]), static_cast<float>(slot));
        break;
    case 2: {
        const float4 d = make_float4(f(p[0]), f(p[1]), f(p[2]), f(p[3]));
        q[0] = d;
        q[1] = make_float4(f(p[4]), f(p[5]), f(p[6]), f(p[7]));
        q[2] = make_float4(slot == 0 ? -(d.x + d.z) : slot == 1 ? d.x : d.z,
                           slot == 0 ? -(d.y + d.w) : slot == 1 ? d.y : d.w, 0.f, 0.f);
        break;
    }
    case 3:
        q[0] = make_float4(f(p[0]), f(p[1]), f(p[2]), f(p[3]));
        q[1] = make_float4(f(p[4]), f(p[slot]), 0.f, 0.f);
        break;
    case 4:
    case 5: q[0] = make_float4(f(p[0]), f(p[1]), (slot == 1 || slot == 3) ? 1.f : 0.f, 0.f); break;
    default: {
        float gm[3];
        for (int k = 0; k < 3; ++k) gm[k] = slot == 0 ? -f(p[k] + p[3 + k] + p[6 + k]) : f(p[3 * (slot - 1) + k]);
        q[0] = make_float4(f(p[0]), f(p[1]), f(p[2]), f(p[3]));
        q[1] = make_float4(f(p[4]), f(p[5]), f(p[6]), f(p[7]));
        q[2] = make_float4(f(p[8]), gm[0], gm[1], gm[2]);
        q[3] = make_float4(f(p[9]), f(p[10]), 0.f, 0.f);
    }
    }
}

// The solve's records of the further energies: per incidence class a fixed-width row per vertex slot (in each row
// the incidences in the reference's order, energy then ascending stencil), each record the element's slots and
// its coefficient planes.
void buildFurtherRecords(spg_vbd_ctx &c, const std::vector<Incidence> &extIncList, const std::vector<int4> &extVh,
                         const std::vector<double> &extPh)
{
    std::vector<Incidence> cls[kXCls];
    for (const auto &in : extIncList) cls[xclsOf(in.local >> 2, in.local & 3)].push_back(in);
    for (int k = 0; k < kXCls; ++k) {
        c.xrW[k] = 0;
        c.xrV[k].release();
        c.xrR[k].release();
        c.xrP[k].release();
        if (cls[k].empty()) continue;
        std::vector<int32_t> ptr;
        const auto order = sortBySlot(cls[k], c.nv, ptr);
        const int32_t w = rowWidth(ptr);
        const int np = xclsPlanes(k);
        const size_t plane = static_cast<size_t>(c.nv) * w;
        std::vector<int4> rv(plane, make_int4(-1, -1, -1, -1));
        std::vector<int2> rr(plane, make_int2(-1, -1));
        std::vector<float4> rp(plane * np, make_float4(0.f, 0.f, 0.f, 0.f));
        for (int64_t v = 0; v < c.nv; ++v)
            for (int32_t i = ptr[v]; i < ptr[v + 1]; ++i) {
                const Incidence &in = cls[k][order[i]];
                const size_t r = static_cast<size_t>(v) * w + (i - ptr[v]);
                float4 q[4];
                furtherCoefficients(k, in.local & 3, &extPh[static_cast<size_t>(kExtParams) * in.elem], q);
                rv[r] = extVh[in.elem];
                rr[r] = make_int2(in.elem, in.local & 3);
                for (int j = 0; j < np; ++j) rp[j * plane + r] = q[j];
            }
        c.xrW[k] = w;
        c.xrV[k].upload(rv, c.stream);
        c.xrR[k].upload(rr, c.stream);
        c.xrP[k].upload(rp, c.stream);
        // the host vectors are released at the end of this iteration; the uploads must have read them
        hipCheck(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
    }
}

void buildContext(spg_vbd_ctx &c, const spg_vbd_object *objs, int32_t nObjects)
{
    if (nObjects < 1 || !objs) throw VbdError(SPG_VBD_ERR_INVALID, "at least one object required");
    c.nObjects = nObjects;
    c.objOffset.assign(nObjects + 1, 0);
    for (int32_t o = 0; o < nObjects; ++o) {
        const auto &ob = objs[o];
        if (ob.n_particles < 1) throw VbdError(SPG_VBD_ERR_INVALID, "object without particles");
        if (!ob.x || !ob.mass) throw VbdError(SPG_VBD_ERR_INVALID, "object without positions or masses");
        if (ob.n_energies < 0 || (ob.n_energies > 0 && !ob.energies))
            throw VbdError(SPG_VBD_ERR_INVALID, "invalid energy list");
        c.objOffset[o + 1] = c.objOffset[o] + ob.n_particles;
    }
    c.nv = c.objOffset[nObjects];
    if (c.nv >= (int64_t(1) << 31) - 1) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "too many particles");

    // validate energies, colour each object (VBD::computeParallelVertexGroups, vbd.cpp:203-229)
    c.localColors.assign(c.nv, 0);
    c.numColors.assign(nObjects, 0);
    for (int32_t o = 0; o < nObjects; ++o) {
        const auto &ob = objs[o];
        std::vector<FlatStencilSet> sets;
        for (int32_t e = 0; e < ob.n_energies; ++e) {
            const auto &en = ob.energies[e];
            const int ss = stencilSizeOf(en.type);
            if (!ss) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "unsupported energy type " + std::to_string(en.type));
            if (en.n_stencils < 0 || (en.n_stencils > 0 && (!en.stencils || !en.params)))
                throw VbdError(SPG_VBD_ERR_INVALID, "invalid energy arrays");
            for (int64_t i = 0; i < en.n_stencils * ss; ++i)
                if (en.stencils[i] < 0 || en.stencils[i] >= ob.n_particles)
                    throw VbdError(SPG_VBD_ERR_INVALID, "stencil index out of range");
            const int np = paramsOf(en.type);
            for (int64_t i = 0; i < en.n_stencils * np; ++i)
                if (!std::isfinite(en.params[i])) throw VbdError(SPG_VBD_ERR_INVALID, "non-finite energy parameter");
            if (en.type == SPG_VBD_MEMBRANE_STVK) {
                for (int64_t i = 0; i < en.n_stencils; ++i) {
                    const double *K = en.params + i * np + 4;
                    if (K[0] != K[4] || K[1] != K[3] || K[2] != 0 || K[5] != 0 || K[6] != 0 || K[7] != 0)
                        throw VbdError(SPG_VBD_ERR_UNSUPPORTED,
                                       "membrane stiffness matrix is not of the StVK form [[a,b,0],[b,a,0],[0,0,c]]");
                }
            }
            if (en.type == SPG_VBD_SPRING || en.type == SPG_VBD_SPRING_CONTINUUM ||
                en.type == SPG_VBD_SPRING_SQUARED_CONTINUUM)
                for (int64_t i = 0; i < en.n_stencils; ++i)
                    if (en.params[i * np] <= 0) throw VbdError(SPG_VBD_ERR_INVALID, "Zero/negative rest length invalid");
            sets.push_back({en.stencils, en.n_stencils * ss, ss});
        }
        const auto col = colorVertices(ob.n_particles, sets);
        int32_t nc = 0;
        for (int64_t i = 0; i < ob.n_particles; ++i) {
            c.localColors[c.objOffset[o] + i] = col[i];
            nc = std::max(nc, col[i] + 1);
        }
        c.numColors[o] = nc;
        c.maxColors = std::max(c.maxColors, nc);
    }

    // colour-major slot numbering
    c.slotOfGid.assign(c.nv, -1);
    c.gidOfSlot.assign(c.nv, -1);
    c.colorBegin.assign(c.maxColors + 1, 0);
    {
        std::vector<int64_t> cnt(c.maxColors, 0);
        for (int64_t g = 0; g < c.nv; ++g) ++cnt[c.localColors[g]];
        for (int32_t k = 0; k < c.maxColors; ++k) c.colorBegin[k + 1] = static_cast<int32_t>(c.colorBegin[k] + cnt[k]);
        std::vector<int32_t> fill(c.colorBegin.begin(), c.colorBegin.end() - 1);
        // objects in order, particles ascending: within colour k the runs are object-major
        for (int64_t g = 0; g < c.nv; ++g) {
            const int32_t s = fill[c.localColors[g]]++;
            c.slotOfGid[g] = s;
            c.gidOfSlot[s] = static_cast<int32_t>(g);
        }
    }
    // object-major slot lists for per-object reductions
    {
        std::vector<int32_t> objSlots(c.nv);
        c.objSlotsOffsets.assign(nObjects + 1, 0);
        for (int32_t o = 0; o < nObjects; ++o) {
            c.objSlotsOffsets[o + 1] = c.objOffset[o + 1];
            for (int64_t g = c.objOffset[o]; g < c.objOffset[o + 1]; ++g) objSlots[g] = c.slotOfGid[g];
        }
        c.objSlots.upload(objSlots, c.stream);
        c.objOffsets.upload(c.objSlotsOffsets, c.stream);
        std::vector<int64_t> seg{0};
        std::vector<int32_t> objSeg{0};
        for (int32_t o = 0; o < nObjects; ++o) {
            for (int64_t b = c.objSlotsOffsets[o]; b < c.objSlotsOffsets[o + 1]; b += kReduceSegment)
                seg.push_back(std::min<int64_t>(b + kReduceSegment, c.objSlotsOffsets[o + 1]));
            objSeg.push_back(static_cast<int32_t>(seg.size() - 1));
        }
        c.nSegments = static_cast<int32_t>(seg.size() - 1);
        c.segBegin.upload(seg, c.stream);
        c.objSeg.upload(objSeg, c.stream);
        c.segPartial.alloc(c.nSegments);
    }

    // per-vertex state
    std::vector<float4> hpos(c.nv), hvel(c.nv), hzero(c.nv, make_float4(0, 0, 0, 0));
    std::vector<float> hmass(c.nv);
    std::vector<double> hmassd(c.nv);
    for (int32_t o = 0; o < nObjects; ++o) {
        const auto &ob = objs[o];
        for (int64_t i = 0; i < ob.n_particles; ++i) {
            const int32_t s = c.slotOfGid[c.objOffset[o] + i];
            if (!finite3(ob.x + 3 * i) || (ob.v && !finite3(ob.v + 3 * i)) || !std::isfinite(ob.mass[i]))
                throw VbdError(SPG_VBD_ERR_NONFINITE, "non-finite particle state");
            hpos[s] = make_float4(float(ob.x[3 * i]), float(ob.x[3 * i + 1]), float(ob.x[3 * i + 2]), 0.f);
            hvel[s] = ob.v ? make_float4(float(ob.v[3 * i]), float(ob.v[3 * i + 1]), float(ob.v[3 * i + 2]), 0.f)
                           : make_float4(0, 0, 0, 0);
            hmass[s] = float(ob.mass[i]);
            hmassd[s] = ob.mass[i];
        }
    }
    c.massd.upload(hmassd, c.stream);
    c.pos.upload(hpos, c.stream);
    c.vel.upload(hvel, c.stream);
    c.xt.upload(hpos, c.stream);
    c.xold.upload(hzero, c.stream);
    c.mass.upload(hmass, c.stream);
    c.counters.upload(std::vector<unsigned long long>(2, 0ull), c.stream);
    c.r2.alloc(c.nv);
    c.norms.alloc(nObjects);
    c.normsAux.alloc(nObjects);

    // incidence streams per type (VBD::computeStencilInfoPerVertex, vbd.cpp:231-268)
    std::array<std::vector<Incidence>, 6> inc;
    std::vector<int2> sprV;
    std::vector<double2> sprP;
    std::vector<double4> ancP;
    std::vector<int4> triV, hinV, tetV;
    std::vector<double> triP, tetP;
    std::vector<double2> hinP;
    std::vector<Incidence> extIncList;  // local = type << 2 | slot in the stencil
    std::vector<int4> extVh;
    std::vector<double> extPh;
    double extResidualBytes = 0.0;
    std::vector<double> colorRecordBytes(c.maxColors, 0.0);
    std::vector<int64_t> colorIncidences(c.maxColors, 0);
    for (int32_t o = 0; o < nObjects; ++o) {
        const auto &ob = objs[o];
        const int64_t base = c.objOffset[o];
        for (int32_t e = 0; e < ob.n_energies; ++e) {
            const auto &en = ob.energies[e];
            const int ss = stencilSizeOf(en.type), np = paramsOf(en.type);
            for (int64_t i = 0; i < en.n_stencils; ++i) {
                const int32_t *st = en.stencils + i * ss;
                const double *p = en.params + i * np;
                int32_t sl[4] = {0, 0, 0, 0};
                for (int k = 0; k < ss; ++k) sl[k] = c.slotOfGid[base + st[k]];
                int32_t elem = 0;
                // SpringContinuumEnergy is a spring: 1/2 Keff (|d| / L0 - 1)^2 = 1/2 (Keff / L0^2) (|d| - L0)^2, so it
                // joins the spring class (closed-form blocks) with stiffness Keff / L0^2
                const int cls = en.type == SPG_VBD_SPRING_CONTINUUM ? SPG_VBD_SPRING : en.type;
                switch (cls) {
                case SPG_VBD_SPRING:
                    elem = static_cast<int32_t>(sprV.size());
                    sprV.push_back(make_int2(sl[0], sl[1]));
                    sprP.push_back(cls == en.type ? make_double2(p[0], p[1]) : make_double2(p[0], p[1] / (p[0] * p[0])));
                    break;
                case SPG_VBD_SPRING_ANCHOR:
                    elem = static_cast<int32_t>(ancP.size());
                    ancP.push_back(make_double4(p[0], p[1], p[2], p[3]));
                    break;
                case SPG_VBD_MEMBRANE_STVK:
                    elem = static_cast<int32_t>(triV.size());
                    triV.push_back(make_int4(sl[0], sl[1], sl[2], 0));
                    triP.insert(triP.end(), {p[0], p[1], p[2], p[3], p[4], p[5], p[12], 0.0});
                    break;
                case SPG_VBD_DISCRETE_BENDING:
                    elem = static_cast<int32_t>(hinV.size());
                    hinV.push_back(make_int4(sl[0], sl[1], sl[2], sl[3]));
                    hinP.push_back(make_double2(p[0], p[1]));
                    break;
                case SPG_VBD_STABLE_NEO_HOOKEAN:
                    elem = static_cast<int32_t>(tetV.size());
                    tetV.push_back(make_int4(sl[0], sl[1], sl[2], sl[3]));
                    tetP.insert(tetP.end(), p, p + 12);
                    break;
                default:  // a further energy
                    extResidualBytes += 16.0 + 8.0 * np;
                    elem = static_cast<int32_t>(extVh.size());
                    extVh.push_back(make_int4(sl[0], ss > 1 ? sl[1] : -1, ss > 2 ? sl[2] : -1, ss > 3 ? sl[3] : -1));
                    extPh.insert(extPh.end(), p, p + np);
                    extPh.insert(extPh.end(), kExtParams - np, 0.0);
                }
                for (int k = 0; k < ss; ++k) {
                    if (isFurther(cls))
                        extIncList.push_back({sl[k], elem, (en.type << 2) | k});
                    else
                        inc[cls].push_back({sl[k], elem, k});
                    const int32_t col = c.localColors[base + st[k]];
                    colorRecordBytes[col] += (isFurther(cls) ? furtherRecordBytes(xclsOf(cls, k)) : recordBytes(cls)) + 4.0;
                    colorIncidences[col] += 1;
                }
            }
        }
    }
    auto checkSize = [](size_t n) {
        if (n >= (size_t(1) << 31) - 1) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "too many incidences");
    };
    // the solve's fixed-width rows and the residual's element-id rows, kept on the host for the neighbourhood layout
    ResidualRowsH resRows;
    std::vector<float4> sprRowsH, triMH;
    std::vector<int2> triOH;
    std::vector<int4> hinVH[2], tetOH;
    std::vector<float2> hinPH[2];
    // springs
    if (!inc[SPG_VBD_SPRING].empty()) {
        checkSize(inc[SPG_VBD_SPRING].size());
        std::vector<int32_t> ptr;
        const auto order = sortBySlot(inc[SPG_VBD_SPRING], c.nv, ptr);
        std::vector<int32_t> el(order.size());
        for (size_t k = 0; k < order.size(); ++k) el[k] = inc[SPG_VBD_SPRING][order[k]].elem;
        c.sprW = rowWidth(ptr);
        sprRowsH = toRows(ptr, c.sprW, make_float4(intBitsAsFloat(-1), 0.f, 0.f, 0.f), [&](int32_t k) {
            const auto &in = inc[SPG_VBD_SPRING][order[k]];
            const int2 v = sprV[in.elem];
            const int32_t other = in.local == 0 ? v.y : v.x;
            return make_float4(intBitsAsFloat(other), float(sprP[in.elem].x), float(sprP[in.elem].y), 0.f);
        });
        c.spr.upload(sprRowsH, c.stream);
        c.sprPtr.upload(ptr, c.stream);
        c.sprElem.upload(el, c.stream);
        resRows.spr = toRows(ptr, c.sprW, -1, [&](int32_t k) { return el[k]; });
        c.sprRowR.upload(resRows.spr, c.stream);
        c.sprV.upload(sprV, c.stream);
        c.sprP.upload(sprP, c.stream);
    }
    // anchors
    if (!inc[SPG_VBD_SPRING_ANCHOR].empty()) {
        checkSize(inc[SPG_VBD_SPRING_ANCHOR].size());
        std::vector<int32_t> ptr;
        const auto order = sortBySlot(inc[SPG_VBD_SPRING_ANCHOR], c.nv, ptr);
        std::vector<float4> rec(order.size());
        std::vector<int32_t> el(order.size());
        for (size_t k = 0; k < order.size(); ++k) {
            const auto &in = inc[SPG_VBD_SPRING_ANCHOR][order[k]];
            const double4 p = ancP[in.elem];
            rec[k] = make_float4(float(p.x), float(p.y), float(p.z), float(p.w));
            el[k] = in.elem;
        }
        c.ancPtr.upload(ptr, c.stream);
        c.anc.upload(rec, c.stream);
        c.ancElem.upload(el, c.stream);
        c.ancP.upload(ancP, c.stream);
    }
    // StVK membrane: rows of Dinv are the shape vectors g_1, g_2; g_0 = -(g_1 + g_2)
    if (!inc[SPG_VBD_MEMBRANE_STVK].empty()) {
        checkSize(inc[SPG_VBD_MEMBRANE_STVK].size());
        std::vector<int32_t> ptr;
        const auto order = sortBySlot(inc[SPG_VBD_MEMBRANE_STVK], c.nv, ptr);
        std::vector<int2> ro(order.size());
        std::vector<float4> rg(order.size()), rk(order.size());
        std::vector<int32_t> el(order.size());
        for (size_t k = 0; k < order.size(); ++k) {
            const auto &in = inc[SPG_VBD_MEMBRANE_STVK][order[k]];
            const int4 v = triV[in.elem];
            const double *p = &triP[8 * static_cast<size_t>(in.elem)];
            const double g[3][2] = {{-(p[0] + p[2]), -(p[1] + p[3])}, {p[0], p[1]}, {p[2], p[3]}};
            const int32_t vv[3] = {v.x, v.y, v.z};
            const int ia = in.local == 0 ? 1 : 0, ib = in.local == 2 ? 1 : 2;
            ro[k] = make_int2(vv[ia], vv[ib]);
            rg[k] = make_float4(float(g[ia][0]), float(g[ia][1]), float(g[ib][0]), float(g[ib][1]));
            rk[k] = make_float4(float(p[4]), float(p[5]), float(p[6]), 0.f);
            el[k] = in.elem;
        }
        c.triW = rowWidth(ptr);
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        triOH = toRows(ptr, c.triW, make_int2(-1, -1), [&](int32_t k) { return ro[k]; });
        triMH = toRows(ptr, c.triW, z4, [&](int32_t k) { return rk[k]; });
        c.triO.upload(triOH, c.stream);
        c.triG.upload(toRows(ptr, c.triW, z4, [&](int32_t k) { return rg[k]; }), c.stream);
        c.triM.upload(triMH, c.stream);
        c.triPtr.upload(ptr, c.stream);
        c.triElem.upload(el, c.stream);
        resRows.tri = toRows(ptr, c.triW, -1, [&](int32_t k) { return el[k]; });
        c.triRowR.upload(resRows.tri, c.stream);
        c.triV.upload(triV, c.stream);
        c.triP.upload(triP, c.stream);
    }
    // discrete bending: two incidence classes by the owner's slot (hinge edge: 0, 1; wing: 2, 3), so
    // that the lanes of a wavefront run one closed form at a time. Records keep the canonical slots;
    // the solve finds its own position in the hinge by comparing slot ids.
    if (!inc[SPG_VBD_DISCRETE_BENDING].empty()) {
        checkSize(inc[SPG_VBD_DISCRETE_BENDING].size());
        // a hinge with a repeated vertex cannot tell its slot from the vertex id
        for (const auto &v : hinV)
            if (v.x == v.y || v.x == v.z || v.x == v.w || v.y == v.z || v.y == v.w || v.z == v.w)
                throw VbdError(SPG_VBD_ERR_INVALID, "bending stencil with a repeated vertex");
        for (int cls = 0; cls < 2; ++cls) {
            std::vector<Incidence> sub;
            for (const auto &in : inc[SPG_VBD_DISCRETE_BENDING])
                if ((in.local >= 2) == (cls == 1)) sub.push_back(in);
            std::vector<int32_t> ptr;
            const auto order = sortBySlot(sub, c.nv, ptr);
            std::vector<int4> rv(order.size());
            std::vector<float2> rp(order.size());
            std::vector<int32_t> el(order.size());
            for (size_t k = 0; k < order.size(); ++k) {
                const auto &in = sub[order[k]];
                rv[k] = hinV[in.elem];
                rp[k] = make_float2(float(hinP[in.elem].x), float(hinP[in.elem].y));
                el[k] = in.elem;
            }
            const int32_t w = rowWidth(ptr);
            (cls == 0 ? c.hinEW : c.hinWW) = w;
            (cls == 0 ? c.hinEPtr : c.hinWPtr).upload(ptr, c.stream);
            hinVH[cls] = toRows(ptr, w, make_int4(-1, -1, -1, -1), [&](int32_t k) { return rv[k]; });
            hinPH[cls] = toRows(ptr, w, make_float2(0.f, 0.f), [&](int32_t k) { return rp[k]; });
            (cls == 0 ? c.hinEV : c.hinWV).upload(hinVH[cls], c.stream);
            (cls == 0 ? c.hinEP : c.hinWP).upload(hinPH[cls], c.stream);
            (cls == 0 ? c.hinEElem : c.hinWElem).upload(el, c.stream);
            resRows.hin[cls] = toRows(ptr, w, -1, [&](int32_t k) { return el[k]; });
            (cls == 0 ? c.hinERowR : c.hinWRowR).upload(resRows.hin[cls], c.stream);
        }
        c.hinVe.upload(hinV, c.stream);
        c.hinPe.upload(hinP, c.stream);
        c.kc.alloc(hinV.size());
    }
    // Stable Neo-Hookean: rows of Dinv are g_1..g_3; g_0 = -(g_1 + g_2 + g_3)
    if (!inc[SPG_VBD_STABLE_NEO_HOOKEAN].empty()) {
        checkSize(inc[SPG_VBD_STABLE_NEO_HOOKEAN].size());
        std::vector<int32_t> ptr;
        const auto order = sortBySlot(inc[SPG_VBD_STABLE_NEO_HOOKEAN], c.nv, ptr);
        std::vector<int4> ro(order.size());
        std::vector<float4> ga(order.size()), gb(order.size()), gc(order.size());
        std::vector<int32_t> el(order.size());
        for (size_t k = 0; k < order.size(); ++k) {
            const auto &in = inc[SPG_VBD_STABLE_NEO_HOOKEAN][order[k]];
            const int4 v = tetV[in.elem];
            const double *p = &tetP[12 * static_cast<size_t>(in.elem)];
            double g[4][3];
            for (int q = 0; q < 3; ++q) {
                g[1][q] = p[q];
                g[2][q] = p[3 + q];
                g[3][q] = p[6 + q];
                g[0][q] = -(p[q] + p[3 + q] + p[6 + q]);
            }
            const int32_t vv[4] = {v.x, v.y, v.z, v.w};
            int oth[3], m = 0;
            for (int q = 0; q < 4; ++q)
                if (q != in.local) oth[m++] = q;
            ro[k] = make_int4(vv[oth[0]], vv[oth[1]], vv[oth[2]], 0);
            auto G = [&](int q, double w) { return make_float4(float(g[q][0]), float(g[q][1]), float(g[q][2]), float(w)); };
            ga[k] = G(oth[0], p[9]);
            gb[k] = G(oth[1], p[10]);
            gc[k] = G(oth[2], p[11]);
            el[k] = in.elem;
        }
        c.tetW = rowWidth(ptr);
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        c.tetPtr.upload(ptr, c.stream);
        tetOH = toRows(ptr, c.tetW, make_int4(-1, -1, -1, -1), [&](int32_t k) { return ro[k]; });
        c.tetO.upload(tetOH, c.stream);
        c.tetGa.upload(toRows(ptr, c.tetW, z4, [&](int32_t k) { return ga[k]; }), c.stream);
        c.tetGb.upload(toRows(ptr, c.tetW, z4, [&](int32_t k) { return gb[k]; }), c.stream);
        c.tetGc.upload(toRows(ptr, c.tetW, z4, [&](int32_t k) { return gc[k]; }), c.stream);
        c.tetElem.upload(el, c.stream);
        resRows.tet = toRows(ptr, c.tetW, -1, [&](int32_t k) { return el[k]; });
        c.tetRowR.upload(resRows.tet, c.stream);
        c.tetV.upload(tetV, c.stream);
        c.tetP.upload(tetP, c.stream);
    }

    // further energies: the per-slot incidence offsets, the element tables and the per-class rows of the solve and
    // the residual
    if (!extIncList.empty()) {
        checkSize(extIncList.size());
        if (extVh.size() >= (size_t(1) << 31) - 1) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "too many elements");
        std::vector<int32_t> ptr;
        (void)sortBySlot(extIncList, c.nv, ptr);
        buildFurtherRecords(c, extIncList, extVh, extPh);
        c.extPtr.upload(ptr, c.stream);
        c.extV.upload(extVh, c.stream);
        c.extP.upload(extPh, c.stream);
    }

    if (extIncList.empty())
        buildNeighbourhoodLayout(c, sprRowsH, triOH, triMH, hinVH, hinPH, tetOH, resRows, triV, hinV, tetV, sprV);

    // compulsory bytes of one FP64 residual: per particle its position, x~ and FP64 mass read and its squared
    // residual written and read back by the reduction (56 B); every element's table entry read once (spring 24,
    // anchor 36, triangle 80, hinge 32 plus its k (theta - theta0) written and read, 48, tetrahedron 112, further
    // element 16 + 8 per parameter); 4 B per incidence for the element reference
    {
        double b = 56.0 * double(c.nv) + 24.0 * double(sprV.size()) + 36.0 * double(ancP.size()) +
                   80.0 * double(triV.size()) + 48.0 * double(hinV.size()) + 112.0 * double(tetV.size());
        int64_t incidences = 0;
        for (int32_t k = 0; k < c.maxColors; ++k) incidences += colorIncidences[k];
        b += 4.0 * double(incidences);
        c.residualBytes = b + extResidualBytes;
    }
    // compulsory bytes per colour sweep: 12 B per vertex of every object that has this colour (all
    // positions read once), 32 B per updated vertex (x write, x_tilde read, mass, row offset), and
    // per incidence a 4 B CSR entry plus the element record.
    c.colorActiveV.assign(c.maxColors, 0);
    c.colorBytes.assign(c.maxColors, 0.0);
    for (int32_t o = 0; o < nObjects; ++o)
        for (int32_t k = 0; k < c.numColors[o]; ++k) c.colorActiveV[k] += objs[o].n_particles;
    for (int32_t k = 0; k < c.maxColors; ++k)
        c.colorBytes[k] = 12.0 * double(c.colorActiveV[k]) + 32.0 * double(c.colorBegin[k + 1] - c.colorBegin[k]) +
                          colorRecordBytes[k];
    hipCheck(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
    buildFeatureTables(c, objs, nObjects);
    c.ensureCheckPipeline();
}

// Static inputs of the learned warm start: one-ring neighbours (vertices sharing a triangle,
// tetrahedron or spring) with rest lengths from x0 (x when x0 is NULL), per-vertex log10 stiffness of
// the incident membrane / solid / spring elements and of the incident hinges (recovered from the
// prepared parameters), the anchored flag, and the slot -> object map of the warm-start gate.
void buildFeatureTables(spg_vbd_ctx &c, const spg_vbd_object *objs, int32_t nObjects)
{
    std::vector<int64_t> deg(c.nv + 1, 0);
    std::vector<double> matY(c.nv, 0.0), matB(c.nv, 0.0);
    std::vector<int32_t> nY(c.nv, 0), nB(c.nv, 0), anchored(c.nv, 0);
    std::vector<std::pair<int32_t, int32_t>> pairs;
    auto rest = [&](const spg_vbd_object &ob) { return ob.x0 ? ob.x0 : ob.x; };
    for (int32_t o = 0; o < nObjects; ++o) {
        const auto &ob = objs[o];
        const int64_t base = c.objOffset[o];
        const double *x0 = rest(ob);
        for (int32_t e = 0; e < ob.n_energies; ++e) {
            const auto &en = ob.energies[e];
            const int ss = stencilSizeOf(en.type), np = paramsOf(en.type);
            for (int64_t i = 0; i < en.n_stencils; ++i) {
                const int32_t *st = en.stencils + i * ss;
                const double *p = en.params + i * np;
                double logY = 0.0;
                bool hasY = false, edges = false;
                switch (en.type) {
                case SPG_VBD_SPRING:
                    logY = std::log10(std::max(p[1], 1e-30));
                    hasY = edges = true;
                    break;
                case SPG_VBD_MEMBRANE_STVK: {
                    // K = A Y / (1 - nu^2) [[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]], A = 1 / (2 |det Dm^-1|)
                    const double area = 0.5 / std::max(std::fabs(p[0] * p[3] - p[1] * p[2]), 1e-300);
                    const double nu = p[4] != 0.0 ? p[5] / p[4] : 0.0;
                    logY = std::log10(std::max(p[4] * (1.0 - nu * nu) / area, 1e-30));
                    hasY = edges = true;
                    break;
                }
                case SPG_VBD_STABLE_NEO_HOOKEAN: {
                    // lambda' = V lambda, mu' = V mu with lambda = mu / (1 - 2 nu), Y = 2 mu (1 + nu)
                    const double d = p[0] * (p[4] * p[8] - p[5] * p[7]) - p[1] * (p[3] * p[8] - p[5] * p[6]) +
                                     p[2] * (p[3] * p[7] - p[4] * p[6]);
                    const double vol = 1.0 / std::max(6.0 * std::fabs(d), 1e-300);
                    const double lam = p[9] / vol, mu = p[10] / vol;
                    const double nu = lam != 0.0 ? 0.5 * (1.0 - mu / lam) : 0.0;
                    logY = std::log10(std::max(2.0 * mu * (1.0 + nu), 1e-30));
                    hasY = edges = true;
                    break;
                }
                case SPG_VBD_DISCRETE_BENDING: {
                    // k_eff = k 12 |e|^2 / (|e x e3| + |e x e4|) at the rest shape
                    double a[3], b[3], d[3];
                    for (int q = 0; q < 3; ++q) {
                        a[q] = x0[3 * st[1] + q] - x0[3 * st[0] + q];
                        b[q] = x0[3 * st[2] + q] - x0[3 * st[0] + q];
                        d[q] = x0[3 * st[3] + q] - x0[3 * st[0] + q];
                    }
                    auto crossNorm = [](const double *u, const double *v) {
                        const double cx = u[1] * v[2] - u[2] * v[1], cy = u[2] * v[0] - u[0] * v[2],
                                     cz = u[0] * v[1] - u[1] * v[0];
                        return std::sqrt(cx * cx + cy * cy + cz * cz);
                    };
                    const double e2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
                    const double k = e2 > 0 ? p[1] * (crossNorm(a, b) + crossNorm(a, d)) / (12.0 * e2) : 0.0;
                    for (int q = 0; q < 4; ++q) {
                        matB[base + st[q]] += std::log10(std::max(k, 1e-30));
                        ++nB[base + st[q]];
                    }
                    break;
                }
                case SPG_VBD_SPRING_ANCHOR:
                    anchored[base + st[0]] = 1;
                    break;
                }
                if (hasY)
                    for (int q = 0; q < ss; ++q) {
                        matY[base + st[q]] += logY;
                        ++nY[base + st[q]];
                    }
                if (edges)
                    for (int q = 0; q < ss; ++q)
                        for (int r = 0; r < ss; ++r)
                            if (q != r && st[q] != st[r])
                                pairs.push_back({static_cast<int32_t>(base + st[q]), static_cast<int32_t>(base + st[r])});
            }
        }
    }
    // neighbour rows in gid order, sorted and unique, then renumbered to slots
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    for (const auto &pr : pairs) ++deg[c.slotOfGid[pr.first] + 1];
    for (int64_t s = 0; s < c.nv; ++s) deg[s + 1] += deg[s];
    if (deg[c.nv] >= (int64_t(1) << 31) - 1) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "too many neighbour pairs");
    std::vector<int32_t> ptr(c.nv + 1), idx(deg[c.nv]);
    std::vector<float> restLen(deg[c.nv]);
    for (int64_t s = 0; s <= c.nv; ++s) ptr[s] = static_cast<int32_t>(deg[s]);
    std::vector<int64_t> fill(deg.begin(), deg.end() - 1);
    std::vector<int32_t> objOfGid(c.nv);
    for (int32_t o = 0; o < nObjects; ++o)
        for (int64_t g = c.objOffset[o]; g < c.objOffset[o + 1]; ++g) objOfGid[g] = o;
    for (const auto &pr : pairs) {
        const int32_t o = objOfGid[pr.first];
        const double *x0 = rest(objs[o]);
        const int64_t ia = pr.first - c.objOffset[o], ib = pr.second - c.objOffset[o];
        double d2 = 0.0;
        for (int q = 0; q < 3; ++q) {
            const double d = x0[3 * ib + q] - x0[3 * ia + q];
            d2 += d * d;
        }
        const int64_t k = fill[c.slotOfGid[pr.first]]++;
        idx[k] = c.slotOfGid[pr.second];
        restLen[k] = static_cast<float>(std::max(std::sqrt(d2), 1e-12));
    }
    std::vector<float4> st(c.nv);
    std::vector<int32_t> objOfSlot(c.nv);
    for (int64_t g = 0; g < c.nv; ++g) {
        const int32_t s = c.slotOfGid[g];
        st[s] = make_float4(nY[g] ? float(matY[g] / nY[g]) : 0.f, nB[g] ? float(matB[g] / nB[g]) : 0.f,
                            float(anchored[g]), 0.f);
        objOfSlot[s] = objOfGid[g];
    }
    c.nbPtr.upload(ptr, c.stream);
    c.nbIdx.upload(idx, c.stream);
    c.nbRest.upload(restLen, c.stream);
    c.stat.upload(st, c.stream);
    c.objOfSlot.upload(objOfSlot, c.stream);
    c.reject.alloc(nObjects);
    c.corr.upload(std::vector<float4>(c.nv, make_float4(0.f, 0.f, 0.f, 0.f)), c.stream);
    hipCheck(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
}

void checkObject(const spg_vbd_ctx *ctx, int32_t object)
{
    if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
    if (object < 0 || object >= ctx->nObjects) throw VbdError(SPG_VBD_ERR_INVALID, "object index out of range");
}

// Writes object `object`'s rows of dst from x (zeros when x is NULL).
void uploadPositions(spg_vbd_ctx &c, int32_t object, const double *x, float4 *dst)
{
    const int64_t n = c.objOffset[object + 1] - c.objOffset[object];
    for (int64_t i = 0; x && i < n; ++i)
        if (!finite3(x + 3 * i)) throw VbdError(SPG_VBD_ERR_NONFINITE, "non-finite positions");
    // read-modify-write of the object's slots
    std::vector<float4> h(c.nv);
    hipCheck(hipMemcpyAsync(h.data(), dst, sizeof(float4) * c.nv, hipMemcpyDeviceToHost, c.stream), "download");
    hipCheck(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
    for (int64_t i = 0; i < n; ++i) {
        const int32_t s = c.slotOfGid[c.objOffset[object] + i];
        h[s] = x ? make_float4(float(x[3 * i]), float(x[3 * i + 1]), float(x[3 * i + 2]), 0.f)
                 : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    hipCheck(hipMemcpyAsync(dst, h.data(), sizeof(float4) * c.nv, hipMemcpyHostToDevice, c.stream), "upload");
    hipCheck(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
}

}  // namespace

extern "C" {

int spg_vbd_create(spg_vbd_ctx **out, const spg_vbd_object *objects, int32_t n_objects, int32_t device)
{
    if (!out) return SPG_VBD_ERR_INVALID;
    *out = nullptr;
    auto c = std::make_unique<spg_vbd_ctx>();
    const int rc = guarded(nullptr, [&] {
        int ndev = 0;
        hipCheck(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        if (device < 0 || device >= ndev) throw VbdError(SPG_VBD_ERR_INVALID, "invalid HIP device");
        c->device = device;
        hipCheck(hipSetDevice(device), "hipSetDevice");
        c->scratchDev = scratchOnDevice(device);
        hipCheck(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        buildContext(*c, objects, n_objects);
    });
    if (rc == SPG_VBD_OK) *out = c.release();
    return rc;
}

void spg_vbd_destroy(spg_vbd_ctx *ctx)
{
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        delete ctx;
    }
}

const char *spg_vbd_last_error(const spg_vbd_ctx *ctx) { return ctx ? ctx->err.c_str() : g_lastError.c_str(); }

int spg_vbd_num_objects(const spg_vbd_ctx *ctx, int32_t *n)
{
    if (!ctx || !n) return SPG_VBD_ERR_INVALID;
    *n = ctx->nObjects;
    return SPG_VBD_OK;
}

int spg_vbd_num_particles(const spg_vbd_ctx *ctx, int32_t object, int64_t *n)
{
    return guarded(const_cast<spg_vbd_ctx *>(ctx), [&] {
        checkObject(ctx, object);
        *n = ctx->objOffset[object + 1] - ctx->objOffset[object];
    });
}

int spg_vbd_num_colors(const spg_vbd_ctx *ctx, int32_t object, int32_t *n)
{
    return guarded(const_cast<spg_vbd_ctx *>(ctx), [&] {
        checkObject(ctx, object);
        *n = ctx->numColors[object];
    });
}

int spg_vbd_get_colors(const spg_vbd_ctx *ctx, int32_t object, int32_t *colors)
{
    return guarded(const_cast<spg_vbd_ctx *>(ctx), [&] {
        checkObject(ctx, object);
        std::copy(ctx->localColors.begin() + ctx->objOffset[object], ctx->localColors.begin() + ctx->objOffset[object + 1],
                  colors);
    });
}

int spg_vbd_set_state(spg_vbd_ctx *ctx, int32_t object, const double *x, const double *v)
{
    return guarded(ctx, [&] {
        checkObject(ctx, object);
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        const int64_t n = ctx->objOffset[object + 1] - ctx->objOffset[object];
        for (int64_t i = 0; i < n; ++i)
            if ((x && !finite3(x + 3 * i)) || (v && !finite3(v + 3 * i)))
                throw VbdError(SPG_VBD_ERR_NONFINITE, "non-finite state");
        if (x) uploadPositions(*ctx, object, x, ctx->pos.p);
        if (v) uploadPositions(*ctx, object, v, ctx->vel.p);
        if (x || v) uploadPositions(*ctx, object, nullptr, ctx->corr.p);  // no substep history
        if (x) ctx->positionsChanged();
    });
}

int spg_vbd_get_state(spg_vbd_ctx *ctx, int32_t object, double *x, double *v)
{
    return guarded(ctx, [&] {
        checkObject(ctx, object);
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        std::vector<float4> hp(ctx->nv), hv(ctx->nv);
        hipCheck(hipMemcpyAsync(hp.data(), ctx->pos.p, sizeof(float4) * ctx->nv, hipMemcpyDeviceToHost, ctx->stream),
                 "download");
        hipCheck(hipMemcpyAsync(hv.data(), ctx->vel.p, sizeof(float4) * ctx->nv, hipMemcpyDeviceToHost, ctx->stream),
                 "download");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        const int64_t n = ctx->objOffset[object + 1] - ctx->objOffset[object];
        for (int64_t i = 0; i < n; ++i) {
            const int32_t s = ctx->slotOfGid[ctx->objOffset[object] + i];
            if (x) {
                x[3 * i] = hp[s].x;
                x[3 * i + 1] = hp[s].y;
                x[3 * i + 2] = hp[s].z;
            }
            if (v) {
                v[3 * i] = hv[s].x;
                v[3 * i + 1] = hv[s].y;
                v[3 * i + 2] = hv[s].z;
            }
        }
    });
}

int spg_vbd_step(spg_vbd_ctx *ctx, double dt, int32_t nsubsteps, int32_t iterations, const double gravity[3])
{
    return guarded(ctx, [&] {
        if (!ctx || !gravity) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->step(dt, nsubsteps, iterations, gravity);
    });
}

int spg_vbd_begin_substep(spg_vbd_ctx *ctx, double h, const double gravity[3])
{
    return guarded(ctx, [&] {
        if (!ctx || !gravity) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        if (ctx->inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "substep already open");
        if (!(h > 0)) throw VbdError(SPG_VBD_ERR_INVALID, "invalid h");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->orderAfterChecks();
        ctx->beginSubstep(h, gravity);
    });
}

int spg_vbd_set_guess(spg_vbd_ctx *ctx, int32_t object, const double *x)
{
    return guarded(ctx, [&] {
        checkObject(ctx, object);
        if (!ctx->inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "set_guess needs an open substep");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        uploadPositions(*ctx, object, x, ctx->pos.p);
        ctx->positionsChanged();
    });
}

int spg_vbd_iterate(spg_vbd_ctx *ctx, int32_t iterations)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        if (!ctx->inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "iterate needs an open substep");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->sweepColors(iterations, ctx->timing);
    });
}

int spg_vbd_end_substep(spg_vbd_ctx *ctx)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        if (!ctx->inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "no open substep");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->endSubstep();
    });
}

int spg_vbd_save_state(spg_vbd_ctx *ctx)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        if (ctx->inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "save_state inside an open substep");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        if (!ctx->savedPos.p) {
            ctx->savedPos.alloc(ctx->nv);
            ctx->savedVel.alloc(ctx->nv);
            ctx->savedCorr.alloc(ctx->nv);
        }
        hipCheck(launchCopy(ctx->pos.p, ctx->savedPos.p, ctx->nv, ctx->stream), "save");
        hipCheck(launchCopy(ctx->vel.p, ctx->savedVel.p, ctx->nv, ctx->stream), "save");
        hipCheck(launchCopy(ctx->corr.p, ctx->savedCorr.p, ctx->nv, ctx->stream), "save");
    });
}

int spg_vbd_restore_state(spg_vbd_ctx *ctx)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        if (ctx->inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "restore_state inside an open substep");
        if (!ctx->savedPos.p) throw VbdError(SPG_VBD_ERR_STATE, "no saved state");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        hipCheck(launchCopy(ctx->savedPos.p, ctx->pos.p, ctx->nv, ctx->stream), "restore");
        hipCheck(launchCopy(ctx->savedVel.p, ctx->vel.p, ctx->nv, ctx->stream), "restore");
        hipCheck(launchCopy(ctx->savedCorr.p, ctx->corr.p, ctx->nv, ctx->stream), "restore");
    });
}

int spg_vbd_residual(spg_vbd_ctx *ctx, double *norms)
{
    return guarded(ctx, [&] {
        if (!ctx || !norms) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->residual(false, norms);
    });
}

int spg_vbd_residual_noalloc(spg_vbd_ctx *ctx, double *norms)
{
    return guarded(ctx, [&] {
        if (!ctx || !norms) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->residualNoAlloc(norms);
    });
}

int spg_vbd_residual_at_inertial(spg_vbd_ctx *ctx, double *norms)
{
    return guarded(ctx, [&] {
        if (!ctx || !norms) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->residual(true, norms);
    });
}

int spg_vbd_iterate_to_tolerance(spg_vbd_ctx *ctx, double rel_tol, int32_t max_iterations, int32_t check_every,
                                 int32_t *iterations, double *rho)
{
    return guarded(ctx, [&] {
        if (!ctx || !iterations || !rho) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        if (!ctx->inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "iterate_to_tolerance needs an open substep");
        if (max_iterations < 0 || check_every < 1) throw VbdError(SPG_VBD_ERR_INVALID, "invalid iteration counts");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->iterateToTolerance(rel_tol, max_iterations, check_every, iterations, rho);
    });
}

int spg_vbd_features(spg_vbd_ctx *ctx, int32_t object, float *features)
{
    return guarded(ctx, [&] {
        if (!ctx || !features) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        if (object < -1 || object >= ctx->nObjects) throw VbdError(SPG_VBD_ERR_INVALID, "object index out of range");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->computeFeatures();
        std::vector<float> all(static_cast<size_t>(ctx->nv) * kMlpFeatures);
        hipCheck(hipMemcpyAsync(all.data(), ctx->feat.p, sizeof(float) * all.size(), hipMemcpyDeviceToHost, ctx->stream),
                 "download");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        const int64_t g0 = object < 0 ? 0 : ctx->objOffset[object];
        const int64_t g1 = object < 0 ? ctx->nv : ctx->objOffset[object + 1];
        for (int64_t gid = g0; gid < g1; ++gid)
            std::memcpy(features + (gid - g0) * kMlpFeatures, all.data() + int64_t(ctx->slotOfGid[gid]) * kMlpFeatures,
                        sizeof(float) * kMlpFeatures);
    });
}

int spg_vbd_warm_start(spg_vbd_ctx *ctx, spg_mlp *net, int32_t smoothing, int32_t gate, int32_t *accepted, double *rho)
{
    return guarded(ctx, [&] {
        if (!ctx || !net) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        if (gate < 0 || gate > 2) throw VbdError(SPG_VBD_ERR_INVALID, "gate must be 0, 1 or 2");
        ctx->warmStart(net, smoothing, gate, accepted, rho);
    });
}

int spg_vbd_history_start(spg_vbd_ctx *ctx, int32_t gate, int32_t *accepted, double *rho)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        if (gate < 0 || gate > 2) throw VbdError(SPG_VBD_ERR_INVALID, "gate must be 0, 1 or 2");
        ctx->historyStart(gate, accepted, rho);
    });
}

int spg_vbd_start_rho(const spg_vbd_ctx *ctx, double *rho)
{
    return guarded(const_cast<spg_vbd_ctx *>(ctx), [&] {
        if (!ctx || !rho) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        for (int32_t o = 0; o < ctx->nObjects; ++o)
            rho[o] = o < static_cast<int32_t>(ctx->startRho.size()) ? ctx->startRho[o]
                                                                    : std::numeric_limits<double>::quiet_NaN();
    });
}

int spg_vbd_get_launch(const spg_vbd_ctx *ctx, spg_vbd_launch *cfg)
{
    if (!ctx || !cfg) return SPG_VBD_ERR_INVALID;
    *cfg = ctx->launch;
    return SPG_VBD_OK;
}

int spg_vbd_set_launch(spg_vbd_ctx *ctx, const spg_vbd_launch *cfg)
{
    return guarded(ctx, [&] {
        if (!ctx || !cfg) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        const int g = cfg->lanes_per_vertex;
        if (g != 1 && g != 2 && g != 3 && g != 4 && g != 6 && g != 8 && g != 16)
            throw VbdError(SPG_VBD_ERR_INVALID, "lanes_per_vertex must be 1, 2, 3, 4, 6, 8 or 16");
        if (cfg->block_size != 64 && cfg->block_size != 128 && cfg->block_size != 256)
            throw VbdError(SPG_VBD_ERR_INVALID, "block_size must be 64, 128 or 256");
        if (cfg->lds_staging < 0 || cfg->lds_staging > 2)
            throw VbdError(SPG_VBD_ERR_INVALID, "lds_staging must be 0, 1 or 2");
        if (cfg->lds_staging == 2 && ctx->nbW == 0)
            throw VbdError(SPG_VBD_ERR_UNSUPPORTED,
                           "neighbourhood staging needs at most 15 distinct neighbours per particle and no further "
                           "energies");
        if (solveStagedLdsBytes(ctx->solveArgs(0, 1.f), g, cfg->block_size, cfg->lds_staging) > kSolveMaxLds)
            throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "the staged data of one workgroup exceed 64 KB of LDS");
        ctx->launch = *cfg;
        ctx->dropGraph();
    });
}

int spg_vbd_autotune(spg_vbd_ctx *ctx, int32_t trials, spg_vbd_launch *best)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        if (ctx->inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "autotune inside an open substep");
        if (trials < 1) trials = 1;
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        // save state, tune inside a scratch substep, restore
        DevBuf<float4> spos, svel;
        spos.alloc(ctx->nv);
        svel.alloc(ctx->nv);
        hipCheck(hipMemcpyAsync(spos.p, ctx->pos.p, sizeof(float4) * ctx->nv, hipMemcpyDeviceToDevice, ctx->stream), "save");
        hipCheck(hipMemcpyAsync(svel.p, ctx->vel.p, sizeof(float4) * ctx->nv, hipMemcpyDeviceToDevice, ctx->stream), "save");
        const spg_vbd_launch keep = ctx->launch;
        const double g[3] = {0.0, -9.81, 0.0};
        ctx->orderAfterChecks();
        ctx->beginSubstep(1.0 / 60.0, g);
        hipEvent_t e0, e1;
        hipCheck(hipEventCreate(&e0), "hipEventCreate");
        hipCheck(hipEventCreate(&e1), "hipEventCreate");
        float bestMs = 1e30f;
        spg_vbd_launch bestCfg = keep;
        const int lanes[] = {1, 2, 3, 4, 6, 8, 16};
        const int blocks[] = {64, 128, 256};
        for (int staged = 0; staged < 3; ++staged)
        for (int l : lanes)
            for (int b : blocks) {
                if (staged == 2 && ctx->nbW == 0) continue;
                if (solveStagedLdsBytes(ctx->solveArgs(0, 1.f), l, b, staged) > kSolveMaxLds) continue;
                ctx->launch = keep;
                ctx->launch.lanes_per_vertex = l;
                ctx->launch.block_size = b;
                ctx->launch.lds_staging = staged;
                ctx->sweepColors(1, false);  // warm
                float ms = 1e30f;
                for (int rep = 0; rep < 2; ++rep) {  // best of two timed runs of `trials` sweeps
                    hipCheck(hipEventRecord(e0, ctx->stream), "hipEventRecord");
                    ctx->sweepColors(trials, false);
                    hipCheck(hipEventRecord(e1, ctx->stream), "hipEventRecord");
                    hipCheck(hipEventSynchronize(e1), "hipEventSynchronize");
                    float t = 0.f;
                    hipCheck(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
                    ms = std::min(ms, t);
                }
                if (ms < bestMs) {
                    bestMs = ms;
                    bestCfg = ctx->launch;
                }
            }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ctx->inSubstep = false;
        hipCheck(hipMemcpyAsync(ctx->pos.p, spos.p, sizeof(float4) * ctx->nv, hipMemcpyDeviceToDevice, ctx->stream), "restore");
        hipCheck(hipMemcpyAsync(ctx->vel.p, svel.p, sizeof(float4) * ctx->nv, hipMemcpyDeviceToDevice, ctx->stream), "restore");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        ctx->launch = bestCfg;
        ctx->dropGraph();
        if (best) *best = bestCfg;
    });
}

int spg_vbd_get_stats(const spg_vbd_ctx *ctx, spg_vbd_stats *stats)
{
    auto *c = const_cast<spg_vbd_ctx *>(ctx);
    return guarded(c, [&] {
        if (!ctx || !stats) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        hipCheck(hipSetDevice(c->device), "hipSetDevice");
        c->drainEvents();
        unsigned long long cnt[2] = {0, 0};
        hipCheck(hipMemcpyAsync(cnt, c->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, c->stream), "download");
        hipCheck(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        *stats = c->stats;
        stats->fp64_fallbacks = static_cast<int64_t>(cnt[0]);
        stats->skipped_blocks = static_cast<int64_t>(cnt[1]);
        stats->prediction_fallbacks = c->predictionFallbacks;
    });
}

int spg_vbd_reset_stats(spg_vbd_ctx *ctx)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->drainEvents();
        ctx->stats = spg_vbd_stats{};
        ctx->predictionFallbacks = 0;
        const unsigned long long zero[2] = {0, 0};
        hipCheck(hipMemcpyAsync(ctx->counters.p, zero, sizeof(zero), hipMemcpyHostToDevice, ctx->stream), "upload");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int spg_vbd_set_timing(spg_vbd_ctx *ctx, int32_t enabled)
{
    if (!ctx) return SPG_VBD_ERR_INVALID;
    ctx->timing = enabled != 0;
    return SPG_VBD_OK;
}

int spg_vbd_synchronize(spg_vbd_ctx *ctx)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int spg_vbd_color_vertices(int64_t n_vertices, int32_t n_sets, const int32_t *const *flat_stencils,
                           const int64_t *lengths, const int32_t *stencil_sizes, int32_t *colors)
{
    return guarded(nullptr, [&] {
        if (n_sets < 0 || (n_sets > 0 && (!flat_stencils || !lengths || !stencil_sizes)) || !colors)
            throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        std::vector<FlatStencilSet> sets;
        for (int32_t i = 0; i < n_sets; ++i) sets.push_back({flat_stencils[i], lengths[i], stencil_sizes[i]});
        const auto col = colorVertices(n_vertices, sets);
        std::copy(col.begin(), col.end(), colors);
    });
}

int spg_vbd_bench_residual(spg_vbd_ctx *ctx, int32_t repeats, double *us, double *bytes)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        if (repeats < 1 || !us) throw VbdError(SPG_VBD_ERR_INVALID, "invalid arguments");
        if (!ctx->inSubstep) throw VbdError(SPG_VBD_ERR_STATE, "residual needs an open substep (begin_substep)");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->enqueueResidual(ctx->posIsInertial, ctx->norms.p);  // warm
        hipEvent_t e0, e1;
        hipCheck(hipEventCreate(&e0), "hipEventCreate");
        hipCheck(hipEventCreate(&e1), "hipEventCreate");
        hipCheck(hipEventRecord(e0, ctx->stream), "hipEventRecord");
        for (int32_t r = 0; r < repeats; ++r) ctx->enqueueResidual(ctx->posIsInertial, ctx->norms.p);
        hipCheck(hipEventRecord(e1, ctx->stream), "hipEventRecord");
        hipCheck(hipEventSynchronize(e1), "hipEventSynchronize");
        float t = 0.f;
        hipCheck(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *us = 1e3 * t / repeats;
        if (bytes) *bytes = ctx->residualBytes;
    });
}

int spg_vbd_bench_copy(int64_t bytes, int32_t repeats, int32_t device, double *ms)
{
    return guarded(nullptr, [&] {
        if (bytes <= 0 || bytes % 16 || repeats < 1 || !ms) throw VbdError(SPG_VBD_ERR_INVALID, "invalid arguments");
        hipCheck(hipSetDevice(device), "hipSetDevice");
        hipStream_t s;
        hipCheck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        const int64_t n = bytes / 16;
        DevBuf<float4> a, b;
        a.alloc(n);
        b.alloc(n);
        hipCheck(launchCopy(b.p, a.p, n, s), "copy launch");  // touch both buffers once
        hipEvent_t e0, e1;
        hipCheck(hipEventCreate(&e0), "hipEventCreate");
        hipCheck(hipEventCreate(&e1), "hipEventCreate");
        hipCheck(hipEventRecord(e0, s), "hipEventRecord");
        for (int32_t r = 0; r < repeats; ++r) hipCheck(launchCopy(a.p, b.p, n, s), "copy launch");
        hipCheck(hipEventRecord(e1, s), "hipEventRecord");
        hipCheck(hipEventSynchronize(e1), "hipEventSynchronize");
        float t = 0.f;
        hipCheck(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
        *ms = t / repeats;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
        (void)hipStreamDestroy(s);
    });
}

int spg_vbd_element_derivatives(int32_t type, int64_t n, const double *xs, const double *params, int32_t precision,
                                int32_t form, int32_t device, double *grad, double *hess_blocks)
{
    return guarded(nullptr, [&] {
        const int ss = stencilSizeOf(type), np = paramsOf(type);
        if (!ss) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "unsupported energy type");
        if (n < 0 || (n > 0 && (!xs || !params || !grad || !hess_blocks)))
            throw VbdError(SPG_VBD_ERR_INVALID, "invalid arguments");
        if (form < 0 || form > 2) throw VbdError(SPG_VBD_ERR_INVALID, "form must be 0, 1 or 2");
        if (n == 0) return;
        if (isFurther(type)) {
            // the further energies' slot blocks exist in the FP32 solve only (device code)
            if (device < 0 || precision != 0 || form != 0)
                throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "further energies: FP32 solve forms on a GPU only");
            hipCheck(hipSetDevice(device), "hipSetDevice");
            hipStream_t s;
            hipCheck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
            std::vector<float4> hp(n * ss);
            for (int64_t k = 0; k < n * ss; ++k)
                hp[k] = make_float4(float(xs[3 * k]), float(xs[3 * k + 1]), float(xs[3 * k + 2]), 0.f);
            DevBuf<float4> dx;
            DevBuf<double> dp, dg, dh;
            dx.upload(hp, s);
            dp.upload(std::vector<double>(params, params + n * np), s);
            dg.alloc(n * ss * 3);
            dh.alloc(n * ss * 9);
            hipCheck(launchExtElementDerivatives(type, n, dx.p, dp.p, np, dg.p, dh.p, s), "element launch");
            hipCheck(hipMemcpyAsync(grad, dg.p, sizeof(double) * n * ss * 3, hipMemcpyDeviceToHost, s), "download");
            hipCheck(hipMemcpyAsync(hess_blocks, dh.p, sizeof(double) * n * ss * 9, hipMemcpyDeviceToHost, s),
                     "download");
            hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
            (void)hipStreamDestroy(s);
            return;
        }
        if (device == -1) {
            // host instantiation of the same closed forms (derivative checks only)
            for (int64_t i = 0; i < n; ++i) {
                if (precision == 0)
                    elementAll<float>(type, xs + i * ss * 3, params + i * np, form, grad + i * ss * 3,
                                      hess_blocks + i * ss * 9);
                else
                    elementAll<double>(type, xs + i * ss * 3, params + i * np, form, grad + i * ss * 3,
                                       hess_blocks + i * ss * 9);
            }
            return;
        }
        hipCheck(hipSetDevice(device), "hipSetDevice");
        hipStream_t s;
        hipCheck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        DevBuf<double> dx, dp, dg, dh;
        dx.upload(std::vector<double>(xs, xs + n * ss * 3), s);
        dp.upload(std::vector<double>(params, params + n * np), s);
        dg.alloc(n * ss * 3);
        dh.alloc(n * ss * 9);
        hipCheck(launchElementDerivatives(type, n, dx.p, dp.p, precision, form, dg.p, dh.p, s), "element launch");
        hipCheck(hipMemcpyAsync(grad, dg.p, sizeof(double) * n * ss * 3, hipMemcpyDeviceToHost, s), "download");
        hipCheck(hipMemcpyAsync(hess_blocks, dh.p, sizeof(double) * n * ss * 9, hipMemcpyDeviceToHost, s), "download");
        hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
        (void)hipStreamDestroy(s);
    });
}

}  // extern "C"
