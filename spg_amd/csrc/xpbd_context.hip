// Host side of the XPBD solver in libspg_vbd.so: builds the device layout of an XPBD context from the
// spg particle groups and energies, orchestrates the step on one HIP stream, and implements the C-ABI
// of include/spg_xpbd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "coloring.h"
#include "host_util.h"
#include "spg_xpbd.h"
#include "vbd_device.h"
#include "xpbd_device.h"
#include "xpbd_element.h"

using namespace spgvbd;

namespace {

thread_local std::string g_lastError;

int stencilSizeOf(int type)
{
    switch (type) {
    case SPG_VBD_SPRING: return 2;
    case SPG_VBD_MEMBRANE_STVK: return 3;
    case SPG_VBD_DISCRETE_BENDING: return 4;
    case SPG_VBD_STABLE_NEO_HOOKEAN: return 4;
    case SPG_VBD_SPRING_ANCHOR: return 1;
    }
    return type >= SPG_VBD_SPRING_CONTINUUM && type < kXpbdNumTypes ? xpbdFurtherStencil(type) : 0;
}
int paramsOf(int type)
{
    switch (type) {
    case SPG_VBD_SPRING: return 2;
    case SPG_VBD_MEMBRANE_STVK: return 13;
    case SPG_VBD_DISCRETE_BENDING: return 2;
    case SPG_VBD_STABLE_NEO_HOOKEAN: return 14;
    case SPG_VBD_SPRING_ANCHOR: return 4;
    case SPG_VBD_SPRING_CONTINUUM:
    case SPG_VBD_SPRING_SQUARED_CONTINUUM:
    case SPG_VBD_BARAFF_WITKIN_BENDING: return 2;
    case SPG_VBD_MEMBRANE_BARAFF_WITKIN: return 7;
    case SPG_VBD_MEMBRANE_CHOI: return 8;
    case SPG_VBD_QUADRATIC_BENDING: return 5;
    case SPG_VBD_STVK:
    case SPG_VBD_STABLE_SQUARED_NEO_HOOKEAN: return 11;
    }
    return 0;
}

bool finite3(const double *p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

// compliances may be +inf (zero stiffness); everything else must be finite
bool complianceOk(double c) { return !std::isnan(c) && c >= 0; }

// symmetric within rounding: |c_ij - c_ji| <= 1e-12 max |c|
bool symmetric(const double *c, int n)
{
    double m = 0;
    for (int k = 0; k < n * n; ++k) m = std::max(m, std::fabs(c[k]));
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (std::fabs(c[i * n + j] - c[j * n + i]) > 1e-12 * m) return false;
    return true;
}

}  // namespace

struct spg_xpbd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int32_t nObjects = 0;
    bool parallel = true;
    std::vector<int64_t> objOffset;
    int64_t nv = 0;
    std::vector<float> invMass;  // per particle, FP32 copy of the w of pos
    std::vector<std::vector<std::vector<int32_t>>> groups;  // [object][energy] group of each stencil
    std::vector<std::vector<int32_t>> numGroups;

    DevBuf<float4> pos, vel, xold, snapPos, snapVel;
    bool haveSnapshot = false;
    DevBuf<int2> sprI;
    DevBuf<float2> sprP;
    DevBuf<int32_t> ancI;
    DevBuf<float4> ancP;
    DevBuf<int4> triI;
    DevBuf<float4> triD, triC0;
    DevBuf<float2> triC1;
    DevBuf<int4> hinI;
    DevBuf<float2> hinP;
    DevBuf<int4> tetI;
    DevBuf<float4> tetD0, tetD1, tetD2;
    DevBuf<float> tetH;
    DevBuf<int4> extI;    // further energies, records of types 6 ... 13 one after the other
    DevBuf<float4> extP;  // 3 per record
    DevBuf<unsigned long long> counters;

    struct Launch {
        int type;
        int32_t begin, end;
        double bytes;  // compulsory bytes of the launch
    };
    std::vector<Launch> launches;  // per substep, in order
    // the launches as enqueued: a group alone, or a run of 2 ... kXpbdRunMax consecutive small groups in one
    // workgroup (the tail colours of the greedy colouring have a few hundred stencils each)
    struct Item {
        int32_t first, count;  // launches[first .. first + count - 1]
    };
    std::vector<Item> plan, single;  // with group runs / one launch per group

    spg_xpbd_launch launch{256, 1, 1};
    bool timing = false;
    spg_xpbd_stats stats{};
    std::vector<std::pair<hipEvent_t, hipEvent_t>> eventPool, events;
    std::vector<int32_t> eventLaunch;

    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    std::array<double, 6> graphKey{};

    ~spg_xpbd_ctx()
    {
        dropGraph();
        for (auto &e : events) eventPool.push_back(e);
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

    XpbdArgs args(float invdt2) const
    {
        XpbdArgs a{};
        a.pos = pos.p;
        a.sprI = sprI.p;
        a.sprP = sprP.p;
        a.ancI = ancI.p;
        a.ancP = ancP.p;
        a.triI = triI.p;
        a.triD = triD.p;
        a.triC0 = triC0.p;
        a.triC1 = triC1.p;
        a.hinI = hinI.p;
        a.hinP = hinP.p;
        a.tetI = tetI.p;
        a.tetD0 = tetD0.p;
        a.tetD1 = tetD1.p;
        a.tetD2 = tetD2.p;
        a.tetH = tetH.p;
        a.extI = extI.p;
        a.extP = extP.p;
        a.counters = counters.p;
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
            const Item &it = (launch.group_runs ? plan : single)[eventLaunch[i]];
            int64_t stencils = 0;
            double bytes = 0.0;
            for (int32_t k = it.first; k < it.first + it.count; ++k) {
                stencils += launches[k].end - launches[k].begin;
                bytes += launches[k].bytes;
            }
            const int t = it.count == 1 ? launches[it.first].type : 0;  // 0: a run of small groups
            stats.projection_ms += ms;
            stats.projection_launches += 1;
            stats.stencil_projections += stencils;
            stats.algorithmic_bytes += bytes;
            stats.type_launches[t] += 1;
            stats.type_ms[t] += ms;
            stats.type_bytes[t] += bytes;
            eventPool.push_back(events[i]);
        }
        events.clear();
        eventLaunch.clear();
    }

    void enqueueStep(double dt, int32_t nsub, const double *g, bool timed)
    {
        const double h = dt / nsub;  // xpbd.cpp:22
        const float fh = static_cast<float>(h), invh = static_cast<float>(1.0 / h);
        const float hg[3] = {static_cast<float>(h * g[0]), static_cast<float>(h * g[1]), static_cast<float>(h * g[2])};
        const XpbdArgs a = args(static_cast<float>(1.0 / (h * h)));
        const int32_t n = static_cast<int32_t>(nv);
        for (int32_t s = 0; s < nsub; ++s) {
            if (s == 0)
                hipCheck(launchXpbdBegin(pos.p, vel.p, xold.p, n, fh, hg[0], hg[1], hg[2], stream), "begin launch");
            else
                hipCheck(launchXpbdBoundary(pos.p, xold.p, n, invh, fh, hg[0], hg[1], hg[2], stream),
                         "boundary launch");
            const std::vector<Item> &items = launch.group_runs ? plan : single;
            for (size_t k = 0; k < items.size(); ++k) {
                const Item &it = items[k];
                std::pair<hipEvent_t, hipEvent_t> ev{};
                if (timed) ev = takeEvents();
                if (it.count == 1) {
                    const Launch &L = launches[it.first];
                    hipCheck(launchXpbdProject(L.type, a, L.begin, L.end, launch.block_size, stream, ev.first, ev.second),
                             "projection launch");
                } else {
                    XpbdRun run{};
                    run.count = it.count;
                    for (int32_t r = 0; r < it.count; ++r) {
                        run.type[r] = launches[it.first + r].type;
                        run.begin[r] = launches[it.first + r].begin;
                        run.end[r] = launches[it.first + r].end;
                    }
                    hipCheck(launchXpbdProjectRun(a, run, stream, ev.first, ev.second), "projection run launch");
                }
                if (timed) {
                    events.push_back(ev);
                    eventLaunch.push_back(static_cast<int32_t>(k));
                    if (events.size() >= 8192) drainEvents();
                }
            }
        }
        hipCheck(launchXpbdEnd(pos.p, xold.p, vel.p, n, invh, stream), "end launch");
    }

    void step(double dt, int32_t nsub, const double *g)
    {
        if (!(dt > 0) || nsub < 1) throw VbdError(SPG_VBD_ERR_INVALID, "invalid dt/substeps");
        if (!std::isfinite(g[0]) || !std::isfinite(g[1]) || !std::isfinite(g[2]))
            throw VbdError(SPG_VBD_ERR_INVALID, "non-finite gravity");
        if (nv == 0) return;
        if (launch.use_graph && !timing) {
            const std::array<double, 6> key{dt / nsub, g[0], g[1], g[2], double(nsub), double(launch.block_size)};
            if (!graphExec || key != graphKey) {
                dropGraph();
                hipCheck(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
                try {
                    enqueueStep(dt, nsub, g, false);
                } catch (...) {
                    hipGraph_t tmp = nullptr;
                    (void)hipStreamEndCapture(stream, &tmp);
                    if (tmp) (void)hipGraphDestroy(tmp);
                    throw;
                }
                hipCheck(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
                hipCheck(hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
                graphKey = key;
            }
            hipCheck(hipGraphLaunch(graphExec, stream), "hipGraphLaunch");
            return;
        }
        enqueueStep(dt, nsub, g, timing);
    }
};

namespace {

template <class F>
int guarded(spg_xpbd_ctx *ctx, F &&f)
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

// One stencil of one energy, with the key that orders the projection records.
struct Entry {
    int32_t energy, group, object;
    int32_t stencil;
};

void validateEnergy(const spg_xpbd_energy &e, int64_t np)
{
    const int ss = stencilSizeOf(e.type), pp = paramsOf(e.type);
    if (!ss) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "unsupported energy type for the GPU XPBD solver");
    if (e.n_stencils < 0 || (e.n_stencils > 0 && (!e.stencils || !e.params)))
        throw VbdError(SPG_VBD_ERR_INVALID, "invalid energy arrays");
    if (e.n_stencils > std::numeric_limits<int32_t>::max())
        throw VbdError(SPG_VBD_ERR_INVALID, "too many stencils");
    for (int64_t i = 0; i < e.n_stencils; ++i) {
        const int32_t *s = e.stencils + i * ss;
        const double *p = e.params + i * pp;
        for (int k = 0; k < ss; ++k) {
            if (s[k] < 0 || s[k] >= np) throw VbdError(SPG_VBD_ERR_INVALID, "stencil index out of range");
            for (int j = 0; j < k; ++j)
                if (s[j] == s[k])
                    throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "a stencil that names one particle twice is not supported");
        }
        switch (e.type) {
        case SPG_VBD_SPRING:
            if (!std::isfinite(p[0]) || !complianceOk(p[1])) throw VbdError(SPG_VBD_ERR_INVALID, "invalid spring data");
            if (p[0] <= 0) throw VbdError(SPG_VBD_ERR_INVALID, "Zero/negative rest length invalid");
            break;
        case SPG_VBD_DISCRETE_BENDING:
            if (!std::isfinite(p[0]) || !complianceOk(p[1]))
                throw VbdError(SPG_VBD_ERR_INVALID, "invalid bending data");
            if (p[0] < -3.12 || p[0] > 3.12) throw VbdError(SPG_VBD_ERR_INVALID, "Invalid restTheta");
            break;
        case SPG_VBD_SPRING_ANCHOR:
            if (!finite3(p) || !complianceOk(p[3])) throw VbdError(SPG_VBD_ERR_INVALID, "invalid anchor data");
            break;
        case SPG_VBD_MEMBRANE_STVK:
            for (int k = 0; k < 13; ++k)
                if (!std::isfinite(p[k])) throw VbdError(SPG_VBD_ERR_INVALID, "invalid membrane data");
            if (!symmetric(p + 4, 3))
                throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "non-symmetric membrane compliance is not supported");
            break;
        case SPG_VBD_STABLE_NEO_HOOKEAN:
            for (int k = 0; k < 14; ++k)
                if (!std::isfinite(p[k])) throw VbdError(SPG_VBD_ERR_INVALID, "invalid SNH data");
            if (!symmetric(p + 9, 2))
                throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "non-symmetric SNH compliance is not supported");
            break;
        default: {  // further energies: finite geometry, compliance diagonal >= 0 (may be +inf)
            const int nc = e.type == SPG_VBD_QUADRATIC_BENDING ? 1 : xpbdFurtherConstraints(e.type);
            for (int k = 0; k < pp - nc; ++k)
                if (!std::isfinite(p[k])) throw VbdError(SPG_VBD_ERR_INVALID, "invalid energy data");
            for (int k = pp - nc; k < pp; ++k)
                if (!complianceOk(p[k])) throw VbdError(SPG_VBD_ERR_INVALID, "invalid compliance");
            if ((e.type == SPG_VBD_SPRING_CONTINUUM || e.type == SPG_VBD_SPRING_SQUARED_CONTINUUM) && p[0] <= 0)
                throw VbdError(SPG_VBD_ERR_INVALID, "Zero/negative rest length invalid");
            if (e.type == SPG_VBD_BARAFF_WITKIN_BENDING && (p[0] < -3.12 || p[0] > 3.12))
                throw VbdError(SPG_VBD_ERR_INVALID, "Invalid restTheta");
        }
        }
    }
}

void buildContext(spg_xpbd_ctx &c, const spg_xpbd_object *objs, int32_t nObjects)
{
    if (nObjects < 1 || !objs) throw VbdError(SPG_VBD_ERR_INVALID, "no objects");
    c.nObjects = nObjects;
    c.objOffset.assign(nObjects + 1, 0);
    for (int32_t o = 0; o < nObjects; ++o) {
        const auto &ob = objs[o];
        if (ob.n_particles < 0 || (ob.n_particles > 0 && (!ob.x || !ob.inv_mass)) || ob.n_energies < 0 ||
            (ob.n_energies > 0 && !ob.energies))
            throw VbdError(SPG_VBD_ERR_INVALID, "invalid object arrays");
        c.objOffset[o + 1] = c.objOffset[o] + ob.n_particles;
    }
    c.nv = c.objOffset[nObjects];
    if (c.nv > std::numeric_limits<int32_t>::max()) throw VbdError(SPG_VBD_ERR_INVALID, "too many particles");

    // particles: float4 (x, w = inverse mass), velocities
    std::vector<float4> hp(c.nv), hv(c.nv);
    c.invMass.assign(c.nv, 0.f);
    for (int32_t o = 0; o < nObjects; ++o) {
        const auto &ob = objs[o];
        for (int64_t i = 0; i < ob.n_particles; ++i) {
            const int64_t g = c.objOffset[o] + i;
            const double wi = ob.inv_mass[i];
            if (!std::isfinite(wi) || wi < 0) throw VbdError(SPG_VBD_ERR_INVALID, "invalid inverse mass");
            if (!finite3(ob.x + 3 * i) || (ob.v && !finite3(ob.v + 3 * i)))
                throw VbdError(SPG_VBD_ERR_NONFINITE, "non-finite state");
            const float w = static_cast<float>(wi);
            c.invMass[g] = w;
            hp[g] = make_float4(float(ob.x[3 * i]), float(ob.x[3 * i + 1]), float(ob.x[3 * i + 2]), w);
            hv[g] = ob.v ? make_float4(float(ob.v[3 * i]), float(ob.v[3 * i + 1]), float(ob.v[3 * i + 2]), 0.f)
                         : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }

    // groups per energy (computeParallelStencilGroups, or the serial levels), and the launch order
    std::vector<Entry> ent[kXpbdNumTypes];
    c.groups.assign(nObjects, {});
    c.numGroups.assign(nObjects, {});
    for (int32_t o = 0; o < nObjects; ++o) {
        const auto &ob = objs[o];
        for (int32_t e = 0; e < ob.n_energies; ++e) {
            const auto &en = ob.energies[e];
            validateEnergy(en, ob.n_particles);
            const int32_t ss = stencilSizeOf(en.type);
            auto grp = c.parallel ? colorStencils(en.stencils, en.n_stencils, ss)
                                  : serialLevels(en.stencils, en.n_stencils, ss);
            int32_t ng = 0;
            for (int64_t i = 0; i < en.n_stencils; ++i) {
                ng = std::max(ng, grp[i] + 1);
                ent[en.type].push_back({e, grp[i], o, static_cast<int32_t>(i)});
            }
            c.groups[o].push_back(std::move(grp));
            c.numGroups[o].push_back(ng);
        }
    }
    struct Range {
        int32_t energy, group, type, begin, end;
    };
    std::vector<Range> ranges;
    // the further energies' records follow each other in one table: type t starts at extBase[t]
    std::array<int32_t, kXpbdNumTypes + 1> extBase{};
    for (int t = SPG_VBD_SPRING_CONTINUUM; t < kXpbdNumTypes; ++t)
        extBase[t + 1] = extBase[t] + static_cast<int32_t>(ent[t].size());
    for (int t = 1; t < kXpbdNumTypes; ++t) {
        auto &v = ent[t];
        const int32_t off = t >= SPG_VBD_SPRING_CONTINUUM ? extBase[t] : 0;
        std::sort(v.begin(), v.end(), [](const Entry &a, const Entry &b) {
            return std::tie(a.energy, a.group, a.object, a.stencil) < std::tie(b.energy, b.group, b.object, b.stencil);
        });
        for (size_t k = 0; k < v.size();) {
            size_t j = k;
            while (j < v.size() && v[j].energy == v[k].energy && v[j].group == v[k].group) ++j;
            ranges.push_back({v[k].energy, v[k].group, t, off + static_cast<int32_t>(k), off + static_cast<int32_t>(j)});
            k = j;
        }
    }
    std::sort(ranges.begin(), ranges.end(), [](const Range &a, const Range &b) {
        return std::tie(a.energy, a.group, a.type) < std::tie(b.energy, b.group, b.type);
    });
    c.launches.clear();
    for (const auto &r : ranges) {
        const int ss = stencilSizeOf(r.type);
        const double bytes = (r.end - r.begin) * (kXpbdRecordBytes[r.type] + 32.0 * ss);
        c.launches.push_back({r.type, r.begin, r.end, bytes});
    }
    // consecutive small groups (whatever their energy) run in one workgroup, in the same order
    c.plan.clear();
    c.single.clear();
    for (int32_t k = 0; k < static_cast<int32_t>(c.launches.size()); ++k) c.single.push_back({k, 1});
    for (int32_t k = 0; k < static_cast<int32_t>(c.launches.size());) {
        int32_t n = 1;
        auto small = [&](int32_t q) { return c.launches[q].end - c.launches[q].begin <= kXpbdSmallGroup; };
        if (small(k))
            while (n < kXpbdRunMax && k + n < static_cast<int32_t>(c.launches.size()) && small(k + n)) ++n;
        c.plan.push_back({k, n});
        k += n;
    }

    // stencil records in launch order
    auto gidx = [&](const Entry &en, int k) {
        const auto &e = objs[en.object].energies[en.energy];
        return static_cast<int32_t>(c.objOffset[en.object] + e.stencils[int64_t(en.stencil) * stencilSizeOf(e.type) + k]);
    };
    auto prm = [&](const Entry &en) {
        const auto &e = objs[en.object].energies[en.energy];
        return e.params + int64_t(en.stencil) * paramsOf(e.type);
    };
    {
        const auto &v = ent[SPG_VBD_SPRING];
        std::vector<int2> I(v.size());
        std::vector<float2> P(v.size());
        for (size_t k = 0; k < v.size(); ++k) {
            const double *p = prm(v[k]);
            I[k] = make_int2(gidx(v[k], 0), gidx(v[k], 1));
            P[k] = make_float2(float(p[0]), float(p[1]));
        }
        c.sprI.upload(I, c.stream);
        c.sprP.upload(P, c.stream);
    }
    {
        const auto &v = ent[SPG_VBD_SPRING_ANCHOR];
        std::vector<int32_t> I(v.size());
        std::vector<float4> P(v.size());
        for (size_t k = 0; k < v.size(); ++k) {
            const double *p = prm(v[k]);
            I[k] = gidx(v[k], 0);
            P[k] = make_float4(float(p[0]), float(p[1]), float(p[2]), float(p[3]));
        }
        c.ancI.upload(I, c.stream);
        c.ancP.upload(P, c.stream);
    }
    {
        const auto &v = ent[SPG_VBD_MEMBRANE_STVK];
        std::vector<int4> I(v.size());
        std::vector<float4> D(v.size()), C0(v.size());
        std::vector<float2> C1(v.size());
        for (size_t k = 0; k < v.size(); ++k) {
            const double *p = prm(v[k]);
            const double *cm = p + 4;
            I[k] = make_int4(gidx(v[k], 0), gidx(v[k], 1), gidx(v[k], 2), 0);
            D[k] = make_float4(float(p[0]), float(p[1]), float(p[2]), float(p[3]));
            C0[k] = make_float4(float(cm[0]), float(cm[1]), float(cm[2]), float(cm[4]));
            C1[k] = make_float2(float(cm[5]), float(cm[8]));
        }
        c.triI.upload(I, c.stream);
        c.triD.upload(D, c.stream);
        c.triC0.upload(C0, c.stream);
        c.triC1.upload(C1, c.stream);
    }
    {
        const auto &v = ent[SPG_VBD_DISCRETE_BENDING];
        std::vector<int4> I(v.size());
        std::vector<float2> P(v.size());
        for (size_t k = 0; k < v.size(); ++k) {
            const double *p = prm(v[k]);
            I[k] = make_int4(gidx(v[k], 0), gidx(v[k], 1), gidx(v[k], 2), gidx(v[k], 3));
            P[k] = make_float2(float(p[0]), float(p[1]));
        }
        c.hinI.upload(I, c.stream);
        c.hinP.upload(P, c.stream);
    }
    {
        const auto &v = ent[SPG_VBD_STABLE_NEO_HOOKEAN];
        std::vector<int4> I(v.size());
        std::vector<float4> D0(v.size()), D1(v.size()), D2(v.size());
        std::vector<float> H(v.size());
        for (size_t k = 0; k < v.size(); ++k) {
            const double *p = prm(v[k]);
            I[k] = make_int4(gidx(v[k], 0), gidx(v[k], 1), gidx(v[k], 2), gidx(v[k], 3));
            D0[k] = make_float4(float(p[0]), float(p[1]), float(p[2]), float(p[3]));
            D1[k] = make_float4(float(p[4]), float(p[5]), float(p[6]), float(p[7]));
            D2[k] = make_float4(float(p[8]), float(p[9]), float(p[10]), float(p[12]));
            H[k] = float(p[13]);
        }
        c.tetI.upload(I, c.stream);
        c.tetD0.upload(D0, c.stream);
        c.tetD1.upload(D1, c.stream);
        c.tetD2.upload(D2, c.stream);
        c.tetH.upload(H, c.stream);
    }
    {
        std::vector<int4> I;
        std::vector<float4> P;
        for (int t = SPG_VBD_SPRING_CONTINUUM; t < kXpbdNumTypes; ++t) {
            const int ss = stencilSizeOf(t), pp = paramsOf(t);
            for (const auto &en : ent[t]) {
                const double *p = prm(en);
                I.push_back(make_int4(gidx(en, 0), ss > 1 ? gidx(en, 1) : -1, ss > 2 ? gidx(en, 2) : -1,
                                      ss > 3 ? gidx(en, 3) : -1));
                float q[12] = {};
                for (int k = 0; k < pp; ++k) q[k] = float(p[k]);
                for (int k = 0; k < 3; ++k) P.push_back(make_float4(q[4 * k], q[4 * k + 1], q[4 * k + 2], q[4 * k + 3]));
            }
        }
        c.extI.upload(I, c.stream);
        c.extP.upload(P, c.stream);
    }
    c.pos.upload(hp, c.stream);
    c.vel.upload(hv, c.stream);
    c.xold.alloc(c.nv);
    c.counters.alloc(1);
    hipCheck(hipMemsetAsync(c.counters.p, 0, sizeof(unsigned long long), c.stream), "memset");
    hipCheck(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
}

void checkObject(const spg_xpbd_ctx *ctx, int32_t object)
{
    if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
    if (object < 0 || object >= ctx->nObjects) throw VbdError(SPG_VBD_ERR_INVALID, "object index out of range");
}

void checkEnergy(const spg_xpbd_ctx *ctx, int32_t object, int32_t energy)
{
    checkObject(ctx, object);
    if (energy < 0 || energy >= static_cast<int32_t>(ctx->groups[object].size()))
        throw VbdError(SPG_VBD_ERR_INVALID, "energy index out of range");
}

}  // namespace

extern "C" {

int spg_xpbd_create(spg_xpbd_ctx **out, const spg_xpbd_object *objects, int32_t n_objects,
                    int32_t parallel_gauss_seidel, int32_t device)
{
    if (!out) return SPG_VBD_ERR_INVALID;
    *out = nullptr;
    auto c = std::make_unique<spg_xpbd_ctx>();
    const int rc = guarded(nullptr, [&] {
        int ndev = 0;
        hipCheck(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        if (device < 0 || device >= ndev) throw VbdError(SPG_VBD_ERR_INVALID, "invalid HIP device");
        c->device = device;
        c->parallel = parallel_gauss_seidel != 0;
        hipCheck(hipSetDevice(device), "hipSetDevice");
        hipCheck(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        buildContext(*c, objects, n_objects);
    });
    if (rc == SPG_VBD_OK) *out = c.release();
    return rc;
}

void spg_xpbd_destroy(spg_xpbd_ctx *ctx)
{
    if (ctx) {
        (void)hipSetDevice(ctx->device);
        delete ctx;
    }
}

const char *spg_xpbd_last_error(const spg_xpbd_ctx *ctx) { return ctx ? ctx->err.c_str() : g_lastError.c_str(); }

int spg_xpbd_num_objects(const spg_xpbd_ctx *ctx, int32_t *n)
{
    if (!ctx || !n) return SPG_VBD_ERR_INVALID;
    *n = ctx->nObjects;
    return SPG_VBD_OK;
}

int spg_xpbd_num_particles(const spg_xpbd_ctx *ctx, int32_t object, int64_t *n)
{
    return guarded(const_cast<spg_xpbd_ctx *>(ctx), [&] {
        checkObject(ctx, object);
        if (!n) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        *n = ctx->objOffset[object + 1] - ctx->objOffset[object];
    });
}

int spg_xpbd_num_groups(const spg_xpbd_ctx *ctx, int32_t object, int32_t energy, int32_t *n)
{
    return guarded(const_cast<spg_xpbd_ctx *>(ctx), [&] {
        checkEnergy(ctx, object, energy);
        if (!n) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        *n = ctx->numGroups[object][energy];
    });
}

int spg_xpbd_get_groups(const spg_xpbd_ctx *ctx, int32_t object, int32_t energy, int32_t *group_of_stencil)
{
    return guarded(const_cast<spg_xpbd_ctx *>(ctx), [&] {
        checkEnergy(ctx, object, energy);
        const auto &g = ctx->groups[object][energy];
        if (!group_of_stencil && !g.empty()) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        std::copy(g.begin(), g.end(), group_of_stencil);
    });
}

int spg_xpbd_num_launches(const spg_xpbd_ctx *ctx, int32_t *n)
{
    if (!ctx || !n) return SPG_VBD_ERR_INVALID;
    *n = static_cast<int32_t>(ctx->launch.group_runs ? ctx->plan.size() : ctx->single.size());
    return SPG_VBD_OK;
}

int spg_xpbd_set_state(spg_xpbd_ctx *ctx, int32_t object, const double *x, const double *v)
{
    return guarded(ctx, [&] {
        checkObject(ctx, object);
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        const int64_t b = ctx->objOffset[object], n = ctx->objOffset[object + 1] - b;
        for (int64_t i = 0; i < n; ++i)
            if ((x && !finite3(x + 3 * i)) || (v && !finite3(v + 3 * i)))
                throw VbdError(SPG_VBD_ERR_NONFINITE, "non-finite state");
        if (n == 0) return;
        std::vector<float4> h(n);
        if (x) {
            for (int64_t i = 0; i < n; ++i)
                h[i] = make_float4(float(x[3 * i]), float(x[3 * i + 1]), float(x[3 * i + 2]), ctx->invMass[b + i]);
            hipCheck(hipMemcpyAsync(ctx->pos.p + b, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice, ctx->stream),
                     "upload");
            hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        }
        if (v) {
            for (int64_t i = 0; i < n; ++i)
                h[i] = make_float4(float(v[3 * i]), float(v[3 * i + 1]), float(v[3 * i + 2]), 0.f);
            hipCheck(hipMemcpyAsync(ctx->vel.p + b, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice, ctx->stream),
                     "upload");
            hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        }
    });
}

int spg_xpbd_get_state(spg_xpbd_ctx *ctx, int32_t object, double *x, double *v)
{
    return guarded(ctx, [&] {
        checkObject(ctx, object);
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        const int64_t b = ctx->objOffset[object], n = ctx->objOffset[object + 1] - b;
        if (n == 0) return;
        std::vector<float4> hp(n), hv(n);
        hipCheck(hipMemcpyAsync(hp.data(), ctx->pos.p + b, sizeof(float4) * n, hipMemcpyDeviceToHost, ctx->stream),
                 "download");
        hipCheck(hipMemcpyAsync(hv.data(), ctx->vel.p + b, sizeof(float4) * n, hipMemcpyDeviceToHost, ctx->stream),
                 "download");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        for (int64_t i = 0; i < n; ++i) {
            if (x) {
                x[3 * i] = hp[i].x;
                x[3 * i + 1] = hp[i].y;
                x[3 * i + 2] = hp[i].z;
            }
            if (v) {
                v[3 * i] = hv[i].x;
                v[3 * i + 1] = hv[i].y;
                v[3 * i + 2] = hv[i].z;
            }
        }
    });
}

int spg_xpbd_step(spg_xpbd_ctx *ctx, double dt, int32_t nsubsteps, const double gravity[3])
{
    return guarded(ctx, [&] {
        if (!ctx || !gravity) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->step(dt, nsubsteps, gravity);
    });
}

int spg_xpbd_save_state(spg_xpbd_ctx *ctx)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        if (!ctx->haveSnapshot) {
            ctx->snapPos.alloc(ctx->nv);
            ctx->snapVel.alloc(ctx->nv);
            ctx->haveSnapshot = true;
        }
        if (ctx->nv == 0) return;
        hipCheck(hipMemcpyAsync(ctx->snapPos.p, ctx->pos.p, sizeof(float4) * ctx->nv, hipMemcpyDeviceToDevice,
                                ctx->stream),
                 "snapshot");
        hipCheck(hipMemcpyAsync(ctx->snapVel.p, ctx->vel.p, sizeof(float4) * ctx->nv, hipMemcpyDeviceToDevice,
                                ctx->stream),
                 "snapshot");
    });
}

int spg_xpbd_restore_state(spg_xpbd_ctx *ctx)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        if (!ctx->haveSnapshot) throw VbdError(SPG_VBD_ERR_STATE, "no saved state");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        if (ctx->nv == 0) return;
        hipCheck(hipMemcpyAsync(ctx->pos.p, ctx->snapPos.p, sizeof(float4) * ctx->nv, hipMemcpyDeviceToDevice,
                                ctx->stream),
                 "restore");
        hipCheck(hipMemcpyAsync(ctx->vel.p, ctx->snapVel.p, sizeof(float4) * ctx->nv, hipMemcpyDeviceToDevice,
                                ctx->stream),
                 "restore");
    });
}

int spg_xpbd_get_launch(const spg_xpbd_ctx *ctx, spg_xpbd_launch *cfg)
{
    if (!ctx || !cfg) return SPG_VBD_ERR_INVALID;
    *cfg = ctx->launch;
    return SPG_VBD_OK;
}

int spg_xpbd_set_launch(spg_xpbd_ctx *ctx, const spg_xpbd_launch *cfg)
{
    return guarded(ctx, [&] {
        if (!ctx || !cfg) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        if (cfg->block_size != 64 && cfg->block_size != 128 && cfg->block_size != 256)
            throw VbdError(SPG_VBD_ERR_INVALID, "block_size must be 64, 128 or 256");
        if (cfg->group_runs != 0 && cfg->group_runs != 1) throw VbdError(SPG_VBD_ERR_INVALID, "group_runs must be 0 or 1");
        ctx->drainEvents();  // timed launches of the previous shape are accounted with its plan
        ctx->launch = *cfg;
        ctx->dropGraph();
    });
}

int spg_xpbd_get_stats(const spg_xpbd_ctx *ctx, spg_xpbd_stats *stats)
{
    auto *c = const_cast<spg_xpbd_ctx *>(ctx);
    return guarded(c, [&] {
        if (!ctx || !stats) throw VbdError(SPG_VBD_ERR_INVALID, "null argument");
        hipCheck(hipSetDevice(c->device), "hipSetDevice");
        c->drainEvents();
        unsigned long long cnt = 0;
        hipCheck(hipMemcpyAsync(&cnt, c->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, c->stream), "download");
        hipCheck(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        *stats = c->stats;
        stats->fp64_solves = static_cast<int64_t>(cnt);
    });
}

int spg_xpbd_reset_stats(spg_xpbd_ctx *ctx)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        ctx->drainEvents();
        ctx->stats = spg_xpbd_stats{};
        const unsigned long long zero = 0;
        hipCheck(hipMemcpyAsync(ctx->counters.p, &zero, sizeof(zero), hipMemcpyHostToDevice, ctx->stream), "upload");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int spg_xpbd_set_timing(spg_xpbd_ctx *ctx, int32_t enabled)
{
    if (!ctx) return SPG_VBD_ERR_INVALID;
    ctx->timing = enabled != 0;
    return SPG_VBD_OK;
}

int spg_xpbd_synchronize(spg_xpbd_ctx *ctx)
{
    return guarded(ctx, [&] {
        if (!ctx) throw VbdError(SPG_VBD_ERR_INVALID, "null context");
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int spg_xpbd_color_stencils(int64_t n_stencils, int32_t stencil_size, const int32_t *flat_stencils, int32_t *colors)
{
    return guarded(nullptr, [&] {
        if (n_stencils < 0 || stencil_size <= 0 || (n_stencils > 0 && (!flat_stencils || !colors)))
            throw VbdError(SPG_VBD_ERR_INVALID, "invalid arguments");
        const auto col = colorStencils(flat_stencils, n_stencils, stencil_size);
        std::copy(col.begin(), col.end(), colors);
    });
}

int spg_xpbd_project_stencils(int32_t type, int64_t n, const double *xs, const double *inv_masses,
                              const double *params, double dt, int32_t precision, int32_t device, double *dx)
{
    return guarded(nullptr, [&] {
        const int ss = stencilSizeOf(type), np = paramsOf(type);
        if (!ss) throw VbdError(SPG_VBD_ERR_UNSUPPORTED, "unsupported energy type");
        if (n < 0 || (n > 0 && (!xs || !inv_masses || !params || !dx)) || !(dt > 0))
            throw VbdError(SPG_VBD_ERR_INVALID, "invalid arguments");
        if (precision != 0 && precision != 1) throw VbdError(SPG_VBD_ERR_INVALID, "precision must be 0 or 1");
        if (n == 0) return;
        if (device == -1) {
            // host instantiation of the same device functions (projection checks only)
            for (int64_t i = 0; i < n; ++i) {
                if (precision == 0)
                    xpbdProjectRecord<true>(type, xs + i * ss * 3, inv_masses + i * ss, params + i * np,
                                            1.0 / (dt * dt), dx + i * ss * 3);
                else
                    xpbdProjectRecord<false>(type, xs + i * ss * 3, inv_masses + i * ss, params + i * np,
                                             1.0 / (dt * dt), dx + i * ss * 3);
            }
            return;
        }
        hipCheck(hipSetDevice(device), "hipSetDevice");
        hipStream_t s;
        hipCheck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        DevBuf<double> dxs, dw, dp, dout;
        dxs.upload(std::vector<double>(xs, xs + n * ss * 3), s);
        dw.upload(std::vector<double>(inv_masses, inv_masses + n * ss), s);
        dp.upload(std::vector<double>(params, params + n * np), s);
        dout.alloc(n * ss * 3);
        hipCheck(launchXpbdProjectCheck(type, n, dxs.p, dw.p, dp.p, 1.0 / (dt * dt), precision, dout.p, s),
                 "projection check launch");
        hipCheck(hipMemcpyAsync(dx, dout.p, sizeof(double) * n * ss * 3, hipMemcpyDeviceToHost, s), "download");
        hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
        (void)hipStreamDestroy(s);
    });
}

}  // extern "C"
