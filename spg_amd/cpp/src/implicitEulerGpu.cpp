// spg::solver::ImplicitEulerGpu: BaseSolver drop-in for spg's implicit Euler family
// (implicitEulerBaraffWitkin.cpp, implicitEulerNewtonDv.cpp, implicitEulerNewtonDx.cpp,
// implicitEulerNewtonRobust.cpp, bdf2.cpp) through the C-ABI of spg_implicit.h.
#include <stdexcept>
#include <string>
#include <vector>

#include <spg/sim/energy/energies.h>
#include <spg/sim/solver/baseSolver.h>

#include "flatten.h"
#include "spg_implicit.h"

namespace spg::solver {

namespace {

void check(int rc, spg_implicit_ctx *ctx)
{
    if (rc != SPG_VBD_OK) throw std::runtime_error(spg_implicit_last_error(ctx));
}

std::vector<double> packVec3(const std::vector<Vector3> &v)
{
    std::vector<double> out(3 * v.size());
    for (size_t i = 0; i < v.size(); ++i)
        for (int d = 0; d < 3; ++d) out[3 * i + d] = v[i][d];
    return out;
}

}  // namespace

ImplicitEulerGpu::ImplicitEulerGpu(Kind kind, int device) : m_kind(kind), m_device(device) {}

ImplicitEulerGpu::~ImplicitEulerGpu()
{
    if (m_ctx) spg_implicit_destroy(m_ctx);
}

void ImplicitEulerGpu::buildContext()
{
    if (m_ctx) {
        spg_implicit_destroy(m_ctx);
        m_ctx = nullptr;
    }
    const size_t n = m_objects.size();
    std::vector<std::vector<FlatEnergy>> flat(n);
    std::vector<std::vector<spg_vbd_energy>> descs(n);
    std::vector<std::vector<double>> xs(n), vs(n);
    std::vector<spg_vbd_object> objs(n);
    for (size_t o = 0; o < n; ++o) {
        const auto &pg = m_objects[o];
        for (const auto &e : pg.energies()) flat[o].push_back(flattenVbd(e.get()));
        for (const auto &f : flat[o])
            descs[o].push_back({f.type, static_cast<int64_t>(f.stencils.size()) / flatStencilSize(f.type),
                                f.stencils.data(), f.params.data()});
        xs[o] = packVec3(pg.positions());
        vs[o] = packVec3(pg.velocities());
        objs[o] = {pg.size(), xs[o].data(), vs[o].data(), pg.masses().data(),
                   static_cast<int32_t>(descs[o].size()), descs[o].data(), nullptr};
    }
    spg_implicit_ctx *ctx = nullptr;
    check(spg_implicit_create(&ctx, objs.data(), static_cast<int32_t>(n), static_cast<int32_t>(m_kind), m_device),
          nullptr);
    m_ctx = ctx;
    m_hostDirty = false;
}

void ImplicitEulerGpu::uploadState()
{
    for (size_t o = 0; o < m_objects.size(); ++o) {
        const auto x = packVec3(m_objects[o].positions());
        const auto v = packVec3(m_objects[o].velocities());
        check(spg_implicit_set_state(m_ctx, static_cast<int32_t>(o), x.data(), v.data()), m_ctx);
    }
    m_hostDirty = false;
}

void ImplicitEulerGpu::syncToHost()
{
    if (!m_ctx) return;
    for (size_t o = 0; o < m_objects.size(); ++o) {
        auto &pg = m_objects[o];
        std::vector<double> x(3 * pg.size()), v(3 * pg.size());
        check(spg_implicit_get_state(m_ctx, static_cast<int32_t>(o), x.data(), v.data()), m_ctx);
        for (int i = 0; i < pg.size(); ++i)
            for (int d = 0; d < 3; ++d) {
                pg.positions()[i][d] = x[3 * i + d];
                pg.velocities()[i][d] = v[3 * i + d];
            }
    }
}

void ImplicitEulerGpu::step()
{
    requireNoRigidBodies("ImplicitEulerGpu");
    if (m_objects.empty()) return;
    if (m_rebuild || !m_ctx) {
        // resident state: the device holds the newest positions unless the host was edited since
        if (m_ctx && m_resident && !m_hostDirty) syncToHost();
        buildContext();
        m_rebuild = false;
    } else if (!m_resident || m_hostDirty) {
        uploadState();
    }
    const double g[3] = {m_gravity[0], m_gravity[1], m_gravity[2]};
    check(spg_implicit_step(m_ctx, m_dtStep, m_nsubsteps, g), m_ctx);
    if (!m_resident) syncToHost();
}

void ImplicitEulerGpu::reset()
{
    // BaseSolver::reset, plus the solver's own state (implicitEulerNewtonRobust.cpp:129-134, bdf2.cpp:78-83),
    // which a rebuilt context starts without
    BaseSolver::reset();
    m_rebuild = true;
}

}  // namespace spg::solver
