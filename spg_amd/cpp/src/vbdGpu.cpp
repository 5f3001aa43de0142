// spg::solver::VBDGpu: BaseSolver drop-in for spg::solver::VBD (vbd.h:14-41, vbd.cpp:12-201) that
// hands the particle groups and energies to the GPU solver through the C-ABI of spg_vbd.h.
#include <stdexcept>
#include <string>

#include <spg/sim/energy/energies.h>
#include <spg/sim/solver/baseSolver.h>

#include "flatten.h"
#include "spg_vbd.h"

namespace spg::solver {

void BaseSolver::reset()
{
    // baseSolver.cpp:27-35
    for (auto &obj : m_objects) obj.reset();
}

VBDGpu::VBDGpu(int device) : m_device(device) {}

VBDGpu::~VBDGpu()
{
    if (m_ctx) spg_vbd_destroy(m_ctx);
}

namespace {

void check(int rc, spg_vbd_ctx *ctx)
{
    if (rc != SPG_VBD_OK) throw std::runtime_error(spg_vbd_last_error(ctx));
}

std::vector<double> packVec3(const std::vector<Vector3> &v)
{
    std::vector<double> out(3 * v.size());
    for (size_t i = 0; i < v.size(); ++i)
        for (int d = 0; d < 3; ++d) out[3 * i + d] = v[i][d];
    return out;
}

}  // namespace

void VBDGpu::buildContext()
{
    if (m_ctx) {
        spg_vbd_destroy(m_ctx);
        m_ctx = nullptr;
    }
    std::vector<std::vector<FlatEnergy>> flat(m_objects.size());
    std::vector<std::vector<spg_vbd_energy>> descs(m_objects.size());
    std::vector<std::vector<double>> xs(m_objects.size()), vs(m_objects.size()), x0s(m_objects.size());
    std::vector<spg_vbd_object> objs(m_objects.size());
    for (size_t o = 0; o < m_objects.size(); ++o) {
        const auto &pg = m_objects[o];
        for (const auto &e : pg.energies()) flat[o].push_back(flattenVbd(e.get()));
        for (const auto &f : flat[o])
            descs[o].push_back({f.type, static_cast<int64_t>(f.stencils.size()) / flatStencilSize(f.type),
                                f.stencils.data(), f.params.data()});
        xs[o] = packVec3(pg.positions());
        vs[o] = packVec3(pg.velocities());
        x0s[o] = packVec3(pg.positions0());
        objs[o] = {pg.size(),         xs[o].data(),   vs[o].data(), pg.masses().data(),
                   static_cast<int32_t>(descs[o].size()), descs[o].data(), x0s[o].data()};
    }
    spg_vbd_ctx *ctx = nullptr;
    check(spg_vbd_create(&ctx, objs.data(), static_cast<int32_t>(objs.size()), m_device), nullptr);
    m_ctx = ctx;
    m_hostDirty = false;
}

void VBDGpu::uploadState()
{
    for (size_t o = 0; o < m_objects.size(); ++o) {
        const auto x = packVec3(m_objects[o].positions());
        const auto v = packVec3(m_objects[o].velocities());
        check(spg_vbd_set_state(m_ctx, static_cast<int32_t>(o), x.data(), v.data()), m_ctx);
    }
    m_hostDirty = false;
}

void VBDGpu::syncToHost()
{
    if (!m_ctx) return;
    for (size_t o = 0; o < m_objects.size(); ++o) {
        auto &pg = m_objects[o];
        std::vector<double> x(3 * pg.size()), v(3 * pg.size());
        check(spg_vbd_get_state(m_ctx, static_cast<int32_t>(o), x.data(), v.data()), m_ctx);
        for (int i = 0; i < pg.size(); ++i)
            for (int d = 0; d < 3; ++d) {
                pg.positions()[i][d] = x[3 * i + d];
                pg.velocities()[i][d] = v[3 * i + d];
            }
    }
}

void VBDGpu::step()
{
    if (m_objects.empty()) return;
    if (m_precomputationUpdateRequired || !m_ctx) {
        buildContext();
        m_precomputationUpdateRequired = false;
    } else if (!m_resident || m_hostDirty) {
        uploadState();
    }
    const double g[3] = {m_gravity[0], m_gravity[1], m_gravity[2]};
    check(spg_vbd_step(m_ctx, m_dtStep, m_nsubsteps, m_iterations, g), m_ctx);
    if (!m_resident) syncToHost();
}

void VBDGpu::reset()
{
    // VBD::reset (vbd.cpp:197-201)
    BaseSolver::reset();
    m_precomputationUpdateRequired = true;
}

}  // namespace spg::solver
