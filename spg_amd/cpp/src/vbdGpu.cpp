// spg::solver::VBDGpu: BaseSolver drop-in for spg::solver::VBD (vbd.h:14-41, vbd.cpp:12-201) that
// hands the particle groups and energies to the GPU solver through the C-ABI of spg_vbd.h.
#include <stdexcept>
#include <string>

#include <spg/sim/energy/energies.h>
#include <spg/sim/energy/rigidBodyEnergies.h>
#include <spg/sim/solver/baseSolver.h>

#include "flatten.h"
#include "spg_rigid.h"
#include "spg_vbd.h"

namespace spg::solver {

void BaseSolver::reset()
{
    // baseSolver.cpp:27-35
    for (auto &obj : m_objects) obj.reset();
    for (auto &obj : m_rbObjects) obj.reset();
}

void BaseSolver::requireNoRigidBodies(const char *solverName) const
{
    if (!m_rbObjects.empty())
        throw std::runtime_error(std::string(solverName) + ": rigid-body groups are solved by VBDGpu only");
}

VBDGpu::VBDGpu(int device) : m_device(device) {}

VBDGpu::~VBDGpu()
{
    if (m_ctx) spg_vbd_destroy(m_ctx);
    if (m_rbCtx) spg_rigid_destroy(m_rbCtx);
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
    if (m_rbCtx) syncRigidToHost();
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
    if (m_objects.empty() && m_rbObjects.empty()) return;
    const bool rebuild = m_precomputationUpdateRequired || (!m_objects.empty() && !m_ctx) ||
                         (!m_rbObjects.empty() && !m_rbCtx);
    if (rebuild) {
        // with resident state the device holds the newest positions unless the host was edited since
        // (markHostStateDirty): bring them back first, so the rebuilt contexts continue from them
        if (m_resident && !m_hostDirty) syncToHost();
        if (!m_objects.empty()) buildContext();
        if (!m_rbObjects.empty()) buildRigidContext();
        m_hostDirty = false;
        m_precomputationUpdateRequired = false;
    } else if (!m_resident || m_hostDirty) {
        if (m_ctx) uploadState();
        if (m_rbCtx) uploadRigidState();
        m_hostDirty = false;
    }
    const double g[3] = {m_gravity[0], m_gravity[1], m_gravity[2]};
    if (m_ctx) check(spg_vbd_step(m_ctx, m_dtStep, m_nsubsteps, m_iterations, g), m_ctx);
    if (m_rbCtx) {
        if (spg_rigid_step(m_rbCtx, m_dtStep, m_nsubsteps, m_iterations, g) != SPG_VBD_OK)
            throw std::runtime_error(spg_rigid_last_error(m_rbCtx));
    }
    if (!m_resident) syncToHost();
}

namespace {

void rigidCheck(int rc, spg_rigid_ctx *ctx)
{
    if (rc != SPG_VBD_OK) throw std::runtime_error(spg_rigid_last_error(ctx));
}

// The addStencil arguments of one rigid-body energy in the spg_rigid_energy layout.
void flattenRigid(const RigidBodyGroup::EnergyT *e, int32_t &type, std::vector<int32_t> &st, std::vector<double> &p)
{
    st = e->flatStencils();
    if (auto *s = dynamic_cast<const RBSpringEnergy *>(e)) {
        type = SPG_RIGID_SPRING;
        for (int i = 0; i < s->nStencils(); ++i) {
            const auto &l = s->localRBPoints()[i];
            p.insert(p.end(), {l[0][0], l[0][1], l[0][2], l[1][0], l[1][1], l[1][2], s->restLengths()[i],
                               s->modelStiffness()[i](0, 0)});
        }
    } else if (auto *q = dynamic_cast<const RBSpringSquaredEnergy *>(e)) {
        type = SPG_RIGID_SPRING_SQUARED;
        for (int i = 0; i < q->nStencils(); ++i) {
            const auto &l = q->localRBPoints()[i];
            p.insert(p.end(), {l[0][0], l[0][1], l[0][2], l[1][0], l[1][1], l[1][2], q->restLengthArguments()[i],
                               q->modelStiffness()[i](0, 0)});
        }
    } else if (auto *a = dynamic_cast<const RBSpringAnchorEnergy *>(e)) {
        type = SPG_RIGID_SPRING_ANCHOR;
        for (int i = 0; i < a->nStencils(); ++i) {
            const auto &l = a->localRBPoints()[i];
            const auto &c = a->anchors()[i];
            p.insert(p.end(), {l[0], l[1], l[2], c[0], c[1], c[2], a->modelStiffness()[i](0, 0)});
        }
    } else if (auto *o = dynamic_cast<const RBOrientationAlignmentEnergy *>(e)) {
        type = SPG_RIGID_ORIENTATION_ALIGNMENT;
        for (int i = 0; i < o->nStencils(); ++i) p.push_back(o->modelStiffness()[i](0, 0));
    } else {
        throw std::runtime_error("VBDGpu: unsupported rigid-body energy type");
    }
}

std::vector<double> packMat3(const std::vector<Matrix3> &m)
{
    std::vector<double> out(9 * m.size());
    for (size_t i = 0; i < m.size(); ++i) std::copy(m[i].m.begin(), m[i].m.end(), out.begin() + 9 * i);
    return out;
}

}  // namespace

void VBDGpu::buildRigidContext()
{
    if (m_rbCtx) {
        spg_rigid_destroy(m_rbCtx);
        m_rbCtx = nullptr;
    }
    const size_t n = m_rbObjects.size();
    std::vector<std::vector<int32_t>> sts;
    std::vector<std::vector<double>> prms;
    std::vector<std::vector<spg_rigid_energy>> descs(n);
    std::vector<std::vector<double>> xs(n), ths(n), vs(n), ws(n), inertia(n);
    std::vector<spg_rigid_group> groups(n);
    for (size_t g = 0; g < n; ++g) {
        const auto &rb = m_rbObjects[g];
        for (const auto &e : rb.energies()) {
            int32_t type = 0;
            sts.emplace_back();
            prms.emplace_back();
            flattenRigid(e.get(), type, sts.back(), prms.back());
            descs[g].push_back({type, static_cast<int64_t>(e->nStencils()), nullptr, nullptr});
        }
        xs[g] = packVec3(rb.positions());
        ths[g] = packVec3(rb.thetas());
        vs[g] = packVec3(rb.velocities());
        ws[g] = packVec3(rb.omegas());
        inertia[g] = packMat3(rb.localInertias());
    }
    // stencil/parameter vectors are complete: take their addresses now
    size_t k = 0;
    for (size_t g = 0; g < n; ++g)
        for (auto &d : descs[g]) {
            d.stencils = sts[k].data();
            d.params = prms[k].data();
            ++k;
        }
    for (size_t g = 0; g < n; ++g) {
        const auto &rb = m_rbObjects[g];
        groups[g] = {rb.size(),         xs[g].data(),      ths[g].data(),
                     vs[g].data(),      ws[g].data(),      rb.masses().data(),
                     inertia[g].data(), static_cast<int32_t>(descs[g].size()), descs[g].data()};
    }
    spg_rigid_ctx *ctx = nullptr;
    rigidCheck(spg_rigid_create(&ctx, groups.data(), static_cast<int32_t>(n), m_device), nullptr);
    m_rbCtx = ctx;
    // create derives each rotation from theta; the group's rotation matrices are composed increments and
    // may differ from aaToR(theta) after steps, so the context starts from the matrices the group holds
    for (size_t g = 0; g < n; ++g) {
        const auto R = packMat3(m_rbObjects[g].rotationMatrices());
        rigidCheck(spg_rigid_set_rotations(m_rbCtx, static_cast<int32_t>(g), R.data()), m_rbCtx);
    }
}

void VBDGpu::uploadRigidState()
{
    for (size_t g = 0; g < m_rbObjects.size(); ++g) {
        const auto &rb = m_rbObjects[g];
        const auto x = packVec3(rb.positions()), th = packVec3(rb.thetas()), v = packVec3(rb.velocities()),
                   w = packVec3(rb.omegas());
        rigidCheck(spg_rigid_set_state(m_rbCtx, static_cast<int32_t>(g), x.data(), th.data(), v.data(), w.data()),
                   m_rbCtx);
        // the rotation matrices are state of their own (composed increments, not recomputed from theta)
        const auto R = packMat3(rb.rotationMatrices());
        rigidCheck(spg_rigid_set_rotations(m_rbCtx, static_cast<int32_t>(g), R.data()), m_rbCtx);
    }
}

void VBDGpu::syncRigidToHost()
{
    for (size_t g = 0; g < m_rbObjects.size(); ++g) {
        auto &rb = m_rbObjects[g];
        const size_t nb = static_cast<size_t>(rb.size());
        std::vector<double> x(3 * nb), th(3 * nb), v(3 * nb), w(3 * nb), R(9 * nb);
        rigidCheck(spg_rigid_get_state(m_rbCtx, static_cast<int32_t>(g), x.data(), th.data(), v.data(), w.data()),
                   m_rbCtx);
        rigidCheck(spg_rigid_get_rotations(m_rbCtx, static_cast<int32_t>(g), R.data()), m_rbCtx);
        // the solver's state as the reference's VBD leaves it in the group: x, theta, v, omega, the composed
        // rotation matrices, and the inertias updated from them (updateInertias)
        for (size_t b = 0; b < nb; ++b) {
            for (int d = 0; d < 3; ++d) {
                rb.positions()[b][d] = x[3 * b + d];
                rb.thetas()[b][d] = th[3 * b + d];
                rb.velocities()[b][d] = v[3 * b + d];
                rb.omegas()[b][d] = w[3 * b + d];
            }
            std::copy(R.begin() + 9 * b, R.begin() + 9 * b + 9, rb.rotationMatrices()[b].m.begin());
        }
        rb.updateInertias();
    }
}

void VBDGpu::reset()
{
    // VBD::reset (vbd.cpp:197-201)
    BaseSolver::reset();
    m_precomputationUpdateRequired = true;
}

}  // namespace spg::solver
