// C binding of the C++ spg mirror (include/spg_host.h).
#include <spg/sim/energy/rigidBodyEnergies.h>
#include <spg/sim/simObject/rigidBodyGroup.h>
#include <algorithm>
#include <memory>
#include <string>

#include <spg/geom/io/msh.h>
#include <spg/geom/io/obj.h>
#include <spg/geom/triangleMesh.h>
#include <spg/sim/energy/energies.h>
#include <spg/sim/solver/baseSolver.h>

#include "flatten.h"
#include "spg_host.h"

struct spg_pgroup {
    spg::ParticleGroup group;
};

struct spg_rbgroup {
    spg::RigidBodyGroup group;
};

struct spg_solver {
    std::unique_ptr<spg::solver::BaseSolver> solver;
    spg::solver::VBDGpu *vbd = nullptr;    // the solver, when it is a VBDGpu
    spg::solver::XPBDGpu *xpbd = nullptr;  // the solver, when it is an XPBDGpu
    spg::solver::ImplicitEulerGpu *implicit = nullptr;  // the solver, when it is an implicit Euler solver
};

namespace {

thread_local std::string g_err;

template <class F>
int guarded(F &&f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return 1;
    }
}

spg::Vector3 v3(const double *p) { return {p[0], p[1], p[2]}; }

template <class E, int N>
std::array<int, N> stencilAt(const int32_t *s, int64_t i)
{
    std::array<int, N> a;
    for (int k = 0; k < N; ++k) a[k] = s[i * N + k];
    return a;
}

// addStencil(stencil, a, b) / addStencil(stencil, a, b, c) over n stencils with their raw arguments
template <class E, int N>
std::shared_ptr<spg::ParticleGroup::EnergyT> addStencils2(int64_t n, const int32_t *st, const double *a)
{
    auto e = std::make_shared<E>();
    for (int64_t i = 0; i < n; ++i) e->addStencil(stencilAt<E, N>(st, i), a[2 * i], a[2 * i + 1]);
    return e;
}
template <class E, int N>
std::shared_ptr<spg::ParticleGroup::EnergyT> addStencils3(int64_t n, const int32_t *st, const double *a)
{
    auto e = std::make_shared<E>();
    for (int64_t i = 0; i < n; ++i) e->addStencil(stencilAt<E, N>(st, i), a[3 * i], a[3 * i + 1], a[3 * i + 2]);
    return e;
}

void checkIndex(const spg_pgroup *pg, int32_t index)
{
    if (!pg || index < 0 || index >= static_cast<int32_t>(pg->group.energies().size()))
        throw std::runtime_error("energy index out of range");
}

}  // namespace

namespace {
template <class Cell, int N, class Loader>
int loadMesh(Loader load, const char *path, int64_t *nv, int64_t *nc, double *vertices, int32_t *cells)
{
    return guarded([&] {
        if (!path || !nv || !nc) throw std::runtime_error("invalid arguments");
        const auto [v, c] = load(std::string(path));
        *nv = static_cast<int64_t>(v.size());
        *nc = static_cast<int64_t>(c.size());
        if (vertices)
            for (size_t i = 0; i < v.size(); ++i)
                for (int d = 0; d < 3; ++d) vertices[3 * i + d] = v[i][d];
        if (cells)
            for (size_t i = 0; i < c.size(); ++i)
                for (int d = 0; d < N; ++d) cells[N * i + d] = c[i][d];
    });
}
}  // namespace

extern "C" {

const char *spg_host_last_error(void) { return g_err.c_str(); }

spg_pgroup *spg_pg_create(void) { return new spg_pgroup; }
void spg_pg_destroy(spg_pgroup *pg) { delete pg; }

int spg_pg_add_particles(spg_pgroup *pg, int64_t n, const double *x, const double *x0, const double *mass)
{
    return guarded([&] {
        if (!pg || n < 0 || (n > 0 && (!x || !x0 || !mass))) throw std::runtime_error("invalid arguments");
        for (int64_t i = 0; i < n; ++i) pg->group.addParticle(v3(x + 3 * i), v3(x0 + 3 * i), mass[i]);
    });
}

int spg_pg_add_energy(spg_pgroup *pg, int32_t type, int64_t n, const int32_t *st, const double *a, int32_t prepare,
                      int32_t *index)
{
    return guarded([&] {
        if (!pg || n < 0 || (n > 0 && (!st || !a))) throw std::runtime_error("invalid arguments");
        std::shared_ptr<spg::ParticleGroup::EnergyT> energy;
        switch (type) {
        case SPG_VBD_SPRING: {
            auto e = std::make_shared<spg::SpringEnergy>();
            for (int64_t i = 0; i < n; ++i) e->addStencil(stencilAt<spg::SpringEnergy, 2>(st, i), a[2 * i], a[2 * i + 1]);
            energy = e;
            break;
        }
        case SPG_VBD_MEMBRANE_STVK: {
            auto e = std::make_shared<spg::MembraneStvkEnergy>();
            for (int64_t i = 0; i < n; ++i)
                e->addStencil(stencilAt<spg::MembraneStvkEnergy, 3>(st, i), a[2 * i], a[2 * i + 1]);
            energy = e;
            break;
        }
        case SPG_VBD_DISCRETE_BENDING: {
            auto e = std::make_shared<spg::DiscreteBendingEnergy>();
            for (int64_t i = 0; i < n; ++i)
                e->addStencil(stencilAt<spg::DiscreteBendingEnergy, 4>(st, i), a[2 * i], a[2 * i + 1]);
            energy = e;
            break;
        }
        case SPG_VBD_STABLE_NEO_HOOKEAN: {
            auto e = std::make_shared<spg::StableNeoHookeanEnergy>();
            for (int64_t i = 0; i < n; ++i)
                e->addStencil(stencilAt<spg::StableNeoHookeanEnergy, 4>(st, i), a[2 * i], a[2 * i + 1]);
            energy = e;
            break;
        }
        case SPG_VBD_SPRING_ANCHOR: {
            auto e = std::make_shared<spg::SpringAnchorEnergy>();
            for (int64_t i = 0; i < n; ++i)
                e->addStencil(stencilAt<spg::SpringAnchorEnergy, 1>(st, i), v3(a + 4 * i), a[4 * i + 3]);
            energy = e;
            break;
        }
        case SPG_VBD_SPRING_CONTINUUM:
            energy = addStencils2<spg::SpringContinuumEnergy, 2>(n, st, a);
            break;
        case SPG_VBD_SPRING_SQUARED_CONTINUUM:
            energy = addStencils2<spg::SpringSquaredContinuumEnergy, 2>(n, st, a);
            break;
        case SPG_VBD_MEMBRANE_BARAFF_WITKIN:
            energy = addStencils3<spg::MembraneBaraffWitkinEnergy, 3>(n, st, a);
            break;
        case SPG_VBD_MEMBRANE_CHOI:
            energy = addStencils3<spg::MembraneChoiEnergy, 3>(n, st, a);
            break;
        case SPG_VBD_QUADRATIC_BENDING:
            energy = addStencils2<spg::QuadraticBendingEnergy, 4>(n, st, a);
            break;
        case SPG_VBD_BARAFF_WITKIN_BENDING:
            energy = addStencils2<spg::BaraffWitkinBendingEnergy, 4>(n, st, a);
            break;
        case SPG_VBD_STVK:
            energy = addStencils2<spg::StvkEnergy, 4>(n, st, a);
            break;
        case SPG_VBD_STABLE_SQUARED_NEO_HOOKEAN:
            energy = addStencils2<spg::StableSquaredNeoHookeanEnergy, 4>(n, st, a);
            break;
        default: throw std::runtime_error("unsupported energy type");
        }
        for (int64_t i = 0; i < n * energy->stencilSize(); ++i)
            if (st[i] < 0 || st[i] >= pg->group.size()) throw std::runtime_error("stencil index out of range");
        if (prepare) energy->preparePrecomputations(pg->group);
        pg->group.addEnergy(energy);
        if (index) *index = static_cast<int32_t>(pg->group.energies().size()) - 1;
    });
}

int spg_pg_prepare_energy(spg_pgroup *pg, int32_t index)
{
    return guarded([&] {
        checkIndex(pg, index);
        pg->group.energies()[index]->preparePrecomputations(pg->group);
    });
}

int spg_pg_num_particles(const spg_pgroup *pg, int64_t *n)
{
    return guarded([&] {
        if (!pg || !n) throw std::runtime_error("invalid arguments");
        *n = pg->group.size();
    });
}

int spg_pg_energy_params(const spg_pgroup *pg, int32_t index, double *out)
{
    return guarded([&] {
        checkIndex(pg, index);
        const auto f = spg::solver::flattenVbd(pg->group.energies()[index].get());
        std::copy(f.params.begin(), f.params.end(), out);
    });
}

int spg_pg_energy_xpbd_params(const spg_pgroup *pg, int32_t index, double *out)
{
    return guarded([&] {
        checkIndex(pg, index);
        const auto f = spg::solver::flattenXpbd(pg->group.energies()[index].get());
        std::copy(f.params.begin(), f.params.end(), out);
    });
}

int spg_pg_inv_masses(const spg_pgroup *pg, double *w)
{
    return guarded([&] {
        if (!pg || !w) throw std::runtime_error("invalid arguments");
        std::copy(pg->group.invMasses().begin(), pg->group.invMasses().end(), w);
    });
}

int spg_pg_get_state(const spg_pgroup *pg, double *x, double *v)
{
    return guarded([&] {
        if (!pg) throw std::runtime_error("invalid arguments");
        for (int i = 0; i < pg->group.size(); ++i)
            for (int d = 0; d < 3; ++d) {
                if (x) x[3 * i + d] = pg->group.positions()[i][d];
                if (v) v[3 * i + d] = pg->group.velocities()[i][d];
            }
    });
}

int spg_pg_set_state(spg_pgroup *pg, const double *x, const double *v)
{
    return guarded([&] {
        if (!pg) throw std::runtime_error("invalid arguments");
        for (int i = 0; i < pg->group.size(); ++i)
            for (int d = 0; d < 3; ++d) {
                if (x) pg->group.positions()[i][d] = x[3 * i + d];
                if (v) pg->group.velocities()[i][d] = v[3 * i + d];
            }
    });
}


int spg_io_load_obj(const char *path, int64_t *n_vertices, int64_t *n_faces, double *vertices, int32_t *faces)
{
    return loadMesh<spg::Int3, 3>(spg::io::loadObj, path, n_vertices, n_faces, vertices, faces);
}

int spg_io_load_msh(const char *path, int64_t *n_vertices, int64_t *n_tets, double *vertices, int32_t *tets)
{
    return loadMesh<spg::Int4, 4>(spg::io::loadMsh, path, n_vertices, n_tets, vertices, tets);
}

int spg_trimesh_bending_stencils(int64_t n_vertices, int64_t n_faces, const int32_t *faces, int32_t *out,
                                 int64_t *count)
{
    return guarded([&] {
        if (n_vertices < 0 || n_faces < 0 || (n_faces > 0 && !faces) || !count)
            throw std::runtime_error("invalid arguments");
        std::vector<spg::Vector3> verts(n_vertices);
        std::vector<spg::Int3> f(n_faces);
        for (int64_t i = 0; i < n_faces; ++i) f[i] = {faces[3 * i], faces[3 * i + 1], faces[3 * i + 2]};
        const spg::TriangleMesh mesh(verts, verts, f);
        const auto st = mesh.bendingStencils();
        *count = static_cast<int64_t>(st.size());
        if (out)
            for (size_t i = 0; i < st.size(); ++i)
                for (int k = 0; k < 4; ++k) out[4 * i + k] = st[i][k];
    });
}

spg_solver *spg_solver_create_vbd_gpu(int32_t device)
{
    auto *s = new spg_solver;
    s->vbd = new spg::solver::VBDGpu(device);
    s->solver.reset(s->vbd);
    return s;
}

spg_solver *spg_solver_create_implicit_gpu(int32_t device, int32_t solver)
{
    if (solver < 0 || solver > 4) {
        g_err = "unknown implicit solver";
        return nullptr;
    }
    auto *s = new spg_solver;
    s->implicit = new spg::solver::ImplicitEulerGpu(static_cast<spg::solver::ImplicitEulerGpu::Kind>(solver), device);
    s->solver.reset(s->implicit);
    return s;
}

spg_solver *spg_solver_create_xpbd_gpu(int32_t device, int32_t parallel_gauss_seidel)
{
    auto *s = new spg_solver;
    s->xpbd = new spg::solver::XPBDGpu(parallel_gauss_seidel != 0, device);
    s->solver.reset(s->xpbd);
    return s;
}

void spg_solver_destroy(spg_solver *s) { delete s; }

int spg_solver_add_object(spg_solver *s, const spg_pgroup *pg)
{
    return guarded([&] {
        if (!s || !pg) throw std::runtime_error("invalid arguments");
        s->solver->addObject(pg->group);
        if (s->vbd) s->vbd->requirePrecomputationUpdate();
        if (s->xpbd) s->xpbd->requirePrecomputationUpdate();
        if (s->implicit) s->implicit->requirePrecomputationUpdate();
    });
}

int spg_solver_set_dt(spg_solver *s, double dt)
{
    return guarded([&] { s->solver->setDt(dt); });
}
int spg_solver_set_num_substeps(spg_solver *s, int32_t n)
{
    return guarded([&] { s->solver->setNumSubsteps(n); });
}
int spg_solver_set_gravity(spg_solver *s, const double g[3])
{
    return guarded([&] { s->solver->setGravity(spg::Vector3{g[0], g[1], g[2]}); });
}
int spg_solver_set_iterations(spg_solver *s, int32_t iterations)
{
    return guarded([&] {
        if (!s->vbd) throw std::runtime_error("set_iterations: not a VBD solver");
        s->vbd->setIterations(iterations);
    });
}
int spg_solver_set_resident_state(spg_solver *s, int32_t resident)
{
    return guarded([&] {
        if (s->vbd) s->vbd->setResidentState(resident != 0);
        if (s->xpbd) s->xpbd->setResidentState(resident != 0);
        if (s->implicit) s->implicit->setResidentState(resident != 0);
    });
}
int spg_solver_step(spg_solver *s)
{
    return guarded([&] { s->solver->step(); });
}
int spg_solver_reset(spg_solver *s)
{
    return guarded([&] { s->solver->reset(); });
}
int spg_solver_require_precomputation_update(spg_solver *s)
{
    return guarded([&] {
        if (s->vbd) s->vbd->requirePrecomputationUpdate();
        if (s->xpbd) s->xpbd->requirePrecomputationUpdate();
        if (s->implicit) s->implicit->requirePrecomputationUpdate();
    });
}
int spg_solver_sync_to_host(spg_solver *s)
{
    return guarded([&] {
        if (s->vbd) s->vbd->syncToHost();
        if (s->xpbd) s->xpbd->syncToHost();
        if (s->implicit) s->implicit->syncToHost();
    });
}

int spg_solver_get_state(spg_solver *s, int32_t object, double *x, double *v)
{
    return guarded([&] {
        if (!s || object < 0 || object >= static_cast<int32_t>(s->solver->particleGroups().size()))
            throw std::runtime_error("invalid object");
        const auto &pg = s->solver->particleGroups()[object];
        for (int i = 0; i < pg.size(); ++i)
            for (int d = 0; d < 3; ++d) {
                if (x) x[3 * i + d] = pg.positions()[i][d];
                if (v) v[3 * i + d] = pg.velocities()[i][d];
            }
    });
}

spg_vbd_ctx *spg_solver_context(spg_solver *s) { return s && s->vbd ? s->vbd->context() : nullptr; }
spg_xpbd_ctx *spg_solver_xpbd_context(spg_solver *s) { return s && s->xpbd ? s->xpbd->context() : nullptr; }
spg_implicit_ctx *spg_solver_implicit_context(spg_solver *s)
{
    return s && s->implicit ? s->implicit->context() : nullptr;
}


/* ---- rigid bodies ---- */

spg_rbgroup *spg_rb_create(void) { return new spg_rbgroup; }
void spg_rb_destroy(spg_rbgroup *rb) { delete rb; }

int spg_rb_add_bodies(spg_rbgroup *rb, int64_t n, const double *x, const double *x0, const double *mass,
                      const double *theta, const double *theta0, const double *inertia)
{
    return guarded([&] {
        if (!rb || n < 0 || (n > 0 && (!x || !mass || !theta || !inertia))) throw std::runtime_error("invalid arguments");
        const spg::TriangleMesh mesh({}, {}, {});
        for (int64_t i = 0; i < n; ++i) {
            const spg::Vector3 p(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
            const spg::Vector3 p0 = x0 ? spg::Vector3(x0[3 * i], x0[3 * i + 1], x0[3 * i + 2]) : p;
            const spg::Vector3 t(theta[3 * i], theta[3 * i + 1], theta[3 * i + 2]);
            const spg::Vector3 t0 = theta0 ? spg::Vector3(theta0[3 * i], theta0[3 * i + 1], theta0[3 * i + 2]) : t;
            spg::Matrix3 I;
            for (int k = 0; k < 9; ++k) I.m[k] = inertia[9 * i + k];
            rb->group.addBody(p, p0, mass[i], t, t0, I, mesh);
        }
    });
}

int spg_rb_add_energy(spg_rbgroup *rb, int32_t type, int64_t n, const int32_t *stencils, const double *args,
                      int32_t *index)
{
    return guarded([&] {
        if (!rb || n < 0 || (n > 0 && (!stencils || !args))) throw std::runtime_error("invalid arguments");
        auto v3 = [](const double *p) { return spg::Vector3(p[0], p[1], p[2]); };
        std::shared_ptr<spg::RigidBodyGroup::EnergyT> e;
        switch (type) {
        case SPG_RIGID_SPRING: {
            auto s = std::make_shared<spg::RBSpringEnergy>();
            for (int64_t i = 0; i < n; ++i)
                s->addStencil({stencils[2 * i], stencils[2 * i + 1]}, {v3(args + 8 * i), v3(args + 8 * i + 3)},
                              args[8 * i + 6], args[8 * i + 7]);
            e = s;
            break;
        }
        case SPG_RIGID_SPRING_SQUARED: {
            auto s = std::make_shared<spg::RBSpringSquaredEnergy>();
            for (int64_t i = 0; i < n; ++i)
                s->addStencil({stencils[2 * i], stencils[2 * i + 1]}, {v3(args + 8 * i), v3(args + 8 * i + 3)},
                              args[8 * i + 6], args[8 * i + 7]);
            e = s;
            break;
        }
        case SPG_RIGID_SPRING_ANCHOR: {
            auto s = std::make_shared<spg::RBSpringAnchorEnergy>();
            for (int64_t i = 0; i < n; ++i)
                s->addStencil({stencils[i]}, v3(args + 7 * i), v3(args + 7 * i + 3), args[7 * i + 6]);
            e = s;
            break;
        }
        case SPG_RIGID_ORIENTATION_ALIGNMENT: {
            auto s = std::make_shared<spg::RBOrientationAlignmentEnergy>();
            for (int64_t i = 0; i < n; ++i) s->addStencil({stencils[2 * i], stencils[2 * i + 1]}, args[i]);
            e = s;
            break;
        }
        default: throw std::runtime_error("unknown rigid-body energy type");
        }
        rb->group.addEnergy(e);
        if (index) *index = static_cast<int32_t>(rb->group.energies().size()) - 1;
    });
}

namespace {
void rbState(const spg::RigidBodyGroup &g, double *x, double *theta, double *v, double *omega, double *rot)
{
    for (int i = 0; i < g.size(); ++i) {
        for (int d = 0; d < 3; ++d) {
            if (x) x[3 * i + d] = g.positions()[i][d];
            if (theta) theta[3 * i + d] = g.thetas()[i][d];
            if (v) v[3 * i + d] = g.velocities()[i][d];
            if (omega) omega[3 * i + d] = g.omegas()[i][d];
        }
        if (rot) std::copy(g.rotationMatrices()[i].m.begin(), g.rotationMatrices()[i].m.end(), rot + 9 * i);
    }
}
}  // namespace

int spg_rb_num_bodies(const spg_rbgroup *rb, int64_t *n)
{
    return guarded([&] {
        if (!rb || !n) throw std::runtime_error("invalid arguments");
        *n = rb->group.size();
    });
}

int spg_rb_get_state(const spg_rbgroup *rb, double *x, double *theta, double *v, double *omega, double *rot)
{
    return guarded([&] {
        if (!rb) throw std::runtime_error("invalid arguments");
        rbState(rb->group, x, theta, v, omega, rot);
    });
}

int spg_rb_set_state(spg_rbgroup *rb, const double *x, const double *theta, const double *v, const double *omega)
{
    return guarded([&] {
        if (!rb) throw std::runtime_error("invalid arguments");
        auto &g = rb->group;
        const size_t nb = static_cast<size_t>(g.size());
        spg::VectorX pos(6 * nb), vel(6 * nb);
        g.getPositions(pos, 0);
        g.getVelocities(vel, 0);
        for (size_t b = 0; b < nb; ++b)
            for (int d = 0; d < 3; ++d) {
                if (x) pos[6 * b + d] = x[3 * b + d];
                if (theta) pos[6 * b + 3 + d] = theta[3 * b + d];
                if (v) vel[6 * b + d] = v[3 * b + d];
                if (omega) vel[6 * b + 3 + d] = omega[3 * b + d];
            }
        g.setPositions(pos, 0);
        g.setVelocities(vel, 0);
    });
}

int spg_solver_add_rb_object(spg_solver *s, const spg_rbgroup *rb)
{
    return guarded([&] {
        if (!s || !rb) throw std::runtime_error("invalid arguments");
        s->solver->addObject(rb->group);
        if (s->vbd) s->vbd->requirePrecomputationUpdate();
        if (s->xpbd) s->xpbd->requirePrecomputationUpdate();
        if (s->implicit) s->implicit->requirePrecomputationUpdate();
    });
}

int spg_solver_get_rb_state(spg_solver *s, int32_t object, double *x, double *theta, double *v, double *omega,
                            double *rot)
{
    return guarded([&] {
        if (!s || object < 0 || object >= static_cast<int32_t>(s->solver->rbGroups().size()))
            throw std::runtime_error("invalid object");
        rbState(s->solver->rbGroups()[object], x, theta, v, omega, rot);
    });
}

spg_rigid_ctx *spg_solver_rigid_context(spg_solver *s) { return s && s->vbd ? s->vbd->rigidContext() : nullptr; }

}  // extern "C"
