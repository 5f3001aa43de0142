/*
 * spg_host.h — C binding of the C++ spg API mirror (libspg_host.so), for callers that are not C++
 * (Python ctypes in this repository, or a cgo/JNI layer). It exposes scene construction through the
 * mirrored spg classes and the VBDGpu drop-in solver:
 *
 *   spg_pg_*          spg::ParticleGroup (spg/include/spg/sim/simObject/particleGroup.h:14-60) with
 *                     addParticle (particleGroup.cpp:74-82), addEnergy and each energy's addStencil +
 *                     preparePrecomputations (the five energy headers under spg/include/spg/sim/energy/).
 *   spg_rb_*          spg::RigidBodyGroup (spg/include/spg/sim/simObject/rigidBodyGroup.h:11-86) with addBody
 *                     (rigidBodyGroup.cpp:136-161), addEnergy and the four rigid-body energies' addStencil.
 *   spg_trimesh_*     spg::TriangleMesh edge/opposing-vertex enumeration (spg/include/spg/geom/triangleMesh.h:16-50)
 *                     and the cloth bending wiring of apps/solverCompare.cpp:441-452.
 *   spg_solver_*      spg::solver::BaseSolver (spg/include/spg/sim/solver/baseSolver.h:14-41) implemented by
 *                     spg::solver::VBDGpu, the GPU drop-in for spg::solver::VBD (vbd.h:14-41), or by
 *                     spg::solver::XPBDGpu, the GPU drop-in for spg::solver::XPBD (xpbd.h:15-30), or by
 *                     spg::solver::ImplicitEulerGpu, the GPU drop-in for the implicit Euler family
 *                     (implicitEulerBaraffWitkin.h, implicitEulerNewtonDv.h, implicitEulerNewtonDx.h,
 *                     implicitEulerNewtonRobust.h, bdf2.h).
 *
 * Errors: 0 on success, nonzero on failure with the message (the reference's exception text where
 * the reference throws) in spg_host_last_error().
 */
#ifndef SPG_HOST_H
#define SPG_HOST_H

#include <stdint.h>

#include "spg_vbd.h"
#include "spg_implicit.h"
#include "spg_rigid.h"
#include "spg_xpbd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct spg_pgroup spg_pgroup;
typedef struct spg_rbgroup spg_rbgroup;
typedef struct spg_solver spg_solver;

const char *spg_host_last_error(void);

spg_pgroup *spg_pg_create(void);
void spg_pg_destroy(spg_pgroup *pg);
/* addParticle(x[i], x0[i], mass[i]) for i < n */
int spg_pg_add_particles(spg_pgroup *pg, int64_t n, const double *x, const double *x0, const double *mass);
/* Creates one energy of `type` (SPG_VBD_* of spg_vbd.h), calls addStencil for each of n stencils with the
   per-stencil arguments `args` and, if prepare != 0, preparePrecomputations(pg); then addEnergy. args:
     SPRING [restLength, stiffness]; MEMBRANE_STVK [young, poisson]; DISCRETE_BENDING [restTheta, stiffness];
     STABLE_NEO_HOOKEAN [young, poisson]; SPRING_ANCHOR [ax, ay, az, stiffness].
   Returns the energy index in *index. */
int spg_pg_add_energy(spg_pgroup *pg, int32_t type, int64_t n, const int32_t *stencils, const double *args,
                      int32_t prepare, int32_t *index);
int spg_pg_prepare_energy(spg_pgroup *pg, int32_t index);
int spg_pg_num_particles(const spg_pgroup *pg, int64_t *n);
/* Prepared data of energy `index` in the params layout of spg_vbd_energy (spg_vbd.h). */
int spg_pg_energy_params(const spg_pgroup *pg, int32_t index, double *params);
/* Same energy in the params layout of spg_xpbd_energy (spg_xpbd.h): the effectiveCompliance() form. */
int spg_pg_energy_xpbd_params(const spg_pgroup *pg, int32_t index, double *params);
/* ParticleGroup::invMasses() (n entries). */
int spg_pg_inv_masses(const spg_pgroup *pg, double *w);
int spg_pg_get_state(const spg_pgroup *pg, double *x, double *v);
int spg_pg_set_state(spg_pgroup *pg, const double *x, const double *v);

/* spg::RigidBodyGroup (spg/include/spg/sim/simObject/rigidBodyGroup.h:11-86): addBody(x[i], x0[i], mass[i],
   theta[i], theta0[i], inertia[i], empty mesh) for i < n (x0 / theta0 may be NULL: the same as x / theta;
   inertia: local tensors, 9 per body, row-major). */
spg_rbgroup *spg_rb_create(void);
void spg_rb_destroy(spg_rbgroup *rb);
int spg_rb_add_bodies(spg_rbgroup *rb, int64_t n, const double *x, const double *x0, const double *mass,
                      const double *theta, const double *theta0, const double *inertia);
/* One rigid-body energy of `type` (SPG_RIGID_* of spg_rigid.h), addStencil per stencil with the arguments of
   spg_rigid_energy, then addEnergy; *index (may be NULL) receives its position. */
int spg_rb_add_energy(spg_rbgroup *rb, int32_t type, int64_t n, const int32_t *stencils, const double *args,
                      int32_t *index);
int spg_rb_num_bodies(const spg_rbgroup *rb, int64_t *n);
/* positions(), thetas(), velocities(), omegas(), rotationMatrices() (9 per body); any pointer may be NULL */
int spg_rb_get_state(const spg_rbgroup *rb, double *x, double *theta, double *v, double *omega, double *rot);
/* setPositions / setVelocities: the given quantities (NULL keeps the current ones); rotations follow theta */
int spg_rb_set_state(spg_rbgroup *rb, const double *x, const double *theta, const double *v, const double *omega);

/* spg::io::loadObj / loadMsh (source/geom/io/obj.cpp:10-35, msh.cpp:10-41): a mesh file read in two calls,
   first with NULL outputs for the counts, then with arrays of those sizes (vertices 3 doubles each, faces
   3 / tetrahedra 4 vertex indices each). */
int spg_io_load_obj(const char *path, int64_t *n_vertices, int64_t *n_faces, double *vertices, int32_t *faces);
int spg_io_load_msh(const char *path, int64_t *n_vertices, int64_t *n_tets, double *vertices, int32_t *tets);

/* TriangleMesh(faces).bendingStencils(): out may be NULL to query the count. */
int spg_trimesh_bending_stencils(int64_t n_vertices, int64_t n_faces, const int32_t *faces, int32_t *out,
                                 int64_t *count);

spg_solver *spg_solver_create_vbd_gpu(int32_t device);
/* XPBDGpu(parallel_gauss_seidel != 0) on `device` (XPBD(true) / XPBD(false) of the reference). */
spg_solver *spg_solver_create_xpbd_gpu(int32_t device, int32_t parallel_gauss_seidel);
/* ImplicitEulerGpu of kind `solver` (SPG_IMPLICIT_* of spg_implicit.h); NULL for an unknown kind. */
spg_solver *spg_solver_create_implicit_gpu(int32_t device, int32_t solver);
void spg_solver_destroy(spg_solver *s);
int spg_solver_add_object(spg_solver *s, const spg_pgroup *pg); /* copies the group, shares its energies */
int spg_solver_set_dt(spg_solver *s, double dt);
int spg_solver_set_num_substeps(spg_solver *s, int32_t n);
int spg_solver_set_gravity(spg_solver *s, const double g[3]);
int spg_solver_set_iterations(spg_solver *s, int32_t iterations); /* VBD only */
int spg_solver_set_resident_state(spg_solver *s, int32_t resident);
int spg_solver_step(spg_solver *s);
int spg_solver_reset(spg_solver *s);
/* VBD::requirePrecomputationUpdate (spg/include/spg/sim/solver/vbd.h:26; XPBD and the implicit solvers alike):
   the next step rebuilds colouring, incidence and device data from the groups' current state. */
int spg_solver_require_precomputation_update(spg_solver *s);
int spg_solver_sync_to_host(spg_solver *s);
int spg_solver_get_state(spg_solver *s, int32_t object, double *x, double *v);
/* BaseSolver::addObject(const RigidBodyGroup &) (baseSolver.cpp:10-14): copies the group, shares its energies.
   Only VBDGpu steps rigid-body groups; the other solvers throw at step(). */
int spg_solver_add_rb_object(spg_solver *s, const spg_rbgroup *rb);
/* rbGroups()[object]: positions, thetas, velocities, omegas, rotation matrices (any pointer may be NULL) */
int spg_solver_get_rb_state(spg_solver *s, int32_t object, double *x, double *theta, double *v, double *omega,
                            double *rot);
/* VBDGpu's rigid-body context (include/spg_rigid.h), NULL before the first step or without rigid bodies. */
spg_rigid_ctx *spg_solver_rigid_context(spg_solver *s);
/* The C-ABI context behind the solver (NULL before the first step, or for the other solver type). */
spg_vbd_ctx *spg_solver_context(spg_solver *s);
spg_xpbd_ctx *spg_solver_xpbd_context(spg_solver *s);
spg_implicit_ctx *spg_solver_implicit_context(spg_solver *s);

#ifdef __cplusplus
}
#endif

#endif /* SPG_HOST_H */
