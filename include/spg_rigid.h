/*
 * spg_rigid.h — C-ABI of the MI355X Vertex Block Descent solver on rigid bodies (in libspg_vbd.so).
 *
 * spg's VBD treats a RigidBodyGroup as a set of 6-DOF blocks (centre of mass and a rotation increment
 * applied on the left of the body's rotation matrix) and descends them colour by colour exactly as it does
 * particles. This library runs that path in FP64 on the GPU. Plain pointers and sizes, no C++ or torch types.
 *
 * Reference interfaces replaced (paths relative to the spg repository):
 *   spg_rigid_create      RigidBodyGroup::addBody / addEnergy (spg/source/sim/simObject/rigidBodyGroup.cpp:130-161)
 *                         with the addStencil of the four rigid-body energies
 *                         (spg/source/sim/energy/rigidBodySpringEnergy.cpp:38-53,
 *                         rigidBodySpringSquaredEnergy.cpp:35-50, rigidBodySpringAnchorEnergy.cpp:63-75,
 *                         rigidBodyOrientationAlignmentEnergy.cpp:33-40), BaseSolver::addObject and VBD's first-step
 *                         precomputation (spg/source/sim/solver/vbd.cpp:203-268: colour groups and per-body
 *                         element lists).
 *   spg_rigid_step        VBD::step (vbd.cpp:12-195) on RigidBodyGroup objects: gyroscopic velocity prediction,
 *                         RigidBodyGroup::integrateVelocities (rigidBodyGroup.cpp:97-113), the 6x6 block descent
 *                         with the angular inertia terms (vbd.cpp:105-145) and
 *                         RigidBodyGroup::updateElementPositionFromDx (rigidBodyGroup.cpp:64-77), then
 *                         computeIntegratedVelocities (rigidBodyGroup.cpp:115-128).
 *   spg_rigid_get_state / set_state  RigidBodyGroup::getPositions/getVelocities and setPositions/setVelocities
 *                         (rigidBodyGroup.cpp:21-62): positions are (x, theta) with theta the axis-angle vector of
 *                         the rotation; setting them rebuilds rotation matrices and world inertias.
 *   spg_rigid_get_rotations / set_rotations  RigidBodyGroup::rotationMatrices() (rigidBodyGroup.h:45-46).
 *   spg_rigid_get_colors  the colour groups of VBD::computeParallelVertexGroups (vbd.cpp:203-229).
 *   spg_rigid_element_derivatives  StencilBlockEnergy::energyGradient / energyHessian of one rigid-body element
 *                         (spg/include/spg/sim/energy/energy.h:268-270; the anchor's own overrides,
 *                         rigidBodySpringAnchorEnergy.cpp:85-100), as the solve evaluates them.
 *
 * Error behaviour and codes are those of spg_vbd.h: SPG_VBD_OK (0) or a nonzero code, with the message from
 * spg_rigid_last_error (per context, or the last context-free failure on this thread when ctx is NULL).
 * Setup checks reproduce the reference's exceptions ("Zero/negative rest length invalid", "Negative rest length
 * invalid"); a stencil index outside the group ("stencil index out of range", which the reference leaves
 * unchecked for single-body stencils) and a singular local inertia are rejected as SPG_VBD_ERR_INVALID. A context
 * is not thread-safe; its work goes to one HIP stream in call order.
 */
#ifndef SPG_RIGID_H
#define SPG_RIGID_H

#include <stdint.h>

#include "spg_vbd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Energy types and their addStencil arguments per stencil (row-major, doubles). */
#define SPG_RIGID_SPRING 21                /* RBSpringEnergy: local points l0 (3), l1 (3), rest length, stiffness */
#define SPG_RIGID_SPRING_SQUARED 22        /* RBSpringSquaredEnergy: l0 (3), l1 (3), rest length, stiffness */
#define SPG_RIGID_SPRING_ANCHOR 23         /* RBSpringAnchorEnergy: local point (3), anchor (3), stiffness */
#define SPG_RIGID_ORIENTATION_ALIGNMENT 24 /* RBOrientationAlignmentEnergy: stiffness */

typedef struct {
    int32_t type;
    int64_t n_stencils;
    const int32_t *stencils; /* n_stencils x stencil size (2, or 1 for the anchor) body indices */
    const double *params;    /* n_stencils x the argument count above */
} spg_rigid_energy;

typedef struct {
    int64_t n_bodies;
    const double *x;       /* centres of mass, 3 per body */
    const double *theta;   /* orientations as axis-angle vectors, 3 per body */
    const double *v;       /* linear velocities, 3 per body, or NULL for zero */
    const double *omega;   /* angular velocities, 3 per body, or NULL for zero */
    const double *mass;    /* 1 per body */
    const double *inertia; /* local inertia tensors, 9 per body, row-major */
    int32_t n_energies;
    const spg_rigid_energy *energies; /* in addEnergy order */
} spg_rigid_group;

typedef struct {
    int32_t n_colors;          /* colour groups (the largest count over groups) */
    int64_t n_bodies;
    int64_t n_elements;
    int64_t incidence_entries; /* (element, slot) pairs read by the solve, over all bodies */
    int64_t skipped_blocks;    /* block solves skipped because det(H) <= 1e-10, since creation */
    int64_t solve_bytes;       /* algorithmic HBM bytes of one sweep over all colours */
} spg_rigid_stats;

typedef struct spg_rigid_ctx spg_rigid_ctx;

/* One context holds any number of RigidBodyGroups (BaseSolver::addObject order); they are solved together
   but never interact. Host arrays are copied. device: HIP device ordinal. */
int spg_rigid_create(spg_rigid_ctx **out, const spg_rigid_group *groups, int32_t n_groups, int32_t device);
void spg_rigid_destroy(spg_rigid_ctx *ctx);
const char *spg_rigid_last_error(const spg_rigid_ctx *ctx);

/* dt split into nsubsteps, `iterations` sweeps over the colours per substep, gravity (3). */
int spg_rigid_step(spg_rigid_ctx *ctx, double dt, int32_t nsubsteps, int32_t iterations, const double *gravity);

/* Any pointer may be NULL (that quantity is not read or written). 3 doubles per body of the group. */
int spg_rigid_get_state(spg_rigid_ctx *ctx, int32_t group, double *x, double *theta, double *v, double *omega);
int spg_rigid_set_state(spg_rigid_ctx *ctx, int32_t group, const double *x, const double *theta, const double *v,
                        const double *omega);

/* RigidBodyGroup::rotationMatrices() (rigidBodyGroup.h:45-46): 9 doubles per body, row-major. The rotation is
   state of its own: the solver composes it with each increment and derives theta from it, so a caller that keeps
   the state on the host reads and writes it beside theta to continue bit for bit. set_rotations leaves theta as it
   is (like assigning rotationMatrices() without updateThetas()). */
int spg_rigid_get_rotations(spg_rigid_ctx *ctx, int32_t group, double *R);
int spg_rigid_set_rotations(spg_rigid_ctx *ctx, int32_t group, const double *R);

int spg_rigid_get_colors(const spg_rigid_ctx *ctx, int32_t group, int32_t *colors, int32_t *n_colors);
int spg_rigid_get_stats(spg_rigid_ctx *ctx, spg_rigid_stats *stats);

/* Capture each step's launches in a HIP graph (default on); replays while dt, substeps, iterations and gravity
   stay the same. */
int spg_rigid_set_graph(spg_rigid_ctx *ctx, int32_t enabled);

/* One substep run eagerly with HIP events around every colour launch of the descent (on the context's stream):
   *solve_ms is the summed duration of those launches, *launches their count. Advances the state like a step. */
int spg_rigid_profile_substep(spg_rigid_ctx *ctx, double h, int32_t iterations, const double *gravity,
                              double *solve_ms, int32_t *launches);

/* Gradient (6 s) and the per-body diagonal Hessian blocks (s x 36, row-major) of n elements of one type at
   theta = 0, with s the stencil size. com: 3 s per element, rot: 9 s per element (row-major rotation matrices),
   params: the addStencil arguments. device >= 0 evaluates on that GPU, -1 on the host (same code). */
int spg_rigid_element_derivatives(int32_t type, int64_t n, const double *com, const double *rot,
                                  const double *params, double *grad, double *hess_blocks, int32_t device);

#ifdef __cplusplus
}
#endif

#endif /* SPG_RIGID_H */
