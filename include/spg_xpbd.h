/*
 * spg_xpbd.h — C-ABI of the MI355X XPBD solver (in libspg_vbd.so, beside the VBD solver of spg_vbd.h).
 *
 * The drop-in boundary for spg's XPBD path: plain pointers and sizes, no C++ or torch types. The C++
 * host mirror (spg_amd/cpp: spg::solver::XPBDGpu, a BaseSolver) calls exactly these entry points.
 *
 * Reference interfaces replaced (paths relative to the spg repository):
 *   spg_xpbd_create      XPBD::XPBD(useParallelGaussSeidel) + computeParallelStencilGroups
 *                        (spg/include/spg/sim/solver/xpbd.h:15-30, spg/source/sim/solver/xpbd.cpp:12-15,
 *                        134-160), reading the ParticleGroup (particleGroup.h:14-60, invMasses() at
 *                        particleGroup.h:33) and the energies' public data: flatStencils()
 *                        (energy.h:102-110), effectiveCompliance() (energy.h:114) and the accessors of
 *                        the five energy headers.
 *   spg_xpbd_step        XPBD::step (xpbd.cpp:16-132) with BaseSolver's dt/substeps/gravity
 *                        (spg/include/spg/sim/solver/baseSolver.h:26-39); per stencil
 *                        StencilBlockEnergy::projectPosition (energy.h:117-153), and for springs
 *                        SpringEnergy::projectPosition (springEnergy.cpp:48-78).
 *   spg_xpbd_get_groups  the stencil groups of computeParallelStencilGroups (xpbd.cpp:147-152).
 *   spg_xpbd_color_stencils  coloring::colorStencils (spg/include/spg/utils/graphColoring.h:32,
 *                        spg/source/utils/graphColoring.cpp:21-70).
 *   spg_xpbd_set_state/get_state  ParticleGroup::positions()/velocities() (particleGroup.h:20-28).
 *
 * XPBD(false), the serial Gauss-Seidel path (xpbd.cpp:99-104), projects the stencils of an energy in
 * index order. Here it runs as levels: level(s) = 1 + the largest level of an earlier stencil sharing a
 * particle with s. Every pair of particle-sharing projections then happens in index order, which is the
 * serial result; the stencils of one level run in parallel.
 *
 * Error codes (SPG_VBD_OK, SPG_VBD_ERR_*) and energy type ids (SPG_VBD_SPRING ...) are those of
 * spg_vbd.h. A context is not thread-safe; all work goes to one HIP stream in call order; host arrays
 * are copied, so callers may free them after the call returns.
 */
#ifndef SPG_XPBD_H
#define SPG_XPBD_H

#include <stdint.h>

#include "spg_vbd.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * One energy of a ParticleGroup, prepared. Per stencil, `params` holds (FP64, row-major matrices):
 *   SPRING              2: restLengths()[i], effectiveCompliance()[i](0,0)
 *   MEMBRANE_STVK      13: inverseReferenceMats()[i] (2x2), effectiveCompliance()[i] (3x3, symmetric)
 *   DISCRETE_BENDING    2: restThetas()[i], effectiveCompliance()[i](0,0)
 *   STABLE_NEO_HOOKEAN 14: inverseReferenceMats()[i] (3x3), effectiveCompliance()[i] (2x2, symmetric),
 *                          hydrostaticCorrectionTerms()[i]
 *   SPRING_ANCHOR       4: anchors()[i] (3), effectiveCompliance()[i](0,0)
 * and the further energies (their compliance matrices are diagonal; the diagonal is passed):
 *   SPRING_CONTINUUM, SPRING_SQUARED_CONTINUUM 2: restLengths()[i], effectiveCompliance()[i](0,0)
 *   MEMBRANE_BARAFF_WITKIN 7: inverseReferenceMats()[i] (2x2), effectiveCompliance()[i] diagonal (3)
 *   MEMBRANE_CHOI      8: inverseReferenceMats()[i] (2x2), effectiveCompliance()[i] diagonal (4)
 *   QUADRATIC_BENDING  5: laplacians()[i] coefficients c0..c3, effectiveCompliance()[i](0,0) (it is c I)
 *   BARAFF_WITKIN_BENDING 2: restThetas()[i], effectiveCompliance()[i](0,0)
 *   STVK, STABLE_SQUARED_NEO_HOOKEAN 11: inverseReferenceMats()[i] (3x3), effectiveCompliance()[i] diagonal (2)
 * The further energies are projected by the generic coupled projection (closed-form constraint gradients,
 * a Cholesky solve of the nC x nC system), FP32 with the same FP64 recomputation rule.
 * `stencils` is Energy::flatStencils(). A stencil may not name one particle twice.
 */
typedef struct {
    int32_t type;
    int64_t n_stencils;
    const int32_t *stencils;
    const double *params;
} spg_xpbd_energy;

/* One ParticleGroup: positions, velocities (3 per particle, AoS), inverse masses (ParticleGroup::invMasses(),
   0 for particles that projections never move; gravity still moves them, xpbd.cpp:58-66), energies in
   addEnergy order. */
typedef struct {
    int64_t n_particles;
    const double *x;
    const double *v; /* may be NULL: zero velocities */
    const double *inv_mass;
    int32_t n_energies;
    const spg_xpbd_energy *energies;
} spg_xpbd_object;

typedef struct {
    int32_t block_size; /* threads per workgroup of the projection kernels: 64, 128 or 256 */
    int32_t use_graph;  /* 1: capture a whole step in a hipGraph and replay it */
    int32_t group_runs; /* 1 (default): consecutive stencil groups of at most 1024 stencils (the tail colours)
                           run in one workgroup, in order with a barrier between them, as one launch; 0: one
                           launch per group. Same result either way. */
} spg_xpbd_launch;

typedef struct {
    int64_t projection_launches; /* projection kernel launches timed */
    double projection_ms;        /* summed duration of those launches (HIP events on the ctx stream) */
    int64_t stencil_projections; /* stencils projected by those launches */
    double algorithmic_bytes;    /* compulsory bytes of those launches: per stencil its record (spring 16,
                                    membrane 56, bending 24, SNH 68, anchor 20 B, further energies 64 B:
                                    4 slots + 12 parameters) + 32 B per particle
                                    (float4 read and written) */
    int64_t fp64_solves;         /* projections recomputed in FP64 because A was ill-conditioned */
    /* the same launches split by energy type (index = SPG_VBD_* type id); entry 0: launches that run several
       consecutive small stencil groups (at most 1024 stencils each, of any type) in one workgroup */
    int64_t type_launches[14];
    double type_ms[14];
    double type_bytes[14];
} spg_xpbd_stats;

typedef struct spg_xpbd_ctx spg_xpbd_ctx;

/* parallel_gauss_seidel: 1 = XPBD(true) (stencil colour groups), 0 = XPBD(false) (serial order). */
int spg_xpbd_create(spg_xpbd_ctx **ctx, const spg_xpbd_object *objects, int32_t n_objects,
                    int32_t parallel_gauss_seidel, int32_t device);
void spg_xpbd_destroy(spg_xpbd_ctx *ctx);
const char *spg_xpbd_last_error(const spg_xpbd_ctx *ctx);

int spg_xpbd_num_objects(const spg_xpbd_ctx *ctx, int32_t *n);
int spg_xpbd_num_particles(const spg_xpbd_ctx *ctx, int32_t object, int64_t *n);
/* Stencil groups of one energy: colours (parallel) or serial levels; group of every stencil. */
int spg_xpbd_num_groups(const spg_xpbd_ctx *ctx, int32_t object, int32_t energy, int32_t *n);
int spg_xpbd_get_groups(const spg_xpbd_ctx *ctx, int32_t object, int32_t energy, int32_t *group_of_stencil);
/* Projection kernel launches per substep (all objects together), with the launch shape's group runs. */
int spg_xpbd_num_launches(const spg_xpbd_ctx *ctx, int32_t *n);

int spg_xpbd_set_state(spg_xpbd_ctx *ctx, int32_t object, const double *x, const double *v);
int spg_xpbd_get_state(spg_xpbd_ctx *ctx, int32_t object, double *x, double *v);

/* XPBD::step: nsubsteps substeps of h = dt / nsubsteps, one projection of every stencil per substep. */
int spg_xpbd_step(spg_xpbd_ctx *ctx, double dt, int32_t nsubsteps, const double gravity[3]);

/* Device checkpoint of positions and velocities; restore-then-step reproduces the saved step bit for bit. */
int spg_xpbd_save_state(spg_xpbd_ctx *ctx);
int spg_xpbd_restore_state(spg_xpbd_ctx *ctx);

int spg_xpbd_get_launch(const spg_xpbd_ctx *ctx, spg_xpbd_launch *cfg);
int spg_xpbd_set_launch(spg_xpbd_ctx *ctx, const spg_xpbd_launch *cfg);
int spg_xpbd_get_stats(const spg_xpbd_ctx *ctx, spg_xpbd_stats *stats);
int spg_xpbd_reset_stats(spg_xpbd_ctx *ctx);
/* 1: time every projection launch with HIP events (default 0). */
int spg_xpbd_set_timing(spg_xpbd_ctx *ctx, int32_t enabled);
int spg_xpbd_synchronize(spg_xpbd_ctx *ctx);

/* coloring::colorStencils of one flat stencil array (n_stencils * stencil_size particle indices). */
int spg_xpbd_color_stencils(int64_t n_stencils, int32_t stencil_size, const int32_t *flat_stencils, int32_t *colors);

/* Check hook (not used by the solver): n independent projections of one type, each from its stencil's
   positions xs (n * size * 3), inverse masses (n * size) and params (spg_xpbd_energy layout), with
   alpha~ = compliance / dt^2, by the device functions of the projection kernels in FP32 (precision 0,
   with their FP64 fallback) or FP64 (1), on GPU `device`, or on the host when device is -1.
   dx: n * size * 3 displacements. */
int spg_xpbd_project_stencils(int32_t type, int64_t n, const double *xs, const double *inv_masses,
                              const double *params, double dt, int32_t precision, int32_t device, double *dx);

#ifdef __cplusplus
}
#endif

#endif /* SPG_XPBD_H */
