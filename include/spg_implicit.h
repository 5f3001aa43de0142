/*
 * spg_implicit.h — C-ABI of the MI355X implicit Euler solvers (in libspg_vbd.so).
 *
 * The drop-in boundary for spg's implicit solver family: the global force vector, mass and stiffness
 * matrices and the conjugate gradient of ImplicitEulerBase, and the steps built on them. All FP64.
 *
 * Reference interfaces replaced (paths relative to the spg repository):
 *   spg_implicit_create          ImplicitEulerBase-derived solvers (spg/include/spg/sim/solver/implicitEulerBase.h:9-29)
 *                                over particle groups: ImplicitEulerBaraffWitkin, ImplicitEulerNewtonDv,
 *                                ImplicitEulerNewtonDx, ImplicitEulerNewtonRobust (implicitEulerNewtonRobust.h:14-26),
 *                                BDF2 (bdf2.h:12-20). The objects are those of spg_vbd_create (prepared energies
 *                                in the stiffness form of spg_vbd_energy).
 *   spg_implicit_step            <solver>::step (implicitEulerBaraffWitkin.cpp:11-59, implicitEulerNewtonDv.cpp:11-67,
 *                                implicitEulerNewtonDx.cpp:11-64, implicitEulerNewtonRobust.cpp:11-127,
 *                                bdf2.cpp:10-76) with BaseSolver's dt/substeps/gravity.
 *   spg_implicit_reset           <solver>::reset (NewtonRobust's dt, BDF2's history; implicitEulerNewtonRobust.cpp:129-134,
 *                                bdf2.cpp:78-83).
 *   spg_implicit_system_force    ImplicitEulerBase::getSystemForce (implicitEulerBase.cpp:102-143).
 *   spg_implicit_stiffness       getSystemStiffnessMatrix (implicitEulerBase.cpp:180-222), as 3x3 blocks.
 *   spg_implicit_cg_solve        solveLinearSystem (implicitEulerBase.cpp:224-239): Eigen's ConjugateGradient with
 *                                Lower|Upper and its diagonal preconditioner, on alpha M + beta K.
 *
 * DOFs are 3 per particle, particle-major, objects concatenated (one global system, as the reference
 * assembles all objects into one matrix). Error codes are those of spg_vbd.h. A context is not
 * thread-safe; all work goes to one HIP stream in call order.
 */
#ifndef SPG_IMPLICIT_H
#define SPG_IMPLICIT_H

#include <stdint.h>

#include "spg_vbd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SPG_IMPLICIT_BARAFF_WITKIN 0
#define SPG_IMPLICIT_NEWTON_DV 1
#define SPG_IMPLICIT_NEWTON_DX 2
#define SPG_IMPLICIT_NEWTON_ROBUST 3
#define SPG_IMPLICIT_BDF2 4

/* Linear solver settings (ImplicitEulerBase::m_cgThreshold / m_cgMaxIterations, implicitEulerBase.h:26-27). */
#define SPG_IMPLICIT_CG_TOLERANCE 1e-5
#define SPG_IMPLICIT_CG_MAX_ITERATIONS 300

typedef struct {
    int64_t steps;              /* spg_implicit_step calls */
    int64_t linear_solves;      /* conjugate-gradient solves */
    int64_t cg_iterations;      /* their iterations (Eigen's count) */
    int64_t newton_iterations;  /* NewtonRobust iterations */
    int64_t failed_substeps;    /* NewtonRobust substeps that were rejected (dt reduced) */
    double current_dt;          /* NewtonRobust's current dt (-1 before its first step) */
    double last_cg_error;       /* relative residual of the last solve */
    /* HIP-event timings when timing is on (spg_implicit_set_timing) */
    double assembly_ms;         /* element derivatives + force + matrix assembly */
    double cg_ms;               /* conjugate-gradient solves */
    int64_t spmv_launches;      /* SpMV launches inside timed solves */
    double spmv_bytes;          /* compulsory bytes of those SpMVs (the system matrices are stored by their
                                   upper 3x3 blocks): 72 B + 4 B column per upper block, 4 B index + 4 B
                                   column per lower block, 24 B of x and 24 B of y per block row, 16 B of
                                   row pointers per block row */
    double spmv_sampled_ms;     /* HIP-event time of the sampled SpMVs inside timed solves (the first of
                                   every chunk of 8 iterations) */
    int64_t spmv_sampled;       /* how many were sampled */
} spg_implicit_stats;

typedef struct spg_implicit_ctx spg_implicit_ctx;

int spg_implicit_create(spg_implicit_ctx **ctx, const spg_vbd_object *objects, int32_t n_objects, int32_t solver,
                        int32_t device);
void spg_implicit_destroy(spg_implicit_ctx *ctx);
const char *spg_implicit_last_error(const spg_implicit_ctx *ctx);

int spg_implicit_num_dofs(const spg_implicit_ctx *ctx, int64_t *n);
/* number of 3x3 blocks of the system matrices (the stiffness pattern: particles sharing a stencil) */
int spg_implicit_num_blocks(const spg_implicit_ctx *ctx, int64_t *n);

int spg_implicit_set_state(spg_implicit_ctx *ctx, int32_t object, const double *x, const double *v);
int spg_implicit_get_state(spg_implicit_ctx *ctx, int32_t object, double *x, double *v);

int spg_implicit_step(spg_implicit_ctx *ctx, double dt, int32_t nsubsteps, const double gravity[3]);
int spg_implicit_reset(spg_implicit_ctx *ctx);

/* Global force at the current positions: 3 * (all particles) doubles. */
int spg_implicit_system_force(spg_implicit_ctx *ctx, const double gravity[3], double *f);
/* K = -Hessian at the current positions as block rows: row_ptr (particles + 1), cols (blocks), blocks
   (9 doubles each, row-major). Any output may be NULL. */
int spg_implicit_stiffness(spg_implicit_ctx *ctx, int64_t *row_ptr, int32_t *cols, double *blocks);
/* Solve (alpha M + beta K) x = b at the current positions with the conjugate gradient (zero initial guess). */
int spg_implicit_cg_solve(spg_implicit_ctx *ctx, double alpha, double beta, const double *b, double tolerance,
                          int32_t max_iterations, double *x, int32_t *iterations, double *error);

/* Measurement hook (not used by the solvers): assembles alpha M + beta K at the current positions and
   times `repeats` back-to-back block-sparse products y = A x with HIP events on the context's stream;
   *ms receives the mean time of one product. */
int spg_implicit_bench_spmv(spg_implicit_ctx *ctx, double alpha, double beta, int32_t repeats, double *ms);

/* Test and measurement hook (no context): gradient (n x 3 size) and Hessian (n x (3 size)^2, row-major) of n
   elements of one energy type at the positions xs (n x 3 size) with their prepared parameters (the layout of
   spg_vbd_energy), as energyGradient / energyHessian (spg/include/spg/sim/energy/energy.h:268-270) give them.
   device >= 0 runs the GPU kernel the solvers use; device -1 evaluates the same closed forms on the host. */
int spg_implicit_element_derivatives(int32_t type, int64_t n, const double *xs, const double *params, double *grad,
                                     double *hess, int32_t device);

int spg_implicit_get_stats(const spg_implicit_ctx *ctx, spg_implicit_stats *stats);
int spg_implicit_reset_stats(spg_implicit_ctx *ctx);
int spg_implicit_set_timing(spg_implicit_ctx *ctx, int32_t enabled);
int spg_implicit_synchronize(spg_implicit_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* SPG_IMPLICIT_H */
