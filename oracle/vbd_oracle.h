/*
 * FP64 CPU oracle for the spg VBD, XPBD and implicit Euler steps — TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker, not the product. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. The product (libspg_vbd.so / libspg_host.so)
 * never links, loads or calls it.
 *
 * It restates, in plain FP64 C++ with OpenMP, the reference algorithm of
 *   spg/source/sim/solver/vbd.cpp:12-268            (VBD::step, colouring groups, incidence lists)
 *   spg/include/spg/sim/energy/energy.h:240-270,336-362 (per-vertex force/Hessian slicing, TinyAD path)
 *   spg/source/sim/energy/{spring,membraneStvk,discreteBending,stableNeoHookean,springAnchor}Energy.cpp
 *   spg/source/sim/simObject/particleGroup.cpp:57-82,97-103
 *   spg/source/utils/graphColoring.cpp:21-49,72-84
 *   spg/include/spg/geom/triangleMesh.h:37-50       (edge map order, opposing vertices)
 *   spg/source/sim/solver/implicitEulerBase.cpp:102-143 (system force, residual convention)
 * and the XPBD step that shares the same element data:
 *   spg/source/sim/solver/xpbd.cpp:16-160            (XPBD::step, computeParallelStencilGroups)
 *   spg/include/spg/sim/energy/energy.h:117-153       (StencilBlockEnergy::projectPosition)
 *   spg/source/sim/energy/springEnergy.cpp:48-78       (the spring's own projectPosition)
 *   the dConstraints lambdas of the five energy sources (constraint gradients by the same dual numbers)
 *   spg/source/utils/graphColoring.cpp:21-70          (colorStencils + greedyColoring)
 * and the implicit Euler family on the same energies:
 *   spg/source/sim/solver/implicitEulerBase.cpp:102-240 (system force, mass and stiffness matrices, CG)
 *   spg/source/sim/solver/implicitEuler{BaraffWitkin,NewtonDv,NewtonDx,NewtonRobust}.cpp, bdf2.cpp
 *   Eigen's ConjugateGradient with DiagonalPreconditioner (an empty submodule of the reference; its
 *   published algorithm is restated: zero initial guess, threshold max(tol^2 |b|^2, DBL_MIN))
 * and VBD on rigid bodies (6-DOF blocks):
 *   spg/source/sim/solver/vbd.cpp:62-72,105-145,173-185 (the RigidBodyGroup branches)
 *   spg/source/sim/simObject/rigidBodyGroup.cpp:6-243   (axis-angle maps, integration, velocities, inertia)
 *   spg/source/sim/energy/rigidBody{Spring,SpringSquared,SpringAnchor,OrientationAlignment}Energy.cpp
 *   (the rotation maps restate Eigen's AngleAxis arithmetic: Rodrigues, and the quaternion of R by the
 *   largest-of-trace-and-diagonal rule, Eigen being an empty submodule of the reference)
 * with two derivative back-ends: closed-form per-vertex blocks, and the oracle's own
 * forward-mode second-order dual numbers that evaluate the full element derivative twice per
 * incidence, as the reference does through TinyAD.
 *
 * Raw per-stencil parameters (what each energy's addStencil receives):
 *   ORC_SPRING           [restLength, stiffness]                 stencil size 2
 *   ORC_MEMBRANE_STVK    [young, poisson]                        stencil size 3
 *   ORC_DISCRETE_BENDING [restTheta, stiffness]                  stencil size 4
 *   ORC_STABLE_NEO_HOOKEAN [young, poisson]                      stencil size 4
 *   ORC_SPRING_ANCHOR    [ax, ay, az, stiffness]                 stencil size 1
 *   ORC_SPRING_CONTINUUM [restLength, stiffness]                 stencil size 2
 *   ORC_SPRING_SQUARED_CONTINUUM [restLength, stiffness]         stencil size 2
 *   ORC_MEMBRANE_BARAFF_WITKIN [kWeft, kWarp, kBias]             stencil size 3
 *   ORC_MEMBRANE_CHOI    [kWeft, kWarp, kBias]                   stencil size 3
 *   ORC_QUADRATIC_BENDING [restTheta (must be 0), stiffness]     stencil size 4
 *   ORC_BARAFF_WITKIN_BENDING [restTheta, stiffness]             stencil size 4
 *   ORC_STVK             [young, poisson]                        stencil size 4
 *   ORC_STABLE_SQUARED_NEO_HOOKEAN [young, poisson]              stencil size 4
 * Prepared per-stencil parameters (what the energy accessors expose after preparePrecomputations):
 *   SPRING   [restLength, k]                       (restLengths(), effectiveStiffness())
 *   STVK     [Dinv 2x2 row-major (4), Keff 3x3 row-major (9)]  (inverseReferenceMats(), effectiveStiffness())
 *   BENDING  [restTheta, keff]                     (restThetas(), effectiveStiffness())
 *   SNH      [Dinv 3x3 row-major (9), lambda, mu, alpha]       (inverseReferenceMats(), effectiveStiffness(),
 *                                                   hydrostaticCorrectionTerms())
 *   ANCHOR   [ax, ay, az, k]                       (anchors(), modelStiffness())
 *   SPRING_CONTINUUM, SPRING_SQUARED_CONTINUUM [restLength, k * restLength]  (restLengths(), effectiveStiffness())
 *   MEMBRANE_BARAFF_WITKIN [Dinv 2x2 (4), Keff diagonal (3)]    (inverseReferenceMats(), effectiveStiffness())
 *   MEMBRANE_CHOI [Dinv 2x2 (4), Keff diagonal (4)]             (inverseReferenceMats(), effectiveStiffness())
 *   QUADRATIC_BENDING [c0, c1, c2, c3, keff]      (laplacians() = [c0 I, c1 I, c2 I, c3 I], effectiveStiffness()
 *                                                   = keff I)
 *   BARAFF_WITKIN_BENDING [restTheta, k]          (restThetas() recomputed from the rest positions,
 *                                                   effectiveStiffness())
 *   STVK, STABLE_SQUARED_NEO_HOOKEAN [Dinv 3x3 (9), lambda, mu]  (inverseReferenceMats(), effectiveStiffness()
 *                                                   diagonal, premultiplied by the rest volume)
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_SPRING = 1,
    ORC_MEMBRANE_STVK = 2,
    ORC_DISCRETE_BENDING = 3,
    ORC_STABLE_NEO_HOOKEAN = 4,
    ORC_SPRING_ANCHOR = 5,
    ORC_SPRING_CONTINUUM = 6,
    ORC_SPRING_SQUARED_CONTINUUM = 7,
    ORC_MEMBRANE_BARAFF_WITKIN = 8,
    ORC_MEMBRANE_CHOI = 9,
    ORC_QUADRATIC_BENDING = 10,
    ORC_BARAFF_WITKIN_BENDING = 11,
    ORC_STVK = 12,
    ORC_STABLE_SQUARED_NEO_HOOKEAN = 13
};
enum { ORC_MODE_CLOSED_FORM = 0, ORC_MODE_AD = 1 };

const char *orc_last_error(void);

/* object lifetime (a ParticleGroup plus its energies, plus the VBD solver state) */
void *orc_new(int64_t nv, const double *x, const double *x0, const double *m);
void orc_free(void *h);
int orc_add_energy(void *h, int type, int64_t n, const int32_t *stencils, const double *raw);
int orc_prepare(void *h);
int orc_num_prepared_params(int type);
int orc_get_prepared(void *h, int energy, double *out);
int orc_set_prepared(void *h, int energy, const double *in);
int orc_set_threads(int n);

/* solver */
int orc_colors(void *h, int32_t *out, int32_t *ncolors);
int orc_incidence_count(void *h, int64_t *count);
int orc_incidence(void *h, int32_t *vertex, int32_t *energy, int32_t *stencil, int32_t *slot);
int orc_step(void *h, double dt, int nsubsteps, int iterations, const double *gravity, int mode);
int orc_substep_curve(void *h, double dt, int iterations, const double *gravity, int mode,
                      const double *start_guess, double *rho_per_iter);
int orc_get_state(void *h, double *x, double *v);
int orc_set_state(void *h, const double *x, const double *v);
int orc_reset(void *h);

/* residual r(x) = -grad E(x) - (M/h^2)(x - xt), returns ||r||_2; r_out may be NULL */
int orc_residual(void *h, const double *x, const double *xt, double dt, int mode, double *norm, double *r_out);
int orc_energy(void *h, const double *x, double *e);

/* single element: value, full gradient (3N) and the N diagonal 3x3 Hessian blocks (N*9, row-major);
   full_hess (3N x 3N) only in AD mode, may be NULL */
int orc_element(int type, const double *xs, const double *prepared, int mode, double *energy,
                double *grad, double *diag_blocks, double *full_hess);

/* TriangleMesh restatement: bending stencils {e0, e1, opp0, opp1} for edges with two opposite vertices,
   in std::map<(min,max)> order. Call with out=NULL to get the count. */
int orc_bending_stencils(int64_t nf, const int32_t *faces, int64_t nv, int32_t *out, int64_t *count);

/* brute force colouring check: 1 if no stencil has two vertices of the same colour */
int orc_coloring_valid(void *h, const int32_t *colors);

/* ---- XPBD ----
 * Compliance per stencil (effectiveCompliance(), nC x nC row-major, nC = orc_num_constraints):
 *   SPRING, ANCHOR 1/k; STVK C^-1 / area; BENDING (1/k) / factor; SNH diag(lambda, mu)^-1 / volume;
 *   continuum springs 1/(k restLength); Baraff-Witkin and Choi membranes diag(k)^-1 / area; quadratic bending
 *   I / keff; Baraff-Witkin bending 1/k; StVK tet and stable squared NH diag(lambda, mu)^-1 / volume.
 * Computed by orc_prepare from the raw addStencil arguments. */
int orc_num_constraints(int type);
/* coloring::colorStencils on one energy's flat stencils; ncolors may be NULL */
int orc_color_stencils(int64_t n_stencils, int stencil_size, const int32_t *flat, int32_t *colors, int32_t *ncolors);
int orc_get_compliance(void *h, int energy, double *out);
/* XPBD::step: nsubsteps substeps of dt / nsubsteps; parallel != 0 is the colour-group path
   (XPBD(true)), 0 the serial index-order path (XPBD(false)). Uses and updates the object's state. */
int orc_xpbd_step(void *h, double dt, int nsubsteps, const double *gravity, int parallel);
/* constraint values c (nC) and gradient grad ((3 size) x nC row-major) by dual numbers */
int orc_constraints(int type, const double *xs, const double *prepared, double *c, double *grad);
/* one projectPosition as a displacement dx (3 size) of the stencil's vertices; generic_spring != 0
   routes a spring through the generic path instead of its own closed form */
int orc_xpbd_project(int type, const double *xs, const double *inv_masses, const double *prepared,
                     const double *compliance, double dt, int generic_spring, double *dx);

/* ---- implicit Euler family ----
 * DOFs: 3 per particle, particle-major. K = -Hessian of the energies (getSystemStiffnessMatrix), M = diag(m). */
int orc_system_force(void *h, const double *x, const double *gravity, double *f);
int orc_stiffness_csr(void *h, const double *x, int64_t *nnz, int64_t *ptr, int64_t *col, double *val);
int orc_cg_solve(void *h, const double *x, double alpha, double beta, const double *b, double tol, int max_iters,
                 double *out, int *iters, double *error);
int orc_cg_csr(int64_t n, const int64_t *ptr, const int64_t *col, const double *val, const double *b, double tol,
               int max_iters, double *out, int *iters, double *error);
void *orc_implicit_new(void);
void orc_implicit_free(void *s);
int orc_implicit_step(void *h, void *state, int solver, double dt, int nsubsteps, const double *gravity,
                      int64_t *counters, double *current_dt);

/* ---- rigid-body VBD (types 21 spring, 22 squared spring, 23 anchor, 24 orientation alignment; raw addStencil
   arguments: 8, 8, 7 and 1 doubles) ---- */
void *orc_rb_new(int64_t nb, const double *x, const double *theta, const double *m, const double *inertia);
void orc_rb_free(void *h);
int orc_rb_add_energy(void *h, int type, int64_t n, const int32_t *stencils, const double *raw);
int orc_rb_step(void *h, double dt, int nsubsteps, int iterations, const double *gravity);
int orc_rb_state(void *h, double *x, double *theta, double *v, double *omega);
int orc_rb_set_state(void *h, const double *x, const double *theta, const double *v, const double *omega);
int orc_rb_colors(void *h, int32_t *colors, int32_t *ncolors);
int orc_rb_element(int type, const double *com, const double *rot, const double *raw, double *grad, double *hess);
int orc_rb_aa_to_r(const double *theta, double *R);
int orc_rb_r_to_aa(const double *R, double *theta);

#ifdef __cplusplus
}
#endif
