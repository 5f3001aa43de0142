/*
 * spg_vbd.h — C-ABI of the MI355X Vertex Block Descent solver (libspg_vbd.so).
 *
 * This is the drop-in boundary for spg's VBD path. Plain pointers and sizes, no C++ or torch types.
 * The C++ host mirror (spg_amd/cpp: spg::solver::VBDGpu, a BaseSolver) calls exactly these entry
 * points; other languages bind them directly (ctypes, cgo, JNI).
 *
 * Reference interfaces replaced (paths relative to the spg repository):
 *   spg_vbd_create       VBD::step's first-call precomputation: computeParallelVertexGroups +
 *                        computeStencilInfoPerVertex (spg/source/sim/solver/vbd.cpp:14-18,203-268),
 *                        reading the ParticleGroup (spg/include/spg/sim/simObject/particleGroup.h:14-60)
 *                        and the energies' public data (spg/include/spg/sim/energy/energy.h:99-115 and
 *                        the accessors of the five energy headers).
 *   spg_vbd_step         VBD::step (vbd.cpp:12-195) with BaseSolver's dt/substeps/gravity
 *                        (spg/include/spg/sim/solver/baseSolver.h:26-39) and the iteration count that
 *                        the reference keeps in VBD::m_iterations (vbd.h:33).
 *   spg_vbd_set_state/get_state  ParticleGroup::positions()/velocities() and setPositions/setVelocities
 *                        (particleGroup.h:20-28).
 *   spg_vbd_get_colors   the colour groups of computeParallelVertexGroups (vbd.cpp:203-229).
 *   spg_vbd_color_vertices  coloring::colorVertices (spg/include/spg/utils/graphColoring.h:34,
 *                        spg/source/utils/graphColoring.cpp:21-49,72-84).
 *   spg_vbd_begin_substep / iterate / end_substep  the three phases inside one substep of VBD::step
 *                        (vbd.cpp:48-83, 87-171, 173-185), exposed separately so a caller can start the
 *                        descent from its own initial guess (VBD::setInitialGuessType, vbd.h:22,25).
 *   spg_vbd_residual     no counterpart in VBD; the force residual of ImplicitEulerBase::getSystemForce
 *                        (spg/source/sim/solver/implicitEulerBase.cpp:102-143) in the x-form of
 *                        implicitEulerNewtonRobust.cpp:65-66, evaluated in FP64 on the GPU.
 *
 * Error behaviour: every entry point returns SPG_VBD_OK (0) or a nonzero code; spg_vbd_last_error
 * gives the message (per context, or the last context-free failure on this thread when ctx is NULL).
 * Setup checks reproduce the reference's exceptions as messages (e.g. "Zero/negative rest length
 * invalid"). A context is not thread-safe; all work goes to one HIP stream in call order; host
 * arrays are copied, so callers may free them after the call returns.
 */
#ifndef SPG_VBD_H
#define SPG_VBD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPG_VBD_OK 0
#define SPG_VBD_ERR_INVALID 1     /* bad argument or invalid scene data */
#define SPG_VBD_ERR_HIP 2         /* HIP runtime error */
#define SPG_VBD_ERR_UNSUPPORTED 3 /* energy type or layout this solver does not handle */
#define SPG_VBD_ERR_STATE 4       /* call out of order (e.g. iterate without begin_substep) */
#define SPG_VBD_ERR_NONFINITE 5   /* non-finite input state */

/* Energy types: the five spg energies of the VBD hot path. */
#define SPG_VBD_SPRING 1              /* SpringEnergy              stencil 2 */
#define SPG_VBD_MEMBRANE_STVK 2       /* MembraneStvkEnergy        stencil 3 */
#define SPG_VBD_DISCRETE_BENDING 3    /* DiscreteBendingEnergy     stencil 4 */
#define SPG_VBD_STABLE_NEO_HOOKEAN 4  /* StableNeoHookeanEnergy    stencil 4 */
#define SPG_VBD_SPRING_ANCHOR 5       /* SpringAnchorEnergy        stencil 1 */
#define SPG_VBD_SPRING_CONTINUUM 6            /* SpringContinuumEnergy            stencil 2 */
#define SPG_VBD_SPRING_SQUARED_CONTINUUM 7    /* SpringSquaredContinuumEnergy     stencil 2 */
#define SPG_VBD_MEMBRANE_BARAFF_WITKIN 8      /* MembraneBaraffWitkinEnergy       stencil 3 */
#define SPG_VBD_MEMBRANE_CHOI 9               /* MembraneChoiEnergy               stencil 3 */
#define SPG_VBD_QUADRATIC_BENDING 10          /* QuadraticBendingEnergy           stencil 4 */
#define SPG_VBD_BARAFF_WITKIN_BENDING 11      /* BaraffWitkinBendingEnergy        stencil 4 */
#define SPG_VBD_STVK 12                       /* StvkEnergy (tetrahedron)         stencil 4 */
#define SPG_VBD_STABLE_SQUARED_NEO_HOOKEAN 13 /* StableSquaredNeoHookeanEnergy    stencil 4 */

/*
 * One energy of a ParticleGroup, as prepared by its preparePrecomputations. Per stencil, `params`
 * holds (FP64, row-major matrices):
 *   SPRING             2: restLengths()[i], effectiveStiffness()[i](0,0)
 *   MEMBRANE_STVK     13: inverseReferenceMats()[i] (2x2), effectiveStiffness()[i] (3x3)
 *   DISCRETE_BENDING   2: restThetas()[i], effectiveStiffness()[i](0,0)
 *   STABLE_NEO_HOOKEAN 12: inverseReferenceMats()[i] (3x3), effectiveStiffness()[i](0,0) (lambda),
 *                         effectiveStiffness()[i](1,1) (mu), hydrostaticCorrectionTerms()[i]
 *   SPRING_ANCHOR      4: anchors()[i] (3), modelStiffness()[i](0,0)
 *   SPRING_CONTINUUM, SPRING_SQUARED_CONTINUUM
 *                      2: restLengths()[i], effectiveStiffness()[i](0,0) (stiffness x rest length)
 *   MEMBRANE_BARAFF_WITKIN 7: inverseReferenceMats()[i] (2x2), effectiveStiffness()[i] diagonal (3)
 *   MEMBRANE_CHOI      8: inverseReferenceMats()[i] (2x2), effectiveStiffness()[i] diagonal (4)
 *   QUADRATIC_BENDING  5: laplacians()[i] as its four coefficients c0..c3 (the 3x12 matrix is
 *                         [c0 I, c1 I, c2 I, c3 I]), effectiveStiffness()[i](0,0)
 *   BARAFF_WITKIN_BENDING 2: restThetas()[i], effectiveStiffness()[i](0,0)
 *   STVK, STABLE_SQUARED_NEO_HOOKEAN
 *                     11: inverseReferenceMats()[i] (3x3), effectiveStiffness()[i](0,0) (lambda),
 *                         effectiveStiffness()[i](1,1) (mu)
 * `stencils` is Energy::flatStencils() (n_stencils * stencil size particle indices).
 */
typedef struct {
    int32_t type;
    int64_t n_stencils;
    const int32_t *stencils;
    const double *params;
} spg_vbd_energy;

/* One ParticleGroup: positions, velocities (3 per particle, AoS) and masses, energies in addEnergy order,
   and optionally the rest positions (ParticleGroup::positions0(); used only by the learned warm start's
   strain features, zero strain features when NULL). */
typedef struct {
    int64_t n_particles;
    const double *x;
    const double *v; /* may be NULL: zero velocities */
    const double *mass;
    int32_t n_energies;
    const spg_vbd_energy *energies;
    const double *x0; /* may be NULL */
} spg_vbd_object;

/* Launch shape of the per-colour vertex solve (chosen by spg_vbd_autotune or set by hand). */
typedef struct {
    int32_t lanes_per_vertex;    /* 1, 2, 3, 4, 6, 8 or 16 lanes of a wavefront cooperate on one vertex */
    int32_t block_size;          /* threads per workgroup: 64, 128 or 256 */
    int32_t use_graph;           /* 1: capture a whole step in a hipGraph and replay it */
    int32_t lds_staging;         /* 1: each workgroup first copies the element-record rows of its vertices
                                    into LDS (coalesced), then reads them from there; available while
                                    they fit 64 KB.
                                    2: each vertex first gathers its own and its distinct neighbours'
                                    positions into LDS once, and its element records index them there;
                                    available when no particle has more than 15 distinct neighbours and
                                    the objects have no further energies (types 6 ... 13) */
} spg_vbd_launch;

/* Counters and HIP-event timings kept by a context since creation (or the last reset_stats). */
typedef struct {
    int64_t solve_launches;      /* per-colour vertex-solve kernel launches timed */
    double solve_ms;             /* summed duration of those launches (HIP events on the ctx stream) */
    int64_t vertex_updates;      /* vertex block updates performed by those launches */
    double algorithmic_bytes;    /* compulsory bytes of those launches: 12 B x every vertex of the objects
                                    with that colour + 32 B x updated vertices + per incidence 4 B plus the
                                    element record (triangle 40, hinge 24, tet 64, spring 16, anchor 20 B;
                                    a further-energy record 16 B of slots and 4 B per FP32 coefficient read) */
    int64_t fp64_fallbacks;      /* block solves that took the FP64 near-singular path */
    int64_t skipped_blocks;      /* block solves skipped by det(H) <= 1e-10 */
    int64_t prediction_fallbacks; /* objects whose learned start (spg_vbd_learned.h) was dropped during
                                    spg_vbd_iterate_to_tolerance because their residual rose above the
                                    inertial guess's or became non-finite */
} spg_vbd_stats;

typedef struct spg_vbd_ctx spg_vbd_ctx;

/* Create a solver for n_objects independent ParticleGroups (colouring and incidence per object, as
   the reference builds them per object). All objects advance together in every launch. */
int spg_vbd_create(spg_vbd_ctx **ctx, const spg_vbd_object *objects, int32_t n_objects, int32_t device);
void spg_vbd_destroy(spg_vbd_ctx *ctx);
const char *spg_vbd_last_error(const spg_vbd_ctx *ctx);

int spg_vbd_num_objects(const spg_vbd_ctx *ctx, int32_t *n);
int spg_vbd_num_particles(const spg_vbd_ctx *ctx, int32_t object, int64_t *n);
int spg_vbd_num_colors(const spg_vbd_ctx *ctx, int32_t object, int32_t *n);
int spg_vbd_get_colors(const spg_vbd_ctx *ctx, int32_t object, int32_t *colors);

int spg_vbd_set_state(spg_vbd_ctx *ctx, int32_t object, const double *x, const double *v);
int spg_vbd_get_state(spg_vbd_ctx *ctx, int32_t object, double *x, double *v);

/* VBD::step: nsubsteps substeps of h = dt / nsubsteps, each with `iterations` colour sweeps. */
int spg_vbd_step(spg_vbd_ctx *ctx, double dt, int32_t nsubsteps, int32_t iterations, const double gravity[3]);

/* The phases of one substep. begin: x_old = x, v += h g, x += h v, x_tilde = x. iterate: colour
   sweeps. end: v = (x - x_old) / h. set_guess replaces x (all objects, host AoS FP64) between begin
   and iterate. */
int spg_vbd_begin_substep(spg_vbd_ctx *ctx, double h, const double gravity[3]);
int spg_vbd_set_guess(spg_vbd_ctx *ctx, int32_t object, const double *x);
int spg_vbd_iterate(spg_vbd_ctx *ctx, int32_t iterations);
int spg_vbd_end_substep(spg_vbd_ctx *ctx);

/* Checkpoint / resume on the device (VBD is Markov in (x, v), particleGroup.cpp:97-103 restores the
   initial state the same way): save copies the current positions and velocities of all objects into
   a device snapshot, restore copies them back. Both are enqueued on the context's stream (no host
   synchronisation) and restore-then-step reproduces the saved step bit for bit. */
int spg_vbd_save_state(spg_vbd_ctx *ctx);
int spg_vbd_restore_state(spg_vbd_ctx *ctx);

/* FP64 force residual of the current substep: ||-grad E(x) - (M/h^2)(x - x_tilde)||_2 per object
   (norms has n_objects entries), computed on the GPU in FP64. Valid between begin and end. */
int spg_vbd_residual(spg_vbd_ctx *ctx, double *norms);
/* spg_vbd_residual without device allocation, for callers that cannot allocate (serving loops, harnesses that
   forbid hipMalloc). No counterpart in the reference. The residual's work memory (per-particle squared
   residuals, per-hinge angle terms, per-segment partials, per-object norms) is a static device array of
   24.06 MiB per GPU, held until the process ends, instead of buffers of the context; the call allocates
   nothing. Limits: at most 1,048,576 particles, 2,097,152 bending hinges, 4096 objects and 4096 reduction
   segments per context (SPG_VBD_ERR_UNSUPPORTED beyond). Same
   norms as spg_vbd_residual, bit for bit. Every call on one GPU shares that array: calls that overlap on one
   GPU (from several threads) race; calls in sequence do not, each call synchronises its stream. */
int spg_vbd_residual_noalloc(spg_vbd_ctx *ctx, double *norms);
/* Same, at x = x_tilde (the inertial guess), for relative residuals. */
int spg_vbd_residual_at_inertial(spg_vbd_ctx *ctx, double *norms);

/* Colour sweeps of the open substep until every object's relative FP64 residual
   rho = ||r(x)|| / ||r(x_tilde)|| is <= rel_tol, checking after every `check_every` sweeps, at most
   max_iterations sweeps. Reports the sweeps done and the final max rho over objects. */
int spg_vbd_iterate_to_tolerance(spg_vbd_ctx *ctx, double rel_tol, int32_t max_iterations, int32_t check_every,
                                 int32_t *iterations, double *rho);

/* spg_vbd_iterate_to_tolerance: iterates (check_every sweeps between FP64 residual checks) until the
   largest per-object |r(x)| / |r(x~)| is at most rel_tol or max_iterations sweeps are done; *rho is that
   ratio (NaN when a residual is not finite). An object that started from a learned prediction whose
   residual rises above |r(x~)| goes back to x~ and continues as plain VBD. */

/* Launch configuration and autotuning. */
int spg_vbd_get_launch(const spg_vbd_ctx *ctx, spg_vbd_launch *cfg);
int spg_vbd_set_launch(spg_vbd_ctx *ctx, const spg_vbd_launch *cfg);
/* Times the candidate shapes (lanes_per_vertex x block_size x lds_staging, where the staged rows fit)
   on this context's data (state is restored afterwards) and keeps the fastest. trials: sweeps per timed
   run; each candidate is timed twice and its faster run counts. */
int spg_vbd_autotune(spg_vbd_ctx *ctx, int32_t trials, spg_vbd_launch *best);

int spg_vbd_get_stats(const spg_vbd_ctx *ctx, spg_vbd_stats *stats);
int spg_vbd_reset_stats(spg_vbd_ctx *ctx);
/* 1: time every vertex-solve launch with HIP events (default 0; adds an event pair per launch). */
int spg_vbd_set_timing(spg_vbd_ctx *ctx, int32_t enabled);
int spg_vbd_synchronize(spg_vbd_ctx *ctx);

/* coloring::colorVertices: greedy colouring over cliques of every stencil (smallest colour unused by
   a lower-index neighbour). n_sets flat-stencil arrays of lengths[i] entries with sizes[i] per stencil. */
int spg_vbd_color_vertices(int64_t n_vertices, int32_t n_sets, const int32_t *const *flat_stencils,
                           const int64_t *lengths, const int32_t *stencil_sizes, int32_t *colors);

/* Time of one FP64 force residual evaluation at the current positions (all its kernels: the per-hinge angle
   pass, the per-vertex residual and the two reduction stages), the mean over `repeats` evaluations queued back to
   back on the context's stream, from HIP events, in *us; *bytes (may be NULL) receives its compulsory bytes: per
   particle position, x~, FP64 mass and the squared residual written and read back (56 B), each element's table
   entry once (spring 24, anchor 36, triangle 80, hinge 48 with its angle term written and read, tetrahedron
   112, further element 16 + 8 per parameter) and 4 B per incidence. Needs an open substep; changes no state.
   Measurement only, no counterpart in the reference. */
int spg_vbd_bench_residual(spg_vbd_ctx *ctx, int32_t repeats, double *us, double *bytes);

/* Plain float4 device-to-device copy of `bytes` (multiple of 16) repeated `repeats` times, for
   calibrating HBM traffic counters against a kernel of known traffic (bytes read + bytes written per
   launch). *ms receives the mean launch time from HIP events. Not part of the solver. */
int spg_vbd_bench_copy(int64_t bytes, int32_t repeats, int32_t device, double *ms);

/* Cost of a barrier between colours inside one persistent (cooperative) launch against back-to-back launches,
   the measurement behind one launch per colour. `rounds` barriers of `grid` blocks of `block` threads, counted in
   groups of `group` consecutive blocks (group == grid: one grid-wide barrier; a smaller group: independent
   barriers per group, as per-instance barriers of a batch would be), each arrival a release add and each wait a
   bounded spin on relaxed loads followed by an acquire fence. *us_per_barrier: mean per round (-1 when a spin hit
   its bound); *us_per_launch: mean of `rounds` back-to-back launches of an empty kernel of the same shape.
   SPG_VBD_ERR_HIP when the runtime refuses the cooperative launch (blocks not co-resident). Not part of the
   solver. */
int spg_vbd_bench_barrier(int32_t grid, int32_t block, int32_t group, int32_t rounds, int32_t device,
                          double *us_per_barrier, double *us_per_launch);

/* Per-element derivatives (check hook for the element closed forms; not used by the solver): for each
   of n stencils of one type (positions xs: n * size * 3, prepared params: n * nparams), the gradient
   w.r.t. each stencil vertex (n*size*3) and its diagonal 3x3 Hessian block (n*size*9), evaluated in
   FP32 (precision 0) or FP64 (1) by the same device functions the solve uses, on GPU `device`, or
   on the host when device is -1. form 0 selects the closed forms of the solve; forms 1 (general
   (a, b, sigma) hinge form) and 2 (single-expression compact hinge form) are independent derivations
   of the hinge derivatives, for cross-checking. The further energies (types 6 ... 13) exist in the FP32
   solve only: precision 0, form 0, device >= 0 (SPG_VBD_ERR_UNSUPPORTED otherwise). */
int spg_vbd_element_derivatives(int32_t type, int64_t n, const double *xs, const double *params,
                                int32_t precision, int32_t form, int32_t device, double *grad,
                                double *hess_blocks);

#ifdef __cplusplus
}
#endif

#endif /* SPG_VBD_H */
