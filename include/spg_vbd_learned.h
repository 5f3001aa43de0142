/* Learned warm start for the GPU VBD solver: a per-vertex MLP that predicts the substep's solution
 * x* from one-ring features of the state, used as the starting point of the vertex block descent.
 *
 * The reference has no counterpart: spg's VBD always starts from the inertial guess
 * (spg/source/sim/solver/vbd.cpp:58-83, InitialGuessType::InertialWithAcceleration in vbd.h:22,25).
 * Plain VBD stays the judge: spg_vbd_warm_start keeps the prediction of an object only when its FP64
 * force residual is no larger than the inertial guess's, and the solve then continues with
 * spg_vbd_iterate / spg_vbd_iterate_to_tolerance of spg_vbd.h exactly as from the inertial guess.
 *
 * Network: 24 inputs -> 128 -> 128 -> 128 -> 3, SiLU after each hidden layer, 36,611 parameters.
 * Inputs are standardised ((f - mean) / std, per feature) before the first layer. Weights and
 * activations are BF16 on the matrix cores with FP32 accumulation; FP32 master weights, Adam.
 *
 * Feature row of a vertex (spg_vbd_features), with s = h^2 a_ref, u = h a_ref, a_ref = |g| (1 when
 * g = 0), c = x* - x~ of the previous substep (0 after create / set_state), v the velocity before this
 * substep's gravity kick, j over the vertices sharing a triangle, tetrahedron or spring with i:
 *   0-2 c_i/s; 3-5 v_i/u; 6-8 mean_j c_j/s; 9-11 max_j |c_j - c_i|/s; 12-14 mean_j v_j/u;
 *   15-17 max_j |v_j - v_i|/u; 18 log10 Young's modulus (membrane or solid; spring stiffness);
 *   19 log10 bending stiffness; 20 anchored flag; 21-23 mean/max/min one-ring strain at x~.
 * Output y: predicted (x* - x~) / s.
 *
 * Error codes and threading are those of spg_vbd.h. A network object and the contexts it serves must
 * live on the same device.
 */
#ifndef SPG_VBD_LEARNED_H
#define SPG_VBD_LEARNED_H

#include <stdint.h>

#include "spg_vbd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SPG_MLP_FEATURES 24
#define SPG_MLP_HIDDEN 128
#define SPG_MLP_OUTPUTS 3
#define SPG_MLP_NUM_PARAMS 36611
#define SPG_MLP_MAX_GRAD_ROWS (1 << 20)

typedef struct spg_mlp spg_mlp;

/* New network on `device`: Glorot-uniform weights from splitmix64(seed), zero biases, identity
 * standardisation (mean 0, std 1). */
int spg_mlp_create(spg_mlp **net, int32_t device, uint64_t seed);
void spg_mlp_destroy(spg_mlp *net);
/* Message of the last failed call on `net` (or of the thread's last failed spg_mlp_create when NULL). */
const char *spg_mlp_last_error(const spg_mlp *net);

/* SPG_MLP_NUM_PARAMS floats: W1 (128x24 row-major, out x in), b1 (128), W2 (128x128), b2, W3 (128x128),
 * b3, W4 (3x128), b4 (3). */
int spg_mlp_get_params(spg_mlp *net, float *params);
int spg_mlp_set_params(spg_mlp *net, const float *params);
/* enabled 1: the network's output y is the change of the previous substep's offset, (x* - x~ - c) / s with c
 * the previous substep's x* - x~ (feature columns 0-2 are c / s), and spg_vbd_warm_start starts from
 * x~ + c + s y, on top of the history start of spg_vbd_history_start. 0 (the default): y is (x* - x~) / s and
 * the start is x~ + s y. A property of how the network was trained; kept with the network. */
int spg_mlp_set_history_base(spg_mlp *net, int32_t enabled);
int spg_mlp_get_history_base(spg_mlp *net, int32_t *enabled);
/* Per-feature standardisation, SPG_MLP_FEATURES floats each; std must be > 0. */
int spg_mlp_get_normalization(spg_mlp *net, float *mean, float *std_dev);
int spg_mlp_set_normalization(spg_mlp *net, const float *mean, const float *std_dev);

/* Inference on host arrays: features n x 24 -> out n x 3. */
int spg_mlp_forward(spg_mlp *net, const float *features, int64_t n, float *out);
/* Loss L = (1/n) sum_rows sum_k (y_k - t_k)^2 over one batch of n <= SPG_MLP_MAX_GRAD_ROWS rows and
 * its gradient with respect to the flat parameters (layout of spg_mlp_get_params). No update. */
int spg_mlp_loss_grad(spg_mlp *net, const float *features, const float *targets, int64_t n, double *loss,
                      float *grad);
/* Minibatch Adam (beta 0.9 / 0.999, eps 1e-8) on the loss above, learning rate cosine-decayed from lr to
 * lr / 20 over the call. Rows are shuffled per epoch with splitmix64(seed); batch must be a positive
 * multiple of 64. When standardize is non-zero the normalisation is first set to the dataset's
 * per-feature mean and standard deviation. epoch_loss (may be NULL) receives the mean loss of each
 * epoch. */
int spg_mlp_train(spg_mlp *net, const float *features, const float *targets, int64_t n, int32_t epochs,
                  int32_t batch, float lr, uint64_t seed, int32_t standardize, double *epoch_loss);

/* Feature rows of `object`'s particles (n_particles x 24, particle order) for the open substep. */
int spg_vbd_features(spg_vbd_ctx *ctx, int32_t object, float *features);
/* Replaces the open substep's start x~ by the network's prediction x~ + s y (x~ + c + s y for a network with
 * spg_mlp_set_history_base). The predicted displacement
 * y is first averaged `smoothing` times over each vertex's one ring (y_i <- (y_i + sum_j y_j) / (1 + deg),
 * 0..64 passes), which removes vertex-to-vertex noise of the prediction and keeps its smooth part.
 * gate 0: the prediction is kept. gate 1: each object keeps it only if rho = |r(pred)| / |r(x~)| <= 1
 * (FP64 residual of spg_vbd_residual) and goes back to x~ otherwise, before this call returns. gate 2
 * (deferred): the same test is the first convergence check of the next spg_vbd_iterate_to_tolerance,
 * which computes it on its check stream while the first sweeps from the prediction run; this call then
 * returns without synchronising and reports rho as NaN (rejections appear in
 * spg_vbd_stats.prediction_fallbacks). accepted and rho (may be NULL) receive one value per object. */
int spg_vbd_warm_start(spg_vbd_ctx *ctx, spg_mlp *net, int32_t smoothing, int32_t gate, int32_t *accepted,
                       double *rho);
/* The non-learned start that the network is measured against: replaces the open substep's start x~ by
 * x~ + (x* - x~) of the previous substep (the solution's offset from its inertial guess carried over, the
 * quantity of feature columns 0-2), with the gates of spg_vbd_warm_start. No reference counterpart. */
int spg_vbd_history_start(spg_vbd_ctx *ctx, int32_t gate, int32_t *accepted, double *rho);
/* Per object, rho = |r(start)| / |r(x~)| of the open (or last) substep's start as its gate measured it: by
 * spg_vbd_warm_start / spg_vbd_history_start with gate 1, or by the first check of
 * spg_vbd_iterate_to_tolerance for a deferred gate; NaN when not measured (plain starts, gate 0). */
int spg_vbd_start_rho(const spg_vbd_ctx *ctx, double *rho);

#ifdef __cplusplus
}
#endif

#endif
