/*
 * drift_mlp.h -- the ladder network of libdriftpt.so: a 16 -> 128 -> 128 -> 128 -> 32 perceptron
 * with SiLU hidden units, BF16 weights and FP32 accumulation, trained and evaluated by the
 * library's own HIP kernels.  It maps 16 instance statistics to a 32-rung temperature ladder.
 *
 * No counterpart in the reference: DRiFT only ever uses the geometric ladder
 * (drift/solvers/parallel_tempering.py:29-37).  The network changes where the rungs sit, never the
 * Metropolis or exchange rules.
 *
 * Output head: z (32 values) -> d_k = softplus(z_k) + 1e-4 -> S_j = d_1 + ... + d_j ->
 * x_j = S_j / S_31 for j = 0..31 (x_0 = 0, x_31 = 1, strictly increasing).  x_j is the position of
 * rung j on the log-temperature axis: log T_j = log T_min + x_j (log T_max - log T_min).
 * Loss: mean over samples and the interior rungs 1..30 of (x_j - target_j)^2.
 *
 * Parameters are one float array of DRIFT_MLP_PARAMS values: W1 (16x128, row-major [in][out]), b1,
 * W2 (128x128), b2, W3 (128x128), b3, W4 (128x32), b4.  The device keeps FP32 master weights (for
 * Adam) and their BF16 copies (for the forward pass).  Handles are not thread-safe.
 */
#ifndef DRIFT_MLP_H
#define DRIFT_MLP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DRIFT_MLP_IN 16
#define DRIFT_MLP_HIDDEN 128
#define DRIFT_MLP_OUT 32
#define DRIFT_MLP_PARAMS 39328

typedef struct drift_mlp drift_mlp;

/* New network with weights drawn from the Philox stream of `seed` (uniform, scaled by
 * sqrt(6 / fan_in)), biases zero, Adam moments zero. */
int drift_mlp_create(uint64_t seed, drift_mlp** out);
int drift_mlp_destroy(drift_mlp* net);

/* FP32 master weights out / in (DRIFT_MLP_PARAMS floats). */
int drift_mlp_get_params(const drift_mlp* net, float* params);
int drift_mlp_set_params(drift_mlp* net, const float* params);

/* Ladder positions x (batch x 32) for inputs X (batch x 16). */
int drift_mlp_forward(drift_mlp* net, const float* X, int32_t batch, float* x_out);

/* Loss and its gradient with respect to the FP32 master weights (through the BF16 forward pass),
 * without updating anything.  grads holds DRIFT_MLP_PARAMS floats. */
int drift_mlp_gradients(drift_mlp* net, const float* X, const float* target, int32_t batch,
                        float* grads, float* loss);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the batch; returns the pre-step loss. */
int drift_mlp_train_step(drift_mlp* net, const float* X, const float* target, int32_t batch,
                         float lr, float* loss);

#ifdef __cplusplus
}
#endif
#endif
