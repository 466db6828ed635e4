/* lymph_mlp.h -- C ABI of liblymph_mlp.so: the learned surrogate of the
 * batched elastodynamics solver (a fully connected network, GELU between
 * layers), trained and evaluated on one MI355X with hand-written BF16 MFMA
 * kernels and FP32 master weights.
 *
 * No reference counterpart: LYMPH3D has no learned component.  The network
 * maps instance parameters (source position, wavelet frequency, material
 * field coefficients) to receiver traces produced by the solver of
 * liblymph_gpu.so, which stays the fallback and the judge of its error.
 *
 * Conventions: every entry point returns 0 on success or a nonzero code
 * (LYMPH_MLP_ERR_*) with a message in lymph_mlp_last_error().  Host arrays
 * are borrowed for the duration of the call.  Parameters cross the ABI in
 * FP32, layer by layer: W_l as a row-major [width[l+1]][width[l]] matrix,
 * then b_l [width[l+1]].  All widths must be multiples of 8.
 *
 * Arithmetic: GEMM operands (weights, activations, output gradients) in
 * bfloat16 (round to nearest even), products accumulated in FP32 on the
 * matrix cores; pre-activations, outputs, gradients of the parameters, the
 * loss and Adam state in FP32.  Loss = mean over batch and outputs of
 * (y - t)^2.
 *
 * Streams: each network runs its kernels and copies on a HIP stream of its
 * own (created with it) and every call returns with that stream drained.
 * Calls on different networks may come from different host threads at the
 * same time and then overlap on the GPU; calls on one network must not
 * overlap.  The error message of lymph_mlp_last_error() is shared.
 */
#ifndef LYMPH_MLP_H
#define LYMPH_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LYMPH_MLP_ERR_ARG 101
#define LYMPH_MLP_ERR_HIP 102
#define LYMPH_MLP_ERR_STATE 103

typedef struct LymphMlp LymphMlp;

const char* lymph_mlp_last_error(void);

/* n_layers linear layers; widths[0..n_layers] (input, hidden..., output);
 * max_batch = largest batch of training steps and gradient calls. */
int lymph_mlp_create(int n_layers, const int32_t* widths, int max_batch, LymphMlp** out);
int lymph_mlp_destroy(LymphMlp* net);
int lymph_mlp_n_params(const LymphMlp* net, int64_t* n);
/* Set (and round to the BF16 working copy) all parameters; resets Adam. */
int lymph_mlp_set_params(LymphMlp* net, const float* params);
int lymph_mlp_get_params(const LymphMlp* net, float* params);

/* Training set kept on the device: x [n][width[0]], y [n][width[L]]. */
int lymph_mlp_set_data(LymphMlp* net, int n, const float* x, const float* y);
/* n_steps Adam steps; step s uses samples order[s*batch .. s*batch+batch-1]
 * of the training set; losses[s] = loss of step s before its update (may be
 * NULL).  Bias-corrected Adam (lr, beta1, beta2, eps) continues its step
 * count across calls. */
int lymph_mlp_train(LymphMlp* net, int n_steps, int batch, const int32_t* order, float lr,
                    float beta1, float beta2, float eps, float* losses);
/* One forward + backward on the given batch without an update: loss and the
 * parameter gradients in the parameter layout. */
int lymph_mlp_gradients(LymphMlp* net, int batch, const float* x, const float* y, float* grads,
                        float* loss);
/* Forward pass of n inputs (any n; processed in batches of max_batch). */
int lymph_mlp_predict(LymphMlp* net, int n, const float* x, float* y);
/* Device time of forward passes over the first n samples of the training
 * set (device-resident inputs), mean over reps, in ms. */
int lymph_mlp_time_predict(LymphMlp* net, int n, int reps, double* ms);

#ifdef __cplusplus
}
#endif
#endif
