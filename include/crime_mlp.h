/*
 * crime_mlp.h -- the C_l emulator's network on one MI355X, in libcrime_gh.so (no counterpart in the
 * reference: GetHI has no emulator; the plain pipeline of crime_gh.h and crime_sht.h stays the method
 * it is judged against).
 *
 * Fixed architecture: 8 inputs -> 5 hidden layers of 256 SiLU units -> 1 output,
 *   h1 = silu(x W0 + b0),  h_{k+1} = silu(h_k W_k + b_k) (k = 1..4),  y = h5 . w5 + b5.
 * The four 256x256 layers run on BF16 MFMA with FP32 accumulation (their weights and the activations
 * h_k are rounded to BF16 at the GEMM inputs); the 8-input and the 1-output layers run in FP32.
 * Training minimises the weighted mean squared error (1/B) sum_b w_b (y_b - t_b)^2 (w_b = 1 when no
 * weights are given; 1/sigma_b^2 makes it a chi^2) with Adam on FP32 master weights.  Parameters
 * travel as one FP32 vector in the order
 *   W0[8][256], b0[256], W1[256][256], b1[256], ..., W4[256][256], b4[256], w5[256], b5[1]
 * (W[in][out], row-major); crime_mlp_param_count() gives its length.
 *
 * A handle owns one HIP stream and its device buffers; handles are not shared across threads.
 * Status codes are those of crime_gh.h.
 */
#ifndef CRIME_MLP_H
#define CRIME_MLP_H

#include <stddef.h>
#include <stdint.h>

#include "crime_gh.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CRIME_MLP_IN 8
#define CRIME_MLP_WIDTH 256
#define CRIME_MLP_HIDDEN 5

typedef struct crime_mlp crime_mlp_t;

int64_t crime_mlp_param_count(void);

/* max_batch: largest batch of any call (a multiple of 64, <= 65536). */
int crime_mlp_create(int max_batch, int device, crime_mlp_t **out);
void crime_mlp_destroy(crime_mlp_t *h);
const char *crime_mlp_last_error(const crime_mlp_t *h); /* NULL handle: the last failed create */

/* FP32 master weights in and out; set_params also clears the Adam moments and the step counter. */
int crime_mlp_set_params(crime_mlp_t *h, const float *params);
int crime_mlp_get_params(crime_mlp_t *h, float *params);

/* y[i] for host inputs x[i][8], any n >= 1 (run in batches of max_batch). */
int crime_mlp_forward(crime_mlp_t *h, int64_t n, const float *x, float *y);

/* The mean of nh networks' y[i] (an ensemble on one device): each batch of inputs goes to the device once
 * and every network runs on it, on the first handle's stream; batches of the smallest max_batch.  The
 * handles' own streams are synchronised first.  No counterpart in the reference (the emulator's). */
int crime_mlp_forward_mean(crime_mlp_t *const *hs, int nh, int64_t n, const float *x, float *y);

/* Loss and its gradient (FP32, parameter order above) for one host batch, n <= max_batch, n % 64 == 0;
 * w may be NULL (unit weights).  The parameters are not changed. */
int crime_mlp_loss_grad(crime_mlp_t *h, int n, const float *x, const float *t, const float *w, double *loss,
                        float *grad);

/* Training set kept in device memory: x[i][8], t[i], optional weights w[i] (NULL: 1), n >= 1. */
int crime_mlp_set_data(crime_mlp_t *h, int64_t n, const float *x, const float *t, const float *w);

/* `steps` Adam steps on minibatches of `batch` rows (a multiple of 64, <= max_batch) drawn uniformly
 * with replacement by Philox-4x32-10 keyed by `seed` and the step number.  Learning rate
 * lr0 * 0.5 (1 + cos(pi s / total_steps)) at global step s (cosine decay over total_steps, counted
 * across calls since set_params).  One step is captured as a HIP graph and replayed.  loss_out
 * (optional, `steps` entries) receives each step's minibatch loss. */
int crime_mlp_train(crime_mlp_t *h, int steps, int batch, float lr0, int total_steps, float beta1, float beta2,
                    float eps, uint32_t seed, float *loss_out);

/* on = 1 (default): crime_mlp_train replays the step as a captured HIP graph; on = 0: it enqueues the
 * same kernels step by step (the same results; for tools that trace kernel launches). */
int crime_mlp_set_graph(crime_mlp_t *h, int on);

#ifdef __cplusplus
}
#endif
#endif
