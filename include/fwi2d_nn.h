/*
 * fwi2d_nn.h -- C ABI of the learned preconditioner: a small convolutional
 * encoder-decoder on the GPU (bf16 weights and activations, fp32 accumulation
 * on the gfx950 matrix cores, hand-written forward and backward kernels,
 * Adam).  The reference (phbastosa/FWI2D) has no learned component; these
 * entry points have no reference counterpart and serve the learned-vs-plain
 * FWI driver and its training.
 *
 * Network (8 convolutions, 3x3, same padding; H and W multiples of 4):
 *   c1 3->32 relu, c2 32->32 relu, pool2, c3 32->64 relu, c4 64->64 relu,
 *   pool2, c5 64->64 relu, up2 + c3 output, c6 64->32 relu,
 *   up2 + c1 output, c7 32->32 relu, c8 32->1 (linear).
 * Inputs  x[b][3][H][W] fp32, outputs y[b][H][W] fp32 (W is the fast axis:
 * for a model [nx][nz] pass H = nx, W = nz).  Loss: mean squared error.
 * Parameters (flat fp32, in this order for c1..c8): W[cout][3][3][cin] then
 * b[cout].  All pointers are host pointers; 0 = success, FWI_ERR_* codes and
 * fwi_last_error() as in fwi2d.h.
 */
#ifndef FWI2D_NN_H
#define FWI2D_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fwi_nn fwi_nn;

/* Allocates a network for up to `batch` images of H x W with seeded
 * He-uniform weights (zero biases). */
int fwi_nn_create(int32_t batch, int32_t height, int32_t width, uint64_t seed, fwi_nn **out);
int fwi_nn_destroy(fwi_nn *nn);
/* The batch and image size the network was created for. */
int fwi_nn_get_shape(fwi_nn *nn, int32_t *batch, int32_t *height, int32_t *width);
int fwi_nn_num_params(fwi_nn *nn, int64_t *count);
int fwi_nn_get_params(fwi_nn *nn, float *params);
int fwi_nn_set_params(fwi_nn *nn, const float *params);

/* Adam state in the flat parameter layout: first and second moments and the
 * number of steps taken, so that training can stop and resume exactly.  The
 * reference keeps no optimizer state (it trains no network); these serve the
 * training checkpoint.  set rejects a negative step count and non-finite or
 * negative second moments. */
int fwi_nn_get_optimizer(fwi_nn *nn, float *m, float *v, int64_t *step);
int fwi_nn_set_optimizer(fwi_nn *nn, const float *m, const float *v, int64_t step);

/* y = net(x) for nb <= batch images. */
int fwi_nn_forward(fwi_nn *nn, int32_t nb, const float *x, float *y);

/* Loss = mean((net(x) - target)^2) and its gradient with respect to every
 * parameter (flat layout above; grads may be NULL). */
int fwi_nn_gradient(fwi_nn *nn, int32_t nb, const float *x, const float *target, float *loss,
                    float *grads);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the loss of a batch;
 * returns the loss before the step. */
int fwi_nn_train_step(fwi_nn *nn, int32_t nb, const float *x, const float *target, float lr,
                      float *loss);

/* Test hook: d[16][16] = a[16][32] b[32][16] (row-major, inputs rounded to
 * bf16) through one v_mfma_f32_16x16x32_bf16 with the kernels' lane layout. */
int fwi_nn_mfma_probe(const float *a, const float *b, float *d);

#ifdef __cplusplus
}
#endif
#endif
