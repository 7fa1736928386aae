/*
 * allium_cnn.h -- learned low-order flux correction for source iteration.
 *
 * No counterpart in the reference.  It stands in for diffusion synthetic
 * acceleration (DiffusionSyntheticAccelerationOperator, src/transport_op.hpp:38-59)
 * in allium_source_iteration: after each transport sweep the network maps
 * the latest element-averaged scalar-flux change, log sigma_t and the
 * scattering ratio on the element-centred grid to an element-constant flux
 * correction.  For callers that want SI with fewer sweeps on structured
 * one-group problems.
 *
 * Network: 5 convolution layers with 3x3x3 kernels (optionally dilated), 3 ->
 * 32 -> 32 -> 32 -> 32 -> 1 channels, ReLU after the first four, zero padding.  Weights and
 * activations are bfloat16 on the GPU with FP32 accumulation (MFMA); the
 * master parameters and the Adam moments are FP32.
 *
 * Parameters are packed per layer l as W_l[27][CO][CI] (offset o =
 * (dz+1)*9 + (dy+1)*3 + (dx+1), x fastest) followed by b_l[CO].
 * Grids are channels-last: x[B][nz][ny][nx][3], out/target[B][nz][ny][nx].
 */
#ifndef ALLIUM_CNN_H
#define ALLIUM_CNN_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct allium_cnn *allium_cnn_t;

/* He-normal weights and zero biases from a seed */
int allium_cnn_create(unsigned long long seed, allium_cnn_t *out);
void allium_cnn_destroy(allium_cnn_t h);
/* number of FP32 parameters in the packed layout */
long allium_cnn_num_params(void);
/* per-layer dilation of the 3x3x3 taps (offsets dilation * {-1, 0, 1} per
 * axis, 1..8; default 1 for every layer): a wider receptive field with the
 * same parameters, e.g. {1, 2, 4, 8, 1} sees 33 cells */
int allium_cnn_set_dilation(allium_cnn_t h, const int *dilation);
int allium_cnn_get_dilation(allium_cnn_t h, int *dilation);
int allium_cnn_get_params(allium_cnn_t h, float *params);
int allium_cnn_set_params(allium_cnn_t h, const float *params);
/* inference on device buffers */
int allium_cnn_forward_device(allium_cnn_t h, int batch, int nz, int ny, int nx, const float *d_x, float *d_out,
                              void *stream);
/* mean squared error against d_target and its gradient (host, packed like
 * the parameters); no parameter update */
int allium_cnn_gradient_device(allium_cnn_t h, int batch, int nz, int ny, int nx, const float *d_x,
                               const float *d_target, float *grad, double *loss);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on a batch; *loss is the
 * batch loss before the step */
int allium_cnn_train_step_device(allium_cnn_t h, int batch, int nz, int ny, int nx, const float *d_x,
                                 const float *d_target, float lr, double *loss);

#ifdef __cplusplus
}
#endif

#endif
