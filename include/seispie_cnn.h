/*
 * seispie_cnn -- C ABI of the learned step conditioner's network: a small
 * fully convolutional network (3x3 convolutions, zero padding, ReLU after
 * every layer but the last) with hand-written gfx950 training and
 * inference kernels (seispie_amd/csrc/cnn_hip.hip).
 *
 * No reference counterpart: icui/seispie's optimizers step along the raw
 * kernel (seispie/optimize/lbfgs.py:23-59, cg.py:32-49, gd.py:4-5).  The
 * network is an optional step conditioner the FWI optimizers call where
 * the reference uses the gradient as it is.
 *
 * Conventions as seispie_hip.h: 0 on success, a message in
 * sp_cnn_last_error otherwise; host buffers are borrowed for the call;
 * *_on_device == 1 takes device pointers ordered on the network's stream.
 *
 * Layout.  Activations float32 [C][H][W] (W contiguous: a solver field with
 * k = i*nz + j is one channel with H = nx, W = nz).  Parameters are one flat
 * float32 array; per layer l the weights [c_{l-1}][3][3][c_l] (output
 * channel contiguous) then the biases [c_l]:
 *   z_l[co][y][x] = b[co] + sum_ci sum_{dy,dx} W[ci][dy][dx][co] * a_{l-1}[ci][y+dy-1][x+dx-1]
 */
#ifndef SEISPIE_CNN_H
#define SEISPIE_CNN_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sp_cnn sp_cnn;

/* channels[0..nlayers]: inputs (1..32), then each layer's outputs (1, 2, 4, 8, 16 or 32) */
int sp_cnn_create(int device, int nlayers, const int *channels, sp_cnn **out);
int sp_cnn_destroy(sp_cnn *net);
const char *sp_cnn_last_error(const sp_cnn *net);
int sp_cnn_set_stream(sp_cnn *net, void *hip_stream); /* NULL: the null stream (the default) */
int sp_cnn_sync(sp_cnn *net);
long long sp_cnn_num_params(const sp_cnn *net);
/* parameters in the layout above; setting them also resets the Adam moments */
int sp_cnn_set_params(sp_cnn *net, const float *params, int on_device);
int sp_cnn_get_params(sp_cnn *net, float *params, int on_device);
/* inference: x [c_0][H][W] -> y [c_L][H][W] */
int sp_cnn_forward(sp_cnn *net, const float *x, float *y, int H, int W, int on_device);
/* loss = sum m (y - t)^2 / sum m over the H*W cells of the one-channel output
 * (m = 1 when mask is NULL); computes the parameter gradient of the loss,
 * kept for sp_cnn_adam_step and readable by sp_cnn_get_grad */
int sp_cnn_loss_grad(sp_cnn *net, const float *x, const float *target, const float *mask, int H, int W,
                     int on_device, double *loss);
int sp_cnn_get_grad(sp_cnn *net, float *grad, int on_device);
/* one Adam update (bias-corrected) with the gradient of the last sp_cnn_loss_grad */
int sp_cnn_adam_step(sp_cnn *net, float lr, float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif
#endif
