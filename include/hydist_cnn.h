/*
 * hydist_cnn.h -- C ABI of the learned flow surrogate in libhydist.so: a
 * small residual convolutional network (3x3 convolutions, width 64) with
 * hand-written BF16 MFMA (FP32 accumulation) forward, backward and Adam
 * kernels for gfx950, plus the state packing that connects it to a hydist
 * context.
 *
 * There is no counterpart in the reference (its only model is the FP64 ADI
 * step, Engine.py:144-366).  The network stands in for the hydrodynamic ADI
 * steps of a bed-change phase: from the state at t it predicts u, v, z at
 * t + K dT; the sediment and bed-change kernels then run on that flow, and
 * the plain FP64 step checks it periodically (hydist_amd.surrogate).
 *
 * Layout: images are NHWC on the case's padded grid, H = N+3 rows, W = M+3
 * columns.  Activations BF16, channels padded to 32 (input) / 64 (hidden) /
 * 32 (output).  Weights BF16 [tap][cout][cin] (tap = 3*(dy+1)+(dx+1)) with
 * FP32 master copies, biases FP32.  Layer 0: conv cin->64, ReLU; layers
 * 1..L-2: a += ReLU(conv(a)); layer L-1: conv 64->cout, linear.  The
 * network's output is added to the normalised input u, v, z (residual).
 *
 * All pointers are device pointers unless named host_*.  Errors as in
 * hydist.h (0 or a negative HD_E* code; hd_cnn_last_error).
 */
#ifndef HYDIST_CNN_H
#define HYDIST_CNN_H

#include <stddef.h>
#include <stdint.h>

#include "hydist.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HD_CNN_CIN 8      /* real input channels */
#define HD_CNN_COUT 3     /* real output channels: du, dv, dz (normalised) */

typedef struct hd_cnn hd_cnn;

/* Input channels, in order: u/s_u, v/s_v, z/s_z, dH/s_b, wet mask, discharge
 * forcing q(t)/s_q, water-level forcing zb(t)/s_z, depth h/s_h (each clipped
 * to [-8, 8]).  The network's three outputs are the increments of u/s_u,
 * v/s_v, z/s_z over the horizon in units of d_u, d_v, d_z, and the loss is
 * the masked mean square of the increment error in those units.  Scales are
 * set once (hd_cnn_set_scales). */
typedef struct hd_cnn_scales {
    float s_u, s_v, s_z, s_b, s_q, s_h;
    float d_u, d_v, d_z;
} hd_cnn_scales;

/* A network for images of H x W pixels, up to `batch` images per call,
 * `layers` conv layers (>= 3), on HIP device `device`.  Weights start from a
 * seeded He-uniform draw. */
int hd_cnn_create(int device, int H, int W, int batch, int layers, uint64_t seed, hd_cnn** out);
int hd_cnn_destroy(hd_cnn* n);
const char* hd_cnn_last_error(const hd_cnn* n);
int hd_cnn_set_scales(hd_cnn* n, const hd_cnn_scales* s);

/* FP32 master weights of layer l: host_w [9][cout][cin] (cin, cout the
 * padded channel counts: 32/64/32), host_b [cout].  set also refreshes the
 * BF16 copies and clears the Adam moments of the layer. */
int hd_cnn_layer_dims(const hd_cnn* n, int l, int* cin, int* cout);
int hd_cnn_get_layer(hd_cnn* n, int l, float* host_w, float* host_b);
int hd_cnn_set_layer(hd_cnn* n, int l, const float* host_w, const float* host_b);
/* Gradients of the last hd_cnn_backward (FP32, same layout). */
int hd_cnn_get_grad(hd_cnn* n, int l, float* host_dw, float* host_db);

/* Input / target images (host, FP32 NHWC): host_x [b][H][W][8] (the real
 * input channels), host_y [b][H][W][3] normalised targets (u, v, z at the
 * horizon, same scales), host_m [b][H][W] loss weights (wet mask).  Images
 * are converted to BF16 input on the device. */
int hd_cnn_load_batch(hd_cnn* n, int b, const float* host_x, const float* host_y, const float* host_m);
/* Forward over the loaded batch; host_out (optional) [b][H][W][3] receives
 * the prediction (normalised, input u, v, z + network output). */
int hd_cnn_forward(hd_cnn* n, float* host_out);
/* Masked mean-squared error of the last forward against the loaded targets
 * and its gradients (into the grad buffers); *loss receives the loss. */
int hd_cnn_backward(hd_cnn* n, double* loss);
/* One Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8) on every layer from
 * the gradient buffers; refreshes the BF16 weights. */
int hd_cnn_adam(hd_cnn* n, float lr);
/* Forward + backward + Adam on the loaded batch (one training step). */
int hd_cnn_train_step(hd_cnn* n, float lr, double* loss);

/* Device-resident training set of `count` samples on the network's grid:
 * per sample the BF16 input image, the FP32 u, v, z it holds (the residual
 * base) and the FP32 normalised u, v, z target.  hd_cnn_capture copies a
 * hydist context's state (same grid: H = N+3, W = M+3) into sample idx:
 * role 0 = input (u, v, z, dH, khouot, h and the forcing values q, zb),
 * role 1 = target (u, v, z). */
int hd_cnn_data_alloc(hd_cnn* n, int count);
int hd_cnn_capture(hd_cnn* n, int idx, int role, hd_ctx* c, float q, float zb);
/* One training step (forward, masked MSE, backward, Adam) on samples
 * host_idx[0..b) (b <= batch); *loss receives the batch loss before the
 * update.  hd_cnn_eval: the same loss without backward / update. */
int hd_cnn_train_samples(hd_cnn* n, const int* host_idx, int b, float lr, double* loss);
int hd_cnn_eval_samples(hd_cnn* n, const int* host_idx, int b, double* loss);

/* Surrogate step on a context: pack its state with forcing q, zb into batch
 * image 0, run the network, and write the predicted u, v, z of the wet
 * cells (khouot == 0) back into the context (u, v, z, t_u, t_v, t_z), then
 * refresh H_moi / Htdu / Htdv (hd_flow_refresh). */
int hd_cnn_predict_into_ctx(hd_cnn* n, hd_ctx* c, float q, float zb);
/* on != 0 (the default): hd_cnn_predict_into_ctx runs the convolutions only
 * over the 4 x 64 image tiles within L - 1 cells of a wet cell, which leaves
 * the outputs at the wet cells those of the dense pass, bit for bit (the
 * river covers about a tenth of its grid); on == 0: every tile. */
int hd_cnn_set_sparse(hd_cnn* n, int on);

/* One-tile check of the MFMA operand layout the kernels use: D = A.B for
 * A [16][32], B [32][16] (host FP32, rounded to BF16), D [16][16] FP32. */
int hd_cnn_mfma_check(int device, const float* host_a, const float* host_b, float* host_d);

#ifdef __cplusplus
}
#endif

#endif /* HYDIST_CNN_H */
