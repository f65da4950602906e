/*
 * gridrec_cnn.h -- C-ABI of the sparse-angle residual CNN in libgridrec_hip.so.
 *
 * No counterpart in the reference (it has no learned component): the network
 * corrects the streaks of gridrec reconstructions from sparse-angle scans and
 * is trained on pairs gridrec itself makes (sparse-angle input, dense-angle
 * target).  Five 3x3 convolutions 1 -> 32 -> 32 -> 32 -> 32 -> 1, ReLU after
 * the first four, output y = x + scale * (conv5(a4) + b5) in the input's units
 * (the network works on x/scale).  BF16 weights and activations, FP32
 * accumulation; MSE loss; Adam on FP32 master weights.
 *
 * All pointers are device pointers; streams are hipStream_t (NULL = the null
 * stream); every call is asynchronous on it.  Return codes are the GR_E* of
 * gridrec_hip.h, gr_last_error() gives the message.
 *
 * Parameters: one FP32 vector of gr_cnn_param_count() floats,
 *   W1[32][9] b1[32] W2[32][32][9] b2[32] W3 b3 W4 b4 W5[32][9] b5[1]
 * (W[co][ci][ky*3+kx], correlation with zero padding).  Images: [n][H][W]
 * float.  Activations a1..a4: [n][H][W][32] bf16 (unsigned short bits).
 */
#ifndef GRIDREC_CNN_H
#define GRIDREC_CNN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif
#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

long gr_cnn_param_count(void);
/* bf16 count of the packed weights of the three 32->32 layers (forward and
 * flipped-transposed for the input gradient). */
long gr_cnn_packed_count(void);
/* Device scratch bytes gr_cnn_forward (slice statistics) and gr_cnn_backward
 * need for a batch of n H x W images. */
size_t gr_cnn_work_bytes(int n, int H, int W);

/* Round the master weights of layers 2-4 to bf16 in both packed layouts; call
 * after every change of params. */
int gr_cnn_pack(const float *params, unsigned short *packed, void *stream);

/* Forward pass.  a1..a4 receive the activations.  For inference a3 may alias
 * a1 and a4 a2: then the first layer runs fused into the second, a1 is never
 * written (a1's buffer ends holding a3, a2's a4), and y, yn and slice_stats
 * are bitwise those of distinct buffers.  y (may be NULL): the corrected images in the input's units;
 * yn (may be NULL): the same in network units (x/scale + residual), needed by
 * gr_cnn_backward.  slice_stats (may be NULL, needs work): per image, the sums
 * of residual^2 and (x/scale)^2 in network units (2n floats) -- the
 * correction norm of the fallback test. */
int gr_cnn_forward(const float *params, const unsigned short *packed, const float *x, int n, int H, int W,
                   float scale, unsigned short *a1, unsigned short *a2, unsigned short *a3,
                   unsigned short *a4, float *y, float *yn, float *slice_stats, void *work, void *stream);

/* Inference: gr_cnn_forward's y (bitwise) with the five layers fused in one
 * kernel and no activation buffers; slice_stats as there (values may differ
 * in the last bits: another summation order).  work: gr_cnn_infer_work_bytes(n,
 * H, W) bytes when slice_stats is given, else may be NULL. */
size_t gr_cnn_infer_work_bytes(int n, int H, int W);
int gr_cnn_infer(const float *params, const unsigned short *packed, const float *x, int n, int H, int W,
                 float scale, float *y, float *slice_stats, void *work, void *stream);

/* Backward pass of the MSE loss mean((yn - target/scale)^2): the gradient of
 * every parameter into grads (param layout, overwritten) and the loss into
 * *loss (device float, may be NULL).  a1..a4 and yn from gr_cnn_forward on
 * the same x (a1..a4 distinct). */
int gr_cnn_backward(const float *params, const unsigned short *packed, const float *x, const float *target,
                    int n, int H, int W, float scale, const unsigned short *a1, const unsigned short *a2,
                    const unsigned short *a3, const unsigned short *a4, const float *yn, float *grads,
                    float *loss, void *work, void *stream);

/* One Adam step (bias-corrected, Kingma & Ba 2015) over count parameters;
 * step counts from 1. */
int gr_cnn_adam(float *params, const float *grads, float *m, float *v, long count, int step, float lr,
                float beta1, float beta2, float eps, void *stream);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif
#ifdef __cplusplus
}
#endif
#endif
