/*
 * gvm_nn.h -- C ABI of the layers of the warm-start network: a small U-Net
 * that maps a dirty image to an initial model for the MEM optimiser.
 *
 * No counterpart in the reference, which always starts the optimiser from
 * a constant image (src/mfs.cu:737-745).  Training and inference both run
 * on these kernels: forward and backward of each layer, the loss and the
 * Adam update.
 *
 * Conventions: tensors are device float arrays in NCHW order (batch,
 * channel, row, column).  All calls enqueue on the stream of `ctx` (the
 * objective's context) and return a gvm_status.  Reductions (weight and bias
 * gradients, the loss) are computed in a fixed order, so training is
 * bitwise reproducible.
 */
#ifndef GVM_NN_H
#define GVM_NN_H

#include "gvm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* out = conv(in, w) + bias, K x K kernel (K = 1 or 3), zero padding K/2,
 * stride 1.  in [B][Ci][H][W], w [Co][Ci][K][K], bias [Co], out [B][Co][H][W].
 * relu != 0 applies max(., 0) to the output. */
gvm_status gvm_nn_conv_fwd(gvm_ctx* ctx, const float* in, const float* w, const float* bias,
                           float* out, int B, int Ci, int Co, int H, int W, int K, int relu);
/* Backward of a convolution whose output passed through ReLU when relu != 0
 * (out is then the forward output, used as the mask; it may be NULL when
 * relu == 0).  dout [B][Co][H][W] is the gradient at the (activated) output.
 * din (may be NULL) receives the input gradient; dw and dbias receive (not
 * accumulate) the parameter gradients. */
gvm_status gvm_nn_conv_bwd(gvm_ctx* ctx, const float* in, const float* w, const float* out,
                           const float* dout, float* din, float* dw, float* dbias, int B, int Ci,
                           int Co, int H, int W, int K, int relu);
/* The same with BF16 weights (uint16 holding the upper half of an FP32,
 * as gvm_nn_to_bf16 writes them); arithmetic and accumulation stay FP32,
 * the weight gradients are FP32. */
gvm_status gvm_nn_conv_fwd_bf16(gvm_ctx* ctx, const float* in, const unsigned short* w,
                                const float* bias, float* out, int B, int Ci, int Co, int H, int W,
                                int K, int relu);
gvm_status gvm_nn_conv_bwd_bf16(gvm_ctx* ctx, const float* in, const unsigned short* w,
                                const float* out, const float* dout, float* din, float* dw,
                                float* dbias, int B, int Ci, int Co, int H, int W, int K, int relu);
/* dst = src rounded to BF16 (nearest even), n values. */
gvm_status gvm_nn_to_bf16(gvm_ctx* ctx, const float* src, unsigned short* dst, long n);
/* 2 x 2 average pooling [B][C][H][W] -> [B][C][H/2][W/2] and its backward. */
gvm_status gvm_nn_pool_fwd(gvm_ctx* ctx, const float* in, float* out, int B, int C, int H, int W);
gvm_status gvm_nn_pool_bwd(gvm_ctx* ctx, const float* dout, float* din, int B, int C, int H,
                           int W);
/* Nearest 2x upsampling [B][C][H][W] -> [B][C][2H][2W] written into channels
 * [c0, c0 + C) of a [B][Cd][2H][2W] tensor, and its backward (H, W: the
 * small size). */
gvm_status gvm_nn_up_fwd(gvm_ctx* ctx, const float* in, float* out, int B, int C, int Cd, int c0,
                         int H, int W);
gvm_status gvm_nn_up_bwd(gvm_ctx* ctx, const float* dout, float* din, int B, int C, int Cd, int c0,
                         int H, int W);
/* Channel-block copy: dst[:, d0 : d0 + C] = src[:, s0 : s0 + C], with Cs and
 * Cd channels in src and dst (skip connections and their gradients). */
gvm_status gvm_nn_copy_channels(gvm_ctx* ctx, const float* src, float* dst, int B, int C, int Cs,
                                int s0, int Cd, int d0, int HW);
/* dst += src (n floats). */
gvm_status gvm_nn_add(gvm_ctx* ctx, const float* src, float* dst, long n);
/* Mean squared error of y against t (n floats): writes sum (y - t)^2 / n to
 * value slot `slot` (read with gvm_values_read) and dy = 2 (y - t) / n. */
gvm_status gvm_nn_mse(gvm_ctx* ctx, const float* y, const float* t, float* dy, long n, int slot);
/* One Adam step on n parameters: m, v moments, step t >= 1. */
gvm_status gvm_nn_adam(gvm_ctx* ctx, float* p, const float* g, float* m, float* v, long n,
                       float lr, float beta1, float beta2, float eps, int t);

#ifdef __cplusplus
}
#endif
#endif
