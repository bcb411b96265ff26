// 3x3 convolution kernels of the learned prior (BF16 storage, FP32 accumulation).
//
// Tensors: images fp32 [B][H][W]; activations bf16 (raw uint16) NHWC
// [B][H][W][32].  Weights: layer 1 w1[9][32], layer 5 w5[9][32] (fp32);
// 32 -> 32 layers wt[9][32 co][32 ci] (bf16), tap t = 3 * (dy + 1) + (dx + 1).
#pragma once

#include "rn_common.h"

#include <cstdint>

namespace rn {

// a1 = relu(conv(x; w1) + b1)
void conv_in(const float* x, const float* w, const float* b, uint16_t* out, int B, int H, int W, hipStream_t st);
// forward (mask == nullptr): out = relu?(conv(in; wt) + bias)
// data gradient (mask != nullptr): out = conv(in; wt) * (mask > 0), with wt the flipped, transposed weights
void conv_mid(const uint16_t* in, const uint16_t* wt, const float* bias, const uint16_t* mask, uint16_t* out,
              int B, int H, int W, bool relu, hipStream_t st);
// y = xres + b5 + conv(a4; w5)   (xres may be null)
void conv_out(const uint16_t* a4, const float* w5, float b5, const float* xres, float* y, int B, int H, int W,
              hipStream_t st);
// The whole network in one launch: y = x + b5 + conv(a4; w5), a4 = relu(conv(a3; w4) + b4), ...,
// a1 = relu(conv(x; w1) + b1) -- the same bits as conv_in, 3 x conv_mid (relu) and conv_out
void cnn_forward(const float* x, const float* w1, const float* b1, const uint16_t* w2, const float* b2,
                 const uint16_t* w3, const float* b3, const uint16_t* w4, const float* b4, const float* w5, float b5,
                 float* y, int B, int H, int W, hipStream_t st);
// dz4 = (a4 > 0) * conv^T(dy; w5)
void conv_out_dgrad(const float* dy, const float* w5, const uint16_t* a4, uint16_t* dz4, int B, int H, int W,
                    hipStream_t st);
// mode 0: grad[t*32 + c] = sum_p x[p + off_t] dz1[p][c], grad[288 + c] = sum_p dz1[p][c]
// mode 1: grad[t*32 + c] = sum_p dy[p] a4[p + off_t][c], grad[288] = sum_p dy[p]
void wgrad_narrow(const float* img, const uint16_t* act, float* grad, int B, int H, int W, int mode, DevBuf& work,
                  hipStream_t st);
// grad[c] = sum_p dz[p][c]
void bias_grad(const uint16_t* dz, float* grad, long npx, DevBuf& work, hipStream_t st);
// grad[t][co][ci] = sum_p dz[p][co] in[p + off_t][ci]   (MFMA)
void wgrad_mid(const uint16_t* in, const uint16_t* dz, float* grad, int B, int H, int W, DevBuf& work, hipStream_t st);

} // namespace rn
