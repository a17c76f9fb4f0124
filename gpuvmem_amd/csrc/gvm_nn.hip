// Layers of the warm-start U-Net (include/gvm_nn.h): convolutions with
// optional ReLU, 2x2 average pooling, nearest 2x upsampling into a channel
// block, channel copies, the MSE loss and Adam, forward and backward.
//
// Direct convolutions, one thread per output element (forward) or input
// element (input gradient).  Parameter gradients take one workgroup per
// (output channel, input channel) pair with the K*K taps in registers,
// accumulated in FP64 and reduced in a fixed order, so training is
// reproducible bit for bit.  Layouts are NCHW.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gvm_internal.h"
#include "gvm_nn.h"

namespace gvm {
namespace {

constexpr int kT = 256;

int nblocks(long n) { return (int)std::max(1L, std::min<long>((n + kT - 1) / kT, 1L << 20)); }

// A weight as float: FP32 storage, or BF16 (the upper half of an FP32).
__device__ __forceinline__ float wload(const float* w, long i) { return w[i]; }
__device__ __forceinline__ float wload(const unsigned short* w, long i) {
  return __uint_as_float((unsigned int)w[i] << 16);
}

__device__ __forceinline__ float relu_grad(const float* out, const float* dout, long i, int relu) {
  const float g = dout[i];
  return relu ? (out[i] > 0.0f ? g : 0.0f) : g;
}

template <int K, class WT>
__global__ __launch_bounds__(kT) void conv_fwd_kernel(const float* __restrict__ in,
                                                      const WT* __restrict__ w,
                                                      const float* __restrict__ bias,
                                                      float* __restrict__ out, int B, int Ci,
                                                      int Co, int H, int W, int relu) {
  constexpr int P = K / 2;
  const long HW = (long)H * W;
  const long n = (long)B * Co * HW;
  for (long o = blockIdx.x * (long)kT + threadIdx.x; o < n; o += (long)gridDim.x * kT) {
    const int x = (int)(o % W);
    const int y = (int)((o / W) % H);
    const int co = (int)((o / HW) % Co);
    const int b = (int)(o / (HW * Co));
    const float* __restrict__ src = in + (long)b * Ci * HW;
    const WT* __restrict__ wc = w + (long)co * Ci * K * K;
    float acc = bias[co];
    for (int ci = 0; ci < Ci; ci++) {
      const float* __restrict__ s = src + (long)ci * HW;
      const WT* __restrict__ wk = wc + ci * K * K;
#pragma unroll
      for (int ky = 0; ky < K; ky++) {
        const int yy = y + ky - P;
        if (yy < 0 || yy >= H) continue;
#pragma unroll
        for (int kx = 0; kx < K; kx++) {
          const int xx = x + kx - P;
          if (xx < 0 || xx >= W) continue;
          acc = fmaf(wload(wk, ky * K + kx), s[(long)yy * W + xx], acc);
        }
      }
    }
    out[o] = relu ? fmaxf(acc, 0.0f) : acc;
  }
}

// din[b][ci][y][x] = sum_co sum_k w[co][ci][k] g[b][co][y - ky + P][x - kx + P]
template <int K, class WT>
__global__ __launch_bounds__(kT) void conv_bwd_data_kernel(const WT* __restrict__ w,
                                                           const float* __restrict__ out,
                                                           const float* __restrict__ dout,
                                                           float* __restrict__ din, int B,
                                                           int Ci, int Co, int H, int W,
                                                           int relu) {
  constexpr int P = K / 2;
  const long HW = (long)H * W;
  const long n = (long)B * Ci * HW;
  for (long o = blockIdx.x * (long)kT + threadIdx.x; o < n; o += (long)gridDim.x * kT) {
    const int x = (int)(o % W);
    const int y = (int)((o / W) % H);
    const int ci = (int)((o / HW) % Ci);
    const int b = (int)(o / (HW * Ci));
    float acc = 0.0f;
    for (int co = 0; co < Co; co++) {
      const long base = ((long)b * Co + co) * HW;
      const WT* __restrict__ wk = w + ((long)co * Ci + ci) * K * K;
#pragma unroll
      for (int ky = 0; ky < K; ky++) {
        const int yy = y - ky + P;
        if (yy < 0 || yy >= H) continue;
#pragma unroll
        for (int kx = 0; kx < K; kx++) {
          const int xx = x - kx + P;
          if (xx < 0 || xx >= W) continue;
          acc = fmaf(wload(wk, ky * K + kx), relu_grad(out, dout, base + (long)yy * W + xx, relu),
                     acc);
        }
      }
    }
    din[o] = acc;
  }
}

// One workgroup per (co, ci): dw[co][ci][ky][kx] = sum_b,y,x g[b][co][y][x]
// in[b][ci][y + ky - P][x + kx - P], FP64 per thread, fixed-order tree.
template <int K>
__global__ __launch_bounds__(kT) void conv_bwd_weight_kernel(const float* __restrict__ in,
                                                             const float* __restrict__ out,
                                                             const float* __restrict__ dout,
                                                             float* __restrict__ dw, int B,
                                                             int Ci, int Co, int H, int W,
                                                             int relu) {
  constexpr int P = K / 2;
  constexpr int KK = K * K;
  __shared__ double red[KK][kT];
  const int co = blockIdx.x / Ci, ci = blockIdx.x % Ci;
  const long HW = (long)H * W;
  double acc[KK];
#pragma unroll
  for (int k = 0; k < KK; k++) acc[k] = 0.0;
  const long n = (long)B * HW;
  for (long i = threadIdx.x; i < n; i += kT) {
    const int b = (int)(i / HW);
    const long p = i % HW;
    const int y = (int)(p / W), x = (int)(p % W);
    const float g = relu_grad(out, dout, ((long)b * Co + co) * HW + p, relu);
    if (g == 0.0f) continue;
    const float* __restrict__ s = in + ((long)b * Ci + ci) * HW;
#pragma unroll
    for (int ky = 0; ky < K; ky++) {
      const int yy = y + ky - P;
      if (yy < 0 || yy >= H) continue;
#pragma unroll
      for (int kx = 0; kx < K; kx++) {
        const int xx = x + kx - P;
        if (xx < 0 || xx >= W) continue;
        acc[ky * K + kx] += (double)g * (double)s[(long)yy * W + xx];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < KK; k++) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int h = kT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int k = 0; k < KK; k++) red[k][threadIdx.x] += red[k][threadIdx.x + h];
    __syncthreads();
  }
  if ((int)threadIdx.x < KK) dw[((long)co * Ci + ci) * KK + threadIdx.x] = (float)red[threadIdx.x][0];
}

// One workgroup per output channel: dbias[co] = sum g[b][co][:, :].
__global__ __launch_bounds__(kT) void conv_bwd_bias_kernel(const float* __restrict__ out,
                                                           const float* __restrict__ dout,
                                                           float* __restrict__ dbias, int B,
                                                           int Co, long HW, int relu) {
  __shared__ double red[kT];
  const int co = blockIdx.x;
  double acc = 0.0;
  const long n = (long)B * HW;
  for (long i = threadIdx.x; i < n; i += kT) {
    const long b = i / HW, p = i % HW;
    acc += (double)relu_grad(out, dout, (b * Co + co) * HW + p, relu);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = kT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) dbias[co] = (float)red[0];
}

__global__ __launch_bounds__(kT) void pool_fwd_kernel(const float* __restrict__ in,
                                                      float* __restrict__ out, long nbc, int H,
                                                      int W) {
  const int h = H / 2, w = W / 2;
  const long n = nbc * h * w;
  for (long o = blockIdx.x * (long)kT + threadIdx.x; o < n; o += (long)gridDim.x * kT) {
    const int x = (int)(o % w), y = (int)((o / w) % h);
    const long c = o / ((long)h * w);
    const float* s = in + c * H * W + (long)(2 * y) * W + 2 * x;
    out[o] = 0.25f * ((s[0] + s[1]) + (s[W] + s[W + 1]));
  }
}

__global__ __launch_bounds__(kT) void pool_bwd_kernel(const float* __restrict__ dout,
                                                      float* __restrict__ din, long nbc, int H,
                                                      int W) {
  const int h = H / 2, w = W / 2;
  const long n = nbc * H * W;
  for (long o = blockIdx.x * (long)kT + threadIdx.x; o < n; o += (long)gridDim.x * kT) {
    const int x = (int)(o % W), y = (int)((o / W) % H);
    const long c = o / ((long)H * W);
    din[o] = 0.25f * dout[c * h * w + (long)(y / 2) * w + x / 2];
  }
}

// out[b][c0 + c][Y][X] = in[b][c][Y/2][X/2]
__global__ __launch_bounds__(kT) void up_fwd_kernel(const float* __restrict__ in,
                                                    float* __restrict__ out, int B, int C, int Cd,
                                                    int c0, int H, int W) {
  const int H2 = 2 * H, W2 = 2 * W;
  const long n = (long)B * C * H2 * W2;
  for (long o = blockIdx.x * (long)kT + threadIdx.x; o < n; o += (long)gridDim.x * kT) {
    const int X = (int)(o % W2), Y = (int)((o / W2) % H2);
    const int c = (int)((o / ((long)H2 * W2)) % C);
    const int b = (int)(o / ((long)H2 * W2 * C));
    out[(((long)b * Cd + c0 + c) * H2 + Y) * W2 + X] =
        in[(((long)b * C + c) * H + Y / 2) * W + X / 2];
  }
}

// din[b][c][y][x] = sum of the four dout[b][c0 + c][2y + dy][2x + dx]
__global__ __launch_bounds__(kT) void up_bwd_kernel(const float* __restrict__ dout,
                                                    float* __restrict__ din, int B, int C, int Cd,
                                                    int c0, int H, int W) {
  const int W2 = 2 * W;
  const long n = (long)B * C * H * W;
  for (long o = blockIdx.x * (long)kT + threadIdx.x; o < n; o += (long)gridDim.x * kT) {
    const int x = (int)(o % W), y = (int)((o / W) % H);
    const int c = (int)((o / ((long)H * W)) % C);
    const int b = (int)(o / ((long)H * W * C));
    const float* s = dout + (((long)b * Cd + c0 + c) * (2 * H) + 2 * y) * W2 + 2 * x;
    din[o] = (s[0] + s[1]) + (s[W2] + s[W2 + 1]);
  }
}

__global__ __launch_bounds__(kT) void copy_channels_kernel(const float* __restrict__ src,
                                                           float* __restrict__ dst, int B, int C,
                                                           int Cs, int s0, int Cd, int d0,
                                                           int HW) {
  const long n = (long)B * C * HW;
  for (long o = blockIdx.x * (long)kT + threadIdx.x; o < n; o += (long)gridDim.x * kT) {
    const int p = (int)(o % HW);
    const int c = (int)((o / HW) % C);
    const int b = (int)(o / ((long)HW * C));
    dst[((long)b * Cd + d0 + c) * HW + p] = src[((long)b * Cs + s0 + c) * HW + p];
  }
}

__global__ __launch_bounds__(kT) void add_kernel(const float* __restrict__ src,
                                                 float* __restrict__ dst, long n) {
  for (long o = blockIdx.x * (long)kT + threadIdx.x; o < n; o += (long)gridDim.x * kT)
    dst[o] += src[o];
}

constexpr int kMseBlocks = 1024;

__global__ __launch_bounds__(kT) void mse_kernel(const float* __restrict__ y,
                                                 const float* __restrict__ t,
                                                 float* __restrict__ dy, long n,
                                                 double* __restrict__ part) {
  __shared__ double red[kT];
  const float inv = 2.0f / (float)n;
  double acc = 0.0;
  for (long o = blockIdx.x * (long)kT + threadIdx.x; o < n; o += (long)gridDim.x * kT) {
    const float d = y[o] - t[o];
    dy[o] = inv * d;
    acc += (double)d * (double)d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = kT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kT) void mse_finish_kernel(const double* __restrict__ part, int nb,
                                                        long n, double* __restrict__ value) {
  __shared__ double red[kT];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kT) acc += part[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = kT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *value = red[0] / (double)n;
}

__global__ __launch_bounds__(kT) void adam_kernel(float* __restrict__ p,
                                                  const float* __restrict__ g,
                                                  float* __restrict__ m, float* __restrict__ v,
                                                  long n, float lr, float b1, float b2,
                                                  float eps, float c1, float c2) {
  for (long o = blockIdx.x * (long)kT + threadIdx.x; o < n; o += (long)gridDim.x * kT) {
    const float gi = g[o];
    const float mi = b1 * m[o] + (1.0f - b1) * gi;
    const float vi = b2 * v[o] + (1.0f - b2) * gi * gi;
    m[o] = mi;
    v[o] = vi;
    p[o] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// Round to BF16, nearest even (finite values).
__global__ __launch_bounds__(kT) void to_bf16_kernel(const float* __restrict__ src,
                                                     unsigned short* __restrict__ dst, long n) {
  for (long o = blockIdx.x * (long)kT + threadIdx.x; o < n; o += (long)gridDim.x * kT) {
    const unsigned int u = __float_as_uint(src[o]);
    dst[o] = (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
  }
}

gvm_status launched(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return GVM_ERR_HIP;
  }
  return GVM_OK;
}

bool dims_ok(long a, long b, long c, long d, long e) {
  return a > 0 && b > 0 && c > 0 && d > 0 && e > 0 && a * b * c * d * e < (1L << 40);
}

}  // namespace
}  // namespace gvm

using namespace gvm;

namespace {

template <class WT>
gvm_status conv_fwd_launch(gvm_ctx* ctx, const float* in, const WT* w, const float* bias,
                           float* out, int B, int Ci, int Co, int H, int W, int K, int relu) {
  GVM_REQUIRE(ctx && in && w && bias && out, "gvm_nn_conv_fwd: null argument");
  GVM_REQUIRE(K == 1 || K == 3, "gvm_nn_conv_fwd: K must be 1 or 3");
  GVM_REQUIRE(dims_ok(B, Ci, Co, H, W), "gvm_nn_conv_fwd: bad sizes");
  const long n = (long)B * Co * H * W;
  if (K == 3)
    hipLaunchKernelGGL((conv_fwd_kernel<3, WT>), dim3(nblocks(n)), dim3(kT), 0, ctx->stream, in, w,
                       bias, out, B, Ci, Co, H, W, relu);
  else
    hipLaunchKernelGGL((conv_fwd_kernel<1, WT>), dim3(nblocks(n)), dim3(kT), 0, ctx->stream, in, w,
                       bias, out, B, Ci, Co, H, W, relu);
  return launched("gvm_nn_conv_fwd");
}

template <class WT>
gvm_status conv_bwd_launch(gvm_ctx* ctx, const float* in, const WT* w, const float* out,
                           const float* dout, float* din, float* dw, float* dbias, int B, int Ci,
                           int Co, int H, int W, int K, int relu) {
  GVM_REQUIRE(ctx && in && w && dout && dw && dbias, "gvm_nn_conv_bwd: null argument");
  GVM_REQUIRE(!relu || out, "gvm_nn_conv_bwd: the ReLU mask needs the forward output");
  GVM_REQUIRE(K == 1 || K == 3, "gvm_nn_conv_bwd: K must be 1 or 3");
  GVM_REQUIRE(dims_ok(B, Ci, Co, H, W), "gvm_nn_conv_bwd: bad sizes");
  GVM_REQUIRE((long)Co * Ci < (1L << 31), "gvm_nn_conv_bwd: too many channel pairs");
  const long nin = (long)B * Ci * H * W;
  if (din) {
    if (K == 3)
      hipLaunchKernelGGL((conv_bwd_data_kernel<3, WT>), dim3(nblocks(nin)), dim3(kT), 0,
                         ctx->stream, w, out, dout, din, B, Ci, Co, H, W, relu);
    else
      hipLaunchKernelGGL((conv_bwd_data_kernel<1, WT>), dim3(nblocks(nin)), dim3(kT), 0,
                         ctx->stream, w, out, dout, din, B, Ci, Co, H, W, relu);
  }
  if (K == 3)
    hipLaunchKernelGGL(conv_bwd_weight_kernel<3>, dim3(Co * Ci), dim3(kT), 0, ctx->stream, in, out,
                       dout, dw, B, Ci, Co, H, W, relu);
  else
    hipLaunchKernelGGL(conv_bwd_weight_kernel<1>, dim3(Co * Ci), dim3(kT), 0, ctx->stream, in, out,
                       dout, dw, B, Ci, Co, H, W, relu);
  hipLaunchKernelGGL(conv_bwd_bias_kernel, dim3(Co), dim3(kT), 0, ctx->stream, out, dout, dbias, B,
                     Co, (long)H * W, relu);
  return launched("gvm_nn_conv_bwd");
}

}  // namespace

extern "C" gvm_status gvm_nn_conv_fwd(gvm_ctx* ctx, const float* in, const float* w,
                                      const float* bias, float* out, int B, int Ci, int Co, int H,
                                      int W, int K, int relu) {
  return conv_fwd_launch(ctx, in, w, bias, out, B, Ci, Co, H, W, K, relu);
}

extern "C" gvm_status gvm_nn_conv_fwd_bf16(gvm_ctx* ctx, const float* in,
                                           const unsigned short* w, const float* bias, float* out,
                                           int B, int Ci, int Co, int H, int W, int K, int relu) {
  return conv_fwd_launch(ctx, in, w, bias, out, B, Ci, Co, H, W, K, relu);
}

extern "C" gvm_status gvm_nn_conv_bwd(gvm_ctx* ctx, const float* in, const float* w,
                                      const float* out, const float* dout, float* din, float* dw,
                                      float* dbias, int B, int Ci, int Co, int H, int W, int K,
                                      int relu) {
  return conv_bwd_launch(ctx, in, w, out, dout, din, dw, dbias, B, Ci, Co, H, W, K, relu);
}

extern "C" gvm_status gvm_nn_conv_bwd_bf16(gvm_ctx* ctx, const float* in,
                                           const unsigned short* w, const float* out,
                                           const float* dout, float* din, float* dw, float* dbias,
                                           int B, int Ci, int Co, int H, int W, int K, int relu) {
  return conv_bwd_launch(ctx, in, w, out, dout, din, dw, dbias, B, Ci, Co, H, W, K, relu);
}

extern "C" gvm_status gvm_nn_to_bf16(gvm_ctx* ctx, const float* src, unsigned short* dst, long n) {
  GVM_REQUIRE(ctx && src && dst && n > 0, "gvm_nn_to_bf16: bad argument");
  hipLaunchKernelGGL(to_bf16_kernel, dim3(nblocks(n)), dim3(kT), 0, ctx->stream, src, dst, n);
  return launched("gvm_nn_to_bf16");
}

extern "C" gvm_status gvm_nn_pool_fwd(gvm_ctx* ctx, const float* in, float* out, int B, int C,
                                      int H, int W) {
  GVM_REQUIRE(ctx && in && out, "gvm_nn_pool_fwd: null argument");
  GVM_REQUIRE(dims_ok(B, C, H, W, 1) && H % 2 == 0 && W % 2 == 0, "gvm_nn_pool_fwd: bad sizes");
  hipLaunchKernelGGL(pool_fwd_kernel, dim3(nblocks((long)B * C * H * W / 4)), dim3(kT), 0,
                     ctx->stream, in, out, (long)B * C, H, W);
  return launched("gvm_nn_pool_fwd");
}

extern "C" gvm_status gvm_nn_pool_bwd(gvm_ctx* ctx, const float* dout, float* din, int B, int C,
                                      int H, int W) {
  GVM_REQUIRE(ctx && dout && din, "gvm_nn_pool_bwd: null argument");
  GVM_REQUIRE(dims_ok(B, C, H, W, 1) && H % 2 == 0 && W % 2 == 0, "gvm_nn_pool_bwd: bad sizes");
  hipLaunchKernelGGL(pool_bwd_kernel, dim3(nblocks((long)B * C * H * W)), dim3(kT), 0, ctx->stream,
                     dout, din, (long)B * C, H, W);
  return launched("gvm_nn_pool_bwd");
}

extern "C" gvm_status gvm_nn_up_fwd(gvm_ctx* ctx, const float* in, float* out, int B, int C, int Cd,
                                    int c0, int H, int W) {
  GVM_REQUIRE(ctx && in && out, "gvm_nn_up_fwd: null argument");
  GVM_REQUIRE(dims_ok(B, C, H, W, 1) && c0 >= 0 && c0 + C <= Cd, "gvm_nn_up_fwd: bad sizes");
  hipLaunchKernelGGL(up_fwd_kernel, dim3(nblocks((long)B * C * H * W * 4)), dim3(kT), 0,
                     ctx->stream, in, out, B, C, Cd, c0, H, W);
  return launched("gvm_nn_up_fwd");
}

extern "C" gvm_status gvm_nn_up_bwd(gvm_ctx* ctx, const float* dout, float* din, int B, int C,
                                    int Cd, int c0, int H, int W) {
  GVM_REQUIRE(ctx && dout && din, "gvm_nn_up_bwd: null argument");
  GVM_REQUIRE(dims_ok(B, C, H, W, 1) && c0 >= 0 && c0 + C <= Cd, "gvm_nn_up_bwd: bad sizes");
  hipLaunchKernelGGL(up_bwd_kernel, dim3(nblocks((long)B * C * H * W)), dim3(kT), 0, ctx->stream,
                     dout, din, B, C, Cd, c0, H, W);
  return launched("gvm_nn_up_bwd");
}

extern "C" gvm_status gvm_nn_copy_channels(gvm_ctx* ctx, const float* src, float* dst, int B, int C,
                                           int Cs, int s0, int Cd, int d0, int HW) {
  GVM_REQUIRE(ctx && src && dst, "gvm_nn_copy_channels: null argument");
  GVM_REQUIRE(dims_ok(B, C, HW, 1, 1) && s0 >= 0 && s0 + C <= Cs && d0 >= 0 && d0 + C <= Cd,
              "gvm_nn_copy_channels: bad sizes");
  hipLaunchKernelGGL(copy_channels_kernel, dim3(nblocks((long)B * C * HW)), dim3(kT), 0,
                     ctx->stream, src, dst, B, C, Cs, s0, Cd, d0, HW);
  return launched("gvm_nn_copy_channels");
}

extern "C" gvm_status gvm_nn_add(gvm_ctx* ctx, const float* src, float* dst, long n) {
  GVM_REQUIRE(ctx && src && dst && n > 0, "gvm_nn_add: bad argument");
  hipLaunchKernelGGL(add_kernel, dim3(nblocks(n)), dim3(kT), 0, ctx->stream, src, dst, n);
  return launched("gvm_nn_add");
}

extern "C" gvm_status gvm_nn_mse(gvm_ctx* ctx, const float* y, const float* t, float* dy, long n,
                                 int slot) {
  GVM_REQUIRE(ctx && y && t && dy && n > 0, "gvm_nn_mse: bad argument");
  GVM_REQUIRE(slot >= 0 && slot < kValueSlots, "gvm_nn_mse: value slot out of range");
  // block partials in the last reduction slot of the context
  double* part = ctx->partials + (long)(kMaxSlots - 1) * kReduceBlocks;
  const int nb = std::min(kMseBlocks, nblocks(n));
  hipLaunchKernelGGL(mse_kernel, dim3(nb), dim3(kT), 0, ctx->stream, y, t, dy, n, part);
  hipLaunchKernelGGL(mse_finish_kernel, dim3(1), dim3(kT), 0, ctx->stream, part, nb, n,
                     ctx->dvalues + slot);
  return launched("gvm_nn_mse");
}

extern "C" gvm_status gvm_nn_adam(gvm_ctx* ctx, float* p, const float* g, float* m, float* v,
                                  long n, float lr, float beta1, float beta2, float eps, int t) {
  GVM_REQUIRE(ctx && p && g && m && v && n > 0 && t >= 1, "gvm_nn_adam: bad argument");
  const float c1 = 1.0f - std::pow(beta1, (float)t);
  const float c2 = 1.0f - std::pow(beta2, (float)t);
  hipLaunchKernelGGL(adam_kernel, dim3(nblocks(n)), dim3(kT), 0, ctx->stream, p, g, m, v, n, lr,
                     beta1, beta2, eps, c1, c2);
  return launched("gvm_nn_adam");
}
