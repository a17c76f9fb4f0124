// Layers of the warm-start U-Net (include/gvm_nn.h): convolutions with
// optional ReLU, 2x2 average pooling, nearest 2x upsampling into a channel
// block, channel copies, the MSE loss and Adam, forward and backward.
//
// Convolutions are implicit GEMMs on the FP32 MFMA units (forward, input
// gradient, and weight gradient in pixel ranges whose partials are summed in
// FP64 in a fixed order); bias gradients and the loss are FP64 fixed-order
// reductions, so training is reproducible bit for bit.  Layouts are NCHW.
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

// ---------------------------------------------------------------------------
// Convolutions as implicit GEMMs on the FP32 MFMA units
// (v_mfma_f32_32x32x2_f32: FP32 products, FP32 accumulation).
//
//   forward      out[co][q]   = sum_{ci,tap} w[co][ci][tap] in[ci][q + tap]
//                M = Co, N = pixels (b, y, x), K = Ci K K
//   input grad   din[ci][q]   = sum_{co,tap} w[co][ci][tap] g[co][q - tap]
//                M = Ci, N = pixels, K = Co K K
//   weight grad  dw[co][ci,tap] = sum_q g[co][q] in[ci][q + tap]
//                M = Co, N = Ci K K, K = pixels, split in ranges of
//                kWgradPixels whose FP32 partials are summed in FP64 in a
//                fixed order (bitwise reproducible)
//
// tap (ky, kx) shifts a pixel by (ky - K/2, kx - K/2), zero outside the
// image; g is the output gradient masked by the ReLU of the forward output.
// A workgroup (4 waves) computes a 64 x 64 tile of M x N, each wave a 32 x
// 32 block; K advances in chunks of 16 staged in LDS (k-major, so the MFMA
// operand reads of a wave are consecutive), the next chunk's values loaded
// into registers while the current one is multiplied.
// ---------------------------------------------------------------------------
typedef float f32x16_t __attribute__((ext_vector_type(16)));
enum { CONV_FWD = 0, CONV_BWD_DATA = 1, CONV_BWD_WEIGHT = 2 };
constexpr int kCTile = 64;
constexpr int kCK = 16;
constexpr int kWgradPixels = 8192;

template <class WT>
struct ConvArgs {
  const float* in;    // [B][Ci][H][W]
  const WT* w;        // [Co][Ci][K][K]
  const float* bias;  // [Co] (forward)
  const float* out;   // forward output (ReLU mask), or NULL
  const float* dout;  // [B][Co][H][W] gradient at the activated output
  float* dst;         // out (forward), din (input grad) or the weight-grad partials
  int B, Ci, Co, H, W, relu;
};

template <int K, class WT>
__device__ __forceinline__ float conv_in_at(const ConvArgs<WT>& a, int ci, int q, int tap) {
  // in[b][ci][y + ky - P][x + kx - P] for pixel q = (b, y, x), 0 outside
  constexpr int P = K / 2;
  const int HW = a.H * a.W;
  const int b = q / HW, p = q - b * HW;
  const int yy = p / a.W + tap / K - P, xx = p % a.W + tap % K - P;
  if (yy < 0 || yy >= a.H || xx < 0 || xx >= a.W) return 0.0f;
  return a.in[((long)b * a.Ci + ci) * HW + (long)yy * a.W + xx];
}

template <class WT>
__device__ __forceinline__ float conv_g_at(const ConvArgs<WT>& a, int co, int q) {
  // the masked output gradient at pixel q = (b, y, x) of output channel co
  const int HW = a.H * a.W;
  const int b = q / HW, p = q - b * HW;
  const long i = ((long)b * a.Co + co) * HW + p;
  return relu_grad(a.out, a.dout, i, a.relu);
}

template <int MODE, int K, class WT>
__device__ __forceinline__ float conv_a(const ConvArgs<WT>& a, int m, int k, int M, int Kd) {
  if (m >= M || k >= Kd) return 0.0f;
  constexpr int KK = K * K;
  if constexpr (MODE == CONV_FWD) return wload(a.w, (long)m * Kd + k);
  if constexpr (MODE == CONV_BWD_DATA) {
    const int co = k / KK, tap = k - co * KK;  // w[co][ci = m][tap]
    return wload(a.w, ((long)co * a.Ci + m) * KK + tap);
  }
  return conv_g_at(a, m, k);  // weight grad: g[co = m][pixel k]
}

template <int MODE, int K, class WT>
__device__ __forceinline__ float conv_b(const ConvArgs<WT>& a, int k, int n, int Kd, int Nd) {
  if (k >= Kd || n >= Nd) return 0.0f;
  constexpr int KK = K * K;
  if constexpr (MODE == CONV_FWD) {
    const int ci = k / KK;
    return conv_in_at<K>(a, ci, n, k - ci * KK);
  }
  if constexpr (MODE == CONV_BWD_DATA) {
    // g[co][q - tap] with k = (co, tap)
    constexpr int P = K / 2;
    const int co = k / KK, tap = k - co * KK;
    const int HW = a.H * a.W;
    const int b = n / HW, p = n - b * HW;
    const int yy = p / a.W - (tap / K - P), xx = p % a.W - (tap % K - P);
    if (yy < 0 || yy >= a.H || xx < 0 || xx >= a.W) return 0.0f;
    return relu_grad(a.out, a.dout, ((long)b * a.Co + co) * HW + (long)yy * a.W + xx, a.relu);
  }
  const int ci = n / KK;  // weight grad: in[ci][pixel k + tap], n = (ci, tap)
  return conv_in_at<K>(a, ci, k, n - ci * KK);
}

template <int MODE, int K, class WT>
__global__ __launch_bounds__(256) void conv_mfma_kernel(ConvArgs<WT> a) {
  constexpr int KK = K * K;
  const int HW = a.H * a.W;
  const int npix = a.B * HW;
  int M, Nd, Kd, kbeg;
  if constexpr (MODE == CONV_FWD) {
    M = a.Co; Nd = npix; Kd = a.Ci * KK; kbeg = 0;
  } else if constexpr (MODE == CONV_BWD_DATA) {
    M = a.Ci; Nd = npix; Kd = a.Co * KK; kbeg = 0;
  } else {
    M = a.Co; Nd = a.Ci * KK; kbeg = blockIdx.z * kWgradPixels;
    Kd = min(npix, kbeg + kWgradPixels);
  }
  __shared__ float As[kCK][kCTile + 4];
  __shared__ float Bs[kCK][kCTile + 4];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m0 = blockIdx.y * kCTile, n0 = blockIdx.x * kCTile;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  // loader coordinates: 4 A and 4 B values per thread per chunk.  The
  // weight grad reads along pixels (its K) for both operands, the others
  // along k for A and along pixels (N) for B.
  float ra[4], rb[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int e = t + 256 * j;
      const int am = e / kCK, ak = e % kCK;
      ra[j] = conv_a<MODE, K>(a, m0 + am, k0 + ak, M, Kd);
      int bk, bn;
      if constexpr (MODE == CONV_BWD_WEIGHT) {
        bk = e % kCK;
        bn = e / kCK;
      } else {
        bk = e / kCTile;
        bn = e % kCTile;
      }
      rb[j] = conv_b<MODE, K>(a, k0 + bk, n0 + bn, Kd, Nd);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int e = t + 256 * j;
      As[e % kCK][e / kCK] = ra[j];
      if constexpr (MODE == CONV_BWD_WEIGHT)
        Bs[e % kCK][e / kCK] = rb[j];
      else
        Bs[e / kCTile][e % kCTile] = rb[j];
    }
  };
  f32x16_t acc;
#pragma unroll
  for (int e = 0; e < 16; e++) acc[e] = 0.0f;
  const int lr = lane & 31, lh = lane >> 5;
  load(kbeg);
  for (int k0 = kbeg; k0 < Kd; k0 += kCK) {
    __syncthreads();  // the previous chunk's reads are done
    store();
    __syncthreads();
    if (k0 + kCK < Kd) load(k0 + kCK);
#pragma unroll
    for (int s = 0; s < kCK / 2; s++)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(As[2 * s + lh][wm + lr], Bs[2 * s + lh][wn + lr],
                                                 acc, 0, 0, 0);
  }
  // 32 x 32 accumulator layout: lane l, register e -> row 8 (e / 4) + 4 (l / 32)
  // + e % 4 of the wave's block, column l % 32
#pragma unroll
  for (int e = 0; e < 16; e++) {
    const int m = m0 + wm + 8 * (e >> 2) + 4 * lh + (e & 3);
    const int n = n0 + wn + lr;
    if (m >= M || n >= Nd) continue;
    if constexpr (MODE == CONV_BWD_WEIGHT) {
      a.dst[((long)blockIdx.z * M + m) * Nd + n] = acc[e];
    } else {
      const int b = n / HW, p = n - b * HW;
      float v = acc[e];
      if constexpr (MODE == CONV_FWD) {
        v = v + a.bias[m];
        if (a.relu) v = fmaxf(v, 0.0f);
      }
      a.dst[((long)b * M + m) * HW + p] = v;
    }
  }
}

// dw = the FP64 sum of the weight-grad partials over the pixel ranges, in
// range order.
__global__ __launch_bounds__(kT) void wgrad_reduce_kernel(const float* __restrict__ part, int nsplit,
                                                          long n, float* __restrict__ dw) {
  for (long i = blockIdx.x * (long)kT + threadIdx.x; i < n; i += (long)gridDim.x * kT) {
    double s = 0.0;
    for (int r = 0; r < nsplit; r++) s += (double)part[(long)r * n + i];
    dw[i] = (float)s;
  }
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
  GVM_REQUIRE((long)B * H * W < (1L << 31) && (long)Ci * K * K < (1L << 24) &&
                  (long)Co * K * K < (1L << 24),
              "gvm_nn_conv_fwd: sizes beyond the kernels' 32-bit indices");
  const ConvArgs<WT> a{in, w, bias, nullptr, nullptr, out, B, Ci, Co, H, W, relu};
  const dim3 grid((unsigned)((B * H * W + kCTile - 1) / kCTile),
                  (unsigned)((Co + kCTile - 1) / kCTile));
  if (K == 3)
    hipLaunchKernelGGL((conv_mfma_kernel<CONV_FWD, 3, WT>), grid, dim3(256), 0, ctx->stream, a);
  else
    hipLaunchKernelGGL((conv_mfma_kernel<CONV_FWD, 1, WT>), grid, dim3(256), 0, ctx->stream, a);
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
  GVM_REQUIRE((long)B * H * W < (1L << 31) && (long)Ci * K * K < (1L << 24) &&
                  (long)Co * K * K < (1L << 24),
              "gvm_nn_conv_bwd: sizes beyond the kernels' 32-bit indices");
  const int npix = B * H * W;
  if (din) {
    const ConvArgs<WT> a{in, w, nullptr, out, dout, din, B, Ci, Co, H, W, relu};
    const dim3 grid((unsigned)((npix + kCTile - 1) / kCTile),
                    (unsigned)((Ci + kCTile - 1) / kCTile));
    if (K == 3)
      hipLaunchKernelGGL((conv_mfma_kernel<CONV_BWD_DATA, 3, WT>), grid, dim3(256), 0, ctx->stream,
                         a);
    else
      hipLaunchKernelGGL((conv_mfma_kernel<CONV_BWD_DATA, 1, WT>), grid, dim3(256), 0, ctx->stream,
                         a);
  }
  // weight gradient: FP32 partials per range of kWgradPixels pixels in the
  // context's network workspace, then their FP64 sum in range order
  const int nsplit = (npix + kWgradPixels - 1) / kWgradPixels;
  const long nw = (long)Co * Ci * K * K;
  const size_t need = (size_t)nsplit * nw;
  if (need > ctx->nn_part_elems) {
    GVM_REQUIRE(!ctx->capturing, "gvm_nn_conv_bwd: the workspace would grow under capture");
    GVM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->nn_part) GVM_HIP_TRY(hipFree(ctx->nn_part));
    ctx->nn_part = nullptr;
    ctx->nn_part_elems = 0;
    GVM_HIP_TRY(hipMalloc((void**)&ctx->nn_part, sizeof(float) * need));
    ctx->nn_part_elems = need;
  }
  {
    const ConvArgs<WT> a{in, w, nullptr, out, dout, ctx->nn_part, B, Ci, Co, H, W, relu};
    const dim3 grid((unsigned)((Ci * K * K + kCTile - 1) / kCTile),
                    (unsigned)((Co + kCTile - 1) / kCTile), (unsigned)nsplit);
    if (K == 3)
      hipLaunchKernelGGL((conv_mfma_kernel<CONV_BWD_WEIGHT, 3, WT>), grid, dim3(256), 0,
                         ctx->stream, a);
    else
      hipLaunchKernelGGL((conv_mfma_kernel<CONV_BWD_WEIGHT, 1, WT>), grid, dim3(256), 0,
                         ctx->stream, a);
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(nblocks(nw)), dim3(kT), 0, ctx->stream,
                       ctx->nn_part, nsplit, nw, dw);
  }
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
