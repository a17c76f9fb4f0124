// Forward model and chi2 on gfx950.
//
// Per slice (channel x stokes) the reference runs calculateInu, apply_beam2I,
// cuFFT C2C (inverse, unnormalised), phase_rotate, vis_mod, residual and
// chi2Vector + a float tree reduction as eight separate launches with a device
// synchronisation after each (src/functions.cu:4323-4454).  Here every step
// covers all slices of the context in one launch (blockIdx.y = slice, one uv
// grid per slice, per-slice descriptors in a device table):
//
//  1. fft_pass<rows>     clip2IWNoise in place on the caller's image, then
//                        I_nu = clip(I0 (nu/nu0)^alpha), times PB and
//                        fg_scale, then the row transforms (Stockham
//                        radix-8/4 in LDS), times the column-frequency
//                        phase-rotation factor;
//  2. fft_pass<cols>     column transforms in LDS, times the row-frequency
//                        phase-rotation factor, in place on the grid;
//  3. degrid_chi2        bilinear 4-tap gather, V_r = V_o - V_m, w |V_r|^2
//                        summed per block in FP64, and max |w V_r| for the
//                        adjoint's FP16 scaling;
//  4. finish             fixed-order FP64 sum of the block partials.
// For up to four slices step 3 and the prior terms' value partials are one
// launch (values_kernel), and step 4 with the priors' finish one more.
// The polar form of c = w conj(V_r) that the adjoint reads is written by the
// adjoint itself (polar_kernel), so value-only evaluations (line-search
// probes) do not pay for it.
//
// The transform is e^{+2 pi i} (CUFFT_INVERSE) and unnormalised.  The phase
// rotation e^{-2 pi i (u x0/M + v y0/N)} (src/functions.cu:2483-2518)
// factorises per axis, so each pass applies its own factor to its output.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gvm_internal.h"
#include "gvm_prior_terms.h"

namespace gvm {

// ---------------------------------------------------------------------------
// clip2IWNoise: src/functions.cu:2694-2719, done by the row pass as it reads
// the image (one pass and one launch less per evaluation)
// ---------------------------------------------------------------------------
struct ClipArgs {
  const float* noise;
  unsigned int* cmax;  // per-slice max |c|, cleared here for the degrid kernel
  float noise_cut;
  float threshold;
  int eta_pos;
  int flag_opt;
  int nsl;
};

// The clipped (I0, alpha) of pixel k; writes them back to the caller's image
// when `store` (one block per pixel does).  Blocks of other slices read the
// same pixel and may see it before or after that store: the clip is
// idempotent, so every block computes the same values.
__device__ __forceinline__ void clip_pixel(float* __restrict__ I, long n, long k,
                                           const ClipArgs& c, float floor_v, bool store,
                                           float* i0, float* alpha) {
  float a = I[k], b = I[n + k];
  if (c.noise[k] > c.noise_cut) {
    a = c.eta_pos ? 0.0f : floor_v;
    b = 0.0f;
    if (store) {
      I[k] = a;
      I[n + k] = b;
    }
  } else if (a < c.threshold && c.flag_opt > 0) {
    b = 0.0f;
    if (store) I[n + k] = b;
  }
  *i0 = a;
  *alpha = b;
}

// ---------------------------------------------------------------------------
// Stockham FFT in LDS
// ---------------------------------------------------------------------------
// Complex product, each operation rounded (no multiply-add contraction), so
// the transform's rounding does not depend on how the compiler schedules the
// unrolled stages of a given batching.
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
#pragma clang fp contract(off)
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// One Stockham stage (radix R, current sub-transform length NS) over B
// transforms of length N held in LDS.  COLS selects the LDS layout:
//   rows: element (t, p) at t * N + p;  cols: element (t, p) at p * B + t.
// All reads of the stage complete (barrier) before any write, so the stage
// runs in place in one buffer.
// Radix-4 butterfly in place: b_m = sum_r a_r e^{+2 pi i r m / 4}.
__device__ __forceinline__ void bfly4(float2& a0, float2& a1, float2& a2, float2& a3) {
#pragma clang fp contract(off)
  const float2 s02 = make_float2(a0.x + a2.x, a0.y + a2.y);
  const float2 d02 = make_float2(a0.x - a2.x, a0.y - a2.y);
  const float2 s13 = make_float2(a1.x + a3.x, a1.y + a3.y);
  const float2 d13 = make_float2(a1.x - a3.x, a1.y - a3.y);
  a0 = make_float2(s02.x + s13.x, s02.y + s13.y);
  a2 = make_float2(s02.x - s13.x, s02.y - s13.y);
  // i * d13 = (-d13.y, d13.x)
  a1 = make_float2(d02.x - d13.y, d02.y + d13.x);
  a3 = make_float2(d02.x + d13.y, d02.y - d13.x);
}

// One Stockham stage (radix R = 2, 4 or 8, current sub-transform length NS)
// over B transforms of length N held in LDS.  A thread takes butterflies
// q = threadIdx.x + b T while q < B N / R.
template <int N, int B, bool COLS, int R, int NS, int T>
__device__ __forceinline__ void stockham_stage(float2* lds, const float2* __restrict__ tw) {
#pragma clang fp contract(off)
  constexpr int NB = B * N / R;          // butterflies in this stage
  constexpr int BPT = (NB + T - 1) / T;  // per thread (the last round may be partial)
  constexpr int STRIDE = N / R;
  float2 v[BPT][R];
  int base_t[BPT], out_j[BPT];
#pragma unroll
  for (int b = 0; b < BPT; b++) {
    const int q = threadIdx.x + b * T;
    if (NB % T != 0 && q >= NB) break;
    const int t = COLS ? (q % B) : (q / STRIDE);
    const int j = COLS ? (q / B) : (q % STRIDE);
    const int k = j % NS;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int p = j + r * STRIDE;
      v[b][r] = COLS ? lds[p * B + t] : lds[t * N + p];
    }
    if (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) {
        const int e = r * k * (N / (NS * R));
        v[b][r] = cmul(v[b][r], tw[e]);
      }
    }
    if constexpr (R == 8) {
      // b_m = sum_r a_r e^{+2 pi i r m / 8}: radix-4 over the even and the
      // odd inputs, then b_m = E_m + W^m O_m, b_{m+4} = E_m - W^m O_m with
      // W = e^{+2 pi i / 8}
      float2 e0 = v[b][0], e1 = v[b][2], e2 = v[b][4], e3 = v[b][6];
      float2 o0 = v[b][1], o1 = v[b][3], o2 = v[b][5], o3 = v[b][7];
      bfly4(e0, e1, e2, e3);
      bfly4(o0, o1, o2, o3);
      constexpr float h = 0.70710678118654752f;  // sqrt(1/2)
      const float2 w1 = make_float2(h * (o1.x - o1.y), h * (o1.x + o1.y));     // W o1
      const float2 w2 = make_float2(-o2.y, o2.x);                              // i o2
      const float2 w3 = make_float2(-h * (o3.x + o3.y), h * (o3.x - o3.y));    // W^3 o3
      v[b][0] = make_float2(e0.x + o0.x, e0.y + o0.y);
      v[b][4] = make_float2(e0.x - o0.x, e0.y - o0.y);
      v[b][1] = make_float2(e1.x + w1.x, e1.y + w1.y);
      v[b][5] = make_float2(e1.x - w1.x, e1.y - w1.y);
      v[b][2] = make_float2(e2.x + w2.x, e2.y + w2.y);
      v[b][6] = make_float2(e2.x - w2.x, e2.y - w2.y);
      v[b][3] = make_float2(e3.x + w3.x, e3.y + w3.y);
      v[b][7] = make_float2(e3.x - w3.x, e3.y - w3.y);
    } else if constexpr (R == 4) {
      bfly4(v[b][0], v[b][1], v[b][2], v[b][3]);
    } else {
      const float2 a0 = v[b][0], a1 = v[b][1];
      v[b][0] = make_float2(a0.x + a1.x, a0.y + a1.y);
      v[b][1] = make_float2(a0.x - a1.x, a0.y - a1.y);
    }
    base_t[b] = t;
    out_j[b] = (j / NS) * NS * R + k;
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BPT; b++) {
    if (NB % T != 0 && threadIdx.x + b * T >= NB) break;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int p = out_j[b] + r * NS;
      if (COLS)
        lds[p * B + base_t[b]] = v[b][r];
      else
        lds[base_t[b] * N + p] = v[b][r];
    }
  }
  __syncthreads();
}

template <int N>
constexpr int log2_of() {
  int l = 0;
  int n = N;
  while (n > 1) {
    n >>= 1;
    l++;
  }
  return l;
}

// Radix of the stage that starts at sub-transform length NS: radix-8 stages
// first, then radix-4 (two when log2 N = 1 mod 3, one when 2 mod 3).
template <int N, int NS>
constexpr int stage_radix() {
  constexpr int left = log2_of<N>() - log2_of<NS>();  // log2 of what remains
  if (left % 3 == 0) return 8;
  if (left % 3 == 1) return left >= 4 ? (left == 4 ? 4 : 8) : 2;
  return left == 2 ? 4 : 8;
}

template <int N, int B, bool COLS, int NS, int T>
__device__ __forceinline__ void fft_stages(float2* lds, const float2* __restrict__ tw) {
  if constexpr (NS < N) {
    constexpr int R = stage_radix<N, NS>();
    stockham_stage<N, B, COLS, R, NS, T>(lds, tw);
    fft_stages<N, B, COLS, NS * R, T>(lds, tw);
  }
}

// N-point transforms (e^{+2 pi i}, unnormalised) of the B sequences in LDS:
// N = 8^a 4^b with b <= 2 (2 for N = 2), e.g. 1024 = 8 8 4 4.
template <int N, int B, bool COLS, int T>
__device__ __forceinline__ void fft_lds(float2* lds, const float2* __restrict__ tw) {
  fft_stages<N, B, COLS, 1, T>(lds, tw);
}

// Row pass: clip2IWNoise on the rows' pixels (written back by the slice-0
// blocks), then I_nu * PB * fg_scale [* GCF] for B rows of slice blockIdx.y
// (calculateInu + apply_beam2I, src/functions.cu:3939-3966, 2424-2444),
// transform each row, multiply by rot_x[column frequency], write the rows of
// that slice's grid.
template <int N, int B, int T>
__global__ __launch_bounds__(T) void fft_rows_kernel(
    float* __restrict__ I, ClipArgs clip, const SliceDev* __restrict__ tab, float fg_scale,
    float floor_v, const float* __restrict__ gcf, const float2* __restrict__ tw,
    const float2* __restrict__ rot_x, float2* __restrict__ grids) {
  __shared__ float2 lds[B * N];
  const long n = (long)N * N;
  const float* __restrict__ pb = tab[blockIdx.y].pb;
  const float nudiv = tab[blockIdx.y].nudiv;
  float2* __restrict__ grid = grids + (long)blockIdx.y * n;
  const int row0 = blockIdx.x * B;
  if (blockIdx.x == 0 && blockIdx.y == 0)
    for (int s = threadIdx.x; s < clip.nsl; s += T) clip.cmax[s] = 0u;
  for (int idx = threadIdx.x; idx < B * N; idx += T) {
    const long k = (long)(row0 + idx / N) * N + (idx % N);
    float i0, alpha;
    clip_pixel(I, n, k, clip, floor_v, blockIdx.y == 0, &i0, &alpha);
    float inu = i0 * powf(nudiv, alpha);
    if (inu < floor_v) inu = floor_v;
    float x = inu * pb[k] * fg_scale;
    if (gcf) x = x * gcf[k];  // apply_GCF (src/functions.cu:2468-2476)
    lds[idx] = make_float2(x, 0.0f);
  }
  __syncthreads();
  fft_lds<N, B, false, T>(lds, tw);
  for (int idx = threadIdx.x; idx < B * N; idx += T) {
    const int j = idx % N;
    grid[(long)(row0 + idx / N) * N + j] = cmul(lds[idx], rot_x[j]);
  }
}

// Column pass: B adjacent columns per block of slice blockIdx.y, in place.
template <int N, int B, int T>
__global__ __launch_bounds__(T) void fft_cols_kernel(float2* __restrict__ grids,
                                                       const float2* __restrict__ tw,
                                                       const float2* __restrict__ rot_y) {
  __shared__ float2 lds[B * N];
  float2* __restrict__ grid = grids + (long)blockIdx.y * N * N;
  // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs
  // (blockIdx.x % 8), each with its own L2; the blocks whose column groups
  // share 128-byte lines are given consecutive groups on one XCD, so a line
  // is fetched into one L2 instead of up to eight.
  const int nbx = gridDim.x;
  const int gx = (nbx % 8 == 0) ? (int)(blockIdx.x % 8) * (nbx / 8) + (int)(blockIdx.x / 8)
                                : (int)blockIdx.x;
  const int col0 = gx * B;
  for (int idx = threadIdx.x; idx < B * N; idx += T) {
    const int p = idx / B, t = idx % B;
    lds[idx] = grid[(long)p * N + col0 + t];
  }
  __syncthreads();
  fft_lds<N, B, true, T>(lds, tw);
  for (int idx = threadIdx.x; idx < B * N; idx += T) {
    const int p = idx / B, t = idx % B;
    grid[(long)p * N + col0 + t] = cmul(lds[idx], rot_y[p]);
  }
}

template <int N, int BR, int TR, int BC, int TC>
gvm_status launch_fft_with(gvm_ctx* ctx, float* I, const ClipArgs& clip, int nsl, float floor_v) {
  hipLaunchKernelGGL((fft_rows_kernel<N, BR, TR>), dim3(N / BR, nsl), dim3(TR), 0, ctx->stream, I,
                     clip, ctx->slice_tab, ctx->g.fg_scale, floor_v, ctx->gcf, ctx->twiddle,
                     ctx->rot_x, ctx->grid);
  GVM_CHECK_LAUNCH();
  hipLaunchKernelGGL((fft_cols_kernel<N, BC, TC>), dim3(N / BC, nsl), dim3(TC), 0, ctx->stream,
                     ctx->grid, ctx->twiddle, ctx->rot_y);
  GVM_CHECK_LAUNCH();
  return GVM_OK;
}

// Two batchings of the same transforms (identical arithmetic, so identical
// results): "wide" blocks of 256 threads over 4096 / N rows and 8192 / N
// columns, for many slices; "narrow" blocks of 1 - 4 rows or 2 - 4 columns
// with 64 - 256 threads, so a single small image still spreads over the
// CUs.  Automatic choice (ctx->fft_mode 0): narrow when the wide row pass
// would launch fewer than 512 blocks.
template <int N>
gvm_status launch_fft(gvm_ctx* ctx, float* I, const ClipArgs& clip, int nsl, float floor_v) {
  constexpr int BR = (4096 / N) > 0 ? (4096 / N) : 1;
  constexpr int BC0 = 8192 / N;
  constexpr int BC = BC0 < 2 ? 2 : (BC0 > N ? N : BC0);
  if constexpr (N <= 1024) {
    constexpr int BRn = 256 / N > 1 ? 256 / N : 1;
    constexpr int TRn = BRn * N / 4 < 256 ? BRn * N / 4 : 256;
    constexpr int BCn = 256 / N > 2 ? 256 / N : 2;
    constexpr int TCn = BCn * N / 4 < 256 ? BCn * N / 4 : 256;
    const bool narrow = ctx->fft_mode == 2 || (ctx->fft_mode == 0 && (long)nsl * (N / BR) < 512);
    if (narrow) return launch_fft_with<N, BRn, TRn, BCn, TCn>(ctx, I, clip, nsl, floor_v);
  }
  return launch_fft_with<N, BR, 256, BC, 256>(ctx, I, clip, nsl, floor_v);
}

// clip2IWNoise on the caller's image and the phase-rotated uv grids of all
// slices (ctx->grid, one N x N grid each).
gvm_status model_grids(gvm_ctx* ctx, float* I, int nsl, int flag_opt) {
  const gvm_geometry& g = ctx->g;
  const float floor_v = (float)(-1.0 * (double)g.eta * (double)g.minpix);
  const ClipArgs clip{ctx->noise, ctx->cmax, g.noise_cut, g.threshold, g.eta > 0.0f ? 1 : 0,
                      flag_opt, nsl};
  switch (g.N) {
    case 64: return launch_fft<64>(ctx, I, clip, nsl, floor_v);
    case 128: return launch_fft<128>(ctx, I, clip, nsl, floor_v);
    case 256: return launch_fft<256>(ctx, I, clip, nsl, floor_v);
    case 512: return launch_fft<512>(ctx, I, clip, nsl, floor_v);
    case 1024: return launch_fft<1024>(ctx, I, clip, nsl, floor_v);
    case 2048: return launch_fft<2048>(ctx, I, clip, nsl, floor_v);
    case 4096: return launch_fft<4096>(ctx, I, clip, nsl, floor_v);
  }
  set_error("model_grid: unsupported N");
  return GVM_ERR_ARG;
}

// ---------------------------------------------------------------------------
// vis_mod + residual + chi2Vector + reduction:
// src/functions.cu:2557-2609, 2663-2671, 2867-2876.  Bytes per visibility:
// cell 4 + frac 8 + Vo 8 + weight 4 read, Vr 8 written.
// ---------------------------------------------------------------------------
constexpr int kDegridThreads = 512;

// One degrid block: block bx of nbx over slice `slice` (blockIdx.x and
// gridDim.x of degrid_chi2_kernel), its partial into partials[slice][bx].
__device__ __forceinline__ void degrid_block(const float2* __restrict__ grids, int N,
                                             const SliceDev* __restrict__ tab,
                                             double* __restrict__ partials,
                                             unsigned int* __restrict__ cmaxs, int bx, int nbx,
                                             int slice, double* sh) {
  const SliceDev& sd = tab[slice];
  const float2* __restrict__ grid = grids + (long)slice * N * N;
  const int* __restrict__ cell = sd.cell;
  const float2* __restrict__ frac = sd.frac;
  const float2* __restrict__ vo = sd.vo;
  const float* __restrict__ weight = sd.weight;
  float2* __restrict__ vr = sd.vr;
  const long nvis = sd.nvis;
  double acc = 0.0;
  float cm = 0.0f;
  const int mask = N - 1;
  // Grid-stride loop, two strides per pass: the loads of two visibilities
  // are issued before their arithmetic, and they are accumulated in stride
  // order, so the per-thread sums are those of the one-at-a-time loop.  (Two
  // rather than four: 46 instead of 108 VGPRs, and with the grid below the
  // fused value launch takes 16.9 instead of 26.0 us at 1.76M visibilities.)
  const long S = (long)nbx * kDegridThreads;
  for (long v0 = bx * (long)kDegridThreads + threadIdx.x; v0 < nvis; v0 += 2 * S) {
    int c[2];
    float2 f[2], o[2];
    float w[2];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const long v = v0 + u * S;
      const bool ok = v < nvis;
      c[u] = ok ? ldg(cell + v) : -1;
      f[u] = ok ? ldg(frac + v) : make_float2(0.f, 0.f);
      o[u] = ok ? ldg(vo + v) : make_float2(0.f, 0.f);
      w[u] = ok ? ldg(weight + v) : 0.0f;
    }
    float2 t11[2], t12[2], t21[2], t22[2];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      t11[u] = t12[u] = t21[u] = t22[u] = make_float2(0.f, 0.f);
      if (c[u] >= 0) {
        const int i1 = c[u] & 0xffff, j1 = c[u] >> 16;
        const int i2 = (i1 + 1) & mask, j2 = (j1 + 1) & mask;
        t11[u] = grid[(long)N * j1 + i1];
        t12[u] = grid[(long)N * j2 + i1];
        t21[u] = grid[(long)N * j1 + i2];
        t22[u] = grid[(long)N * j2 + i2];
      }
    }
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const long v = v0 + u * S;
      if (v >= nvis) break;
      float2 m = make_float2(0.f, 0.f);
      if (c[u] >= 0) {
        const float2 fu = f[u];
        const float w11 = (1.0f - fu.x) * (1.0f - fu.y);
        const float w12 = (1.0f - fu.x) * fu.y;
        const float w21 = fu.x * (1.0f - fu.y);
        const float w22 = fu.x * fu.y;
        m.x = w11 * t11[u].x + w12 * t12[u].x + w21 * t21[u].x + w22 * t22[u].x;
        m.y = w11 * t11[u].y + w12 * t12[u].y + w21 * t21[u].y + w22 * t22[u].y;
      }
      const float2 r = make_float2(o[u].x - m.x, o[u].y - m.y);
      stg(vr + v, r);
      const float a2 = resid_abs2(r);
      acc += (double)(w[u] * a2);
      cm = fmaxf(cm, w[u] * sqrtf(a2));  // |c|, as polar_kernel computes it
    }
  }
  const double tot = block_sum(acc, sh);
  __syncthreads();
  const double mx = block_max((double)cm, sh);
  if (threadIdx.x == 0) {
    partials[(long)slice * kReduceBlocks + bx] = tot;
    atomicMax(cmaxs + slice, __float_as_uint((float)mx));
  }
}

__global__ __launch_bounds__(kDegridThreads) void degrid_chi2_kernel(
    const float2* __restrict__ grids, int N, const SliceDev* __restrict__ tab,
    double* __restrict__ partials, unsigned int* __restrict__ cmaxs) {
  __shared__ double sh[kDegridThreads / 64];
  degrid_block(grids, N, tab, partials, cmaxs, blockIdx.x, gridDim.x, blockIdx.y, sh);
}

// 0.5 * the sum over slices of the per-slice sums (each divided by its
// visibility count if normalize), in slice order, in FP64 (src/functions.cu:
// 4377-4453): one thread, so the order is fixed.
__global__ void chi2_combine_kernel(const double* __restrict__ sums,
                                    const SliceDev* __restrict__ tab, int nsl, int normalize,
                                    double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double total = 0.0;
  for (int s = 0; s < nsl; s++) {
    double r = sums[s];
    if (normalize) r /= (double)tab[s].nvis;
    total += r;
  }
  *out = 0.5 * total;
}

// Sum over the 256 threads of group threadIdx.x / 256 of a 512-thread block
// with block_sum's arithmetic for a block of 256 (wave sums, then the group's
// four wave sums in order from 0.0); valid in the group's first thread.  sh
// holds kDegridThreads / 64 doubles; every thread of the block calls it.
__device__ __forceinline__ double group_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if ((threadIdx.x & (kReduceThreads - 1)) == 0)
    for (int w = 0; w < kReduceThreads / 64; ++w) r += sh[wid + w];
  return r;
}

struct ValuesArgs {
  const float2* grids;
  const SliceDev* tab;
  double* partials;
  unsigned int* cmaxs;
  const float* I;
  const float* noise;
  PriorTermsDev T;
  float noise_cut;
  int N;
  int nb_vis;      // degrid blocks per slice (degrid_chi2_kernel's gridDim.x)
  int nb_prior;    // prior blocks of 256 threads (priors_value_kernel's grid)
  int prior_row0;  // partial row of the first prior term
};

// The block partials of a value evaluation after the two FFT passes, in one
// launch:
//  - blocks [0, ceil(nb_prior / 2)): two of priors_value_kernel's 256-thread
//    blocks each (the value partials of every prior term on the clipped
//    image, which the row pass has written), with its arithmetic;
//  - the rest: degrid_chi2_kernel's blocks, slice after slice.
// The partials are those of the separate launches bit for bit.  (Finishing
// them in the last block to end, found by a ticket counter, took the kernel
// from 34 to 56 us: each block's release fence writes back its L2's dirty
// lines.)
__global__ __launch_bounds__(kDegridThreads) void values_kernel(ValuesArgs a) {
  __shared__ double sh[kDegridThreads / 64];
  const int np = (a.nb_prior + 1) / 2;
  const int b = blockIdx.x;
  if (b >= np) {
    const int d = b - np;
    degrid_block(a.grids, a.N, a.tab, a.partials, a.cmaxs, d % a.nb_vis, a.nb_vis, d / a.nb_vis,
                 sh);
    return;
  }
  const int g = threadIdx.x >> 8, vt = threadIdx.x & (kReduceThreads - 1);
  const int vb = 2 * b + g;
  const long n = (long)a.N * a.N;
  double acc[GVM_MAX_FUSED_PRIORS];
#pragma unroll
  for (int t = 0; t < GVM_MAX_FUSED_PRIORS; t++) acc[t] = 0.0;
  if (vb < a.nb_prior) {
    for (long k = vb * (long)kReduceThreads + vt; k < n; k += (long)a.nb_prior * kReduceThreads) {
      if (!(a.noise[k] < a.noise_cut)) continue;
      const int i = pix_row(k, a.N), j = pix_col(k, a.N);
#pragma unroll
      for (int t = 0; t < GVM_MAX_FUSED_PRIORS; t++)
        if (t < a.T.n && a.T.t[t].active)
          acc[t] += (double)prior_term(a.T.t[t].kind, a.I + n * a.T.t[t].index, a.N, i, j,
                                       a.T.t[t].prior, a.T.t[t].eta);
    }
  }
#pragma unroll
  for (int t = 0; t < GVM_MAX_FUSED_PRIORS; t++) {
    if (t >= a.T.n) break;
    const double tot = group_sum(acc[t], sh);
    if (vt == 0 && vb < a.nb_prior) a.partials[(long)(a.prior_row0 + t) * kReduceBlocks + vb] = tot;
    __syncthreads();  // sh is reused by the next term
  }
}

// chi2 (block 0: finish_kernel's arithmetic on each slice's partials,
// combined as chi2_combine_kernel does) and prior values (block 1 + t:
// finish_kernel's arithmetic on partial row row0 + t into out[t]) in one
// launch, bit-identical to the separate finishing launches.
__global__ __launch_bounds__(kReduceThreads) void values_finish_kernel(
    const double* __restrict__ partials, int nblocks_chi2, const SliceDev* __restrict__ tab,
    int nsl, int normalize, double* __restrict__ chi2_out, int row0, int nblocks_prior,
    double* __restrict__ prior_out) {
  __shared__ double sh[kReduceThreads / 64];
  if (blockIdx.x == 0) {
    double total = 0.0;
    for (int s = 0; s < nsl; s++) {
      const double* p = partials + (long)s * kReduceBlocks;
      double v = 0.0;
      for (int i = threadIdx.x; i < nblocks_chi2; i += blockDim.x) v = v + p[i];
      const double r = block_sum(v, sh);
      if (threadIdx.x == 0) total += normalize ? r / (double)tab[s].nvis : r;
      __syncthreads();  // sh is reused by the next slice
    }
    if (threadIdx.x == 0) *chi2_out = 0.5 * total;
    return;
  }
  const int t = blockIdx.x - 1;
  const double* p = partials + (long)(row0 + t) * kReduceBlocks;
  double v = 0.0;
  for (int i = threadIdx.x; i < nblocks_prior; i += blockDim.x) v = v + p[i];
  const double r = block_sum(v, sh);
  if (threadIdx.x == 0) prior_out[t] = r;
}

// The degrid blocks per slice: one visibility per thread, at most 1.5 blocks
// per CU (measured at 1024^2 / 1.76M visibilities: 384 blocks 16.9 us, 256
// 18.2, 512 19.2, 1024 26.0 for the fused value launch) and kReduceBlocks.
static int degrid_blocks(const gvm_ctx* ctx) {
  long vmax = 0;
  for (const Slice& sl : ctx->slices) vmax = std::max(vmax, sl.nvis);
  const long cap = std::min<long>(kReduceBlocks, std::max(1, 3 * ctx->cu_count / 2));
  return (int)std::min<long>(cap, std::max<long>(1, (vmax + kDegridThreads - 1) / kDegridThreads));
}

// chi2 into *chi2_out and n prior values into prior_out[0, n) for up to four
// slices: the two FFT passes, values_kernel and values_finish_kernel.
static gvm_status values_fused(gvm_ctx* ctx, float* I, int flag_opt, int normalize,
                               double* chi2_out, const gvm_prior_term* terms, int n,
                               double* prior_out) {
  const int nsl = (int)ctx->slices.size();
  ValuesArgs a{};
  gvm_status st = check_terms(ctx, terms, n, &a.T);
  if (st) return st;
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  if ((st = model_grids(ctx, I, nsl, flag_opt))) return st;
  a.grids = ctx->grid;
  a.tab = ctx->slice_tab;
  a.partials = ctx->partials;
  a.cmaxs = ctx->cmax;
  a.I = I;
  a.noise = ctx->noise;
  a.noise_cut = ctx->g.noise_cut;
  a.N = ctx->g.N;
  a.nb_vis = degrid_blocks(ctx);
  a.nb_prior = n > 0 ? reduce_grid((long)ctx->g.N * ctx->g.N) : 0;
  a.prior_row0 = nsl;
  const int nblocks = (a.nb_prior + 1) / 2 + a.nb_vis * nsl;
  hipLaunchKernelGGL(values_kernel, dim3(nblocks), dim3(kDegridThreads), 0, ctx->stream, a);
  GVM_CHECK_LAUNCH();
  hipLaunchKernelGGL(values_finish_kernel, dim3(1 + n), dim3(kReduceThreads), 0, ctx->stream,
                     ctx->partials, a.nb_vis, ctx->slice_tab, nsl, normalize, chi2_out, nsl,
                     a.nb_prior, prior_out);
  GVM_CHECK_LAUNCH();
  ctx->have_residuals = true;
  return GVM_OK;
}

// The forward pass up to the chi2 block partials (slice s in partial row s);
// returns the blocks per slice.
static gvm_status forward_chi2_partials(gvm_ctx* ctx, float* I, int flag_opt, int* nb_out) {
  const gvm_geometry& g = ctx->g;
  const int nsl = (int)ctx->slices.size();
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  gvm_status st = model_grids(ctx, I, nsl, flag_opt);
  if (st) return st;
  const int nb = degrid_blocks(ctx);
  hipLaunchKernelGGL(degrid_chi2_kernel, dim3(nb, nsl), dim3(kDegridThreads), 0, ctx->stream,
                     ctx->grid, g.N, ctx->slice_tab, ctx->partials, ctx->cmax);
  GVM_CHECK_LAUNCH();
  *nb_out = nb;
  return GVM_OK;
}

gvm_status values_enqueue(gvm_ctx* ctx, float* I, int flag_opt, int normalize, int chi2_slot,
                          const gvm_prior_term* terms, int n, int first_slot) {
  if ((int)ctx->slices.size() <= 4)
    return values_fused(ctx, I, flag_opt, normalize, ctx->dvalues + chi2_slot, terms, n,
                        ctx->dvalues + first_slot);
  // many slices: chi2 and the priors take their own finishing launches
  gvm_status st = forward_chi2_enqueue(ctx, I, flag_opt, normalize, chi2_slot);
  if (st) return st;
  return priors_value_partials_finish(ctx, I, terms, n, first_slot);
}

gvm_status forward_chi2_enqueue(gvm_ctx* ctx, float* I, int flag_opt, int normalize, int slot) {
  const int nsl = (int)ctx->slices.size();
  if (nsl <= 4)
    return values_fused(ctx, I, flag_opt, normalize, ctx->dvalues + slot, nullptr, 0,
                        nullptr);
  int nb = 0;
  gvm_status st = forward_chi2_partials(ctx, I, flag_opt, &nb);
  if (st) return st;
  if ((st = finish_reductions(ctx, nsl, nb))) return st;
  hipLaunchKernelGGL(chi2_combine_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->dscalars,
                     ctx->slice_tab, nsl, normalize, ctx->dvalues + slot);
  GVM_CHECK_LAUNCH();
  ctx->have_residuals = true;
  return GVM_OK;
}

gvm_status forward_chi2(gvm_ctx* ctx, float* I, int flag_opt, int normalize, double* out) {
  gvm_status st = forward_chi2_enqueue(ctx, I, flag_opt, normalize, 0);
  if (st) return st;
  return values_read(ctx, 1, out);
}

}  // namespace gvm
