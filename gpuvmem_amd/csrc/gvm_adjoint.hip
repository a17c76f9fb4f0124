// DChi2 adjoint on gfx950 MFMA.
//
// Reference: DChi2 (src/functions.cu:3698-3791) evaluates, for every pixel
// inside the mask,
//     d(i,j) = -fg_scale PB(i,j) sum_v w_v (Re Vr_v cos phi + Im Vr_v sin phi),
//     phi    = 2 pi (x_j u_v + y_i v_v + (z_ij - 1) w_v),
// i.e. d = -fg PB Re sum_v c_v e^{i phi} with c_v = w_v conj(Vr_v): a direct
// DFT over all visibilities for each pixel (M N Nvis sincos).
//
// Here the phase is split per axis, e^{i phi} = Ex_v(j) Ey_v(i) with
//     Ex_v(j) = e^{2 pi i (u_v x_j + w_v f(x_j))},  f(t) = sqrt(1 - t^2) - 1,
// (the neglected cross term of the w phase is about w x^2 y^2 / 4 turns), so
//     G(i,j) = Re sum_v [c_v Ey_v(i)] Ex_v(j) = sum_k P(i,k) Q(k,j)
// is a real GEMM over k = (v, re/im): P = [Re(c Ey), -Im(c Ey)],
// Q = [Re Ex, Im Ex].  P and Q are generated per chunk of visibilities into
// LDS (they are never materialised in HBM) and multiplied on the MFMA units:
//   F32    v_mfma_f32_32x32x2_f32 on FP32 operands;
//   F16X3  FP16 hi/lo split of both operands, P_hi Q_hi + P_hi Q_lo + P_lo Q_hi
//          with v_mfma_f32_32x32x16_f16, FP32 accumulation;
//   BF16X3 the same with BF16 (v_mfma_f32_32x32x16_bf16).
// P carries c scaled by a power of two (from the forward pass' max |c|) so the
// FP16 operands stay in range.  The visibilities are split over K into
// `nsplit` ranges; each workgroup owns one 256x256 output tile and one range
// and writes an FP32 partial tile; the epilogue sums the partials in FP64,
// applies -fg PB (and 1/nvis) on the mask and the chain rule of
// DChi2_total_I_nu_0 / DChi2_total_alpha (src/functions.cu:4000-4024,
// 3968-3998).
//
// Phase arguments are kept exact in turns: u dx = a_hi + a_lo with a_hi on a
// 2^-(23 - log2 N) grid, so a_hi * n is exact in FP32 and its fraction is
// exact; the remainder a_lo n + w f(x) is small.  The FP64 split is done once
// at set-up.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gvm_internal.h"

namespace gvm {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

enum { PREC_F32 = 0, PREC_F16 = 1, PREC_BF16 = 2 };

constexpr int kTile = 256;  // output tile edge (rows and columns)
constexpr int kKV = 16;     // visibilities per chunk
constexpr int kKP = 2 * kKV;

// sin and cos of 2 pi t (t in turns).  v_sin_f32 / v_cos_f32 take their
// argument in turns, valid on [-256, 256]; |t| < 2 here.
__device__ __forceinline__ void sincos_turns(float t, float* s, float* c) {
  *s = __builtin_amdgcn_sinf(t);
  *c = __builtin_amdgcn_cosf(t);
}

__device__ __forceinline__ float fract(float x) { return __builtin_amdgcn_fractf(x); }

__device__ __forceinline__ f32x16_t mfma16(const u16x8_t& a, const u16x8_t& b, f32x16_t c,
                                           int prec) {
  if (prec == PREC_F16)
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(half8_t, a),
                                                  __builtin_bit_cast(half8_t, b), c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                 __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
}

// LDS of the GEMM kernel: two buffers of P (rows) and Q (columns) tiles,
// 256 rows x kKP values each (16-bit precisions keep hi and lo parts in
// separate arrays), addressed in 16-byte groups with an XOR swizzle so the
// 32 rows read by one MFMA operand load spread over the banks; and two
// buffers of visibility records.
template <int PREC>
struct AdjLds {
  static constexpr int kRowBytes = 64;            // 32 x 16-bit
  static constexpr int kGroups = kRowBytes / 16;  // 4
  unsigned char ah[2][kTile * kRowBytes];
  unsigned char al[2][kTile * kRowBytes];
  unsigned char bh[2][kTile * kRowBytes];
  unsigned char bl[2][kTile * kRowBytes];
  float4 rec_s[2][kKV];  // (a_hi, a_lo, b_hi, b_lo)
  float4 rec_d[2][kKV];  // (scale |c|, arg c in turns, w, 0)
  unsigned int mask[kTile / 32][kTile];  // the output tile's mask bits, word-major
};

template <>
struct AdjLds<PREC_F32> {
  static constexpr int kRowBytes = 128;           // 32 x float
  static constexpr int kGroups = kRowBytes / 16;  // 8
  unsigned char a[2][kTile * kRowBytes];
  unsigned char b[2][kTile * kRowBytes];
  float4 rec_s[2][kKV];
  float4 rec_d[2][kKV];
  unsigned int mask[kTile / 32][kTile];
};

// Byte offset of 16-byte group g of a row, with an XOR swizzle of the group
// by row bits.  For 64-byte rows (NG = 4) the pattern (row ^ row >> 2) & 3
// was measured free of LDS bank conflicts (SQ_LDS_BANK_CONFLICT = 0) for
// both access patterns of this kernel: 16-byte reads of 32 consecutive rows
// per half-wave (MFMA operands) and 16-byte writes of 64 consecutive rows
// per wave (generation).
template <int NG>
__device__ __forceinline__ int swz(int row, int g) {
  const int f = NG == 4 ? (row ^ (row >> 2)) & 3 : (row ^ (row >> 3)) & 7;
  return row * (NG * 16) + ((g ^ f) << 4);
}

// hi/lo split of a pair of floats into packed 16-bit pairs.
// FP16: hi = x truncated to FP16 (v_cvt_pkrtz; for normal FP16 values that is
// x with the low 13 mantissa bits cleared), lo = x - hi exact in FP32, then
// converted.  BF16: both parts by the hardware round-to-nearest-even
// conversion; the FP32 value of a BF16 number is its bits shifted up by 16.
template <int PREC>
__device__ __forceinline__ void split_pair(float x0, float x1, unsigned int* hi, unsigned int* lo) {
  if constexpr (PREC == PREC_F16) {
    const float h0 = __uint_as_float(__float_as_uint(x0) & 0xffffe000u);
    const float h1 = __uint_as_float(__float_as_uint(x1) & 0xffffe000u);
    *hi = __builtin_bit_cast(unsigned int, __builtin_amdgcn_cvt_pkrtz(x0, x1));
    *lo = __builtin_bit_cast(unsigned int, __builtin_amdgcn_cvt_pkrtz(x0 - h0, x1 - h1));
  } else {
    bf16x2_t h;
    h.x = (__bf16)x0;
    h.y = (__bf16)x1;
    unsigned int hb = __builtin_bit_cast(unsigned int, h);
    asm("" : "+v"(hb));  // the hi parts below come from these packed bits
    bf16x2_t l;
    l.x = (__bf16)(x0 - __uint_as_float(hb << 16));
    l.y = (__bf16)(x1 - __uint_as_float(hb & 0xffff0000u));
    *hi = hb;
    *lo = __builtin_bit_cast(unsigned int, l);
  }
}

// Workgroup barrier that orders LDS accesses only: global loads issued
// before it (prefetches into registers) stay in flight across it, where
// __syncthreads would wait for them.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

constexpr int kGemmThreads = 512;
constexpr int kPScale = 15;  // max |P| = 2^15 < 65504 (FP16 max)
constexpr int kPScaleLimit = 100;

// Exponent of the power-of-two scale of P for a slice whose largest |c| is
// cm: 2^e cm <= 2^kPScale.  Clamped to [-kPScaleLimit, kPScaleLimit] so that
// the scale and its inverse stay finite and normal in FP32 for any cm (an
// unclamped exponent overflows for cm below about 2^-113, and 0 * inf would
// turn the zero amplitudes of off-grid samples into NaN); 0 when cm is 0.
__device__ __forceinline__ int p_scale_exponent(float cm) {
  if (!(cm > 0.0f)) return 0;
  const int e = kPScale - (int)ceilf(log2f(cm));
  return min(max(e, -kPScaleLimit), kPScaleLimit);
}
constexpr int kHalfKV = kKV / 2;

// One workgroup (8 waves, 2 per SIMD) owns a 256x256 tile of G for one
// K-range of visibilities.  Wave w accumulates rows (w&1)*128.. and columns
// (w>>1)*64.. as 4x2 blocks of 32x32 MFMA outputs (128 FP32 accumulators per
// lane).  Per chunk of kKV visibilities, thread t generates P (row t%256) and
// Q (column t%256) for half of the chunk's visibilities (t/256 picks the
// half), into the other LDS buffer while the chunk before is multiplied.
// Waves 0-3 generate first and multiply second, waves 4-7 the other way
// round, so on each SIMD one wave issues VALU work while the other keeps the
// MFMA pipe busy; one barrier per chunk.  Visibility records are loaded from
// HBM two chunks ahead by wave 0 and staged in LDS.
//
// P uses c in polar form (from the forward pass): c Ey = |c| e^{2 pi i
// (phi_y + arg c)}, so the residual enters as one extra phase term and one
// scale per value.
template <int PREC>
__global__ __launch_bounds__(kGemmThreads, 1) void adjoint_gemm_kernel(
    const SliceDev* __restrict__ tab, const int4* __restrict__ splits,
    const float2* __restrict__ xtab, const float2* __restrict__ ytab, int N, int tiles_x,
    int ntiles, const unsigned int* __restrict__ cmaxs, const unsigned int* __restrict__ mbits,
    float* __restrict__ part) {
  __shared__ AdjLds<PREC> sm;
  constexpr int NG = AdjLds<PREC>::kGroups;

  // XCD-aware order: the workgroups of one K-range (which read the same
  // visibilities) are dispatched to the same XCD (blockIdx % 8) and share its L2.
  const int nwg = gridDim.x;
  const int per_xcd = nwg / 8;
  const int lin = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  const int split = lin / ntiles;
  const int tile = lin % ntiles;
  const int ti = tile / tiles_x, tj = tile % tiles_x;

  // this workgroup's visibility range: slice e.x, [e.y, e.z)
  const int4 e = splits[split];
  const SliceDev& sd = tab[e.x];
  const float4* __restrict__ phs = sd.phs;
  const float* __restrict__ wcoord = sd.wcoord;
  const float2* __restrict__ cpol = sd.cpol;
  const long nvis = sd.nvis;
  const long kbeg = e.y;
  const long kend = e.z;
  const int nchunks = kend > kbeg ? (int)((kend - kbeg + kKV - 1) / kKV) : 0;
  const unsigned int* __restrict__ cmax = cmaxs + e.x;

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const bool group_b = wave >= 4;  // multiply first
  const int gidx = t & (kTile - 1);
  const int half = t >> 8;  // which half of the chunk's visibilities this thread generates
  const int row = ti * kTile + gidx, col = tj * kTile + gidx;
  const bool rvalid = row < N, cvalid = col < N;
  const float2 yt = rvalid ? ytab[row] : make_float2(0.f, 0.f);
  const float2 xt = cvalid ? xtab[col] : make_float2(0.f, 0.f);

  const float cm = __uint_as_float(*cmax);
  // P is scaled by a power of two (undone in the epilogue) so |P| <= 2^kPScale
  // fits FP16; the scale is folded into the visibility amplitudes as they
  // are staged.  Rows and columns past the image produce finite values that
  // only reach outputs past the image, which are not written.
  const float scale = ldexpf(1.0f, p_scale_exponent(cm));
  // Odd ranges accumulate -G (P negated, which the split and every product
  // follow exactly); the epilogue negates their partials back.  With the
  // 16-bit MFMAs each range's partial carries an error that does not change
  // sign with the terms, as a downward rounding of their accumulation would:
  // on a 2048^2 image it was nearly uniform over the pixels (mean / rms 0.87
  // for bf16x3, 0.92 for f16x3, 0.0002 for FP32) and the same in every
  // gradient of an optimisation.  Alternating the sign flips that error from
  // one range to the next, so it cancels in the sum.
  const float pscale = (split & 1) ? -scale : scale;

  const int lr = lane & 31, lh = lane >> 5;
  const int wr = (wave & 1) * 128, wc = (wave >> 1) * 64;

  // ---- visibility records (wave 0; lane l handles slot l % kKV): loaded
  //      into registers one chunk before they are staged; the arithmetic on
  //      them waits until then, so the loads stay in flight ----
  const int slot = lane & (kKV - 1);
  float4 rs;
  float2 rp;
  float rw;
  bool rin;
  auto load_rec = [&](int c) {
    const long k = kbeg + (long)c * kKV + slot;
    const long kc = min(k, nvis - 1);
    rs = ldg(phs + kc);
    rp = ldg(cpol + kc);
    rw = ldg(wcoord + kc);
    rin = k < kend;
  };
  auto store_rec = [&](int buf) {
    sm.rec_s[buf][slot] = rs;
    sm.rec_d[buf][slot] = make_float4(rin ? rp.x * pscale : 0.0f, rp.y, rw, 0.f);
  };

  // ---- generation: P row gidx and Q column gidx for half of the chunk ----
  auto generate = [&](int buf) {
    float pv[2 * kHalfKV], qv[2 * kHalfKV];
#pragma unroll
    for (int q = 0; q < kHalfKV; q++) {
      const int kk = half * kHalfKV + q;
      const float4 s = sm.rec_s[buf][kk];
      const float4 d = sm.rec_d[buf][kk];
      float sy, cy, sx, cx;
      sincos_turns(fract(s.z * yt.x) + fmaf(s.w, yt.x, fmaf(d.z, yt.y, d.y)), &sy, &cy);
      sincos_turns(fmaf(d.z, xt.y, fmaf(s.y, xt.x, fract(s.x * xt.x))), &sx, &cx);
      pv[2 * q] = d.x * cy;
      pv[2 * q + 1] = -(d.x * sy);
      qv[2 * q] = cx;
      qv[2 * q + 1] = sx;
    }
    if constexpr (PREC == PREC_F32) {
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const int o = swz<NG>(gidx, 4 * half + g);
        *(float4*)&sm.a[buf][o] = make_float4(pv[4 * g], pv[4 * g + 1], pv[4 * g + 2], pv[4 * g + 3]);
        *(float4*)&sm.b[buf][o] = make_float4(qv[4 * g], qv[4 * g + 1], qv[4 * g + 2], qv[4 * g + 3]);
      }
    } else {
#pragma unroll
      for (int g = 0; g < 2; g++) {
        uint4 ph, pl, qh, ql;
        split_pair<PREC>(pv[8 * g + 0], pv[8 * g + 1], &ph.x, &pl.x);
        split_pair<PREC>(pv[8 * g + 2], pv[8 * g + 3], &ph.y, &pl.y);
        split_pair<PREC>(pv[8 * g + 4], pv[8 * g + 5], &ph.z, &pl.z);
        split_pair<PREC>(pv[8 * g + 6], pv[8 * g + 7], &ph.w, &pl.w);
        split_pair<PREC>(qv[8 * g + 0], qv[8 * g + 1], &qh.x, &ql.x);
        split_pair<PREC>(qv[8 * g + 2], qv[8 * g + 3], &qh.y, &ql.y);
        split_pair<PREC>(qv[8 * g + 4], qv[8 * g + 5], &qh.z, &ql.z);
        split_pair<PREC>(qv[8 * g + 6], qv[8 * g + 7], &qh.w, &ql.w);
        const int o = swz<NG>(gidx, 2 * half + g);
        *(uint4*)&sm.ah[buf][o] = ph;
        *(uint4*)&sm.al[buf][o] = pl;
        *(uint4*)&sm.bh[buf][o] = qh;
        *(uint4*)&sm.bl[buf][o] = ql;
      }
    }
  };

  bool active = false;  // this wave's sub-tile has a pixel on the mask
  f32x16_t acc[4][2];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int e = 0; e < 16; e++) acc[a][b][e] = 0.0f;

  // ---- multiply the chunk in buffer buf ----
  auto multiply = [&](int buf) {
    if constexpr (PREC == PREC_F32) {
      // lane (row lr, group lh) takes k = lh*16 + s at step s: 16-byte groups 4 lh .. 4 lh + 3
#pragma unroll
      for (int g = 0; g < 4; g++) {
        float4 av[4], bv[2];
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
          av[rb] = *(const float4*)&sm.a[buf][swz<NG>(wr + rb * 32 + lr, 4 * lh + g)];
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
          bv[cb] = *(const float4*)&sm.b[buf][swz<NG>(wc + cb * 32 + lr, 4 * lh + g)];
#pragma unroll
        for (int e = 0; e < 4; e++)
#pragma unroll
          for (int rb = 0; rb < 4; rb++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++)
              acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[rb][e], bv[cb][e],
                                                                 acc[rb][cb], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int s = 0; s < kKP / 16; s++) {
        // lane (row lr, group lh) takes k = s*16 + lh*8 + e: 16-byte group 2 s + lh
        const int g = 2 * s + lh;
        u16x8_t ahi[4], alo[4], bhi[2], blo[2];
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
          const int o = swz<NG>(wr + rb * 32 + lr, g);
          ahi[rb] = *(const u16x8_t*)&sm.ah[buf][o];
          alo[rb] = *(const u16x8_t*)&sm.al[buf][o];
        }
#pragma unroll
        for (int cb = 0; cb < 2; cb++) {
          const int o = swz<NG>(wc + cb * 32 + lr, g);
          bhi[cb] = *(const u16x8_t*)&sm.bh[buf][o];
          blo[cb] = *(const u16x8_t*)&sm.bl[buf][o];
        }
        // The three products of a block are issued eight MFMAs apart (all
        // blocks' hi*hi, then hi*lo, then lo*hi): back to back on one
        // accumulator, each waited for the one before (3.5 % slower).  Each
        // accumulator still sees the same sequence of products.
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
          for (int cb = 0; cb < 2; cb++) acc[rb][cb] = mfma16(ahi[rb], bhi[cb], acc[rb][cb], PREC);
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
          for (int cb = 0; cb < 2; cb++) acc[rb][cb] = mfma16(ahi[rb], blo[cb], acc[rb][cb], PREC);
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
          for (int cb = 0; cb < 2; cb++) acc[rb][cb] = mfma16(alo[rb], bhi[cb], acc[rb][cb], PREC);
      }
    }
  };

  // the output tile's mask bits (pixel k of the image in bit k % 32 of word
  // k / 32; N is a multiple of 32), read from LDS when the tile is written
  {
    const int mr = t >> 1, mw = (t & 1) * 4;
    const int gr = ti * kTile + mr;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int gw = tj * (kTile / 32) + mw + i;
      sm.mask[mw + i][mr] = (gr < N && gw < N / 32) ? mbits[(long)gr * (N / 32) + gw] : 0u;
    }
  }

  if (nchunks > 0) {
    // prologue: records of chunks 0 and 1, tiles of chunk 0
    if (wave == 0) {
      load_rec(0);
      store_rec(0);
      load_rec(1);
      store_rec(1);
      load_rec(2);
    }
    __syncthreads();
    // does this wave's 128 x 64 sub-tile hold any pixel on the mask?  (lane l
    // looks at rows l and l + 64, both words of the sub-tile's columns)
    {
      bool any = false;
#pragma unroll
      for (int h = 0; h < 2; h++)
#pragma unroll
        for (int w2 = 0; w2 < 2; w2++) any |= sm.mask[(wc >> 5) + w2][wr + 64 * h + lane] != 0u;
      active = __ballot(any) != 0ull;
    }
    generate(0);
    __syncthreads();
    // main loop: one barrier per chunk
    for (int c = 0; c < nchunks; c++) {
      const int cur = c & 1, nxt = cur ^ 1;
      // chunk c+1 (past the end: zero weights, never multiplied) is generated
      // before or after the multiply by the wave's group.  The multiply is one
      // instance outside the branches, so the accumulators keep one register
      // assignment through the loop (two instances, one per branch, cost
      // about 4 % in copies between them).  A wave whose sub-tile has no
      // pixel on the mask skips its MFMAs (and writes nothing): at 1024^2
      // with a disc-shaped mask 20 % of the sub-tiles, 9.6 % of the time.
      if (!group_b) generate(nxt);
      if (active) multiply(cur);
      if (group_b) generate(nxt);
      if (wave == 0) {
        store_rec(cur);  // records of chunk c+2
        load_rec(c + 3);
      }
      lds_barrier();
    }
  }

  // ---- write the partial tile on the mask (the epilogue reads only those
  //      pixels; 36 % fewer bytes written at 1024^2): 32x32 accumulator
  //      layout, lane l, reg e: row 8*(e/4) + 4*(l/32) + e%4, column l%32 ----
  if (nchunks == 0) __syncthreads();  // the mask tile, when the loop has not ordered it
  float* out = part + (long)split * N * N;
  // nothing of this wave's sub-tile is on the mask (a range with no
  // visibilities, nchunks == 0, still writes its zeros: the epilogue reads it)
  if (!active && nchunks > 0) return;
#pragma unroll
  for (int rb = 0; rb < 4; rb++)
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
      // the mask words of this lane's 16 rows (4 groups of 4 consecutive rows),
      // read before the stores
      const int mw = (wc + cb * 32) >> 5, row0 = wr + rb * 32 + 4 * lh;
      uint4 m4[4];
#pragma unroll
      for (int q = 0; q < 4; q++) m4[q] = *(const uint4*)&sm.mask[mw][row0 + 8 * q];
#pragma unroll
      for (int e = 0; e < 16; e++) {
        const int r = ti * kTile + row0 + 8 * (e >> 2) + (e & 3);
        const int c = tj * kTile + wc + cb * 32 + lr;
        const uint4 m = m4[e >> 2];
        const unsigned int word = (e & 3) == 0 ? m.x : (e & 3) == 1 ? m.y : (e & 3) == 2 ? m.z : m.w;
        if (r < N && c < N && ((word >> lr) & 1u)) out[(long)r * N + c] = acc[rb][cb][e];
      }
    }
}

// c = w conj(V_r) of every visibility of slice blockIdx.y in polar form, (|c|,
// arg(c) in turns), from the residuals of the last forward pass: the GEMM's
// per-visibility input.  |c| is computed as the degrid kernel computes the
// per-slice max that sets the P scale.
__global__ __launch_bounds__(256) void polar_kernel(const SliceDev* __restrict__ tab) {
  const SliceDev& sd = tab[blockIdx.y];
  const float2* __restrict__ vr = sd.vr;
  const float* __restrict__ weight = sd.weight;
  float2* __restrict__ cpol = sd.cpol;
  const long nvis = sd.nvis;
  for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < nvis;
       v += (long)gridDim.x * blockDim.x) {
    const float2 r = ldg(vr + v);
    const float w = ldg(weight + v);
    stg(cpol + v, make_float2(w * sqrtf(resid_abs2(r)), -atan2f(r.y, r.x) * 0.15915494309189535f));
  }
}

// Direct sum (the reference's algorithm, exact w term), one thread per pixel,
// visibilities staged through LDS.  FP32 terms, FP64 accumulation per 256
// visibilities.  Writes G to part[0] on the mask; a block with no pixel on
// the mask returns at once (the reference's DChi2 skips those pixels too).
__global__ __launch_bounds__(256) void adjoint_direct_kernel(
    const float4* __restrict__ phs, const float* __restrict__ wcoord,
    const float* __restrict__ weight, const float2* __restrict__ vr, long nvis,
    const float2* __restrict__ xtab, const float2* __restrict__ ytab, int N, double dx_rad,
    double dy_rad, const unsigned int* __restrict__ mbits, float* __restrict__ part) {
  __shared__ float4 sp[256];
  __shared__ float4 sc[256];  // (w, cr, ci, 0)
  const long pix = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const bool valid = pix < (long)N * N;
  const bool on_mask = valid && ((mbits[pix >> 5] >> (pix & 31)) & 1u);
  if (!__syncthreads_or(on_mask)) return;
  const int i = valid ? (int)(pix / N) : 0, j = valid ? (int)(pix % N) : 0;
  const float2 xt = xtab[j], yt = ytab[i];
  const double x = xt.x * dx_rad, y = yt.x * dy_rad;
  const float zm1 = (float)(-(x * x + y * y) / (1.0 + sqrt(1.0 - x * x - y * y)));
  double acc = 0.0;
  for (long k0 = 0; k0 < nvis; k0 += 256) {
    const long k = k0 + threadIdx.x;
    __syncthreads();
    if (k < nvis) {
      sp[threadIdx.x] = phs[k];
      const float2 r = vr[k];
      const float w = weight[k];
      sc[threadIdx.x] = make_float4(wcoord[k], w * r.x, -w * r.y, 0.f);
    } else {
      sp[threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
      sc[threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    float part_acc = 0.0f;
    for (int q = 0; q < 256; q++) {
      const float4 p = sp[q];
      const float4 c = sc[q];
      const float t = fract(p.x * xt.x) + fract(p.z * yt.x) +
                      fmaf(p.y, xt.x, fmaf(p.w, yt.x, c.x * zm1));
      float s, co;
      sincos_turns(t, &s, &co);
      part_acc += c.y * co - c.z * s;
    }
    acc += (double)part_acc;
  }
  if (on_mask) part[pix] = (float)acc;
}

// Per pixel inside the mask and per slice in slice order: the FP64 sum of
// the slice's split-K partials, times -fg PB (and 1/nvis if normalize), then
// the chain rule of DChi2_total_I_nu_0 / DChi2_total_alpha accumulated into
// result in float, one slice after the other as the reference's per-channel
// launches do (src/functions.cu:4456-4558).
__global__ void adjoint_epilogue_kernel(const float* __restrict__ part,
                                        const int* __restrict__ p0,
                                        const SliceDev* __restrict__ tab, int nsl, long n,
                                        const unsigned int* __restrict__ cmaxs, int scaled,
                                        const float* __restrict__ noise, float noise_cut,
                                        float fg_scale, const float* __restrict__ gcf,
                                        int normalize, float threshold, int flag_opt,
                                        const float* __restrict__ I,
                                        float* __restrict__ result) {
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    if (!(noise[k] < noise_cut)) continue;
    const float i0 = I[k], alpha = I[n + k];
    const float gk = gcf ? gcf[k] : 1.0f;
    float r0 = result[k], r1 = result[n + k];
    for (int sl = 0; sl < nsl; sl++) {
      // FP64 sum over the slice's splits in split order; eight loads are
      // issued before their adds so each thread keeps several HBM requests
      // in flight.  The GEMM's odd splits hold -G (their sign is flipped
      // back here); the direct form's are not flipped.
      const int pa = p0[sl], pe = p0[sl + 1];
      const int alt = scaled ? 1 : 0;
      double g = 0.0;
      int s = pa;
      for (; s + 8 <= pe; s += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = part[(long)(s + u) * n + k];
#pragma unroll
        for (int u = 0; u < 8; u++) g += (double)(((s + u) & alt) ? -v[u] : v[u]);
      }
      for (; s < pe; s++) {
        const float v = part[(long)s * n + k];
        g += (double)((s & alt) ? -v : v);
      }
      double inv_scale = 1.0;
      if (scaled) {
        const float cm = __uint_as_float(cmaxs[sl]);
        inv_scale = ldexp(1.0, -p_scale_exponent(cm));
      }
      const SliceDev& sd = tab[sl];
      const double inv_nvis = normalize ? 1.0 / (double)sd.nvis : 1.0;
      // DChi2: dchi2 = -(sum * fg_scale * atten [* gcf]) [/ nvis]
      // (src/functions.cu:3698-3791; with gcf :3793-3888)
      float sc = fg_scale * sd.pb[k];
      if (gcf) sc = sc * gk;
      const float d = (float)(-(g * inv_scale) * (double)sc * inv_nvis);
      const float dI = powf(sd.nudiv, alpha);
      if (flag_opt % 2 == 0) {
        r0 += d * dI;
      } else if (i0 > threshold) {
        r1 += d * (i0 * dI * fg_scale * logf(sd.nudiv));
      }
    }
    if (flag_opt % 2 == 0)
      result[k] = r0;
    else
      result[n + k] = r1;
  }
}

namespace {

// Splits the visibilities of every slice into ranges for the GEMM
// workgroups.  Each range accumulates in FP32 MFMA accumulators and the
// ranges are combined in FP64, so ranges are kept near ctx->split_vis
// visibilities (about 10^3 accumulation steps: the FP32 rounding error grows
// with the square root of the accumulation length), with at least enough
// ranges to fill the GPU, within a workspace of at most kMaxWorkspace bytes
// (and half the free memory); the total is padded to a multiple of 8 for
// the XCD remap with empty ranges.  The direct form takes one range per
// slice.
gvm_status build_splits(gvm_ctx* ctx, int mode, int ntiles, long n) {
  const long kSplitVis = ctx->split_vis;
  constexpr long kMaxWorkspace = 32L << 30;
  const int nsl = (int)ctx->slices.size();
  std::vector<int4> tab;
  std::vector<int> p0(nsl + 1, 0);
  if (mode == GVM_ADJOINT_DIRECT) {
    for (int s = 0; s < nsl; s++) {
      p0[s] = s;
      tab.push_back(make_int4(s, 0, (int)ctx->slices[s].nvis, 0));
    }
    p0[nsl] = nsl;
  } else {
    size_t free_b = 0, total_b = 0;
    GVM_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    const long avail = (long)(free_b + sizeof(float) * ctx->adj_part_elems) / 2;
    const long budget = std::max<long>(8, std::min(kMaxWorkspace, avail) / (long)(sizeof(float) * n));
    const long fill = (4L * std::max(ctx->cu_count, 64) + ntiles - 1) / ntiles;
    long kvis = kSplitVis;
    std::vector<long> ns(nsl);
    for (;;) {
      long total = 0;
      for (int s = 0; s < nsl; s++) {
        const long nv = ctx->slices[s].nvis;
        long k = (nv + kvis - 1) / kvis;
        // small observations: more, shorter ranges so the GPU fills
        k = std::max(k, std::min((fill + nsl - 1) / nsl, std::max(1L, nv / 512)));
        ns[s] = std::max(1L, k);
        total += ns[s];
      }
      if (((total + 7) / 8) * 8 <= budget || kvis >= (1L << 30)) break;
      kvis *= 2;
    }
    for (int s = 0; s < nsl; s++) {
      const long nv = ctx->slices[s].nvis;
      const long kper = ((nv + ns[s] - 1) / ns[s] + kKV - 1) / kKV * kKV;
      p0[s] = (int)tab.size();
      for (long j = 0; j < ns[s]; j++) {
        const long kb = std::min(nv, j * kper), ke = std::min(nv, kb + kper);
        tab.push_back(make_int4(s, (int)kb, (int)ke, 0));
      }
    }
    while (tab.size() % 8) {
      const int last = nsl - 1;
      const int nv = (int)ctx->slices[last].nvis;
      tab.push_back(make_int4(last, nv, nv, 0));  // empty range: a zero partial tile
    }
    p0[nsl] = (int)tab.size();
  }
  if ((int)tab.size() > ctx->split_cap) {
    GVM_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->split_tab) GVM_HIP_TRY(hipFree(ctx->split_tab));
    ctx->split_tab = nullptr;
    ctx->split_cap = 0;
    GVM_HIP_TRY(hipMalloc((void**)&ctx->split_tab, sizeof(int4) * tab.size()));
    ctx->split_cap = (int)tab.size();
  }
  GVM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  GVM_HIP_TRY(hipMemcpy(ctx->split_tab, tab.data(), sizeof(int4) * tab.size(),
                        hipMemcpyHostToDevice));
  GVM_HIP_TRY(hipMemcpy(ctx->slice_p0, p0.data(), sizeof(int) * p0.size(),
                        hipMemcpyHostToDevice));
  ctx->nsplit_total = (int)tab.size();
  ctx->split_mode = mode;
  const size_t need = (size_t)ctx->nsplit_total * n;
  if (need > ctx->adj_part_elems) {
    if (ctx->adj_part) GVM_HIP_TRY(hipFree(ctx->adj_part));
    ctx->adj_part = nullptr;
    ctx->adj_part_elems = 0;
    GVM_HIP_TRY(hipMalloc((void**)&ctx->adj_part, sizeof(float) * need));
    ctx->adj_part_elems = need;
  }
  return GVM_OK;
}

}  // namespace

gvm_status rebuild_splits(gvm_ctx* ctx) {
  const int N = ctx->g.N;
  const int tiles_x = (N + kTile - 1) / kTile;
  gvm_status st = build_splits(ctx, ctx->adjoint_mode, tiles_x * tiles_x, (long)N * N);
  if (st) return st;
  ctx->epoch++;
  return GVM_OK;
}

gvm_status adjoint(gvm_ctx* ctx, const float* I, float* result, int flag_opt, int normalize) {
  const gvm_geometry& g = ctx->g;
  const int N = g.N;
  const long n = (long)N * N;
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  const int tiles_x = (N + kTile - 1) / kTile;
  const int ntiles = tiles_x * tiles_x;
  const int mode = ctx->adjoint_mode;
  const int nsl = (int)ctx->slices.size();
  if (ctx->split_mode != mode) {
    GVM_REQUIRE(!ctx->capturing,
                "gvm_dchi2: the workspace is sized by the first gvm_dchi2; run one before "
                "capturing");
    gvm_status st = build_splits(ctx, mode, ntiles, n);
    if (st) return st;
    ctx->epoch++;
  }
  if (mode != GVM_ADJOINT_DIRECT) {
    long vmax = 0;
    for (const Slice& s : ctx->slices) vmax = std::max(vmax, s.nvis);
    const unsigned nbp = (unsigned)std::min<long>(2048, std::max<long>(1, (vmax + 255) / 256));
    hipLaunchKernelGGL(polar_kernel, dim3(nbp, nsl), dim3(256), 0, ctx->stream, ctx->slice_tab);
    GVM_CHECK_LAUNCH();
  }
  // launch timing (not recorded inside graphs)
  const bool timed = !ctx->capturing && !ctx->timing_paused;
  int pair = 0;
  if (timed) {
    if (ctx->tev_pending == gvm_ctx::kTimingPairs) {
      gvm_status st = timing_fold(ctx, 1);
      if (st) return st;
    }
    pair = (ctx->tev_head + ctx->tev_pending) % gvm_ctx::kTimingPairs;
    GVM_HIP_TRY(hipEventRecord(ctx->tev[2 * pair], ctx->stream));
  }
  if (mode == GVM_ADJOINT_DIRECT) {
    const double dxr = g.deltax * 3.14159265358979323846 / 180.0;
    const double dyr = g.deltay * 3.14159265358979323846 / 180.0;
    for (int si = 0; si < nsl; si++) {
      const Slice& s = ctx->slices[si];
      hipLaunchKernelGGL(adjoint_direct_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                         ctx->stream, s.phs, s.wcoord, s.weight, s.vr, s.nvis, ctx->xtab,
                         ctx->ytab, N, dxr, dyr, ctx->mask_bits, ctx->adj_part + (long)si * n);
      GVM_CHECK_LAUNCH();
    }
  } else {
    const dim3 grid(ntiles * ctx->nsplit_total);
    if (mode == GVM_ADJOINT_F32)
      hipLaunchKernelGGL(adjoint_gemm_kernel<PREC_F32>, grid, dim3(kGemmThreads), 0, ctx->stream,
                         ctx->slice_tab, ctx->split_tab, ctx->xtab, ctx->ytab, N, tiles_x, ntiles,
                         ctx->cmax, ctx->mask_bits, ctx->adj_part);
    else if (mode == GVM_ADJOINT_F16X3)
      hipLaunchKernelGGL(adjoint_gemm_kernel<PREC_F16>, grid, dim3(kGemmThreads), 0, ctx->stream,
                         ctx->slice_tab, ctx->split_tab, ctx->xtab, ctx->ytab, N, tiles_x, ntiles,
                         ctx->cmax, ctx->mask_bits, ctx->adj_part);
    else
      hipLaunchKernelGGL(adjoint_gemm_kernel<PREC_BF16>, grid, dim3(kGemmThreads), 0, ctx->stream,
                         ctx->slice_tab, ctx->split_tab, ctx->xtab, ctx->ytab, N, tiles_x, ntiles,
                         ctx->cmax, ctx->mask_bits, ctx->adj_part);
    GVM_CHECK_LAUNCH();
  }
  if (timed) {
    GVM_HIP_TRY(hipEventRecord(ctx->tev[2 * pair + 1], ctx->stream));
    ctx->tev_pending++;
  }
  hipLaunchKernelGGL(adjoint_epilogue_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     ctx->stream, ctx->adj_part, ctx->slice_p0, ctx->slice_tab, nsl, n, ctx->cmax,
                     mode == GVM_ADJOINT_DIRECT ? 0 : 1, ctx->noise, g.noise_cut, g.fg_scale,
                     ctx->gcf, normalize, g.threshold, flag_opt, I, result);
  GVM_CHECK_LAUNCH();
  return GVM_OK;
}

gvm_status timing_fold(gvm_ctx* ctx, int n) {
  while (ctx->tev_pending > 0 && n != 0) {
    const int pair = ctx->tev_head;
    GVM_HIP_TRY(hipEventSynchronize(ctx->tev[2 * pair + 1]));
    float ms = 0.f;
    GVM_HIP_TRY(hipEventElapsedTime(&ms, ctx->tev[2 * pair], ctx->tev[2 * pair + 1]));
    ctx->adj_ms_total += ms;
    ctx->adj_count += 1;
    ctx->tev_head = (ctx->tev_head + 1) % gvm_ctx::kTimingPairs;
    ctx->tev_pending--;
    if (n > 0) n--;
  }
  return GVM_OK;
}

}  // namespace gvm

using namespace gvm;

extern "C" gvm_status gvm_set_split_vis(gvm_ctx* ctx, long split_vis) {
  GVM_REQUIRE(ctx, "gvm_set_split_vis: null context");
  GVM_REQUIRE(split_vis >= 256 && split_vis <= 16384,
              "gvm_set_split_vis: the range length must be in [256, 16384]");
  GVM_REQUIRE(!ctx->capturing, "gvm_set_split_vis: capturing");
  if (ctx->split_vis != split_vis) {
    ctx->split_vis = split_vis;
    ctx->split_mode = -1;  // rebuilt by the next gvm_dchi2
    ctx->epoch++;
  }
  return GVM_OK;
}

extern "C" gvm_status gvm_set_fft_mode(gvm_ctx* ctx, int mode) {
  GVM_REQUIRE(ctx, "gvm_set_fft_mode: null context");
  GVM_REQUIRE(mode >= 0 && mode <= 2, "gvm_set_fft_mode: mode must be 0, 1 or 2");
  GVM_REQUIRE(!ctx->capturing, "gvm_set_fft_mode: capturing");
  if (ctx->fft_mode != mode) ctx->epoch++;
  ctx->fft_mode = mode;
  return GVM_OK;
}

extern "C" gvm_status gvm_adjoint_computed_pixels(const gvm_ctx* ctx, long* pixels) {
  GVM_REQUIRE(ctx && pixels, "gvm_adjoint_computed_pixels: null argument");
  *pixels = ctx->adjoint_mode == GVM_ADJOINT_DIRECT ? ctx->adj_px_direct : ctx->adj_px_gemm;
  return GVM_OK;
}

extern "C" gvm_status gvm_get_split_vis(const gvm_ctx* ctx, long* split_vis) {
  GVM_REQUIRE(ctx && split_vis, "gvm_get_split_vis: null argument");
  *split_vis = ctx->split_vis;
  return GVM_OK;
}

namespace {

std::mutex g_tune_mutex;
std::map<std::string, long> g_tune_cache;

std::string tune_key(const gvm_ctx* ctx) {
  long nv = 0;
  for (const auto& s : ctx->slices) nv += s.nvis;
  char buf[160];
  std::snprintf(buf, sizeof buf, "adjoint M=%d nvis=%ld slices=%d mode=%d cu=%d", ctx->g.M, nv,
                (int)ctx->slices.size(), ctx->adjoint_mode, ctx->cu_count);
  return buf;
}

// The cached range length of `key` (process table, then the GVM_TUNE_CACHE
// file: lines "<key> = <value>", the last one for a key wins), or 0.
long tune_lookup(const std::string& key) {
  std::lock_guard<std::mutex> lk(g_tune_mutex);
  auto it = g_tune_cache.find(key);
  if (it != g_tune_cache.end()) return it->second;
  const char* path = std::getenv("GVM_TUNE_CACHE");
  if (!path || !*path) return 0;
  FILE* f = std::fopen(path, "r");
  if (!f) return 0;
  long found = 0;
  char line[256];
  while (std::fgets(line, sizeof line, f)) {
    std::string l(line);
    const size_t eq = l.rfind(" = ");
    if (eq == std::string::npos || l.compare(0, eq, key) != 0 || eq != key.size()) continue;
    const long v = std::strtol(l.c_str() + eq + 3, nullptr, 10);
    if (v >= 256 && v <= 16384) found = v;
  }
  std::fclose(f);
  if (found) g_tune_cache[key] = found;
  return found;
}

void tune_store(const std::string& key, long value) {
  std::lock_guard<std::mutex> lk(g_tune_mutex);
  g_tune_cache[key] = value;
  const char* path = std::getenv("GVM_TUNE_CACHE");
  if (!path || !*path) return;
  FILE* f = std::fopen(path, "a");
  if (!f) return;
  std::fprintf(f, "%s = %ld\n", key.c_str(), value);
  std::fclose(f);
}

}  // namespace

extern "C" gvm_status gvm_autotune_adjoint(gvm_ctx* ctx, const long* candidates, int ncand,
                                           long* chosen, double* ms_out, int* cached) {
  static const long kDefault[] = {1536, 2048, 3072, 4096, 6144, 8192};
  GVM_REQUIRE(ctx && chosen, "gvm_autotune_adjoint: null argument");
  GVM_REQUIRE(!ctx->capturing, "gvm_autotune_adjoint: capturing");
  GVM_REQUIRE(ctx->adjoint_mode != GVM_ADJOINT_DIRECT,
              "gvm_autotune_adjoint: the direct form has no range length");
  if (!candidates) {
    // the default list up to the longest range this image size allows
    candidates = kDefault;
    ncand = 0;
    while (ncand < (int)(sizeof kDefault / sizeof kDefault[0]) &&
           kDefault[ncand] <= max_tuned_split_vis(ctx->g.N))
      ncand++;
  }
  GVM_REQUIRE(ncand >= 1 && ncand <= 64, "gvm_autotune_adjoint: 1 to 64 candidates");
  for (int c = 0; c < ncand; c++)
    GVM_REQUIRE(candidates[c] >= 256 && candidates[c] <= 16384,
                "gvm_autotune_adjoint: candidates must be in [256, 16384]");
  if (!ctx->have_residuals) {
    set_error("gvm_autotune_adjoint: no residuals (call gvm_chi2 first)");
    return GVM_ERR_STATE;
  }
  if (cached) *cached = 0;
  const std::string key = tune_key(ctx);
  const long hit = tune_lookup(key);
  if (hit) {
    if (cached) *cached = 1;
    *chosen = hit;
    return gvm_set_split_vis(ctx, hit);
  }
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  const long n2 = 2L * ctx->g.M * ctx->g.N;
  float *img = nullptr, *res = nullptr;
  hipEvent_t ea = nullptr, eb = nullptr;
  auto cleanup = [&]() {
    if (img) (void)hipFree(img);
    if (res) (void)hipFree(res);
    if (ea) (void)hipEventDestroy(ea);
    if (eb) (void)hipEventDestroy(eb);
  };
  if (hipMalloc((void**)&img, sizeof(float) * n2) != hipSuccess ||
      hipMalloc((void**)&res, sizeof(float) * n2) != hipSuccess ||
      hipMemset(img, 0, sizeof(float) * n2) != hipSuccess ||
      hipMemset(res, 0, sizeof(float) * n2) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess ||  // the context's stream does not wait for the null stream
      hipEventCreate(&ea) != hipSuccess || hipEventCreate(&eb) != hipSuccess) {
    cleanup();
    set_error("gvm_autotune_adjoint: scratch allocation failed");
    return GVM_ERR_OOM;
  }
  // the tuning launches are not part of the adjoint timing
  ctx->timing_paused = true;
  long best = candidates[0];
  double best_ms = 1e300;
  gvm_status st = GVM_OK;
  for (int c = 0; c < ncand && st == GVM_OK; c++) {
    ctx->split_vis = candidates[c];
    if ((st = rebuild_splits(ctx))) break;
    // one launch to warm up, then three timed (range table, kernels and epilogue)
    if ((st = adjoint(ctx, img, res, 0, 0))) break;
    if (hipEventRecord(ea, ctx->stream) != hipSuccess) { st = GVM_ERR_HIP; break; }
    for (int r = 0; r < 3 && st == GVM_OK; r++) st = adjoint(ctx, img, res, 0, 0);
    if (st) break;
    float ms = 0.f;
    if (hipEventRecord(eb, ctx->stream) != hipSuccess || hipEventSynchronize(eb) != hipSuccess ||
        hipEventElapsedTime(&ms, ea, eb) != hipSuccess) {
      st = GVM_ERR_HIP;
      break;
    }
    const double avg = ms / 3.0;
    if (ms_out) ms_out[c] = avg;
    if (avg < best_ms) {
      best_ms = avg;
      best = candidates[c];
    }
  }
  ctx->timing_paused = false;
  if (st) {
    cleanup();
    if (st == GVM_ERR_HIP) set_error("gvm_autotune_adjoint: timing failed");
    return st;
  }
  cleanup();
  ctx->split_vis = best;
  ctx->split_mode = -1;
  ctx->epoch++;
  tune_store(key, best);
  *chosen = best;
  return GVM_OK;
}
