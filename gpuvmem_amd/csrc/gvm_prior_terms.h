// Per-pixel prior terms (Entropy, TSV, Laplacian) shared by the kernels
// that evaluate them: the one-term and fused kernels of gvm_ops.hip and the
// fused value kernel of gvm_forward.hip.
#pragma once

#include "gvm_internal.h"

namespace gvm {

enum { PRIOR_ENTROPY = 0, PRIOR_TSV = 1, PRIOR_LAPLACIAN = 2 };

// Row and column of pixel k of an N x N image; N is a power of two
// (gvm_create), so no 64-bit division.
__device__ __forceinline__ int pix_row(long k, int N) { return (int)(k >> (31 - __clz(N))); }
__device__ __forceinline__ int pix_col(long k, int N) { return (int)k & (N - 1); }

// calculateS / calculateTSV / calculateL for pixel (i, j) of image `im`.
__device__ __forceinline__ float prior_term(int kind, const float* __restrict__ im, int N, int i,
                                            int j, float G, float eta) {
  const long k = (long)N * i + j;
  const float c = im[k];
  if (kind == PRIOR_ENTROPY) return c * logf((c / G) + (eta + 1.0f));
  if (kind == PRIOR_TSV) {
    if (i < N - 1 && j < N - 1) {
      const float dx = c - im[k + 1], dy = c - im[k + N];
      return dx * dx + dy * dy;
    }
    return c;
  }
  if (i > 0 && i < N - 1 && j > 0 && j < N - 1) {
    const float Dx = im[k - 1] - 2.0f * c + im[k + 1];
    const float Dy = im[k - N] - 2.0f * c + im[k + N];
    return 0.5f * (Dx + Dy) * (Dx + Dy);
  }
  return c;
}

// calculateDS / calculateDTSV / calculateDL (without lambda).
__device__ __forceinline__ float prior_grad(int kind, const float* __restrict__ im, int N, int i,
                                            int j, float G, float eta) {
  const long k = (long)N * i + j;
  const float c = im[k];
  if (kind == PRIOR_ENTROPY)
    return logf((c / G) + (eta + 1.0f)) + 1.0f / (1.0f + (((eta + 1.0f) * G) / c));
  if (kind == PRIOR_TSV) {
    if (i > 0 && i < N - 1 && j > 0 && j < N - 1)
      return 8.0f * c - 2.0f * (im[k - N] + im[k - 1] + im[k + N] + im[k + 1]);
    return c;
  }
  if (i > 1 && i < N - 2 && j > 1 && j < N - 2) {
    const float d = im[k + N], u = im[k - N], r = im[k + 1], l = im[k - 1];
    const float corners = im[k + N - 1] + im[k + N + 1] + im[k - N - 1] + im[k - N + 1];
    const float far = im[k + 2 * N] + im[k + 2] + im[k - 2 * N] + im[k - 2];
    // the reference's stencil: -8 (d - r - u - l)
    return 20.0f * c - 8.0f * (d - r - u - l) + 2.0f * corners + far;
  }
  return 0.0f;
}

// The terms of one fused pass (the objective's terms after Chi2), passed to
// the kernels by value.
struct PriorTermsDev {
  gvm_prior_term t[GVM_MAX_FUSED_PRIORS];
  int n;
};

// Validates n <= GVM_MAX_FUSED_PRIORS terms against the context's images and
// copies them into *out.
gvm_status check_terms(const gvm_ctx* ctx, const gvm_prior_term* terms, int n,
                       PriorTermsDev* out);

}  // namespace gvm
