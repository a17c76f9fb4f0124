// Gridding kernels and their correction images (include/gvm/ckernel.h).
// Float arithmetic in the reference's order (host code, no contraction), so
// the taps and images are the ones the reference's CKernel classes build.
#include "gvm/ckernel.h"

#include <math.h>

#include <cmath>

namespace {

constexpr float kPiF = 3.14159265358979323846f;  // PI (include/MSFITSIO.cuh:52)

// distance (src/MSFITSIO.cu:47-51)
float distance2(float x, float y, float x0, float y0) {
  const float s = (x - x0) * (x - x0) + (y - y0) * (y - y0);
  return sqrtf(s);
}

// pillBox1D / pillBox2D (src/pillBox2D.cu:3-16)
float pillbox1(float amp, float x, float limit) { return fabsf(x) < limit ? amp : 0.0f; }

// gaussian2D (src/gaussian2D.cu:17-38): zero outside |x| < w sigma_x,
// |y| < w sigma_y; exp(-(fx^alpha + fy^alpha)) inside.
float gaussian2(float amp, float x, float y, float x0, float y0, float sx, float sy, float w,
                float alpha) {
  const float rx = distance2(x, 0.0f, x0, 0.0f);
  const float ry = distance2(0.0f, y, 0.0f, y0);
  if (rx < w * sx && ry < w * sy) {
    const float fx = rx / (w * sx);
    const float fy = ry / (w * sy);
    const float vx = powf(fx, alpha);
    const float vy = powf(fy, alpha);
    return amp * expf(-1.0f * (vx + vy));
  }
  return 0.0f;
}

// sincf / sinc1D / sinc2D (src/sinc2D.cu:3-35); sinc2 takes (x, x0, y, y0).
float sinc0(float x) { return x == 0.0f ? 1.0f : sinf(kPiF * x) / (kPiF * x); }

float sinc1(float amp, float x, float x0, float sigma, float w) {
  const float r = distance2(x, 0.0f, x0, 0.0f);
  const float v = r / (w * sigma);
  return r < w * sigma ? amp * sinc0(v) : 0.0f;
}

float sinc2(float amp, float x, float x0, float y, float y0, float sx, float sy, float w) {
  const float sxv = sinc1(1.0f, x, x0, sx, w);
  const float syv = sinc1(1.0f, y, y0, sy, w);
  return amp * sxv * syv;
}

// The rational approximation of the prolate spheroidal function of Schwab
// (F. R. Schwab, "Optimal gridding of visibility data in radio
// interferometry", Indirect Imaging, 1984) for support m = 6, alpha = 1, as
// tabulated in src/pswf_12D.cu:3-53: two pieces split at |nu| = 0.75,
// numerator degree 4 and denominator degree 2 in nu^2 - nu_end^2.
float pswf_rational(float nu) {
  static const float P[2][5] = {
      {8.203343e-2f, -3.644705e-1f, 6.278660e-1f, -5.335581e-1f, 2.312756e-1f},
      {4.028559e-3f, -3.697768e-2f, 1.021332e-1f, -1.201436e-1f, 6.412774e-2f}};
  static const float Q[2][3] = {{1.0000000e0f, 8.212018e-1f, 2.078043e-1f},
                                {1.0000000e0f, 9.599102e-1f, 2.918724e-1f}};
  const float a = fabsf(nu);
  if (a > 1.0f) return 0.0f;
  const int k = (a >= 0.0f && a < 0.75) ? 0 : 1;
  const float end = k == 0 ? 0.75f : 1.0f;
  const float d = a * a - end * end;
  float top = P[k][0], bottom = Q[k][0];
  for (int i = 1; i < 5; i++) top += P[k][i] * powf(d, (float)i);
  for (int i = 1; i < 3; i++) bottom += Q[k][i] * powf(d, (float)i);
  return bottom > 0.0f ? top / bottom : 0.0f;
}

// pswf_11D / pswf_12D (src/pswf_12D.cu:55-84): (1 - nu^2) psi(nu), 1 at nu = 0.
float pswf1(float amp, float x, float x0, float sigma, float w) {
  const float r = distance2(x, 0.0f, x0, 0.0f);
  const float nu = r / (w * sigma);
  if (nu == 0.0f) return 1.0f;
  const float ps = pswf_rational(nu);
  const float nu2 = nu * nu;
  return amp * (1.0f - nu2) * ps;
}

float pswf2(float amp, float x, float y, float x0, float y0, float sx, float sy, float w) {
  const float xv = pswf1(1.0f, x, x0, sx, w);
  const float yv = pswf1(1.0f, y, y0, sy, w);
  return amp * xv * yv;
}

// One tap of kind k at (x, y) (CKernel::buildKernel of each class; amp, x0,
// y0 as the constructors set them: 1, 0, 0).
float tap(const gvm_ckernel_params& p, float x, float y, float sx, float sy) {
  const float x0 = 0.0f, y0 = 0.0f;
  switch (p.kind) {
    case GVM_CKERNEL_PILLBOX: {
      // src/pillBox2D.cu:169-186: limits (m / 2) sigma_x and (n / 2) sigma_y
      const float lx = (p.m / 2.0f) * sx, ly = (p.n / 2.0f) * sy;
      return pillbox1(p.amp, x, lx) * pillbox1(p.amp, y, ly);
    }
    case GVM_CKERNEL_GAUSSIAN:
      return gaussian2(p.amp, x, y, x0, y0, sx, sy, p.w, p.alpha);
    case GVM_CKERNEL_SINC:
      // Sinc2D::buildKernel (src/sinc2D.cu:154-167) passes (x, y, x0, y0) to
      // sinc2D(amp, x, x0, y, y0, ...): the tap is sinc(|x - y| / (w sigma_x)).
      return sinc2(p.amp, x, y, x0, y0, sx, sy, p.w);
    case GVM_CKERNEL_GAUSSIAN_SINC: {
      // gaussianSinc2D (src/gaussianSinc2D.cu:14-27)
      const float G = gaussian2(1.0f, x, y, x0, y0, sx, sy, p.w, p.alpha);
      const float S = sinc2(1.0f, x, x0, y, y0, sx, sy, p.w2);
      return p.amp * G * S;
    }
    case GVM_CKERNEL_PSWF:
      return pswf2(p.amp, x, y, x0, y0, sx, sy, p.w);
    default:
      return 0.0f;
  }
}

}  // namespace

extern "C" int gvm_ckernel_defaults(int kind, gvm_ckernel_params* p) {
  if (!p || kind < GVM_CKERNEL_PILLBOX || kind > GVM_CKERNEL_PSWF) return 1;
  p->kind = kind;
  p->m = p->n = kind == GVM_CKERNEL_PILLBOX ? 1 : 7;
  p->amp = 1.0f;
  p->w = kind == GVM_CKERNEL_GAUSSIAN_SINC ? 2.52f : (kind == GVM_CKERNEL_PSWF ? 6.0f : 1.0f);
  p->w2 = 1.55f;
  p->alpha = 2.0f;
  return 0;
}

extern "C" int gvm_ckernel_build(const gvm_ckernel_params* p, float sigma_x, float sigma_y,
                                 float* taps, int* support_x, int* support_y) {
  if (!p || !taps || p->m < 1 || p->n < 1 || p->kind < GVM_CKERNEL_PILLBOX ||
      p->kind > GVM_CKERNEL_PSWF)
    return 1;
  const int s = (int)floorf(p->m / 2.0f);  // CKernel::setSupports: both from m
  for (int i = 0; i < p->m; i++)
    for (int j = 0; j < p->n; j++) {
      const float y = (i - s) * sigma_y;
      const float x = (j - s) * sigma_x;
      taps[(long)p->n * i + j] = tap(*p, x, y, sigma_x, sigma_y);
    }
  if (support_x) *support_x = s;
  if (support_y) *support_y = s;
  return 0;
}

extern "C" int gvm_ckernel_gcf(const gvm_ckernel_params* p, int M, int N, float dx, float dy,
                               float* gcf) {
  if (!p || !gcf || M < 1 || N < 1) return 1;
  if (p->kind != GVM_CKERNEL_PILLBOX && p->kind != GVM_CKERNEL_GAUSSIAN &&
      p->kind != GVM_CKERNEL_PSWF)
    return 1;
  // the clone: m = M, n = N, sigmas (dx, dy), w = M, supports floor(M / 2)
  const float w = (float)M;
  const int s = (int)floorf(M / 2.0f);
  for (int i = 0; i < M; i++)
    for (int j = 0; j < N; j++) {
      const float y = (i - s) * dy;
      const float x = (j - s) * dx;
      float v = 1.0f;  // PillBox2D::GCF (src/pillBox2D.cu:206-214)
      if (p->kind == GVM_CKERNEL_GAUSSIAN)
        v = gaussian2(p->amp, kPiF * x, kPiF * y, kPiF * 0.0f, kPiF * 0.0f, dx, dy, 2.0f * w,
                      p->alpha);
      else if (p->kind == GVM_CKERNEL_PSWF)
        v = 1.0f / pswf2(p->amp, x, y, 0.0f, 0.0f, dx, dy, w);
      gcf[(long)N * i + j] = v;
    }
  return 0;
}
