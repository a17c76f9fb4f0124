/*
 * gvm_oracle_grid.c -- CPU restatement of gpuvmem's gridded-visibility mode:
 * convolutional gridding (do_gridding) and the weighting schemes.
 *
 * TEST INFRASTRUCTURE ONLY (see gvm_oracle.h).  Compiled with
 * -ffp-contract=off: float and double expressions round after every
 * operation, in source order, as the reference's host code does on x86-64.
 *
 * Unlike the rest of the oracle this file is not FP64: the reference grids
 * in float on the host, one row after the other (one OpenMP thread), and
 * these functions repeat that float arithmetic in the same order, so that
 * the device gridder can be held to it bit for bit.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gvm_oracle.h"

#define ORC_RPDEG (3.14159265358979323846 / 180.0)

/* freq_to_wavelength (src/MSFITSIO.cu:36-39): a float. */
static float wavelength(float freq) { return 2.99792458E8f / freq; }

/* src/functions.cu:1339-1653 (do_gridding) for one channel x stokes slice,
 * with one gridding thread: rows z = 0 .. 2 nvis - 1 (z >= nvis: the
 * Hermitian twin of row z - nvis, u, v, w and Im Vo negated), each at cell
 * int(u / deltau + floor(N/2) + 0.5) and spread over the kernel's taps
 * (kernel value taps[kn * ki + kj]); float sums per cell; normalisation
 * (:1518-1559) and row-major compaction of the cells with weight > 0
 * (:1561-1607).  uvw_m in metres; out_* sized M*N. */
void orc_grid_visibilities(int M, int N, double deltax_deg, double deltay_deg, float freq,
                           long nvis, const double* uvw_m, const float* vo, const float* weight,
                           int km, int kn, int supx, int supy, const float* taps,
                           double* out_uvw, float* out_vo, float* out_w, long* nout) {
  double deltax = ORC_RPDEG * deltax_deg, deltay = ORC_RPDEG * deltay_deg;
  double deltau = 1.0 / (M * deltax), deltav = 1.0 / (N * deltay);
  long n = (long)M * N;
  float* gw = (float*)calloc(n, sizeof(float));
  float* gw2 = (float*)calloc(n, sizeof(float));
  float* gre = (float*)calloc(n, sizeof(float));
  float* gim = (float*)calloc(n, sizeof(float));
  double center_j = floor(N / 2.0), center_k = floor(M / 2.0);
  float lambda = wavelength(freq);
  for (long z = 0; z < 2 * nvis; z++) {
    long vi = z < nvis ? z : z - nvis;
    double u = uvw_m[3 * vi], v = uvw_m[3 * vi + 1];
    float w = weight[vi];
    float vre = vo[2 * vi], vim = vo[2 * vi + 1];
    if (z >= nvis) {
      u *= -1.0;
      v *= -1.0;
      vim *= -1.0f;
    }
    u = u / lambda; /* metres_to_lambda, src/MSFITSIO.cu:41-45 */
    v = v / lambda;
    double gx = u / deltau, gy = v / deltav;
    double j_fp = gx + center_j + 0.5, k_fp = gy + center_k + 0.5;
    if (!(fabs(j_fp) < 1073741824.0 && fabs(k_fp) < 1073741824.0)) continue;
    int j = (int)j_fp, k = (int)k_fp;
    for (int m = -supy; m <= supy; m++) {
      for (int q = -supx; q <= supx; q++) {
        int sj = j + q, sk = k + m;
        int ki = m + supy, kj = q + supx;
        if (sk >= 0 && sk < M && sj >= 0 && sj < N && ki >= 0 && ki < km && kj >= 0 &&
            kj < kn) {
          float kv = taps[kn * ki + kj];
          float kv2 = kv * kv;
          long gi = (long)N * sk + sj;
          gw[gi] += w * kv;
          gw2[gi] += w * kv2;
          gre[gi] += w * vre * kv;
          gim[gi] += w * vim * kv;
        }
      }
    }
  }
  long o = 0;
  for (int gk = 0; gk < M; gk++) {
    for (int gj = 0; gj < N; gj++) {
      long gi = (long)N * gk + gj;
      double ul = (gj - center_j) * deltau, vl = (gk - center_k) * deltav;
      float ws = gw[gi], aux = gw2[gi];
      float weff = 0.0f, re = 0.0f, im = 0.0f;
      if (aux != 0.0f && ws != 0.0f) {
        weff = ws * ws / aux;
        re = gre[gi] / ws;
        im = gim[gi] / ws;
      }
      if (weff > 0.0f) {
        out_uvw[3 * o] = ul * lambda;
        out_uvw[3 * o + 1] = vl * lambda;
        out_uvw[3 * o + 2] = 0.0;
        out_vo[2 * o] = re;
        out_vo[2 * o + 1] = im;
        out_w[o] = weff;
        o++;
      }
    }
  }
  *nout = o;
  free(gw);
  free(gw2);
  free(gre);
  free(gim);
}

/* UVTaper::getValue (include/classes/uvtaper.cuh:101-114). */
static float taper_value(const orc_taper* t, double u, double v) {
  double x = u - t->u_0, y = v - t->v_0;
  float cos_bpa = cosf(t->bpa), sin_bpa = sinf(t->bpa), sin_bpa_2 = sinf(2.0f * t->bpa);
  float a = (cos_bpa * cos_bpa) / (2.0f * t->sigma_maj * t->sigma_maj) +
            (sin_bpa * sin_bpa) / (2.0f * t->sigma_min * t->sigma_min);
  float b = sin_bpa_2 / (2.0f * t->sigma_maj * t->sigma_maj) -
            sin_bpa_2 / (2.0f * t->sigma_min * t->sigma_min);
  float c = (sin_bpa * sin_bpa) / (2.0f * t->sigma_maj * t->sigma_maj) +
            (cos_bpa * cos_bpa) / (2.0f * t->sigma_min * t->sigma_min);
  float value = t->amplitude * exp(-a * x * x - b * x * y - c * y * y);
  return value;
}

/* The cell of the weighting schemes: wavelengths, rows with u < 0 mirrored,
 * cells of |deltau| x |deltav| (src/uniformweightingscheme.cu:35-51,
 * src/briggsweightingscheme.cu:71-91); -1 off the grid. */
static long weight_cell(double* u, double* v, int M, int N, double adu, double adv) {
  if (*u < 0.0) {
    *u *= -1.0;
    *v *= -1.0;
  }
  double gx = *u / adu, gy = *v / adv;
  double xf = gx + (int)floor(N / 2) + 0.5, yf = gy + (int)floor(M / 2) + 0.5;
  if (!(fabs(xf) < 1073741824.0 && fabs(yf) < 1073741824.0)) return -1;
  int x = (int)xf, y = (int)yf;
  if (x >= 0 && y >= 0 && x < N && y < M) return (long)N * y + x;
  return -1;
}

/* The weighting schemes (WeightingScheme::apply):
 *   0 Natural  src/naturalweightingscheme.cu:9-48   (taper only)
 *   1 Uniform  src/uniformweightingscheme.cu:9-99   w / G(cell)
 *   2 Briggs   src/briggsweightingscheme.cu:30-187  w / (1 + G(cell) f^2)
 *   3 Radial   src/radialweightingscheme.cu:8-43    w * |(u, v)|
 * then the taper (NULL: none).  Briggs: f^2 = (5 10^-robust)^2 / (S2 / S1),
 * S1 the sum of all weights, S2 the sum over slices of sum G^2 over the
 * columns n >= N/2 of a grid that pass 1 does not clear between slices.
 * float_totals = 1 sums S1 and S2 in float in the reference's order;
 * 0 sums them in FP64 (the device's choice).  Weights are updated in place. */
void orc_weighting(int scheme, float robust, const orc_taper* taper, int M, int N,
                   double deltax_deg, double deltay_deg, int nsl, const float* freq,
                   const long* nvis, const double* const* uvw_m, float* const* weight,
                   int float_totals) {
  double deltax = ORC_RPDEG * deltax_deg, deltay = ORC_RPDEG * deltay_deg;
  double adu = fabs(1.0 / (M * deltax)), adv = fabs(1.0 / (N * deltay));
  long n = (long)M * N;
  float* g = (float*)calloc(n, sizeof(float));
  float f_squared = 0.0f;
  if (scheme == 2) {
    float s1f = 0.0f, s2f = 0.0f;
    double s1d = 0.0, s2d = 0.0;
    for (int s = 0; s < nsl; s++) {
      float acc = 0.0f;
      for (long z = 0; z < nvis[s]; z++) {
        acc = acc + weight[s][z];
        s1d += (double)weight[s][z];
      }
      s1f += acc;
    }
    for (int s = 0; s < nsl; s++) {
      float lambda = wavelength(freq[s]);
      for (long z = 0; z < nvis[s]; z++) {
        double u = uvw_m[s][3 * z] / lambda, v = uvw_m[s][3 * z + 1] / lambda;
        long c = weight_cell(&u, &v, M, N, adu, adv);
        if (c >= 0) g[c] += weight[s][z];
      }
      for (int m = 0; m < M; m++)
        for (int q = N / 2; q < N; q++) {
          float gv = g[(long)N * m + q];
          s2f += gv * gv;
          s2d += (double)(gv * gv);
        }
    }
    float average_weights = float_totals ? s2f / s1f : (float)s2d / (float)s1d;
    f_squared = (5.0f * powf(10.0f, -robust)) * (5.0f * powf(10.0f, -robust)) / average_weights;
  }
  for (int s = 0; s < nsl; s++) {
    float lambda = wavelength(freq[s]);
    if (scheme == 1 || scheme == 2) {
      memset(g, 0, n * sizeof(float));
      for (long z = 0; z < nvis[s]; z++) {
        double u = uvw_m[s][3 * z] / lambda, v = uvw_m[s][3 * z + 1] / lambda;
        long c = weight_cell(&u, &v, M, N, adu, adv);
        if (c >= 0) g[c] += weight[s][z];
      }
    }
    for (long z = 0; z < nvis[s]; z++) {
      double u = uvw_m[s][3 * z] / lambda, v = uvw_m[s][3 * z + 1] / lambda;
      float* w = &weight[s][z];
      if (scheme == 3) {
        float fu = (float)u, fv = (float)v;
        float sumsqr = (fu - 0.0f) * (fu - 0.0f) + (fv - 0.0f) * (fv - 0.0f);
        *w *= sqrtf(sumsqr);
      } else {
        long c = weight_cell(&u, &v, M, N, adu, adv);
        if (scheme == 1 || scheme == 2) {
          if (c < 0)
            *w = 0.0f;
          else if (scheme == 1)
            *w /= g[c];
          else
            *w /= (1.0 + g[c] * f_squared);
        }
      }
      if (taper) *w *= taper_value(taper, u, v);
    }
  }
  free(g);
}
