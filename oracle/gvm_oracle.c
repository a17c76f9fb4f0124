/*
 * gvm_oracle.c -- FP64 CPU restatement of the gpuvmem maximum-entropy hot
 * path.  TEST INFRASTRUCTURE ONLY (see gvm_oracle.h): the product path never
 * loads this library.
 *
 * Each function cites the reference lines it follows (paths relative to the
 * reference repository root).  Arithmetic is double; semantics (masks, gating,
 * border rules, quirks) follow the reference exactly.
 */
#define _GNU_SOURCE
#include "gvm_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846
#define ORC_RPDEG (ORC_PI / 180.0)
#define ORC_LIGHTSPEED 2.99792458e8
/* first zero of J1 (reference src/functions.cu AiryDiskBeam uses RZ) */
#define ORC_RZ 1.2196698912665045

/* The gridding correction factor of pixel k: apply_GCF and the DChi2 variant
 * with gcf (src/functions.cu:2468-2476, 3793-3888); 1 without gridding. */
static inline double gcf_at(const orc_geom* g, long k) {
  return g->gcf ? (double)g->gcf[k] : 1.0;
}

/* Thread count of the OpenMP regions that follow (the environment's
 * OMP_NUM_THREADS is read once, when the OpenMP runtime loads). */
void orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}

int orc_max_threads(void) { return omp_get_max_threads(); }

/* sin/cos of 2*pi*t with the argument reduced exactly in turns first. */
static inline void sincos_turns(double t, double* s, double* c) {
  double r = t - nearbyint(t);
  sincos(2.0 * ORC_PI * r, s, c);
}

/* ------------------------------------------------------------------------
 * Set-up (row a19)
 * ---------------------------------------------------------------------- */

/* src/functions.cu:2256-2273 (hermitianSymmetry) and src/MSFITSIO.cu:36-45
 * (freq_to_wavelength / metres_to_lambda: the wavelength is a float). */
void orc_hermitian_to_lambda(long n, double* uvw, float* vo, float freq) {
  float lambda = (float)(2.99792458E8f / freq);
  for (long i = 0; i < n; i++) {
    if (uvw[3 * i] > 0.0) {
      uvw[3 * i] *= -1.0;
      uvw[3 * i + 1] *= -1.0;
      vo[2 * i + 1] *= -1.0f;
    }
    uvw[3 * i] = uvw[3 * i] / lambda;
    uvw[3 * i + 1] = uvw[3 * i + 1] / lambda;
    uvw[3 * i + 2] = uvw[3 * i + 2] / lambda;
  }
}

/* src/functions.cu:1655-1700 (calc_sBeam, calc_beamSize) and :1700-1814
 * (calculateNoiseAndBeam): weighted second moments of (u,v) in wavelengths
 * give the clean-beam axes; the visibility noise is 0.5*sqrt(1/sum w).
 * ch_metres holds uvw in metres (before the Hermitian flip). */
void orc_beam_and_noise(const orc_chan* ch, int nch, double deltax_deg,
                        double* bmaj_deg, double* bmin_deg, double* bpa_deg,
                        double* sum_weights, double* vis_noise) {
  (void)deltax_deg;
  double suu = 0, svv = 0, suv = 0, sw = 0;
  for (int c = 0; c < nch; c++) {
    float lambda = (float)(2.99792458E8f / ch[c].freq);
    for (long i = 0; i < ch[c].nvis; i++) {
      double u = ch[c].uvw[3 * i] / lambda;
      double v = ch[c].uvw[3 * i + 1] / lambda;
      double w = ch[c].weight[i];
      suu += u * u * w;
      svv += v * v * w;
      suv += u * v * w;
      sw += w;
    }
  }
  suu /= sw;
  svv /= sw;
  suv /= sw;
  double d = suu - svv;
  double s = suu + svv;
  double root = sqrt(d * d + 4.0 * suv * suv);
  double bmaj = 1.0 / sqrt(2.0) / ORC_PI / sqrt(s - root);
  double bmin = 1.0 / sqrt(2.0) / ORC_PI / sqrt(s + root);
  double bpa = -0.5 * atan2(2.0 * suv, d);
  *bmaj_deg = bmaj / ORC_RPDEG;
  *bmin_deg = bmin / ORC_RPDEG;
  *bpa_deg = bpa / ORC_RPDEG;
  *sum_weights = sw;
  *vis_noise = 0.5 * sqrt(1.0 / sw);
}

/* src/mfs.cu:640-643: noise per pixel at the beam FWHM; beam axes in pixels
 * are the degree values over |DELTAX|. */
double orc_noise_jypix(double vis_noise, double bmaj_deg, double bmin_deg,
                       double deltax_deg) {
  double bmaj = bmaj_deg / fabs(deltax_deg);
  double bmin = bmin_deg / fabs(deltax_deg);
  return vis_noise / (ORC_PI * bmaj * bmin / (4.0 * log(2.0)));
}

/* ------------------------------------------------------------------------
 * Primary beam (row a3): src/functions.cu:2275-2333 (AiryDiskBeam,
 * GaussianBeam, attenuation).  Centre truncated to int, distance in radians,
 * zero beyond pb_cutoff.
 * ---------------------------------------------------------------------- */
double orc_attenuation(const orc_beam* b, double freq, int i, int j,
                       float xobs, float yobs, double dx_deg, double dy_deg) {
  int x0 = (int)xobs;
  int y0 = (int)yobs;
  double x = (j - x0) * dx_deg * ORC_RPDEG;
  double y = (i - y0) * dy_deg * ORC_RPDEG;
  double arc = sqrt(x * x + y * y);
  double lambda = ORC_LIGHTSPEED / freq;
  double atten;
  if (b->kind == ORC_GAUSSIAN) {
    double fwhm = b->pb_factor * lambda / b->diameter;
    double r = arc / fwhm;
    atten = exp(-4.0 * log(2.0) * r * r);
  } else {
    atten = 1.0;
    if (arc != 0.0) {
      double z = ORC_PI * arc * b->diameter / lambda * (ORC_RZ / b->pb_factor);
      double j1 = jn(1, z);
      atten = 4.0 * (j1 / z) * (j1 / z);
    }
  }
  return (arc <= b->pb_cutoff) ? atten : 0.0;
}

void orc_pb_image(const orc_beam* b, double freq, int M, int N, float xobs,
                  float yobs, double dx_deg, double dy_deg, double* out) {
  for (int i = 0; i < N; i++)
    for (int j = 0; j < M; j++)
      out[(long)N * i + j] =
          orc_attenuation(b, freq, i, j, xobs, yobs, dx_deg, dy_deg);
}

/* src/mfs.cu:849-916 with src/functions.cu:2382-2422 (total_attenuation,
 * weight_image, noise_image): one field, attenuation at nu_0,
 * noise = noise_jypix * sqrt(max(PB^2) / PB^2); noise_min is the smallest
 * noise (it becomes fg_scale, and noise_cut is scaled by it). */
void orc_noise_image(const orc_beam* b, float nu_0, int M, int N, double dx,
                     double dy, float x0_ref, float y0_ref, double noise_jypix,
                     double* noise_out, double* noise_min) {
  long n = (long)M * N;
  double* w = (double*)malloc(sizeof(double) * n);
  double wmax = 0.0;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < M; j++) {
      double a = orc_attenuation(b, nu_0, i, j, x0_ref, y0_ref, dx, dy);
      w[(long)N * i + j] = a * a;
      if (a * a > wmax) wmax = a * a;
    }
  double nmin = INFINITY;
  for (long k = 0; k < n; k++) {
    double nw = (w[k] / wmax) / (noise_jypix * noise_jypix);
    double v = sqrt(1.0 / nw);
    noise_out[k] = v;
    if (v < nmin) nmin = v;
  }
  *noise_min = nmin;
  free(w);
}

/* ------------------------------------------------------------------------
 * FFT (row a4).  The reference calls cuFFT C2C with CUFFT_INVERSE
 * (src/functions.cu:2153-2178): an unnormalised 2-D DFT with the e^{+2 pi i}
 * sign.  Restated as an iterative radix-2 FFT with twiddles from sincos of
 * the exact angle (no recurrences).  Power-of-two sizes only.
 * ---------------------------------------------------------------------- */
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void fft1d(ocplx* a, int n, int stride, int sign, const ocplx* tw) {
  /* bit reversal */
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      ocplx t = a[(long)i * stride];
      a[(long)i * stride] = a[(long)j * stride];
      a[(long)j * stride] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1;
    int step = n / len;
    for (int s = 0; s < n; s += len) {
      for (int k = 0; k < half; k++) {
        ocplx w = tw[(long)k * step];
        if (sign < 0) w.im = -w.im;
        ocplx* x = &a[(long)(s + k) * stride];
        ocplx* y = &a[(long)(s + k + half) * stride];
        double tr = y->re * w.re - y->im * w.im;
        double ti = y->re * w.im + y->im * w.re;
        y->re = x->re - tr;
        y->im = x->im - ti;
        x->re += tr;
        x->im += ti;
      }
    }
  }
}

int orc_fft2d(ocplx* a, int rows, int cols, int sign) {
  if (!is_pow2(rows) || !is_pow2(cols)) return -1;
  int nmax = rows > cols ? rows : cols;
  ocplx* tw = (ocplx*)malloc(sizeof(ocplx) * nmax);
  /* rows: length cols */
  for (int k = 0; k < cols; k++) sincos_turns((double)k / cols, &tw[k].im, &tw[k].re);
#pragma omp parallel for schedule(static)
  for (int r = 0; r < rows; r++) fft1d(a + (long)r * cols, cols, 1, sign, tw);
  for (int k = 0; k < rows; k++) sincos_turns((double)k / rows, &tw[k].im, &tw[k].re);
#pragma omp parallel
  {
    ocplx* col = (ocplx*)malloc(sizeof(ocplx) * rows);
#pragma omp for schedule(static)
    for (int c = 0; c < cols; c++) {
      for (int r = 0; r < rows; r++) col[r] = a[(long)r * cols + c];
      fft1d(col, rows, 1, sign, tw);
      for (int r = 0; r < rows; r++) a[(long)r * cols + c] = col[r];
    }
    free(col);
  }
  free(tw);
  return 0;
}

/* ------------------------------------------------------------------------
 * Forward chi2 (rows a1-a8)
 * ---------------------------------------------------------------------- */

/* src/functions.cu:2694-2719 (clip2IWNoise) via :4594-4599: masked pixels
 * are set to the positivity floor (image 0) and alpha to 0; with a schedule
 * (flag_opt > 0) alpha is also zeroed where I_nu_0 < threshold.  In place. */
void orc_clip2IWNoise(const orc_geom* g, double* I, int flag_opt) {
  long n = (long)g->M * g->N;
  for (long k = 0; k < n; k++) {
    if (g->noise[k] > g->noise_cut) {
      if (g->eta > 0.0f)
        I[k] = 0.0;
      else
        I[k] = -1.0 * (double)g->eta * (double)g->minpix;
      I[n + k] = 0.0;
    } else {
      if (I[k] < g->threshold && flag_opt > 0) I[n + k] = 0.0;
    }
  }
}

/* src/functions.cu:3939-3966 (calculateInu), :2424-2444 (apply_beam2I),
 * :2153-2178 (FFT2D, inverse, no shift) and :2483-2518 (phase_rotate).
 * Produces the phase-rotated uv grid V[row][col] for one channel. */
void orc_model_grid(const orc_geom* g, const orc_beam* b, const double* I,
                    float freq, ocplx* grid) {
  int M = g->M, N = g->N;
  long n = (long)M * N;
  double nudiv = (double)freq / (double)g->nu_0;
  double floor_v = -1.0 * (double)g->eta * (double)g->minpix;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    for (int j = 0; j < M; j++) {
      long k = (long)N * i + j;
      double inu = I[k] * pow(nudiv, I[n + k]);
      if (inu < floor_v) inu = floor_v;
      double pb = orc_attenuation(b, freq, i, j, g->x0_ref, g->y0_ref,
                                  g->deltax, g->deltay);
      grid[k].re = inu * pb * (double)g->fg_scale * gcf_at(g, k);
      grid[k].im = 0.0;
    }
  }
  orc_fft2d(grid, N, M, +1);
  double upix = (double)g->x0_phs / (double)M;
  double vpix = (double)g->y0_phs / (double)N;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double v = vpix * (i < N / 2 ? i : i - N);
    for (int j = 0; j < M; j++) {
      double u = upix * (j < M / 2 ? j : j - M);
      double s, c;
      sincos_turns(-(u + v), &s, &c);
      long k = (long)N * i + j;
      double re = grid[k].re * c - grid[k].im * s;
      double im = grid[k].re * s + grid[k].im * c;
      grid[k].re = re;
      grid[k].im = im;
    }
  }
}

/* src/functions.cu:2557-2609 (vis_mod), :2663-2671 (residual),
 * :2867-2876 (chi2Vector) and the reduction: bilinear interpolation of the
 * wrap-around grid (column from u, row from v), V_r = V_o - V_m,
 * sum w |V_r|^2.  Samples off the grid get weight 0 (permanently) and V_m 0. */
double orc_degrid_chi2(const orc_geom* g, const ocplx* grid, orc_chan* ch,
                       ocplx* vm_out, ocplx* vr_out) {
  long N = g->N;
  double sum = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sum)
  for (long i = 0; i < ch->nvis; i++) {
    double ux = ch->uvw[3 * i] / g->deltau;
    double uy = ch->uvw[3 * i + 1] / g->deltav;
    if (ux < 0.0) ux += N;
    if (uy < 0.0) uy += N;
    long i1 = (long)floor(ux);
    long i2 = (i1 + 1) % N;
    double du = ux - i1;
    long j1 = (long)floor(uy);
    long j2 = (j1 + 1) % N;
    double dv = uy - j1;
    ocplx vm = {0.0, 0.0};
    if (i1 >= 0 && i1 < N && j1 >= 0 && j1 < N) {
      ocplx v11 = grid[N * j1 + i1], v12 = grid[N * j2 + i1];
      ocplx v21 = grid[N * j1 + i2], v22 = grid[N * j2 + i2];
      double w11 = (1.0 - du) * (1.0 - dv), w12 = (1.0 - du) * dv;
      double w21 = du * (1.0 - dv), w22 = du * dv;
      vm.re = w11 * v11.re + w12 * v12.re + w21 * v21.re + w22 * v22.re;
      vm.im = w11 * v11.im + w12 * v12.im + w21 * v21.im + w22 * v22.im;
    } else {
      ch->weight[i] = 0.0f;
    }
    ocplx vr = {(double)ch->vo[2 * i] - vm.re, (double)ch->vo[2 * i + 1] - vm.im};
    if (vm_out) vm_out[i] = vm;
    if (vr_out) vr_out[i] = vr;
    sum += (double)ch->weight[i] * (vr.re * vr.re + vr.im * vr.im);
  }
  return sum;
}

/* src/functions.cu:4323-4454 (chi2): clip in place, then per channel the
 * model grid and the degrid/residual/chi2 sum; optional per-slice division by
 * nvis (normalize); returns 0.5 * total.  vr_out receives the residuals of
 * all channels back to back. */
double orc_chi2(const orc_geom* g, const orc_beam* b, orc_chan* ch, int nch,
                double* I, int flag_opt, int normalize, ocplx* vr_out) {
  long n = (long)g->M * g->N;
  ocplx* grid = (ocplx*)malloc(sizeof(ocplx) * n);
  double total = 0.0;
  long off = 0;
  orc_clip2IWNoise(g, I, flag_opt);
  for (int c = 0; c < nch; c++) {
    if (ch[c].nvis <= 0) continue;
    orc_model_grid(g, b, I, ch[c].freq, grid);
    double r = orc_degrid_chi2(g, grid, &ch[c], NULL, vr_out ? vr_out + off : NULL);
    if (normalize) r /= (double)ch[c].nvis;
    total += r;
    off += ch[c].nvis;
  }
  free(grid);
  return 0.5 * total;
}

/* ------------------------------------------------------------------------
 * Adjoint (row a10): src/functions.cu:3698-3791 (DChi2).  Per pixel inside
 * the mask: x = (j - int(x0_phs)) dx, y = (i - int(y0_phs)) dy (radians),
 * z = sqrt(1 - x^2 - y^2), phase = 2 pi (x u + y v + (z - 1) w),
 * d = -fg_scale * PB * sum_v w_v (Re Vr cos + Im Vr sin) [/ nvis].
 * ---------------------------------------------------------------------- */
static double pixel_dchi2(const orc_geom* g, const orc_chan* ch,
                          const ocplx* vr, int i, int j) {
  double x = (j - (int)g->x0_phs) * g->deltax * ORC_RPDEG;
  double y = (i - (int)g->y0_phs) * g->deltay * ORC_RPDEG;
  double zm1 = sqrt(1.0 - x * x - y * y) - 1.0;
  double s = 0.0;
  for (long v = 0; v < ch->nvis; v++) {
    double t = x * ch->uvw[3 * v] + y * ch->uvw[3 * v + 1] + zm1 * ch->uvw[3 * v + 2];
    double sn, cs;
    sincos_turns(t, &sn, &cs);
    s += (double)ch->weight[v] * (vr[v].re * cs + vr[v].im * sn);
  }
  return s;
}

void orc_dchi2_direct_pixels(const orc_geom* g, const orc_beam* b,
                             const orc_chan* ch, const ocplx* vr, int normalize,
                             const int* pix, int npix, double* out) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int p = 0; p < npix; p++) {
    int i = pix[2 * p], j = pix[2 * p + 1];
    long k = (long)g->N * i + j;
    if (!(g->noise[k] < g->noise_cut)) {
      out[p] = 0.0;
      continue;
    }
    double atten = orc_attenuation(b, ch->freq, i, j, g->x0_ref, g->y0_ref,
                                   g->deltax, g->deltay);
    double d = pixel_dchi2(g, ch, vr, i, j) * (double)g->fg_scale * atten * gcf_at(g, k);
    if (normalize) d /= (double)ch->nvis;
    out[p] = -d;
  }
}

/* orc_dchi2 (direct sums, every channel, chain rule) at sampled pixels only:
 * out[2 p] = result[0] and out[2 p + 1] = result[1] of pixel p.  The
 * chain rule of DChi2_total_I_nu_0 / DChi2_total_alpha
 * (src/functions.cu:3968-4024), accumulated channel after channel as in
 * orc_dchi2. */
void orc_dchi2_pixels(const orc_geom* g, const orc_beam* b, const orc_chan* ch, int nch,
                      const double* I, const ocplx* vr, int flag_opt, int normalize,
                      const int* pix, int npix, double* out) {
  long n = (long)g->M * g->N;
  double* d = (double*)malloc(sizeof(double) * (size_t)(npix > 0 ? npix : 1));
  for (int p = 0; p < 2 * npix; p++) out[p] = 0.0;
  long off = 0;
  for (int c = 0; c < nch; c++) {
    if (ch[c].nvis <= 0) continue;
    orc_dchi2_direct_pixels(g, b, &ch[c], vr + off, normalize, pix, npix, d);
    off += ch[c].nvis;
    double nudiv = (double)ch[c].freq / (double)g->nu_0;
    for (int p = 0; p < npix; p++) {
      long k = (long)g->N * pix[2 * p] + pix[2 * p + 1];
      if (!(g->noise[k] < g->noise_cut)) continue;
      double i0 = I[k], alpha = I[n + k];
      double dI = pow(nudiv, alpha);
      if (flag_opt % 2 == 0)
        out[2 * p] += d[p] * dI;
      else if (i0 > g->threshold)
        out[2 * p + 1] += d[p] * i0 * dI * (double)g->fg_scale * log(nudiv);
    }
  }
  free(d);
}

/* Separable evaluation of the same sum.  exp(2 pi i (x u + y v + (z-1) w))
 * is split as exp(2 pi i (u x + w f(x))) * exp(2 pi i (v y + w f(y))) with
 * f(t) = sqrt(1 - t^2) - 1; the neglected cross term is about w x^2 y^2 / 4
 * turns.  Then d(i,j) = Re sum_k [w_k conj(Vr_k) Ey_k(i)] Ex_k(j): a complex
 * matrix product over the visibilities, done in K-blocks. */
#define ORC_KB 64
void orc_dchi2_channel(const orc_geom* g, const orc_beam* b,
                       const orc_chan* ch, const ocplx* vr, int normalize,
                       int separable, double* dchi2) {
  int M = g->M, N = g->N;
  long n = (long)M * N;
  if (!separable) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < N; i++)
      for (int j = 0; j < M; j++) {
        long k = (long)N * i + j;
        if (!(g->noise[k] < g->noise_cut)) {
          dchi2[k] = 0.0;
          continue;
        }
        double atten = orc_attenuation(b, ch->freq, i, j, g->x0_ref, g->y0_ref,
                                       g->deltax, g->deltay);
        double d = pixel_dchi2(g, ch, vr, i, j) * (double)g->fg_scale * atten * gcf_at(g, k);
        if (normalize) d /= (double)ch->nvis;
        dchi2[k] = -d;
      }
    return;
  }
  double* xs = (double*)malloc(sizeof(double) * M);
  double* fx = (double*)malloc(sizeof(double) * M);
  double* ys = (double*)malloc(sizeof(double) * N);
  double* fy = (double*)malloc(sizeof(double) * N);
  for (int j = 0; j < M; j++) {
    xs[j] = (j - (int)g->x0_phs) * g->deltax * ORC_RPDEG;
    fx[j] = sqrt(1.0 - xs[j] * xs[j]) - 1.0;
  }
  for (int i = 0; i < N; i++) {
    ys[i] = (i - (int)g->y0_phs) * g->deltay * ORC_RPDEG;
    fy[i] = sqrt(1.0 - ys[i] * ys[i]) - 1.0;
  }
  double* acc = (double*)calloc(n, sizeof(double));
  long nblk = (ch->nvis + ORC_KB - 1) / ORC_KB;
  /* tables for one block: Qr/Qi [kb][M], Pr/Pi [kb][N] */
  double* Qr = (double*)malloc(sizeof(double) * ORC_KB * M);
  double* Qi = (double*)malloc(sizeof(double) * ORC_KB * M);
  double* Pr = (double*)malloc(sizeof(double) * ORC_KB * N);
  double* Pi = (double*)malloc(sizeof(double) * ORC_KB * N);
  for (long blk = 0; blk < nblk; blk++) {
    long k0 = blk * ORC_KB;
    int kb = (int)((ch->nvis - k0) < ORC_KB ? (ch->nvis - k0) : ORC_KB);
#pragma omp parallel for schedule(static)
    for (int kk = 0; kk < kb; kk++) {
      long k = k0 + kk;
      double u = ch->uvw[3 * k], v = ch->uvw[3 * k + 1], w = ch->uvw[3 * k + 2];
      double cr = (double)ch->weight[k] * vr[k].re;
      double ci = -(double)ch->weight[k] * vr[k].im;
      for (int j = 0; j < M; j++) {
        double s, c;
        sincos_turns(u * xs[j] + w * fx[j], &s, &c);
        Qr[(long)kk * M + j] = c;
        Qi[(long)kk * M + j] = s;
      }
      for (int i = 0; i < N; i++) {
        double s, c;
        sincos_turns(v * ys[i] + w * fy[i], &s, &c);
        Pr[(long)kk * N + i] = cr * c - ci * s;
        Pi[(long)kk * N + i] = cr * s + ci * c;
      }
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
      double* row = acc + (long)N * i;
      for (int kk = 0; kk < kb; kk++) {
        double pr = Pr[(long)kk * N + i], pi = -Pi[(long)kk * N + i];
        const double* qr = Qr + (long)kk * M;
        const double* qi = Qi + (long)kk * M;
        for (int j = 0; j < M; j++) row[j] += pr * qr[j] + pi * qi[j];
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++)
    for (int j = 0; j < M; j++) {
      long k = (long)N * i + j;
      if (!(g->noise[k] < g->noise_cut)) {
        dchi2[k] = 0.0;
        continue;
      }
      double atten = orc_attenuation(b, ch->freq, i, j, g->x0_ref, g->y0_ref,
                                     g->deltax, g->deltay);
      double d = acc[k] * (double)g->fg_scale * atten * gcf_at(g, k);
      if (normalize) d /= (double)ch->nvis;
      dchi2[k] = -d;
    }
  free(acc);
  free(Qr);
  free(Qi);
  free(Pr);
  free(Pi);
  free(xs);
  free(fx);
  free(ys);
  free(fy);
}

/* src/functions.cu:4456-4558 (dchi2) with the chain rule kernels
 * :4000-4024 (DChi2_total_I_nu_0) and :3968-3998 (DChi2_total_alpha,
 * including the extra fg_scale factor of the reference). */
void orc_dchi2(const orc_geom* g, const orc_beam* b, const orc_chan* ch,
               int nch, const double* I, const ocplx* vr, double* result,
               int flag_opt, int normalize, int separable) {
  long n = (long)g->M * g->N;
  double* d = (double*)malloc(sizeof(double) * n);
  long off = 0;
  for (int c = 0; c < nch; c++) {
    if (ch[c].nvis <= 0) continue;
    orc_dchi2_channel(g, b, &ch[c], vr + off, normalize, separable, d);
    off += ch[c].nvis;
    double nudiv = (double)ch[c].freq / (double)g->nu_0;
    for (long k = 0; k < n; k++) {
      if (!(g->noise[k] < g->noise_cut)) continue;
      double i0 = I[k], alpha = I[n + k];
      double dI = pow(nudiv, alpha);
      if (flag_opt % 2 == 0) {
        result[k] += d[k] * dI;
      } else if (i0 > g->threshold) {
        result[n + k] += d[k] * i0 * dI * (double)g->fg_scale * log(nudiv);
      }
    }
  }
  free(d);
}

/* ------------------------------------------------------------------------
 * Priors (rows a12-a14).  All evaluate only inside the mask
 * (noise < noise_cut) on image `index`.
 * ---------------------------------------------------------------------- */

/* src/functions.cu:3036-3055, 3081-3095 (calculateS, SVector) */
double orc_entropy_value(const orc_geom* g, const double* I, int index,
                         double G, double eta) {
  long n = (long)g->M * g->N;
  const double* im = I + n * index;
  double s = 0.0;
  for (long k = 0; k < n; k++)
    if (g->noise[k] < g->noise_cut) s += im[k] * log(im[k] / G + (eta + 1.0));
  return s;
}

/* src/functions.cu:3057-3079, 3097-3112 (calculateDS, DS) */
void orc_entropy_grad(const orc_geom* g, const double* I, int index, double G,
                      double eta, double lambda, double* dgi) {
  long n = (long)g->M * g->N;
  const double* im = I + n * index;
  for (long k = 0; k < n; k++) {
    double d = 0.0;
    double c = im[k];
    if (g->noise[k] < g->noise_cut)
      d = log(c / G + (eta + 1.0)) + 1.0 / (1.0 + ((eta + 1.0) * G) / c);
    dgi[k] = d * lambda;
  }
}

/* src/functions.cu:3353-3393 (calculateTSV, TSVVector): forward differences;
 * on the last row/column the raw pixel value is used. */
double orc_tsv_value(const orc_geom* g, const double* I, int index) {
  int N = g->N;
  long n = (long)g->M * N;
  const double* im = I + n * index;
  double s = 0.0;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      long k = (long)N * i + j;
      if (!(g->noise[k] < g->noise_cut)) continue;
      double c = im[k];
      if (i < N - 1 && j < N - 1) {
        double dx = c - im[k + 1], dy = c - im[k + N];
        s += dx * dx + dy * dy;
      } else {
        s += c;
      }
    }
  return s;
}

/* src/functions.cu:3395-3441 (calculateDTSV, DTSV): 8c - 2(u+l+d+r) in the
 * interior, lambda * c on the border. */
void orc_tsv_grad(const orc_geom* g, const double* I, int index, double lambda,
                  double* dgi) {
  int N = g->N;
  long n = (long)g->M * N;
  const double* im = I + n * index;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      long k = (long)N * i + j;
      double d = 0.0;
      if (g->noise[k] < g->noise_cut) {
        double c = im[k];
        if (i > 0 && i < N - 1 && j > 0 && j < N - 1)
          d = 8.0 * c - 2.0 * (im[k - N] + im[k - 1] + im[k + N] + im[k + 1]);
        else
          d = c;
      }
      dgi[k] = d * lambda;
    }
}

/* src/functions.cu:3443-3486 (calculateL, LVector): 0.5 (Dx + Dy)^2 in the
 * interior, raw pixel on the 1-pixel border. */
double orc_laplacian_value(const orc_geom* g, const double* I, int index) {
  int N = g->N;
  long n = (long)g->M * N;
  const double* im = I + n * index;
  double s = 0.0;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      long k = (long)N * i + j;
      if (!(g->noise[k] < g->noise_cut)) continue;
      double c = im[k];
      if (i > 0 && i < N - 1 && j > 0 && j < N - 1) {
        double Dx = im[k - 1] - 2.0 * c + im[k + 1];
        double Dy = im[k - N] - 2.0 * c + im[k + N];
        s += 0.5 * (Dx + Dy) * (Dx + Dy);
      } else {
        s += c;
      }
    }
  return s;
}

/* src/functions.cu:3488-3545 (calculateDL, DL): the reference's stencil,
 * including its sign pattern -8 (d - r - u - l), on 2 <= i,j <= N-3; zero
 * elsewhere. */
void orc_laplacian_grad(const orc_geom* g, const double* I, int index,
                        double lambda, double* dgi) {
  int N = g->N;
  long n = (long)g->M * N;
  const double* im = I + n * index;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      long k = (long)N * i + j;
      double d = 0.0;
      if (g->noise[k] < g->noise_cut && i > 1 && i < N - 2 && j > 1 && j < N - 2) {
        double c = im[k];
        double dn = im[k + N], up = im[k - N], r = im[k + 1], l = im[k - 1];
        double corners = im[k + N - 1] + im[k + N + 1] + im[k - N - 1] + im[k - N + 1];
        double far = im[k + 2 * N] + im[k + 2] + im[k - 2 * N] + im[k - 2];
        d = 20.0 * c - 8.0 * (dn - r - up - l) + 2.0 * corners + far;
      }
      dgi[k] = d * lambda;
    }
}

/* ------------------------------------------------------------------------
 * Objective function (rows a9, a15): include/classes/objectivefunction.cuh
 * :15-57, src/chi2.cu:44-70, src/entropy.cu:42-62,
 * src/totalsquaredvariation.cu:13-33, src/laplacian.cu, and the host
 * launchers src/functions.cu:4722-4964 (iter > 0 and flag_opt gating).
 * ---------------------------------------------------------------------- */
int orc_of_init(orc_objective* of) {
  long tot = 0;
  for (int c = 0; c < of->nch; c++) tot += of->ch[c].nvis;
  long n = (long)of->g->M * of->g->N;
  of->vr = (ocplx*)calloc(tot > 0 ? tot : 1, sizeof(ocplx));
  of->dphi = (double*)calloc(n * of->image_count, sizeof(double));
  of->dgi = (double*)calloc(n * of->image_count, sizeof(double));
  of->n_fcalls = of->n_gcalls = 0;
  return (of->vr && of->dphi && of->dgi) ? 0 : -1;
}

void orc_of_free(orc_objective* of) {
  free(of->vr);
  free(of->dphi);
  free(of->dgi);
  of->vr = NULL;
  of->dphi = of->dgi = NULL;
}

static double fi_value(orc_objective* of, orc_fi* fi, double* p) {
  const orc_geom* g = of->g;
  if (fi->kind == ORC_FI_CHI2) {
    fi->value = orc_chi2(g, of->b, of->ch, of->nch, p, of->flag_opt,
                         of->normalize, of->vr);
    return fi->lambda * fi->value;
  }
  double v = 0.0;
  if (fi->iteration > 0 && fi->lambda != 0.0) {
    if (fi->kind == ORC_FI_ENTROPY)
      v = orc_entropy_value(g, p, fi->imageIndex, fi->prior, fi->eta);
    else if (fi->kind == ORC_FI_TSV)
      v = orc_tsv_value(g, p, fi->imageIndex);
    else if (fi->kind == ORC_FI_LAPLACIAN)
      v = orc_laplacian_value(g, p, fi->imageIndex);
  }
  fi->value = v;
  return fi->lambda * v;
}

double orc_of_calc_function(orc_objective* of, double* p) {
  double value = 0.0;
  of->n_fcalls++;
  for (int f = 0; f < of->nfi; f++) value += fi_value(of, &of->fi[f], p);
  return value;
}

void orc_of_calc_gradient(orc_objective* of, double* p, double* xi, int iter) {
  long n = (long)of->g->M * of->g->N;
  long tot = n * of->image_count;
  of->n_gcalls++;
  memset(of->dphi, 0, sizeof(double) * tot);
  for (int f = 0; f < of->nfi; f++) {
    orc_fi* fi = &of->fi[f];
    fi->iteration = iter;
    memset(of->dgi, 0, sizeof(double) * tot);
    if (fi->kind == ORC_FI_CHI2) {
      /* Chi2::calcGi + addToDphi (src/chi2.cu:51-70): with more than one
       * image dphi is overwritten by the chi2 gradient. */
      orc_dchi2(of->g, of->b, of->ch, of->nch, p, of->vr, of->dgi,
                of->flag_opt, of->normalize, of->separable);
      if (of->image_count > 1)
        memcpy(of->dphi, of->dgi, sizeof(double) * tot);
      else
        for (long k = 0; k < n; k++) of->dphi[k] += of->dgi[k];
      continue;
    }
    if (!(iter > 0 && fi->lambda != 0.0)) continue;
    if (of->flag_opt % 2 != fi->imageIndex) continue;
    if (fi->kind == ORC_FI_ENTROPY)
      orc_entropy_grad(of->g, p, fi->imageIndex, fi->prior, fi->eta, fi->lambda, of->dgi);
    else if (fi->kind == ORC_FI_TSV)
      orc_tsv_grad(of->g, p, fi->imageIndex, fi->lambda, of->dgi);
    else if (fi->kind == ORC_FI_LAPLACIAN)
      orc_laplacian_grad(of->g, p, fi->imageIndex, fi->lambda, of->dgi);
    double* dst = of->dphi + n * fi->imageToAdd;
    for (long k = 0; k < n; k++) dst[k] += of->dgi[k];
  }
  memcpy(xi, of->dphi, sizeof(double) * tot);
}

static double of_value_cb(void* ctx, double* p) {
  return orc_of_calc_function((orc_objective*)ctx, p);
}
static void of_gradient_cb(void* ctx, double* p, double* xi, int iter) {
  orc_of_calc_gradient((orc_objective*)ctx, p, xi, iter);
}
orc_problem orc_of_problem(orc_objective* of) {
  orc_problem pr;
  pr.n = (long)of->g->M * of->g->N;
  pr.image_count = of->image_count;
  pr.nopositivity = of->nopositivity;
  pr.floor_v = -1.0 * (double)of->g->eta * (double)of->g->minpix;
  pr.value = of_value_cb;
  pr.gradient = of_gradient_cb;
  pr.ctx = of;
  return pr;
}

/* ------------------------------------------------------------------------
 * Line search (row a18).  Bracketing by golden-ratio expansion with
 * parabolic extrapolation, then Brent's parabolic/golden-section
 * minimisation (Brent 1973), with the reference's constants and decision
 * order: src/mnbrak.cu:36-98 (GOLD 1.618034, GLIMIT 100, TINY 1e-20) and
 * src/brent.cu:36-121 (ITMAX 500, CGOLD 0.3819660, ZEPS 1e-10).
 * ---------------------------------------------------------------------- */
#define MB_GOLD 1.618034
#define MB_GLIMIT 100.0
#define MB_TINY 1.0e-20
static inline double sign_of(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

void orc_mnbrak(double* ax, double* bx, double* cx, double* fa, double* fb,
                double* fc, orc_f1 f, void* ctx) {
  *fa = f(*ax, ctx);
  *fb = f(*bx, ctx);
  if (*fb > *fa) {
    double t = *ax; *ax = *bx; *bx = t;
    t = *fb; *fb = *fa; *fa = t;
  }
  *cx = *bx + MB_GOLD * (*bx - *ax);
  *fc = f(*cx, ctx);
  while (*fb > *fc) {
    double r = (*bx - *ax) * (*fb - *fc);
    double q = (*bx - *cx) * (*fb - *fa);
    double den = 2.0 * sign_of(fmax(fabs(q - r), MB_TINY), q - r);
    double u = *bx - ((*bx - *cx) * q - (*bx - *ax) * r) / den;
    double ulim = *bx + MB_GLIMIT * (*cx - *bx);
    double fu;
    if ((*bx - u) * (u - *cx) > 0.0) {
      fu = f(u, ctx);
      if (fu < *fc) {
        *ax = *bx; *bx = u;
        *fa = *fb; *fb = fu;
        return;
      } else if (fu > *fb) {
        *cx = u; *fc = fu;
        return;
      }
      u = *cx + MB_GOLD * (*cx - *bx);
      fu = f(u, ctx);
    } else if ((*cx - u) * (u - ulim) > 0.0) {
      fu = f(u, ctx);
      if (fu < *fc) {
        /* shift (b, c, u) <- (c, u, u + GOLD (u - c)) */
        *bx = *cx; *cx = u; u = *cx + MB_GOLD * (*cx - *bx);
        *fb = *fc; *fc = fu; fu = f(u, ctx);
      }
    } else if ((u - ulim) * (ulim - *cx) >= 0.0) {
      u = ulim;
      fu = f(u, ctx);
    } else {
      u = *cx + MB_GOLD * (*cx - *bx);
      fu = f(u, ctx);
    }
    *ax = *bx; *bx = *cx; *cx = u;
    *fa = *fb; *fb = *fc; *fc = fu;
  }
}

#define BR_ITMAX 500
#define BR_CGOLD 0.3819660
#define BR_ZEPS 1.0e-10
double orc_brent(double ax, double bx, double cx, double tol, double* xmin,
                 orc_f1 f, void* ctx) {
  double a = ax < cx ? ax : cx;
  double b = ax > cx ? ax : cx;
  double x = bx, w = bx, v = bx;
  double fx = f(x, ctx), fw = fx, fv = fx;
  double d = 0.0, e = 0.0;
  for (int it = 1; it <= BR_ITMAX; it++) {
    double xm = 0.5 * (a + b);
    double tol1 = tol * fabs(x) + BR_ZEPS;
    double tol2 = 2.0 * tol1;
    if (fabs(x - xm) <= (tol2 - 0.5 * (b - a))) {
      *xmin = x;
      return fx;
    }
    if (fabs(e) > tol1) {
      double r = (x - w) * (fx - fv);
      double q = (x - v) * (fx - fw);
      double p = (x - v) * q - (x - w) * r;
      q = 2.0 * (q - r);
      if (q > 0.0) p = -p;
      q = fabs(q);
      double etemp = e;
      e = d;
      if (fabs(p) >= fabs(0.5 * q * etemp) || p <= q * (a - x) || p >= q * (b - x)) {
        e = (x >= xm ? a - x : b - x);
        d = BR_CGOLD * e;
      } else {
        d = p / q;
        double u = x + d;
        if (u - a < tol2 || b - u < tol2) d = sign_of(tol1, xm - x);
      }
    } else {
      e = (x >= xm ? a - x : b - x);
      d = BR_CGOLD * e;
    }
    double u = (fabs(d) >= tol1 ? x + d : x + sign_of(tol1, d));
    double fu = f(u, ctx);
    if (fu <= fx) {
      if (u >= x) a = x; else b = x;
      v = w; w = x; x = u;
      fv = fw; fw = fx; fx = fu;
    } else {
      if (u < x) a = u; else b = u;
      if (fu <= fw || w == x) {
        v = w; w = u;
        fv = fw; fw = fu;
      } else if (fu <= fv || v == x || v == w) {
        v = u; fv = fu;
      }
    }
  }
  *xmin = x;
  return fx;
}

/* src/f1dim.cu:49-80 and the evaluateXt kernels src/functions.cu:2832-2865:
 * x_t = pcom + x xicom; image 0 is floored at -eta*MINPIX (positivity), the
 * other images are not (unless nopositivity, where none is). */
typedef struct {
  const orc_problem* pr;
  const double* pcom;
  const double* xicom;
  double* xt;
  long calls;
} f1dim_ctx;

static double f1dim(double x, void* vctx) {
  f1dim_ctx* c = (f1dim_ctx*)vctx;
  const orc_problem* pr = c->pr;
  long n = pr->n;
  double floor_v = pr->floor_v;
  c->calls++;
  for (int im = 0; im < pr->image_count; im++) {
    const double* pc = c->pcom + n * im;
    const double* xc = c->xicom + n * im;
    double* xt = c->xt + n * im;
    if (im == 0 && !pr->nopositivity) {
      for (long k = 0; k < n; k++) {
        double t = pc[k] + x * xc[k];
        xt[k] = t > floor_v ? t : floor_v;
      }
    } else {
      for (long k = 0; k < n; k++) xt[k] = pc[k] + x * xc[k];
    }
  }
  return pr->value(pr->ctx, c->xt);
}

/* src/linmin.cu:52-110 and the newP kernels src/functions.cu:2779-2812:
 * bracket from (0, 1), Brent with TOL 1e-7, then xi <- xmin xi and
 * p <- p + xi with the image-0 floor (xi set to 0 where floored). */
static long g_linmin_calls = 0;

/* Optional log of the accepted step lengths (xmin of each linmin), for
 * replaying this line search's decisions in another implementation. */
static double* g_xmin_log = NULL;
static int g_xmin_cap = 0, g_xmin_count = 0;

void orc_set_xmin_log(double* buf, int cap) {
  g_xmin_log = buf;
  g_xmin_cap = buf ? cap : 0;
  g_xmin_count = 0;
}

int orc_xmin_log_count(void) { return g_xmin_count; }

double orc_linmin(const orc_problem* pr, double* p, double* xi) {
  long n = pr->n;
  long tot = n * pr->image_count;
  double* pcom = (double*)malloc(sizeof(double) * tot);
  double* xicom = (double*)malloc(sizeof(double) * tot);
  double* xt = (double*)malloc(sizeof(double) * tot);
  memcpy(pcom, p, sizeof(double) * tot);
  memcpy(xicom, xi, sizeof(double) * tot);
  f1dim_ctx c = {pr, pcom, xicom, xt, 0};
  double ax = 0.0, xx = 1.0, bx, fa, fx, fb, xmin;
  orc_mnbrak(&ax, &xx, &bx, &fa, &fx, &fb, f1dim, &c);
  double fret = orc_brent(ax, xx, bx, 1.0e-7, &xmin, f1dim, &c);
  g_linmin_calls += c.calls;
  if (g_xmin_log && g_xmin_count < g_xmin_cap) g_xmin_log[g_xmin_count] = xmin;
  if (g_xmin_log) g_xmin_count++;
  double floor_v = pr->floor_v;
  for (int im = 0; im < pr->image_count; im++) {
    double* pp = p + n * im;
    double* xx2 = xi + n * im;
    if (im == 0 && !pr->nopositivity) {
      for (long k = 0; k < n; k++) {
        xx2[k] *= xmin;
        if (pp[k] + xx2[k] > floor_v) {
          pp[k] += xx2[k];
        } else {
          pp[k] = floor_v;
          xx2[k] = 0.0;
        }
      }
    } else {
      for (long k = 0; k < n; k++) {
        xx2[k] *= xmin;
        pp[k] += xx2[k];
      }
    }
  }
  free(pcom);
  free(xicom);
  free(xt);
  return fret;
}

/* ------------------------------------------------------------------------
 * Conjugate gradient (row a16): src/frprmn.cu:85-195 with the kernels
 * searchDirection, CGGradCondition, getGGandDGG, newXi
 * (src/functions.cu:3599-3612, 2729-2749, 3655-3687).
 * ---------------------------------------------------------------------- */
#define OPT_EPS 1.0e-10
int orc_frprmn(const orc_problem* pr, double* p, int total_iterations, double ftol,
               double gtol, orc_opt_result* res, double* traj, int cap) {
  long tot = pr->n * pr->image_count;
  long ng = 0, nv0 = g_linmin_calls;
  double* g = (double*)calloc(tot, sizeof(double));
  double* h = (double*)calloc(tot, sizeof(double));
  double* xi = (double*)calloc(tot, sizeof(double));
  memset(res, 0, sizeof(*res));
  long nv = 1;
  double fp = pr->value(pr->ctx, p);
  pr->gradient(pr->ctx, p, xi, 0);
  ng++;
  for (long k = 0; k < tot; k++) {
    g[k] = -xi[k];
    xi[k] = h[k] = g[k];
  }
  int reason = 0, it;
  double fret = fp;
  for (it = 1; it <= total_iterations; it++) {
    fret = orc_linmin(pr, p, xi);
    if (traj && it - 1 < cap) traj[it - 1] = fret;
    if (2.0 * fabs(fret - fp) <= ftol * (fabs(fret) + fabs(fp) + OPT_EPS)) {
      reason = 1;
      break;
    }
    fp = pr->value(pr->ctx, p);
    nv++;
    pr->gradient(pr->ctx, p, xi, it);
    ng++;
    double den = fp > 1.0 ? fp : 1.0;
    double gmax = 0.0;
    for (long k = 0; k < tot; k++) {
      double t = fabs(xi[k]) * fmax(fabs(p[k]), 1.0) / den;
      if (t > gmax) gmax = t;
    }
    if (gmax < gtol) {
      reason = 2;
      break;
    }
    double gg = 0.0, dgg = 0.0;
    for (long k = 0; k < tot; k++) {
      gg += g[k] * g[k];
      dgg += (xi[k] + g[k]) * xi[k];
    }
    if (gg == 0.0) {
      reason = 3;
      break;
    }
    double gam = dgg / gg > 0.0 ? dgg / gg : 0.0;
    for (long k = 0; k < tot; k++) {
      g[k] = -xi[k];
      xi[k] = h[k] = g[k] + gam * h[k];
    }
  }
  res->iterations = it > total_iterations ? total_iterations : it;
  res->exit_reason = reason;
  res->fret = fret;
  res->final_value = pr->value(pr->ctx, p);
  res->n_values = nv + 1 + (g_linmin_calls - nv0);
  res->n_gradients = ng;
  res->n_fcalls = res->n_values;
  res->n_gcalls = ng;
  free(g);
  free(h);
  free(xi);
  return 0;
}

/* ------------------------------------------------------------------------
 * L-BFGS (row a17): src/lbfgs.cu:99-341 with the kernels
 * src/functions.cu:3564-3653.  Restated with the reference's storage
 * addressing: y and s of image `im` for slot `it` are written at block
 * (im*it + im) of K*image_count image-sized blocks; dot products read block
 * (im*k + im); the q/r updates read block (im + k).  For image 0 every
 * write and dot product uses block 0, so image 0 behaves as memory-1 BFGS.
 * y is the new gradient plus the previous search direction.
 * ---------------------------------------------------------------------- */
static double dotn(const double* a, const double* b, long n) {
  double s = 0.0;
  for (long k = 0; k < n; k++) s += a[k] * b[k];
  return s;
}

static void lbfgs_recursion(const orc_problem* pr, double* d_y, double* d_s,
                            double* xi, int par_M, int lbfgs_it) {
  long n = pr->n;
  int ic = pr->image_count;
  long tot = n * ic;
  double* q = (double*)malloc(sizeof(double) * tot);
  double* r = (double*)calloc(tot, sizeof(double));
  double* alpha = (double*)calloc((size_t)ic * par_M, sizeof(double));
  memcpy(q, xi, sizeof(double) * tot);
#define BLK_DOT(buf, im, k) ((buf) + n * ((long)(im) * (k) + (im)))
#define BLK_UPD(buf, im, k) ((buf) + n * ((long)(im) + (k)))
  for (int im = 0; im < ic; im++) {
    for (int k = par_M - 1; k >= 0; k--) {
      double rho_den = dotn(BLK_DOT(d_y, im, k), BLK_DOT(d_s, im, k), n);
      double rho = rho_den != 0.0 ? 1.0 / rho_den : 0.0;
      double a = rho * dotn(BLK_DOT(d_s, im, k), q + n * im, n);
      alpha[im * par_M + k] = a;
      const double* yk = BLK_UPD(d_y, im, k);
      double* qi = q + n * im;
      for (long t = 0; t < n; t++) qi[t] += -a * yk[t];
    }
  }
  double sy_yy = 0.0;
  for (int im = 0; im < ic; im++) {
    double sy = dotn(BLK_DOT(d_y, im, lbfgs_it), BLK_DOT(d_s, im, lbfgs_it), n);
    double yy = dotn(BLK_DOT(d_y, im, lbfgs_it), BLK_DOT(d_y, im, lbfgs_it), n);
    if (yy != 0.0) sy_yy += sy / yy;
  }
  for (int im = 0; im < ic; im++) {
    double* ri = r + n * im;
    const double* qi = q + n * im;
    for (long t = 0; t < n; t++) ri[t] = qi[t] * sy_yy;
    for (int k = 0; k < par_M; k++) {
      double rho_den = dotn(BLK_DOT(d_y, im, k), BLK_DOT(d_s, im, k), n);
      double rho = rho_den != 0.0 ? 1.0 / rho_den : 0.0;
      double beta = rho * dotn(BLK_DOT(d_y, im, k), ri, n);
      double coef = alpha[im * par_M + k] - beta;
      const double* sk = BLK_UPD(d_s, im, k);
      for (long t = 0; t < n; t++) ri[t] += coef * sk[t];
    }
  }
  for (long t = 0; t < tot; t++) xi[t] = -r[t];
#undef BLK_DOT
#undef BLK_UPD
  free(q);
  free(r);
  free(alpha);
}

int orc_lbfgs(const orc_problem* pr, double* p, int K, int total_iterations,
              double ftol, double gtol, orc_opt_result* res, double* traj, int cap) {
  long n = pr->n;
  int ic = pr->image_count;
  long ng = 0, nv0 = g_linmin_calls, nv = 0;
  long tot = n * ic;
  double* d_y = (double*)calloc((size_t)tot * K, sizeof(double));
  double* d_s = (double*)calloc((size_t)tot * K, sizeof(double));
  double* p_old = (double*)calloc(tot, sizeof(double));
  double* xi = (double*)calloc(tot, sizeof(double));
  double* xi_old = (double*)calloc(tot, sizeof(double));
  memset(res, 0, sizeof(*res));
  double fp = pr->value(pr->ctx, p);
  nv++;
  pr->gradient(pr->ctx, p, xi, 0);
  ng++;
  for (long k = 0; k < tot; k++) xi[k] = -xi[k];
  int reason = 0, it;
  double fret = fp;
  for (it = 1; it <= total_iterations; it++) {
    memcpy(p_old, p, sizeof(double) * tot);
    memcpy(xi_old, xi, sizeof(double) * tot);
    fret = orc_linmin(pr, p, xi);
    if (traj && it - 1 < cap) traj[it - 1] = fret;
    double den = fmax(fmax(fabs(fret), fabs(fp)), 1.0);
    if ((fp - fret) / den <= ftol) {
      reason = 1;
      break;
    }
    double mx = 0.0;
    for (long k = 0; k < tot; k++)
      if (fabs(xi[k]) > mx) mx = fabs(xi[k]);
    if (mx <= gtol) {
      reason = 2;
      break;
    }
    fp = pr->value(pr->ctx, p);
    nv++;
    pr->gradient(pr->ctx, p, xi, it);
    ng++;
    int slot = (it - 1) % K;
    for (int im = 0; im < ic; im++) {
      double* y = d_y + n * ((long)im * slot + im);
      double* s = d_s + n * ((long)im * slot + im);
      for (long k = 0; k < n; k++) {
        y[k] = xi[n * im + k] - (-1.0 * xi_old[n * im + k]);
        s[k] = p[n * im + k] - p_old[n * im + k];
      }
    }
    lbfgs_recursion(pr, d_y, d_s, xi, K < it ? K : it, slot);
  }
  res->iterations = it > total_iterations ? total_iterations : it;
  res->exit_reason = reason;
  res->fret = fret;
  res->final_value = pr->value(pr->ctx, p);
  res->n_values = nv + 1 + (g_linmin_calls - nv0);
  res->n_gradients = ng;
  res->n_fcalls = res->n_values;
  res->n_gcalls = ng;
  free(d_y);
  free(d_s);
  free(p_old);
  free(xi);
  free(xi_old);
  return 0;
}

/* ------------------------------------------------------------------------
 * CPU gridding comparator: src/functions.cu:1339-1653 (do_gridding) with the
 * default 7x7 PillBox kernel (include/classes/ckernel.cuh:101-114,
 * src/pillBox2D.cu:149-170: every tap inside the support is 1).  Each
 * visibility is gridded twice (as is and Hermitian-conjugated) to the nearest
 * cell of a centred grid with the same atomic / critical structure; then the
 * cells are normalised by the weight sum.  uvw in metres.
 * ---------------------------------------------------------------------- */
void orc_do_gridding_pillbox(int M, int N, double deltau, double deltav,
                             const orc_chan* ch, int nch, int support,
                             int nthreads, double* grid_re, double* grid_im,
                             double* grid_w) {
  long n = (long)M * N;
  float* gw = (float*)calloc(n, sizeof(float));
  float* gw2 = (float*)calloc(n, sizeof(float));
  float* gre = (float*)calloc(n, sizeof(float));
  float* gim = (float*)calloc(n, sizeof(float));
  double cj = floor(N / 2.0), ck = floor(M / 2.0);
  for (int c = 0; c < nch; c++) {
    float freq = ch[c].freq;
    float lambda = (float)(2.99792458E8f / freq);
    long nv = ch[c].nvis;
    memset(gw, 0, n * sizeof(float));
    memset(gw2, 0, n * sizeof(float));
    memset(gre, 0, n * sizeof(float));
    memset(gim, 0, n * sizeof(float));
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long z = 0; z < 2 * nv; z++) {
      long vi = z < nv ? z : z - nv;
      double u = ch[c].uvw[3 * vi], v = ch[c].uvw[3 * vi + 1];
      float w = ch[c].weight[vi];
      float vre = ch[c].vo[2 * vi], vim = ch[c].vo[2 * vi + 1];
      if (z >= nv) {
        u = -u;
        v = -v;
        vim = -vim;
      }
      u = u / lambda;
      v = v / lambda;
      int j = (int)(u / deltau + cj + 0.5);
      int k = (int)(v / deltav + ck + 0.5);
      for (int m = -support; m <= support; m++)
        for (int q = -support; q <= support; q++) {
          int sj = j + q, sk = k + m;
          if (sk >= 0 && sk < M && sj >= 0 && sj < N) {
            float kv = 1.0f;
            long gi = (long)N * sk + sj;
#pragma omp atomic
            gw[gi] += w * kv;
#pragma omp atomic
            gw2[gi] += w * kv * kv;
#pragma omp critical
            {
              gre[gi] += w * vre * kv;
              gim[gi] += w * vim * kv;
            }
          }
        }
    }
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long gi = 0; gi < n; gi++) {
      float ws = gw[gi], a = gw2[gi];
      if (a != 0.0f && ws != 0.0f) {
        grid_re[gi] = gre[gi] / ws;
        grid_im[gi] = gim[gi] / ws;
        grid_w[gi] = ws * ws / a;
      } else {
        grid_re[gi] = grid_im[gi] = grid_w[gi] = 0.0;
      }
    }
  }
  free(gw);
  free(gw2);
  free(gre);
  free(gim);
}
