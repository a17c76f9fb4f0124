// Set-up of the hot path's inputs (host, FP64): Hermitian flip and
// wavelength conversion, clean-beam and noise estimates, the noise image and
// the fg_scale / noise_cut it implies.  Restates the reference's one-time
// set-up (src/mfs.cu:530-943, src/functions.cu:1655-1814, 2256-2273,
// 2382-2422); not part of the timed evaluation.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "gvm/setup.h"
#include "gvm_hip.h"

namespace {
constexpr double kPi = 3.14159265358979323846;
constexpr double kRad = kPi / 180.0;
// freq_to_wavelength (src/MSFITSIO.cu:36-39) works in float.
inline float wavelength_f(float freq) { return 2.99792458E8f / freq; }
}  // namespace

extern "C" void gvm_setup_hermitian_lambda(long n, double* uvw, float* vo, float freq) {
  const float lambda = wavelength_f(freq);
  for (long i = 0; i < n; i++) {
    if (uvw[3 * i] > 0.0) {
      uvw[3 * i] *= -1.0;
      uvw[3 * i + 1] *= -1.0;
      vo[2 * i + 1] *= -1.0f;
    }
    uvw[3 * i] /= lambda;
    uvw[3 * i + 1] /= lambda;
    uvw[3 * i + 2] /= lambda;
  }
}

extern "C" int gvm_setup_beam_noise(int nch, const float* freq, const long* nvis,
                                    const double* const* uvw_m, const float* const* weight,
                                    double deltax_deg, double* bmaj_deg, double* bmin_deg,
                                    double* bpa_deg, double* sum_weights, double* vis_noise,
                                    double* noise_jypix) {
  double suu = 0, svv = 0, suv = 0, sw = 0;
  for (int c = 0; c < nch; c++) {
    const float lambda = wavelength_f(freq[c]);
    for (long i = 0; i < nvis[c]; i++) {
      const double u = uvw_m[c][3 * i] / lambda, v = uvw_m[c][3 * i + 1] / lambda;
      const double w = weight[c][i];
      suu += u * u * w;
      svv += v * v * w;
      suv += u * v * w;
      sw += w;
    }
  }
  if (!(sw > 0.0)) return GVM_ERR_ARG;  // "The sum of the visibility weights cannot be zero"
  return gvm_setup_beam_from_moments(suu, svv, suv, sw, deltax_deg, bmaj_deg, bmin_deg, bpa_deg,
                                     sum_weights, vis_noise, noise_jypix);
}

extern "C" double gvm_setup_pb_cutoff(double pb_factor, double max_wavelength, double diameter) {
  return pb_factor * (max_wavelength / diameter);
}

extern "C" int gvm_setup_noise_image(int beam_kind, double diameter, double pb_factor,
                                     double pb_cutoff, double nu_0, int M, int N,
                                     double deltax_deg, double deltay_deg, float x0_ref,
                                     float y0_ref, double noise_jypix, double noise_cut_factor,
                                     float* noise_out, float* fg_scale, float* noise_cut) {
  const long n = (long)M * N;
  std::vector<float> pb(n);
  int st = gvm_primary_beam(beam_kind, diameter, pb_factor, pb_cutoff, nu_0, M, N, x0_ref, y0_ref,
                            deltax_deg, deltay_deg, pb.data());
  if (st) return st;
  double wmax = 0.0;
  for (long k = 0; k < n; k++) wmax = std::max(wmax, (double)pb[k] * pb[k]);
  if (!(wmax > 0.0)) return GVM_ERR_ARG;
  double nmin = std::numeric_limits<double>::infinity();
  for (long k = 0; k < n; k++) {
    const double w = (double)pb[k] * pb[k];
    const double nw = (w / wmax) / (noise_jypix * noise_jypix);
    const double v = std::sqrt(1.0 / nw);  // inf where the beam is 0
    noise_out[k] = (float)v;
    nmin = std::min(nmin, v);
  }
  *fg_scale = (float)nmin;
  *noise_cut = (float)(noise_cut_factor * nmin);
  return GVM_OK;
}
