// Synthetic interferometric observations for the imaging workloads.
//
// Produces what the reference reads from a Measurement Set for one field and
// one channel (src/MSFITSIO.cu:398-754): uvw in metres, Vo (complex64),
// weights 1/sigma^2, plus the true sky image.  Everything is computed in FP64
// and seeded, so a workload is reproducible from its config alone.
//
//  * antennas: uniform in a disk, or Gaussian-radial clipped to the disk;
//  * uvw: local ENU baselines -> equatorial XYZ -> (u, v, w) over hour angle
//    (Thompson, Moran & Swenson, eq. 4.1);
//  * sky: Gaussian components and point sources; the visibility of a Gaussian
//    times the Gaussian primary beam is again a Gaussian, so V is analytic;
//  * noise: complex Gaussian with the configured sigma per real/imag part.
//
// The sign convention matches the forward model of the reference
// (V(u,v) = sum I(x,y) e^{+2 pi i (u x + v y)}, src/functions.cu:2153-2178 and
// :2483-2518), with x = (j - x0) * CDELT1 in radians.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "gvm/simdata.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kRad = kPi / 180.0;

// splitmix64 seeding + xoshiro256** (Blackman & Vigna 2018).
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull;
    for (auto& v : s) {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      v = x ^ (x >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  // Box-Muller
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
  }
};

}  // namespace

extern "C" long gvm_sim_count_visibilities(const gvm_sim_config* c) {
  if (!c || c->n_antennas < 2 || c->dump_seconds <= 0) return -1;
  long nbl = (long)c->n_antennas * (c->n_antennas - 1) / 2;
  long ndump = (long)std::llround((c->ha_end_hours - c->ha_start_hours) * 3600.0 /
                                  c->dump_seconds);
  return nbl * ndump;
}

extern "C" int gvm_sim_generate(const gvm_sim_config* c, double* uvw_m,
                                float* vo, float* weight, float* sky,
                                gvm_sim_component* comps_out, int* ncomps_out) {
  long nvis = gvm_sim_count_visibilities(c);
  if (nvis <= 0 || !uvw_m || !vo || !weight) return 1;
  const int npix = c->npix;
  const int nant = c->n_antennas;
  const double radius = 0.5 * c->array_diameter_m;

  // ---- antennas (ENU, metres) ----
  Rng rl(c->seed_layout);
  std::vector<double> east(nant), north(nant);
  for (int a = 0; a < nant; a++) {
    double r;
    if (c->layout == GVM_SIM_LAYOUT_UNIFORM) {
      r = radius * std::sqrt(rl.uniform());
    } else {
      do {
        r = std::fabs(rl.normal() * c->layout_sigma_m);
      } while (r > radius);
    }
    double th = 2.0 * kPi * rl.uniform();
    east[a] = r * std::cos(th);
    north[a] = r * std::sin(th);
  }

  // ---- uvw over hour angle, baseline-major ----
  const double lat = c->latitude_deg * kRad;
  const double dec = c->declination_deg * kRad;
  const long ndump = nvis / ((long)nant * (nant - 1) / 2);
  long row = 0;
  for (int a = 0; a < nant; a++) {
    for (int b = a + 1; b < nant; b++) {
      double de = east[b] - east[a];
      double dn = north[b] - north[a];
      double lx = -std::sin(lat) * dn;
      double ly = de;
      double lz = std::cos(lat) * dn;
      for (long t = 0; t < ndump; t++) {
        double ha_h = c->ha_start_hours + (t + 0.5) * c->dump_seconds / 3600.0;
        double H = ha_h * 15.0 * kRad;
        double sh = std::sin(H), chh = std::cos(H);
        uvw_m[3 * row + 0] = sh * lx + chh * ly;
        uvw_m[3 * row + 1] = -std::sin(dec) * chh * lx + std::sin(dec) * sh * ly + std::cos(dec) * lz;
        uvw_m[3 * row + 2] = std::cos(dec) * chh * lx - std::cos(dec) * sh * ly + std::sin(dec) * lz;
        row++;
      }
    }
  }

  // ---- sky ----
  Rng rs(c->seed_sky);
  const double cell = c->cell_arcsec / 3600.0 * kRad;  // radians
  const double dx = -cell, dy = cell;                   // CDELT1 < 0, CDELT2 > 0
  const double x0 = npix / 2, y0 = npix / 2;
  const double lambda = 2.99792458e8 / c->freq_hz;
  const double fwhm_pb = c->pb_factor * lambda / c->dish_diameter_m;  // radians
  const double sig_pb = fwhm_pb / (2.0 * std::sqrt(2.0 * std::log(2.0)));
  const double pos_r = c->position_radius_frac * npix;
  int ncomp = c->n_gaussians + c->n_points;
  std::vector<gvm_sim_component> comps(ncomp);
  for (int k = 0; k < ncomp; k++) {
    gvm_sim_component& s = comps[k];
    double r = pos_r * std::sqrt(rs.uniform());
    double th = 2.0 * kPi * rs.uniform();
    s.x_pix = x0 + std::round(r * std::cos(th));
    s.y_pix = y0 + std::round(r * std::sin(th));
    double fmin = k < c->n_gaussians ? c->flux_min : c->point_flux_min;
    double fmax = k < c->n_gaussians ? c->flux_max : c->point_flux_max;
    if (c->flux_law == GVM_SIM_FLUX_POWERLAW && k < c->n_gaussians) {
      double a = 1.0 + c->flux_index;  // dN/dS ~ S^index
      double u = rs.uniform();
      s.flux = std::pow(std::pow(fmin, a) + u * (std::pow(fmax, a) - std::pow(fmin, a)), 1.0 / a);
    } else {
      s.flux = std::exp(rs.uniform(std::log(fmin), std::log(fmax)));
    }
    s.sigma_pix = k < c->n_gaussians ? rs.uniform(c->sigma_min_pix, c->sigma_max_pix) : 0.0;
    s.alpha = rs.uniform(c->alpha_min, c->alpha_max);
  }
  // flux at this frequency (the sky image and comps_out keep the nu0 fluxes)
  std::vector<double> flux_nu(ncomp);
  for (int k = 0; k < ncomp; k++)
    flux_nu[k] = c->nu0_hz > 0 ? comps[k].flux * std::pow(c->freq_hz / c->nu0_hz, comps[k].alpha)
                               : comps[k].flux;
  if (comps_out) std::memcpy(comps_out, comps.data(), sizeof(gvm_sim_component) * ncomp);
  if (ncomps_out) *ncomps_out = ncomp;

  if (sky) {
    std::memset(sky, 0, sizeof(float) * (size_t)npix * npix);
    for (const auto& s : comps) {
      if (s.sigma_pix <= 0) {
        sky[(long)npix * (long)s.y_pix + (long)s.x_pix] += (float)s.flux;
        continue;
      }
      double peak = s.flux / (2.0 * kPi * s.sigma_pix * s.sigma_pix);
      int rad = (int)std::ceil(7.0 * s.sigma_pix);
      for (int i = (int)s.y_pix - rad; i <= (int)s.y_pix + rad; i++) {
        if (i < 0 || i >= npix) continue;
        for (int j = (int)s.x_pix - rad; j <= (int)s.x_pix + rad; j++) {
          if (j < 0 || j >= npix) continue;
          double r2 = (i - s.y_pix) * (i - s.y_pix) + (j - s.x_pix) * (j - s.x_pix);
          sky[(long)npix * i + j] += (float)(peak * std::exp(-r2 / (2.0 * s.sigma_pix * s.sigma_pix)));
        }
      }
    }
  }

  // ---- visibilities at the float wavelength the set-up divides by ----
  const float lam_f = (float)(2.99792458E8f / (float)c->freq_hz);
  struct Term {
    double amp, s2, cx, cy;
  };
  std::vector<Term> terms(ncomp);
  for (int k = 0; k < ncomp; k++) {
    const auto& s = comps[k];
    double cxr = (s.x_pix - x0) * dx, cyr = (s.y_pix - y0) * dy;
    double c2 = cxr * cxr + cyr * cyr;
    if (s.sigma_pix <= 0) {
      terms[k] = {flux_nu[k] * std::exp(-c2 / (2.0 * sig_pb * sig_pb)), 0.0, cxr, cyr};
    } else {
      double ss = s.sigma_pix * cell;
      double sum2 = ss * ss + sig_pb * sig_pb;
      double sp2 = ss * ss * sig_pb * sig_pb / sum2;
      double shrink = sig_pb * sig_pb / sum2;
      double amp = flux_nu[k] / (2.0 * kPi * ss * ss) * std::exp(-c2 / (2.0 * sum2)) * 2.0 * kPi * sp2;
      terms[k] = {amp, sp2, cxr * shrink, cyr * shrink};
    }
  }
  Rng rn(c->seed_noise);
  std::vector<double> noise(2 * nvis);
  for (long i = 0; i < 2 * nvis; i++) noise[i] = rn.normal() * c->noise_sigma;
  const float wgt = (float)(1.0 / (c->noise_sigma * c->noise_sigma));
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nvis; i++) {
    double u = uvw_m[3 * i] / lam_f, v = uvw_m[3 * i + 1] / lam_f;
    double uu = u * u + v * v;
    double re = 0.0, im = 0.0;
    for (const auto& t : terms) {
      double g = t.s2 > 0 ? std::exp(-2.0 * kPi * kPi * t.s2 * uu) : 1.0;
      if (g < 1e-30) continue;
      double ph = u * t.cx + v * t.cy;
      ph -= std::nearbyint(ph);
      re += t.amp * g * std::cos(2.0 * kPi * ph);
      im += t.amp * g * std::sin(2.0 * kPi * ph);
    }
    vo[2 * i] = (float)(re + noise[2 * i]);
    vo[2 * i + 1] = (float)(im + noise[2 * i + 1]);
    weight[i] = wgt;
  }
  return 0;
}

extern "C" int gvm_sim_generate_flagged(const gvm_sim_config* c, double* uvw_m, float* vo,
                                        float* weight, float* sky, gvm_sim_component* comps_out,
                                        int* ncomps_out, long* nkept) {
  if (!nkept || c->flag_fraction < 0.0 || c->flag_fraction >= 1.0) return 1;
  const int rc = gvm_sim_generate(c, uvw_m, vo, weight, sky, comps_out, ncomps_out);
  if (rc != 0) return rc;
  const long nvis = gvm_sim_count_visibilities(c);
  Rng rf(c->seed_flag);
  long k = 0;
  for (long i = 0; i < nvis; i++) {
    if (c->flag_fraction > 0.0 && rf.uniform() < c->flag_fraction) continue;
    if (k != i) {
      for (int d = 0; d < 3; d++) uvw_m[3 * k + d] = uvw_m[3 * i + d];
      vo[2 * k] = vo[2 * i];
      vo[2 * k + 1] = vo[2 * i + 1];
      weight[k] = weight[i];
    }
    k++;
  }
  *nkept = k;
  return 0;
}

extern "C" void gvm_sim_default_config(int which, gvm_sim_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->cell_arcsec = 0.05;
  c->declination_deg = -30.0;
  c->latitude_deg = -23.03;
  c->dish_diameter_m = 12.0;
  c->pb_factor = 1.13;
  c->noise_sigma = 1e-3;
  c->position_radius_frac = 0.22;
  c->point_flux_min = 0.01;
  c->point_flux_max = 1.0;
  c->alpha_min = 0.0;
  c->alpha_max = 0.0;
  if (which == 2) {
    // 2048^2 precision workload: 64 antennas uniform in a 5 km disk, 6 h at
    // 10 s (2016 x 2160 = 4,354,560 rows), 60 GHz so the longest baselines
    // reach about 1e6 wavelengths; 200 Gaussian components (power-law fluxes).
    c->npix = 2048;
    c->n_antennas = 64;
    c->layout = GVM_SIM_LAYOUT_UNIFORM;
    c->array_diameter_m = 5000.0;
    c->ha_start_hours = -3.0;
    c->ha_end_hours = 3.0;
    c->dump_seconds = 10.0;
    c->freq_hz = 60e9;
    c->n_gaussians = 200;
    c->n_points = 0;
    c->sigma_min_pix = 1.0;
    c->sigma_max_pix = 8.0;
    c->flux_law = GVM_SIM_FLUX_POWERLAW;
    c->flux_index = -1.5;
    c->flux_min = 1e-3;
    c->flux_max = 1.0;
    c->seed_layout = 20;
    c->seed_sky = 21;
    c->seed_noise = 22;
  } else if (which == 3) {
    // One channel of the 512^2 multi-frequency workload: 35 antennas
    // Gaussian-radial in a 2 km disk, 504 dumps of 10 s (595 x 504 = 299,880
    // rows); 30 Gaussian components (power-law fluxes) + 5 points with
    // spectral indices U(-1, 3) about nu0 = 225 GHz.  The caller sets
    // freq_hz (210-240 GHz), flag_fraction and the per-channel seeds.
    c->npix = 512;
    c->n_antennas = 35;
    c->layout = GVM_SIM_LAYOUT_GAUSSIAN;
    c->array_diameter_m = 2000.0;
    c->layout_sigma_m = 500.0;
    c->ha_start_hours = -0.7;
    c->ha_end_hours = 0.7;
    c->dump_seconds = 10.0;
    c->freq_hz = 225e9;
    c->nu0_hz = 225e9;
    c->n_gaussians = 30;
    c->n_points = 5;
    c->sigma_min_pix = 1.0;
    c->sigma_max_pix = 6.0;
    c->flux_law = GVM_SIM_FLUX_POWERLAW;
    c->flux_index = -1.5;
    c->flux_min = 1e-3;
    c->flux_max = 1.0;
    c->alpha_min = -1.0;
    c->alpha_max = 3.0;
    c->seed_layout = 30;
    c->seed_sky = 31;
    c->seed_noise = 32;
    c->seed_flag = 33;
  } else if (which == 4) {
    // 512^2 warm-start workload: the 256^2 workload's skies scaled up (5
    // Gaussian blobs of 4-20 px, log-uniform fluxes 0.01-1 Jy, 3 points) seen
    // by 50 antennas uniform in a 1 km disk, 2 h 20 min at 10 s (1225 x 840 =
    // 1,029,000 rows), 230 GHz.
    c->npix = 512;
    c->n_antennas = 50;
    c->layout = GVM_SIM_LAYOUT_UNIFORM;
    c->array_diameter_m = 1000.0;
    c->ha_start_hours = -7.0 / 6.0;
    c->ha_end_hours = 7.0 / 6.0;
    c->dump_seconds = 10.0;
    c->freq_hz = 230e9;
    c->n_gaussians = 5;
    c->n_points = 3;
    c->sigma_min_pix = 4.0;
    c->sigma_max_pix = 20.0;
    c->flux_law = GVM_SIM_FLUX_LOGUNIFORM;
    c->flux_min = 0.01;
    c->flux_max = 1.0;
    c->seed_layout = 40;
    c->seed_sky = 41;
    c->seed_noise = 42;
  } else if (which == 1) {
    // 1024^2; 50 antennas Gaussian-radial in a 2 km disk; 4 h at 10 s;
    // 200 Gaussian components with power-law fluxes.
    c->npix = 1024;
    c->n_antennas = 50;
    c->layout = GVM_SIM_LAYOUT_GAUSSIAN;
    c->array_diameter_m = 2000.0;
    c->layout_sigma_m = 500.0;
    c->ha_start_hours = -2.0;
    c->ha_end_hours = 2.0;
    c->dump_seconds = 10.0;
    c->freq_hz = 230e9;
    c->n_gaussians = 200;
    c->n_points = 0;
    c->sigma_min_pix = 1.0;
    c->sigma_max_pix = 8.0;
    c->flux_law = GVM_SIM_FLUX_POWERLAW;
    c->flux_index = -1.5;
    c->flux_min = 1e-3;
    c->flux_max = 1.0;
    c->seed_layout = 10;
    c->seed_sky = 11;
    c->seed_noise = 12;
  } else {
    // 256^2; 25 antennas uniform in a 500 m disk; 1 h at 30 s; 5 Gaussian
    // blobs + 3 point sources.
    c->npix = 256;
    c->n_antennas = 25;
    c->layout = GVM_SIM_LAYOUT_UNIFORM;
    c->array_diameter_m = 500.0;
    c->ha_start_hours = -0.5;
    c->ha_end_hours = 0.5;
    c->dump_seconds = 30.0;
    c->freq_hz = 230e9;
    c->n_gaussians = 5;
    c->n_points = 3;
    c->sigma_min_pix = 2.0;
    c->sigma_max_pix = 10.0;
    c->flux_law = GVM_SIM_FLUX_LOGUNIFORM;
    c->flux_min = 0.01;
    c->flux_max = 1.0;
    c->seed_layout = 0;
    c->seed_sky = 1;
    c->seed_noise = 2;
  }
}
