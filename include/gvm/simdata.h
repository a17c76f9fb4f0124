/* Synthetic observation generator (host C, FP64, seeded).  It stands in for
 * the Measurement Set reader of the reference (src/MSFITSIO.cu:398-754),
 * which is not part of the accelerated path: it yields, for one field and one
 * channel, uvw in metres, Vo, weights and the true sky. */
#ifndef GVM_SIMDATA_H
#define GVM_SIMDATA_H

#ifdef __cplusplus
extern "C" {
#endif

enum { GVM_SIM_LAYOUT_UNIFORM = 0, GVM_SIM_LAYOUT_GAUSSIAN = 1 };
enum { GVM_SIM_FLUX_LOGUNIFORM = 0, GVM_SIM_FLUX_POWERLAW = 1 };

typedef struct {
  int npix;
  double cell_arcsec;
  int n_antennas;
  int layout;
  double array_diameter_m;
  double layout_sigma_m;
  double latitude_deg;
  double declination_deg;
  double ha_start_hours, ha_end_hours;
  double dump_seconds;
  double freq_hz;
  double dish_diameter_m;
  double pb_factor;
  int n_gaussians;
  int n_points;
  double sigma_min_pix, sigma_max_pix;
  int flux_law;
  double flux_index;
  double flux_min, flux_max;
  double point_flux_min, point_flux_max;
  double position_radius_frac;
  double alpha_min, alpha_max;
  double noise_sigma;
  unsigned long long seed_layout, seed_sky, seed_noise;
  /* multi-frequency: component fluxes scale as (freq_hz / nu0_hz)^alpha
   * (nu0_hz <= 0: no scaling); each row is flagged (dropped) with
   * probability flag_fraction, drawn from seed_flag. */
  double nu0_hz;
  double flag_fraction;
  unsigned long long seed_flag;
} gvm_sim_config;

typedef struct {
  double x_pix, y_pix;
  double flux;
  double sigma_pix; /* 0 for a point source */
  double alpha;
} gvm_sim_component;

/* which: 0 = 256^2 / 36,000 vis workload, 1 = 1024^2 / 1,764,000 vis,
 * 2 = 2048^2 / 4,354,560 vis at 60 GHz (baselines to about 1e6 wavelengths),
 * 3 = one channel of the 512^2 multi-frequency workload (35 antennas,
 * 595 baselines x 504 dumps = 299,880 rows per channel, nu0 = 225 GHz,
 * spectral indices U(-1, 3); the caller sets freq_hz, flag_fraction and the
 * per-channel seeds), 4 = the 512^2 / 1,029,000 vis warm-start workload (the
 * 256^2 skies scaled up, 50 antennas in a 1 km disk, 1225 baselines x 840
 * dumps). */
void gvm_sim_default_config(int which, gvm_sim_config* cfg);
long gvm_sim_count_visibilities(const gvm_sim_config* cfg);
/* uvw_m: 3*nvis doubles (metres); vo: 2*nvis floats; weight: nvis floats;
 * sky (optional): npix*npix floats (Jy/pixel, unattenuated);
 * comps_out (optional): n_gaussians + n_points entries.  Returns 0 on success. */
int gvm_sim_generate(const gvm_sim_config* cfg, double* uvw_m, float* vo,
                     float* weight, float* sky, gvm_sim_component* comps_out,
                     int* ncomps_out);
/* gvm_sim_generate with flagging: the rows that survive flag_fraction are
 * packed to the front of uvw_m / vo / weight (sized for
 * gvm_sim_count_visibilities rows) and their number is written to *nkept. */
int gvm_sim_generate_flagged(const gvm_sim_config* cfg, double* uvw_m, float* vo,
                             float* weight, float* sky, gvm_sim_component* comps_out,
                             int* ncomps_out, long* nkept);

#ifdef __cplusplus
}
#endif
#endif
