/* Set-up of the hot path's inputs (host, FP64), restating what
 * MFS::configure / MFS::setDevice compute before the optimiser runs.
 * Plain C ABI (libgvm). */
#ifndef GVM_SETUP_H
#define GVM_SETUP_H

#ifdef __cplusplus
extern "C" {
#endif

/* hermitianSymmetry (src/functions.cu:2256-2273) with metres_to_lambda
 * (src/MSFITSIO.cu:36-45, float wavelength): u > 0 rows become (-u, -v) with
 * conjugated Vo; uvw (3*n doubles) are converted to wavelengths in place. */
void gvm_setup_hermitian_lambda(long n, double* uvw, float* vo, float freq);

/* calculateNoiseAndBeam (src/functions.cu:1700-1814, calc_sBeam /
 * calc_beamSize :1655-1700) and the noise per pixel of src/mfs.cu:640-643,
 * for nch slices with uvw in metres.  Outputs: beam axes and angle in
 * degrees, sum of weights, visibility noise 0.5 sqrt(1/sum w), noise_jypix. */
int gvm_setup_beam_noise(int nch, const float* freq, const long* nvis,
                         const double* const* uvw_m, const float* const* weight,
                         double deltax_deg, double* bmaj_deg, double* bmin_deg,
                         double* bpa_deg, double* sum_weights, double* vis_noise,
                         double* noise_jypix);

/* pb_cutoff = pb_factor * max_wavelength / D (src/MSFITSIO.cu:550-551). */
double gvm_setup_pb_cutoff(double pb_factor, double max_wavelength, double diameter);

/* total_attenuation -> weight_image -> noise_image (src/mfs.cu:849-916,
 * src/functions.cu:2382-2422) for one field at nu_0:
 * noise = noise_jypix sqrt(max(PB^2) / PB^2); fg_scale = min(noise);
 * noise_cut = noise_cut_factor * min(noise) (default factor 10, -N). */
int gvm_setup_noise_image(int beam_kind, double diameter, double pb_factor,
                          double pb_cutoff, double nu_0, int M, int N,
                          double deltax_deg, double deltay_deg, float x0_ref,
                          float y0_ref, double noise_jypix, double noise_cut_factor,
                          float* noise_out, float* fg_scale, float* noise_cut);

#ifdef __cplusplus
}
#endif
#endif
