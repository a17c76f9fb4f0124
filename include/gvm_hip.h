/*
 * gvm_hip.h -- C ABI of the MI355X (gfx950) implementation of gpuvmem's
 * maximum-entropy objective and gradient.
 *
 * Plain C: opaque context, plain pointers and sizes, int status codes.  No
 * HIP, torch or C++ types cross this boundary.  The C++ Fi / ObjectiveFunction
 * / Optimizer classes (include/gvm/) are thin forwards to these calls.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the gpuvmem repository root).
 *
 * Conventions
 *  - Images are device pointers to row-major float[image_count][N][M]
 *    (index M*N*k + N*i + j), M == N, a power of two in [64, 4096].
 *    image 0 is I_nu_0, image 1 is the spectral index alpha.
 *  - uvw are in wavelengths after the Hermitian flip (u <= 0), as the
 *    reference keeps them on the device after set-up.
 *  - Calls on one context are stream-ordered on the context's own HIP stream
 *    and are not re-entrant.  Calls that return a host scalar synchronise that
 *    stream; the others only enqueue work.
 *  - Device pointers passed in must stay valid until the call's work has
 *    completed (gvm_sync).
 */
#ifndef GVM_HIP_H
#define GVM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gvm_ctx gvm_ctx;
typedef int gvm_status;

/* Status codes.  The reference aborts on errors (checkCudaErrors, exit(-1));
 * this ABI reports them instead. */
#define GVM_OK 0
#define GVM_ERR_ARG 1      /* invalid argument or unsupported size */
#define GVM_ERR_OOM 2      /* device allocation failed */
#define GVM_ERR_HIP 3      /* HIP runtime error (see gvm_last_error) */
#define GVM_ERR_STATE 4    /* call out of order (e.g. dchi2 before chi2) */
#define GVM_ERR_NODEV 5    /* no usable gfx950 device */

/* Primary beam models: reference include/MSFITSIO.cuh:56 (AIRYDISK, GAUSSIAN). */
#define GVM_BEAM_AIRYDISK 0
#define GVM_BEAM_GAUSSIAN 1

/* Adjoint (DChi2) evaluation forms. */
#define GVM_ADJOINT_F32 0      /* separable complex GEMM, FP32 MFMA            */
#define GVM_ADJOINT_F16X3 1    /* FP16 hi/lo split, 3 MFMA products, FP32 acc  */
#define GVM_ADJOINT_BF16X3 2   /* BF16 hi/lo split, 3 MFMA products, FP32 acc  */
#define GVM_ADJOINT_DIRECT 3   /* direct sum over visibilities per pixel (VALU) */

/* Geometry and the scalars the reference keeps in globals
 * (src/mfs.cu:4-38, src/functions.cu:37-71): image size, cell size,
 * pointing/phase centres, nu_0, fg_scale, noise_cut, MINPIX
 * (= initial_values[0]), eta, threshold, and the antenna beam
 * (src/MSFITSIO.cu:522-552). */
typedef struct {
  int M, N;
  int image_count;         /* must be 2 (I_nu_0, alpha), as src/mfs.cu:175-179 forces */
  double deltax, deltay;   /* CDELT1/CDELT2 in degrees */
  float x0_ref, y0_ref;    /* pointing centre in pixels */
  float x0_phs, y0_phs;    /* phase centre in pixels */
  float nu_0;              /* reference frequency (Hz) */
  float fg_scale;
  float noise_cut;         /* absolute (already scaled by the minimum noise) */
  float minpix;            /* initial_values[0] */
  float eta;               /* global eta (default -1) */
  float threshold;         /* -T (default 0) */
  int beam_kind;           /* GVM_BEAM_* */
  float antenna_diameter;  /* metres */
  float pb_factor;
  float pb_cutoff;         /* radians */
} gvm_geometry;

/* ---- context ---------------------------------------------------------- */

/* Replaces the device state built by MFS::setDevice (src/mfs.cu:530-943):
 * copies the noise image (float[M*N], host) used as the mask, allocates the
 * uv grid and work buffers, creates the context's stream. */
gvm_status gvm_create(const gvm_geometry* geom, const float* noise_host,
                      gvm_ctx** out);
/* Set-up on the device, for visibilities already in device memory (no
 * counterpart entry point in the reference, whose set-up kernels these are):
 * hermitianSymmetry (src/functions.cu:2256-2273: rows with u > 0 get u, v and
 * Im Vo negated; metres to wavelengths with the float wavelength), then what
 * gvm_add_visibilities does, in FP64 on the device.  uvw_m_dev: 3*nvis
 * doubles in metres; vo_dev: 2*nvis floats; weight_dev: nvis floats (all
 * device, left unchanged).  The slice equals the one gvm_add_visibilities
 * builds from the same rows after gvm_setup_hermitian_lambda. */
gvm_status gvm_add_visibilities_dev(gvm_ctx* ctx, float freq, long nvis,
                                    const double* uvw_m_dev, const float* vo_dev,
                                    const float* weight_dev, int* slice_out);
/* calculateNoiseAndBeam (src/functions.cu:1700-1814) with the weighted uv
 * moments summed on the device (device uvw in metres and weights per
 * channel); outputs as gvm_setup_beam_noise. */
int gvm_setup_beam_noise_dev(int nch, const float* freq, const long* nvis,
                             const double* const* uvw_m_dev, const float* const* weight_dev,
                             double deltax_deg, double* bmaj_deg, double* bmin_deg,
                             double* bpa_deg, double* sum_weights, double* vis_noise,
                             double* noise_jypix);
/* The clean beam and the noise estimates from the weighted moments
 * sum w u^2, sum w v^2, sum w u v, sum w (src/functions.cu:1760-1814). */
int gvm_setup_beam_from_moments(double suu, double svv, double suv, double sw,
                                double deltax_deg, double* bmaj_deg, double* bmin_deg,
                                double* bpa_deg, double* sum_weights, double* vis_noise,
                                double* noise_jypix);
/* The noise image (src/mfs.cu:849-916, noise_image src/functions.cu:2382-2422)
 * computed on the device and copied to noise_out (host, float[M*N]), with
 * fg_scale = its minimum and noise_cut = noise_cut_factor * fg_scale. */
int gvm_setup_noise_image_dev(int beam_kind, double diameter, double pb_factor,
                              double pb_cutoff, double nu_0, int M, int N, double deltax_deg,
                              double deltay_deg, float x0_ref, float y0_ref, double noise_jypix,
                              double noise_cut_factor, float* noise_out, float* fg_scale,
                              float* noise_cut);

/* ---- gridded-visibility mode (set-up) ----------------------------------- */
/* A convolution (anti-aliasing) kernel as the reference's CKernel holds it
 * (include/classes/ckernel.cuh:415-511): m rows by n columns of taps,
 * taps[n*i + j], centred at (support_y, support_x).  The reference's default
 * is PillBox2D() (src/pillBox2D.cu:22-25): m = n = 1, support 0, tap 1. */
typedef struct {
  int m, n;
  int support_x, support_y;
  const float* taps; /* host, m*n floats */
} gvm_ckernel;

/* Convolutional gridding of one channel x stokes slice onto the M x N uv
 * grid of ctx: do_gridding (src/functions.cu:1339-1653) for one slice.  Each
 * row and its Hermitian twin (u, v, w and Im Vo negated) is placed at the
 * cell int(u / deltau + floor(N/2) + 0.5) (u in wavelengths, FP64) and spread
 * over the kernel's taps; per cell, sum w k, sum w k^2 and sum w k Vo are
 * accumulated in float in row order (the reference's order with one gridding
 * thread), then Vo = sum w k Vo / sum w k and weight = (sum w k)^2 /
 * sum w k^2.  Cells with a positive weight are written in row-major cell
 * order as visibilities at the cell centre (uvw in metres, w = 0).
 * uvw_m_dev / vo_dev / weight_dev: the raw rows on the device (metres,
 * before the Hermitian flip; 3, 2 and 1 values per row); the outputs hold
 * `capacity` rows (M*N always suffices); *nout receives the count.
 * Synchronises the device. */
gvm_status gvm_grid_visibilities_dev(gvm_ctx* ctx, const gvm_ckernel* kernel, float freq,
                                     long nvis, const double* uvw_m_dev, const float* vo_dev,
                                     const float* weight_dev, double* uvw_out_dev,
                                     float* vo_out_dev, float* weight_out_dev, long capacity,
                                     long* nout);
/* The gridded rows of one slice added to ctx as a new slice: the
 * do_gridding step of MFS::configure (src/mfs.cu:510-527) followed by the
 * device set-up of gvm_add_visibilities_dev.  *ngridded receives the number
 * of gridded rows (may be NULL). */
gvm_status gvm_add_visibilities_gridded_dev(gvm_ctx* ctx, const gvm_ckernel* kernel, float freq,
                                            long nvis, const double* uvw_m_dev,
                                            const float* vo_dev, const float* weight_dev,
                                            long* ngridded, int* slice_out);
/* The gridding correction function image (CKernel::getGCFGPU, host float[M*N],
 * or NULL for none): once set, the model image is multiplied by it before
 * the transform (apply_GCF, src/functions.cu:2468-2476, called at :4226-4229
 * and :4358-4362) and the gradient by it per pixel (the DChi2 variant with
 * gcf, src/functions.cu:3793-3888). */
gvm_status gvm_set_gcf(gvm_ctx* ctx, const float* gcf_host);

/* Weighting schemes (the WeightingScheme::apply plugins run by
 * MFS::configure before gridding, src/mfs.cu:499-508). */
#define GVM_WEIGHTING_NATURAL 0 /* src/naturalweightingscheme.cu:9-48 */
#define GVM_WEIGHTING_UNIFORM 1 /* src/uniformweightingscheme.cu:9-99 */
#define GVM_WEIGHTING_BRIGGS 2  /* src/briggsweightingscheme.cu:30-187 */
#define GVM_WEIGHTING_RADIAL 3  /* src/radialweightingscheme.cu:8-43 */

/* UVTaper (include/classes/uvtaper.cuh): amplitude * exp(-a x^2 - b x y -
 * c y^2) at (x, y) = (u - u_0, v - v_0) in wavelengths. */
typedef struct {
  float amplitude;
  double u_0, v_0;
  float sigma_maj, sigma_min, bpa;
} gvm_uvtaper;

/* Applies `scheme` to the weights of nsl slices in place on the device
 * (weight_dev[s]: nvis[s] floats; uvw_m_dev[s]: raw rows in metres), with
 * the grid of ctx (cells of |deltau| x |deltav|, rows with u < 0 mirrored).
 * Uniform: w / (sum of w in the cell); Briggs: w / (1 + G f^2) with f^2 =
 * (5 10^-robust)^2 / (sum G^2 / sum w), robust in [-2, 2]; Radial: w * |uv|;
 * then each scheme multiplies by the taper (NULL: none).  The per-cell sums
 * are float sums in row order, as the reference's with one thread; the two
 * Briggs totals are FP64 sums (the reference sums them in float). */
gvm_status gvm_weighting_apply_dev(gvm_ctx* ctx, int scheme, float robust,
                                   const gvm_uvtaper* taper, int nsl, const float* freq,
                                   const long* nvis, const double* const* uvw_m_dev,
                                   float* const* weight_dev);

/* Replaces MFS::unSetDevice (src/mfs.cu:1157-1253). */
gvm_status gvm_destroy(gvm_ctx* ctx);

/* Replaces the per channel x stokes H2D copies of MFS::setDevice
 * (src/mfs.cu:555-617) and the off-grid weight zeroing of vis_mod
 * (src/functions.cu:2605-2606), done once here.  uvw: 3*nvis doubles
 * (wavelengths, u <= 0), vo: 2*nvis floats, weight: nvis floats, all host.
 * Returns the slice index.  Slices are evaluated in insertion order. */
gvm_status gvm_add_visibilities(gvm_ctx* ctx, float freq, long nvis,
                                const double* uvw, const float* vo,
                                const float* weight, int* slice_out);
gvm_status gvm_num_slices(const gvm_ctx* ctx, int* n);
/* Copy of the geometry the context was created with. */
gvm_status gvm_get_geometry(const gvm_ctx* ctx, gvm_geometry* out);

/* Adjoint form used by gvm_dchi2 (GVM_ADJOINT_*; default GVM_ADJOINT_BF16X3: on
 * MI355X it is the faster split form and the more accurate one at 1024^2). */
gvm_status gvm_set_adjoint(gvm_ctx* ctx, int mode);

/* ---- objective: chi2 ----------------------------------------------------- */

/* Replaces chi2() (src/functions.cu:4323-4454) as called by Chi2::calcFi
 * (src/chi2.cu:44-49): clip2IWNoise in place on I (:2694-2719), then per
 * slice calculateInu (:3939-3966), apply_beam2I (:2424-2444), the inverse 2-D
 * FFT (:2153-2178), phase_rotate (:2483-2518), vis_mod (:2557-2609),
 * residual (:2663-2671) and chi2Vector + reduction (:2867-2876).
 * *chi2_out = 0.5 * sum_slices [sum_v w |Vo - Vm|^2 (/ nvis if normalize)],
 * summed in FP64.  The residuals stay on the device for gvm_dchi2.
 * flag_opt is the optimisation schedule flag (src/frprmn.cu:46). */
gvm_status gvm_chi2(gvm_ctx* ctx, float* I_dev, int flag_opt, int normalize,
                    double* chi2_out);

/* Asynchronous value forms, with no counterpart in the reference: they
 * enqueue an evaluation on the context's stream and leave its FP64 value in
 * value slot `slot` (0..15) instead of returning it, so a caller can
 * evaluate several terms of an objective with one synchronisation
 * (gvm_values_read).  The slots are pinned host memory that the kernels
 * write directly.  Results are identical to the synchronous forms.  A slot
 * keeps its value until it is written again. */
#define GVM_VALUE_SLOTS 16
gvm_status gvm_chi2_enqueue(gvm_ctx* ctx, float* I_dev, int flag_opt, int normalize, int slot);
gvm_status gvm_entropy_value_enqueue(gvm_ctx* ctx, const float* I_dev, int index, float prior,
                                     float eta, int slot);
gvm_status gvm_tsv_value_enqueue(gvm_ctx* ctx, const float* I_dev, int index, int slot);
gvm_status gvm_laplacian_value_enqueue(gvm_ctx* ctx, const float* I_dev, int index, int slot);
/* Waits for the stream and copies value slots [0, n) to out. */
gvm_status gvm_values_read(gvm_ctx* ctx, int n, double* out);
/* gvm_values_read in two halves: the first enqueues nothing (the kernels
 * write the slots in host memory themselves; it is kept so that callers and
 * captured graphs written for a copy still work), the second waits for the
 * stream and reads slots [0, n). */
gvm_status gvm_values_fetch_enqueue(gvm_ctx* ctx, int n);
gvm_status gvm_values_host(gvm_ctx* ctx, int n, double* out);

/* ---- HIP graphs of enqueued work ----------------------------------------- */
/* No counterpart in the reference, which launches every kernel from the host
 * each time.  gvm_capture_begin starts capturing the context's stream
 * (thread-local capture mode); the calls that follow only record their work;
 * gvm_capture_end instantiates it as a graph that gvm_graph_launch replays on
 * the stream.  Only calls that enqueue without waiting may be captured: the
 * *_enqueue value forms, gvm_values_fetch_enqueue, gvm_dchi2 (after one
 * uncaptured gvm_dchi2 has sized its workspace), the prior gradients and the
 * vector operations; a call that waits fails the capture.  A graph holds the
 * device pointers it was recorded with and reads the data behind them when it
 * runs.  Adding visibilities, changing the adjoint form or the GCF moves the
 * context to a new epoch; a graph from an older epoch is refused
 * (GVM_ERR_STATE). */
typedef struct gvm_graph gvm_graph;
gvm_status gvm_capture_begin(gvm_ctx* ctx);
gvm_status gvm_capture_end(gvm_ctx* ctx, gvm_graph** graph);
gvm_status gvm_graph_launch(gvm_ctx* ctx, gvm_graph* graph);
gvm_status gvm_graph_destroy(gvm_graph* graph);
/* A counter that changes whenever graphs of the context become invalid. */
gvm_status gvm_epoch(const gvm_ctx* ctx, long* epoch);

/* Replaces dchi2() (src/functions.cu:4456-4558) as called by Chi2::calcGi
 * (src/chi2.cu:51-53): per slice the DChi2 adjoint (:3698-3791) at the
 * residuals of the last gvm_chi2, then DChi2_total_I_nu_0 (:4000-4024) when
 * flag_opt is even or DChi2_total_alpha (:3968-3998) when odd, accumulated
 * into result_dev (float[image_count*M*N], device).  result_dev is not
 * cleared first (Chi2::restartDGi does that). */
gvm_status gvm_dchi2(gvm_ctx* ctx, const float* I_dev, float* result_dev,
                     int flag_opt, int normalize);

/* Residuals V_r of one slice from the last gvm_chi2 (host float[2*nvis]);
 * replaces the Vr readback of MFS::writeResiduals (src/mfs.cu:1115-1155). */
gvm_status gvm_get_residuals(gvm_ctx* ctx, int slice, float* vr_host);
/* The model and residual visibilities of one slice from the last gvm_chi2, in
 * the frame of the observation: modelToHost (src/MSFITSIO.cu:1114-1138)
 * conjugates V_m back for the rows that had u > 0 before the Hermitian flip
 * (uvw_obs: the observed uvw, 3*nvis doubles, host; NULL when no row was
 * flipped); the residuals V_o - V_m are conjugated likewise.  model_host /
 * residual_host: 2*nvis floats each, either may be NULL. */
gvm_status gvm_get_visibilities_out(gvm_ctx* ctx, int slice, const double* uvw_obs,
                                    float* model_host, float* residual_host);

/* ---- priors ------------------------------------------------------------- */
/* Values: sum over the mask of the per-pixel term on image `index`, FP64.
 * Gradients: overwrite dgi_dev (float[M*N]) with lambda * d(term).
 * The iter > 0 and flag_opt gating of the reference host launchers is the
 * caller's (the Fi classes') job. */

/* SEntropy -> SVector/calculateS (src/functions.cu:4722-4742, 3081-3095,
 * 3036-3055). */
gvm_status gvm_entropy_value(gvm_ctx* ctx, const float* I_dev, int index,
                             float prior, float eta, double* out);
/* DEntropy -> DS/calculateDS (src/functions.cu:4744-4763, 3097-3112,
 * 3057-3079). */
gvm_status gvm_entropy_grad(gvm_ctx* ctx, const float* I_dev, float* dgi_dev,
                            int index, float prior, float eta, float lambda);
/* TotalSquaredVariation -> TSVVector/calculateTSV (src/functions.cu:4927-4945,
 * 3382-3393, 3353-3380). */
gvm_status gvm_tsv_value(gvm_ctx* ctx, const float* I_dev, int index,
                         double* out);
/* DTSVariation -> DTSV/calculateDTSV (src/functions.cu:4947-4964, 3427-3441,
 * 3395-3425). */
gvm_status gvm_tsv_grad(gvm_ctx* ctx, const float* I_dev, float* dgi_dev,
                        int index, float lambda);
/* laplacian -> LVector/calculateL (src/functions.cu:4808-4826, 3475-3486,
 * 3443-3473). */
gvm_status gvm_laplacian_value(gvm_ctx* ctx, const float* I_dev, int index,
                               double* out);
/* DLaplacian -> DL/calculateDL (src/functions.cu:4828-4845, 3532-3545,
 * 3488-3530), including the reference's stencil sign pattern. */
gvm_status gvm_laplacian_grad(gvm_ctx* ctx, const float* I_dev,
                              float* dgi_dev, int index, float lambda);

/* ---- several priors in one pass ------------------------------------------ */
/* No counterpart in the reference, which launches each term's value, each
 * term's gradient and each AddToDPhi separately (src/functions.cu:4722-4964,
 * 4560-4564).  A term: its kind, the image it reads, the image of dphi its
 * gradient adds to, its value slot's order, entropy's prior G and eta
 * (ignored by the others), the gradient factor lambda, and whether it is
 * active (the reference's iter > 0 and flag_opt gating, decided by the
 * caller; an inactive value is 0 and an inactive gradient adds +0).  Results
 * are bit-identical to the one-term calls in the same order. */
#define GVM_PRIOR_ENTROPY 0
#define GVM_PRIOR_TSV 1
#define GVM_PRIOR_LAPLACIAN 2
#define GVM_MAX_FUSED_PRIORS 4
typedef struct {
  int kind;
  int index;
  int image_to_add;
  int active;
  float prior;
  float eta;
  float lambda;
  float pad;
} gvm_prior_term;
/* The values of n <= GVM_MAX_FUSED_PRIORS terms into value slots
 * [first_slot, first_slot + n) (read with gvm_values_read / values_host);
 * one pass over the image for all terms. */
gvm_status gvm_priors_value_enqueue(gvm_ctx* ctx, const float* I_dev, const gvm_prior_term* terms,
                                    int n, int first_slot);
/* gvm_chi2_enqueue into chi2_slot followed by gvm_priors_value_enqueue into
 * [first_slot, first_slot + n): for up to 4 slices the prior partials share
 * the degrid's launch and one finishing launch writes all the values (more
 * slices take chi2's own finishing launches).  Same values bit for bit. */
gvm_status gvm_values_enqueue(gvm_ctx* ctx, float* I_dev, int flag_opt, int normalize,
                              int chi2_slot, const gvm_prior_term* terms, int n, int first_slot);
/* ObjectiveFunction's gradient assembly after Chi2::calcGi in one pass:
 * dphi = chi2_grad (NULL: zero), then dphi[term.image_to_add] += lambda *
 * dterm for each term in order, then xi = dphi; every buffer
 * float[image_count*M*N]. */
gvm_status gvm_dphi_assemble(gvm_ctx* ctx, const float* I_dev, const float* chi2_grad_dev,
                             float* dphi_dev, float* xi_dev, const gvm_prior_term* terms, int n);

/* linkAddToDPhi -> AddToDPhi (src/functions.cu:4560-4564, 3890-3895):
 * dphi[index] += dgi. */
gvm_status gvm_add_to_dphi(gvm_ctx* ctx, float* dphi_dev, const float* dgi_dev,
                           int index);

/* ---- optimiser vector operations (all on float[nimg*M*N] device vectors;
 *      reductions are FP64) ------------------------------------------------ */

/* cudaMemset of the reference, as a fill kernel (graph-capture safe). */
gvm_status gvm_vec_zero(gvm_ctx* ctx, float* x_dev, long n);
/* cudaMemcpy device-to-device of the reference. */
gvm_status gvm_vec_copy(gvm_ctx* ctx, float* dst_dev, const float* src_dev,
                        long n);
/* searchDirection (src/functions.cu:3655-3668): g = -xi; xi = h = g, image k. */
gvm_status gvm_cg_search_direction(gvm_ctx* ctx, float* g, float* xi, float* h,
                                   int image);
/* CGGradCondition + deviceMaxReduce (src/functions.cu:3599-3612, 743-879):
 * max over all nimg images of |xi| * max(|p|, 1) / den. */
gvm_status gvm_cg_grad_condition(gvm_ctx* ctx, const float* xi, const float* p,
                                 double den, int nimg, double* gmax);
/* getGGandDGG + 2 x deviceReduce (src/functions.cu:2729-2749):
 * gg = sum g^2, dgg = sum (xi + g) xi over all nimg images. */
gvm_status gvm_cg_gg_dgg(gvm_ctx* ctx, const float* xi, const float* g,
                         int nimg, double* gg, double* dgg);
/* newXi (src/functions.cu:3678-3687): g = -xi; xi = h = g + gam h, image k. */
gvm_status gvm_cg_new_xi(gvm_ctx* ctx, float* g, float* xi, float* h,
                         double gam, int image);
/* evaluateXt / evaluateXtNoPositivity (src/functions.cu:2832-2865):
 * xt = pcom + x xicom on image k, floored at -eta*minpix if positivity. */
gvm_status gvm_evaluate_xt(gvm_ctx* ctx, float* xt, const float* pcom,
                           const float* xicom, double x, int image,
                           int positivity, float minpix, float eta);
/* newP / newPNoPositivity (src/functions.cu:2779-2812): xi *= xmin;
 * p += xi, or with positivity p = floor and xi = 0 where p + xi <= floor. */
gvm_status gvm_new_p(gvm_ctx* ctx, float* p, float* xi, double xmin, int image,
                     int positivity, float minpix, float eta);
/* normArray + deviceMaxReduce (src/functions.cu:3587-3597): max |x| over n. */
gvm_status gvm_vec_absmax(gvm_ctx* ctx, const float* x, long n, double* out);
/* searchDirection_LBFGS / getR (src/functions.cu:3564-3569, 3623-3633):
 * dst = src * scale over n elements (dst may alias src). */
gvm_status gvm_vec_scale(gvm_ctx* ctx, float* dst, const float* src,
                         double scale, long n);
/* getDot_LBFGS_ff + deviceReduce (src/functions.cu:3571-3585): sum a*b, FP64. */
gvm_status gvm_vec_dot(gvm_ctx* ctx, const float* a, const float* b, long n,
                       double* out);
/* updateQ (src/functions.cu:3614-3621): y += alpha x over n elements. */
gvm_status gvm_vec_axpy(gvm_ctx* ctx, float* y, const float* x, double alpha,
                        long n);
/* Enqueue-only forms of the L-BFGS recursion's scalars (no counterpart in the
 * reference, which reads every dot product back to the host: src/lbfgs.cu:
 * 270-341): the scalars stay in device doubles (caller-owned device memory)
 * and the vector updates read them there, so the two-loop recursion runs
 * without a host round trip.  Each scalar is computed with the host form's
 * double operations, rounded once each, so the coefficients are identical.
 *   gvm_vec_dot_dev:      *out_dev = sum a*b (FP64, the order of gvm_vec_dot);
 *   gvm_scalar_rho_times: r = (*rho_den != 0 ? 1 / *rho_den : 0) * *dot;
 *                         *out = minus_from ? *minus_from - r : r;
 *   gvm_scalar_h0:        *out = sum over i < nimg of sy[i] / yy[i] (yy[i] != 0);
 *   gvm_vec_axpy_dev:     y += (float)(negate ? -*coef : *coef) * x;
 *   gvm_vec_scale_dev:    dst = src * (float)*coef. */
gvm_status gvm_vec_dot_dev(gvm_ctx* ctx, const float* a, const float* b, long n, double* out_dev);
/* gvm_vec_dot_dev with no context and no allocation (no counterpart in the
 * reference).  For callers that cannot allocate device memory: code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc.  The
 * FP64 block partials live in a static device array of 1024 doubles (8 KiB)
 * that each GPU the library runs on holds until the process ends; its device
 * address is looked up on the first call on each device.  The result equals
 * gvm_vec_dot_dev's bit for bit.  Enqueued on `stream` (a hipStream_t; NULL:
 * the null stream), nothing waits.  Calls that overlap on one GPU (other
 * streams, other threads) share the array and race; the partials need no
 * clearing (each call writes every partial it reads).  n in
 * [1, GVM_NOALLOC_DOT_MAX_N], otherwise GVM_ERR_ARG and nothing is launched. */
#define GVM_NOALLOC_DOT_MAX_N (1L << 37)
gvm_status gvm_vec_dot_dev_noalloc(const float* a, const float* b, long n, double* out_dev,
                                   void* stream);
gvm_status gvm_scalar_rho_times(gvm_ctx* ctx, const double* rho_den, const double* dot,
                                const double* minus_from, double* out);
gvm_status gvm_scalar_h0(gvm_ctx* ctx, const double* sy, const double* yy, int nimg,
                         double* out);
gvm_status gvm_vec_axpy_dev(gvm_ctx* ctx, float* y, const float* x, const double* coef,
                            int negate, long n);
gvm_status gvm_vec_scale_dev(gvm_ctx* ctx, float* dst, const float* src, const double* coef,
                             long n);
/* calculateSandY (src/functions.cu:3635-3653) for one image:
 * y = xi + xi_old; s = p - p_old over n elements. */
gvm_status gvm_lbfgs_s_and_y(gvm_ctx* ctx, float* y, float* s, const float* p,
                             const float* xi, const float* p_old,
                             const float* xi_old, long n);

/* ---- set-up helpers (host, FP64) ---------------------------------------- */

/* attenuation (src/functions.cu:2304-2333) with GaussianBeam / AiryDiskBeam
 * (:2293-2302, 2275-2291) for every pixel, host float[M*N]. */
gvm_status gvm_primary_beam(int beam_kind, double antenna_diameter,
                            double pb_factor, double pb_cutoff, double freq,
                            int M, int N, float xobs, float yobs,
                            double deltax, double deltay, float* out);

/* ---- runtime ------------------------------------------------------------ */
gvm_status gvm_sync(gvm_ctx* ctx);
/* Device memory helpers for callers without their own allocator. */
gvm_status gvm_malloc(void** ptr, size_t bytes);
gvm_status gvm_free(void* ptr);
gvm_status gvm_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes);
gvm_status gvm_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes);
/* Text of the last error on this thread. */
const char* gvm_last_error(void);
/* Device name, CU count and clock of the current device; 0 if none. */
gvm_status gvm_device_info(char* name, int name_len, int* cu_count,
                           int* clock_khz);

/* ---- timing of the adjoint (for measurement) ---------------------------- */
/* Average duration in ms of the adjoint GEMM kernel over the gvm_dchi2 calls
 * since the last reset, measured with HIP events on the context's stream. */
gvm_status gvm_timing_reset(gvm_ctx* ctx);
gvm_status gvm_timing_adjoint_ms(gvm_ctx* ctx, double* avg_ms, long* count);

/* ---- adjoint K tile: the visibility-range length ------------------------ */
/* No counterpart in the reference.  Each workgroup of the split-precision
 * adjoint accumulates one 256 x 256 output tile over a range of about
 * `split_vis` visibilities in FP32 (the ranges are summed in FP64): shorter
 * ranges keep the FP32 accumulation short, longer ones write fewer partial
 * tiles.  Default 6144, and 3072 for images of 2048^2 and larger (where
 * ranges of 6144 bring the gradient's relative error near 1e-5); allowed
 * 256 .. 16384. */
gvm_status gvm_set_split_vis(gvm_ctx* ctx, long split_vis);
/* Batching of the image-to-grid transform (no counterpart in the reference,
 * which calls cuFFT): 0 automatic (default), 1 wide blocks (many rows or
 * columns per block, for many slices), 2 narrow blocks (one to four, for a
 * single small image).  Both give identical grids. */
gvm_status gvm_set_fft_mode(gvm_ctx* ctx, int mode);
gvm_status gvm_get_split_vis(const gvm_ctx* ctx, long* split_vis);
/* No counterpart in the reference (whose DChi2 computes the pixels on the
 * mask, noise < noise_cut): the pixels whose sums the adjoint of the current
 * form computes.  The split-precision GEMM computes the 128 x 64 sub-tiles of
 * its workgroup tiles that hold a pixel on the mask and skips the others; the
 * direct form, its blocks of 256 consecutive pixels that hold one.  The work
 * of one adjoint launch is 4 x pixels x visibilities flop. */
gvm_status gvm_adjoint_computed_pixels(const gvm_ctx* ctx, long* pixels);
/* Autotuner: times the adjoint of the current residuals (a gvm_chi2 must
 * have run) for each candidate range length (NULL: 1536, 2048, 3072, 4096,
 * 6144, 8192, up to 3072 for images of 2048^2 and larger), with HIP events,
 * and keeps the fastest.  The choice is cached
 * per shape (M, total visibilities, slices, adjoint form, CU count) for the
 * process, and in the text file named by the GVM_TUNE_CACHE environment
 * variable when it is set; a cached shape is not timed again (*cached = 1).
 * ms_out (optional, ncand values) receives the timings. */
gvm_status gvm_autotune_adjoint(gvm_ctx* ctx, const long* candidates, int ncand, long* chosen,
                                double* ms_out, int* cached);

#ifdef __cplusplus
}
#endif
#endif /* GVM_HIP_H */
