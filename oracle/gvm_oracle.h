/*
 * gvm_oracle.h -- FP64 CPU restatement of gpuvmem's maximum-entropy
 * objective, gradient and optimiser trajectory.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only to check
 * results against (or as the timed CPU baseline).  The product path
 * (gpuvmem_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * double precision throughout; inputs that are float in the reference (images,
 * visibilities, weights, channel frequencies) are read as float and promoted.
 *
 * Conventions (reference src/functions.cu, src/mfs.cu):
 *   image memory is row-major [image k][row i][col j] at M*N*k + N*i + j,
 *   with M == N (the reference's vis_mod/phase_rotate assume it);
 *   uvw are in wavelengths after the Hermitian flip (all u <= 0);
 *   the uv grid is the unshifted output of an unnormalised C2C transform
 *   with the e^{+2 pi i} sign.
 */
#ifndef GVM_ORACLE_H
#define GVM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double re, im;
} ocplx;

enum { ORC_AIRYDISK = 0, ORC_GAUSSIAN = 1 };

/* Geometry + the scalars the hot path consumes (reference globals,
 * src/mfs.cu:4-38).  noise is the per-pixel noise image (M*N floats). */
typedef struct {
  int M, N;
  double deltax, deltay;  /* CDELT1/CDELT2, degrees */
  double deltau, deltav;  /* 1/(M*deltax_rad), 1/(N*deltay_rad) */
  float x0_ref, y0_ref;   /* pointing centre, pixels */
  float x0_phs, y0_phs;   /* phase centre, pixels */
  float fg_scale;
  float noise_cut;
  float minpix;    /* initial_values[0] */
  float eta;       /* global eta (default -1) */
  float threshold; /* -T, default 0 */
  float nu_0;
  const float* noise;
  const float* gcf; /* gridding correction image (M*N floats) or NULL */
} orc_geom;

typedef struct {
  int kind; /* ORC_GAUSSIAN or ORC_AIRYDISK */
  double diameter;
  double pb_factor;
  double pb_cutoff;
} orc_beam;

/* One channel x stokes slice of visibilities (after set-up). weight is
 * writable: vis_mod zeroes the weight of samples that fall off the grid. */
typedef struct {
  float freq;
  long nvis;
  const double* uvw; /* 3*nvis, wavelengths */
  const float* vo;   /* 2*nvis */
  float* weight;     /* nvis */
} orc_chan;

/* UVTaper (include/classes/uvtaper.cuh). */
typedef struct {
  float amplitude;
  double u_0, v_0;
  float sigma_maj, sigma_min, bpa;
} orc_taper;

/* ---- gridded-visibility mode (gvm_oracle_grid.c) ---- */
void orc_grid_visibilities(int M, int N, double deltax_deg, double deltay_deg, float freq,
                           long nvis, const double* uvw_m, const float* vo, const float* weight,
                           int km, int kn, int supx, int supy, const float* taps,
                           double* out_uvw, float* out_vo, float* out_w, long* nout);
void orc_weighting(int scheme, float robust, const orc_taper* taper, int M, int N,
                   double deltax_deg, double deltay_deg, int nsl, const float* freq,
                   const long* nvis, const double* const* uvw_m, float* const* weight,
                   int float_totals);

/* ---- set-up (row a19) ---- */
void orc_hermitian_to_lambda(long n, double* uvw, float* vo, float freq);
void orc_beam_and_noise(const orc_chan* ch_metres, int nch, double deltax_deg,
                        double* bmaj_deg, double* bmin_deg, double* bpa_deg,
                        double* sum_weights, double* vis_noise);
double orc_noise_jypix(double vis_noise, double bmaj_deg, double bmin_deg,
                       double deltax_deg);
void orc_noise_image(const orc_beam* b, float nu_0, int M, int N, double dx,
                     double dy, float x0_ref, float y0_ref, double noise_jypix,
                     double* noise_out, double* noise_min);

/* ---- primary beam (row a3) ---- */
double orc_attenuation(const orc_beam* b, double freq, int i, int j,
                       float xobs, float yobs, double dx_deg, double dy_deg);
void orc_pb_image(const orc_beam* b, double freq, int M, int N, float xobs,
                  float yobs, double dx_deg, double dy_deg, double* out);

/* ---- FFT (row a4): in place, rows x cols, sign +1 = e^{+2 pi i}, unnormalised */
int orc_fft2d(ocplx* a, int rows, int cols, int sign);

/* ---- forward chi2 (rows a1-a8) ---- */
void orc_clip2IWNoise(const orc_geom* g, double* I, int flag_opt);
void orc_model_grid(const orc_geom* g, const orc_beam* b, const double* I,
                    float freq, ocplx* grid);
double orc_degrid_chi2(const orc_geom* g, const ocplx* grid, orc_chan* ch,
                       ocplx* vm_out, ocplx* vr_out);
double orc_chi2(const orc_geom* g, const orc_beam* b, orc_chan* ch, int nch,
                double* I, int flag_opt, int normalize, ocplx* vr_out);

/* ---- adjoint (rows a10-a11) ---- */
void orc_dchi2_channel(const orc_geom* g, const orc_beam* b,
                       const orc_chan* ch, const ocplx* vr, int normalize,
                       int separable, double* dchi2);
/* omp_set_num_threads for the calling thread's later parallel regions. */
void orc_set_threads(int n);
int orc_max_threads(void);
void orc_dchi2_direct_pixels(const orc_geom* g, const orc_beam* b,
                             const orc_chan* ch, const ocplx* vr, int normalize,
                             const int* pix, int npix, double* out);
/* orc_dchi2 with direct sums at sampled pixels (pix: npix (row, col) pairs);
 * out: npix x 2 (I_nu_0 and alpha components). */
void orc_dchi2_pixels(const orc_geom* g, const orc_beam* b, const orc_chan* ch, int nch,
                      const double* I, const ocplx* vr, int flag_opt, int normalize,
                      const int* pix, int npix, double* out);
void orc_dchi2(const orc_geom* g, const orc_beam* b, const orc_chan* ch,
               int nch, const double* I, const ocplx* vr, double* result,
               int flag_opt, int normalize, int separable);

/* ---- priors (rows a12-a14) ---- */
double orc_entropy_value(const orc_geom* g, const double* I, int index,
                         double G, double eta);
void orc_entropy_grad(const orc_geom* g, const double* I, int index, double G,
                      double eta, double lambda, double* dgi);
double orc_tsv_value(const orc_geom* g, const double* I, int index);
void orc_tsv_grad(const orc_geom* g, const double* I, int index, double lambda,
                  double* dgi);
double orc_laplacian_value(const orc_geom* g, const double* I, int index);
void orc_laplacian_grad(const orc_geom* g, const double* I, int index,
                        double lambda, double* dgi);

/* ---- objective function (rows a9, a15) ---- */
enum { ORC_FI_CHI2 = 0, ORC_FI_ENTROPY = 1, ORC_FI_TSV = 2, ORC_FI_LAPLACIAN = 3 };
#define ORC_MAX_FI 8
typedef struct {
  int kind;
  double lambda; /* penalization factor */
  int imageIndex, imageToAdd;
  double prior, eta; /* entropy only */
  int iteration;
  double value; /* last fi value (without lambda) */
} orc_fi;

typedef struct {
  const orc_geom* g;
  const orc_beam* b;
  orc_chan* ch;
  int nch;
  int image_count;
  int flag_opt;
  int normalize;
  int separable; /* gradient form: 0 direct sum, 1 separable */
  int nopositivity;
  int nfi;
  orc_fi fi[ORC_MAX_FI];
  ocplx* vr; /* residual storage: sum of nvis over channels */
  double* dphi;
  double* dgi;
  long n_fcalls, n_gcalls;
} orc_objective;

int orc_of_init(orc_objective* of);
void orc_of_free(orc_objective* of);
double orc_of_calc_function(orc_objective* of, double* p);
void orc_of_calc_gradient(orc_objective* of, double* p, double* xi, int iter);

/* ---- line search + optimisers (rows a16-a18) ---- */
typedef double (*orc_f1)(double x, void* ctx);

/* A generic objective for the optimisers: image_count images of n values;
 * image 0 is floored at floor_v during the line search unless nopositivity. */
typedef struct {
  long n;
  int image_count;
  int nopositivity;
  double floor_v;
  double (*value)(void* ctx, double* p);
  void (*gradient)(void* ctx, double* p, double* xi, int iter);
  void* ctx;
} orc_problem;

/* The imaging objective as an orc_problem (ctx = the orc_objective). */
orc_problem orc_of_problem(orc_objective* of);
void orc_mnbrak(double* ax, double* bx, double* cx, double* fa, double* fb,
                double* fc, orc_f1 f, void* ctx);
double orc_brent(double ax, double bx, double cx, double tol, double* xmin,
                 orc_f1 f, void* ctx);
/* Log of linmin's accepted step lengths: buf[0..cap) receives the xmin of
 * each later linmin call (NULL stops logging); count is the number of calls. */
void orc_set_xmin_log(double* buf, int cap);
int orc_xmin_log_count(void);
double orc_linmin(const orc_problem* pr, double* p, double* xi);

typedef struct {
  int iterations;  /* iterations completed */
  long n_values, n_gradients;
  int exit_reason; /* 0 iter limit, 1 ftol, 2 gtol, 3 gg==0 */
  double fret;
  double final_value;
  long n_fcalls, n_gcalls;
} orc_opt_result;

/* traj (optional, capacity cap) receives the value after each line search. */
int orc_frprmn(const orc_problem* pr, double* p, int total_iterations, double ftol,
               double gtol, orc_opt_result* res, double* traj, int cap);
int orc_lbfgs(const orc_problem* pr, double* p, int K, int total_iterations,
              double ftol, double gtol, orc_opt_result* res, double* traj, int cap);

/* ---- CPU gridding comparator (reference src/functions.cu:1339-1653) ---- */
void orc_do_gridding_pillbox(int M, int N, double deltau, double deltav,
                             const orc_chan* ch, int nch, int support,
                             int nthreads, double* grid_re, double* grid_im,
                             double* grid_w);

#ifdef __cplusplus
}
#endif
#endif
