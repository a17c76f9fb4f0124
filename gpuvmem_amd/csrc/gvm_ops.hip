// Prior values/gradients and the optimiser's vector operations on gfx950.
//
// Priors: reference src/functions.cu:3036-3112 (entropy), 3353-3441 (TSV),
// 3443-3545 (Laplacian).  Values are FP32 per pixel, summed in FP64 per
// block and then in a fixed order (no float tree, no host loop); gradients
// are written in one pass.  Optimiser vector kernels: src/functions.cu
// :2729-2865, 3564-3687; the reductions that the reference runs as a float
// reduction + D2H per call (deviceReduce / deviceMaxReduce, :605-879) are
// FP64 here.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "gvm_internal.h"
#include "gvm_prior_terms.h"

namespace gvm {

// ---------------------------------------------------------------------------
// priors
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void prior_value_kernel(int kind, const float* __restrict__ im,
                                                          const float* __restrict__ noise,
                                                          float noise_cut, int N, float G,
                                                          float eta, double* __restrict__ partials) {
  __shared__ double sh[4];
  const long n = (long)N * N;
  double acc = 0.0;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    if (noise[k] < noise_cut)
      acc += (double)prior_term(kind, im, N, pix_row(k, N), pix_col(k, N), G, eta);
  }
  const double tot = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void prior_grad_kernel(int kind, const float* __restrict__ im,
                                                         const float* __restrict__ noise,
                                                         float noise_cut, int N, float G, float eta,
                                                         float lambda, float* __restrict__ dgi) {
  const long n = (long)N * N;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    float d = 0.0f;
    if (noise[k] < noise_cut) d = prior_grad(kind, im, N, pix_row(k, N), pix_col(k, N), G, eta);
    dgi[k] = d * lambda;
  }
}

// ---------------------------------------------------------------------------
// Several priors in one pass (the objective's terms after Chi2).  Each kernel
// repeats, per term, the arithmetic of the one-term kernels above in the same
// order, so values and gradients are bit-identical to the one-term path.
// ---------------------------------------------------------------------------

// Block partials of every term's value: row t of partials for term t (grid
// and per-thread order of prior_value_kernel).
__global__ __launch_bounds__(256) void priors_value_kernel(PriorTermsDev T,
                                                           const float* __restrict__ I,
                                                           const float* __restrict__ noise,
                                                           float noise_cut, int N,
                                                           double* __restrict__ partials) {
  __shared__ double sh[4];
  const long n = (long)N * N;
  double acc[GVM_MAX_FUSED_PRIORS];
#pragma unroll
  for (int t = 0; t < GVM_MAX_FUSED_PRIORS; t++) acc[t] = 0.0;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    if (!(noise[k] < noise_cut)) continue;
    const int i = pix_row(k, N), j = pix_col(k, N);
#pragma unroll
    for (int t = 0; t < GVM_MAX_FUSED_PRIORS; t++)
      if (t < T.n && T.t[t].active)
        acc[t] += (double)prior_term(T.t[t].kind, I + n * T.t[t].index, N, i, j, T.t[t].prior,
                                     T.t[t].eta);
  }
#pragma unroll
  for (int t = 0; t < GVM_MAX_FUSED_PRIORS; t++) {
    if (t >= T.n) break;
    const double tot = block_sum(acc[t], sh);
    if (threadIdx.x == 0) partials[(long)t * kReduceBlocks + blockIdx.x] = tot;
    __syncthreads();  // sh is reused by the next term
  }
}

// dphi and xi from the chi2 gradient and every prior term's gradient, per
// element in the order of the one-term path: dphi = chi2 gradient (copied,
// or added to the cleared dphi with one image), then dphi[image_to_add] +=
// lambda * dterm for each term in turn (an inactive term adds the +0 of its
// cleared buffer), then xi = dphi.
__global__ __launch_bounds__(256) void dphi_assemble_kernel(
    PriorTermsDev T, const float* __restrict__ I, const float* __restrict__ noise,
    float noise_cut, int N, int image_count, const float* __restrict__ chi2g,
    float* __restrict__ dphi, float* __restrict__ xi) {
#pragma clang fp contract(off)
  const long n = (long)N * N;
  const long total = n * image_count;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < total;
       k += (long)gridDim.x * blockDim.x) {
    const int img = (int)(k >> (2 * (31 - __clz(N))));  // k / n
    const long kk = k - (long)img * n;
    float r = 0.0f;
    if (chi2g) r = image_count == 1 ? 0.0f + chi2g[k] : chi2g[k];
    const bool inside = noise[kk] < noise_cut;
#pragma unroll
    for (int t = 0; t < GVM_MAX_FUSED_PRIORS; t++) {
      if (t >= T.n) break;
      if (T.t[t].image_to_add != img) continue;
      float ds = 0.0f;
      if (T.t[t].active) {
        float d = 0.0f;
        if (inside)
          d = prior_grad(T.t[t].kind, I + n * T.t[t].index, N, pix_row(kk, N), pix_col(kk, N),
                         T.t[t].prior, T.t[t].eta);
        ds = d * T.t[t].lambda;
      }
      r = r + ds;
    }
    dphi[k] = r;
    xi[k] = r;
  }
}

gvm_status check_terms(const gvm_ctx* ctx, const gvm_prior_term* terms, int n,
                              PriorTermsDev* out) {
  GVM_REQUIRE(n >= 0 && n <= GVM_MAX_FUSED_PRIORS,
              "fused priors: at most GVM_MAX_FUSED_PRIORS terms");
  GVM_REQUIRE(n == 0 || terms, "fused priors: null term list");
  *out = PriorTermsDev{};
  out->n = n;
  for (int t = 0; t < n; t++) {
    const gvm_prior_term& p = terms[t];
    GVM_REQUIRE(p.kind >= GVM_PRIOR_ENTROPY && p.kind <= GVM_PRIOR_LAPLACIAN,
                "fused priors: unknown kind");
    GVM_REQUIRE(p.index >= 0 && p.index < ctx->g.image_count &&
                    p.image_to_add >= 0 && p.image_to_add < ctx->g.image_count,
                "fused priors: bad image index");
    out->t[t] = p;
  }
  return GVM_OK;
}

static gvm_status prior_value_enqueue(gvm_ctx* ctx, int kind, const float* I, int index,
                                      float G, float eta, int slot) {
  GVM_REQUIRE(ctx && I, "prior value: null argument");
  GVM_REQUIRE(index >= 0 && index < ctx->g.image_count, "prior value: bad image index");
  GVM_REQUIRE(slot >= 0 && slot < kValueSlots, "prior value: value slot out of range");
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  const int N = ctx->g.N;
  const long n = (long)N * N;
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(prior_value_kernel, dim3(nb), dim3(256), 0, ctx->stream, kind, I + n * index,
                     ctx->noise, ctx->g.noise_cut, N, G, eta, ctx->partials);
  GVM_CHECK_LAUNCH();
  return finish_reductions_to(ctx, 1, nb, ctx->dvalues + slot);
}

static gvm_status prior_value(gvm_ctx* ctx, int kind, const float* I, int index, float G,
                              float eta, double* out) {
  GVM_REQUIRE(out, "prior value: null argument");
  gvm_status st = prior_value_enqueue(ctx, kind, I, index, G, eta, 0);
  if (st) return st;
  return values_read(ctx, 1, out);
}

static gvm_status prior_grad(gvm_ctx* ctx, int kind, const float* I, float* dgi, int index,
                             float G, float eta, float lambda) {
  GVM_REQUIRE(ctx && I && dgi, "prior gradient: null argument");
  GVM_REQUIRE(index >= 0 && index < ctx->g.image_count, "prior gradient: bad image index");
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  const int N = ctx->g.N;
  const long n = (long)N * N;
  hipLaunchKernelGGL(prior_grad_kernel, dim3(reduce_grid(n)), dim3(256), 0, ctx->stream, kind,
                     I + n * index, ctx->noise, ctx->g.noise_cut, N, G, eta, lambda, dgi);
  GVM_CHECK_LAUNCH();
  return GVM_OK;
}

// ---------------------------------------------------------------------------
// vector kernels
// ---------------------------------------------------------------------------
#define GRID_LOOP(k, n) \
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < (n); k += (long)gridDim.x * blockDim.x)

__global__ void add_kernel(float* __restrict__ y, const float* __restrict__ x, long n) {
  GRID_LOOP(k, n) y[k] += x[k];
}
__global__ void fill_kernel(float* __restrict__ y, float v, long n) {
  GRID_LOOP(k, n) y[k] = v;
}
__global__ void copy_kernel(float* __restrict__ y, const float* __restrict__ x, long n) {
  GRID_LOOP(k, n) y[k] = x[k];
}
__global__ void scale_kernel(float* y, const float* x, float a, long n) {
  GRID_LOOP(k, n) y[k] = x[k] * a;
}
__global__ void axpy_kernel(float* __restrict__ y, const float* __restrict__ x, float a, long n) {
  GRID_LOOP(k, n) y[k] += a * x[k];
}
__global__ void search_direction_kernel(float* __restrict__ g, float* __restrict__ xi,
                                        float* __restrict__ h, long n) {
  GRID_LOOP(k, n) {
    const float v = -xi[k];
    g[k] = v;
    xi[k] = v;
    h[k] = v;
  }
}
__global__ void new_xi_kernel(float* __restrict__ g, float* __restrict__ xi, float* __restrict__ h,
                              float gam, long n) {
  GRID_LOOP(k, n) {
    const float gv = -xi[k];
    g[k] = gv;
    const float v = gv + gam * h[k];
    xi[k] = v;
    h[k] = v;
  }
}
__global__ void evaluate_xt_kernel(float* __restrict__ xt, const float* __restrict__ pcom,
                                   const float* __restrict__ xicom, float x, int positivity,
                                   float floor_v, long n) {
  GRID_LOOP(k, n) {
    const float v = pcom[k] + x * xicom[k];
    xt[k] = (positivity && !(v > floor_v)) ? floor_v : v;
  }
}
__global__ void new_p_kernel(float* __restrict__ p, float* __restrict__ xi, float xmin,
                             int positivity, float floor_v, long n) {
  GRID_LOOP(k, n) {
    const float d = xi[k] * xmin;
    if (!positivity || p[k] + d > floor_v) {
      xi[k] = d;
      p[k] += d;
    } else {
      xi[k] = 0.0f;
      p[k] = floor_v;
    }
  }
}
__global__ void s_and_y_kernel(float* __restrict__ y, float* __restrict__ s,
                               const float* __restrict__ p, const float* __restrict__ xi,
                               const float* __restrict__ p_old, const float* __restrict__ xi_old,
                               long n) {
  GRID_LOOP(k, n) {
    y[k] = xi[k] - (-1.0f * xi_old[k]);
    s[k] = p[k] - p_old[k];
  }
}

__global__ __launch_bounds__(256) void dot_kernel(const float* __restrict__ a,
                                                  const float* __restrict__ b, long n,
                                                  double* __restrict__ partials) {
  __shared__ double sh[4];
  double acc = 0.0;
  GRID_LOOP(k, n) acc += (double)a[k] * (double)b[k];
  const double tot = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void gg_dgg_kernel(const float* __restrict__ xi,
                                                     const float* __restrict__ g, long n,
                                                     double* __restrict__ partials) {
  __shared__ double sh[4];
  double gg = 0.0, dgg = 0.0;
  GRID_LOOP(k, n) {
    const double gv = g[k], xv = xi[k];
    gg += gv * gv;
    dgg += (xv + gv) * xv;
  }
  const double t1 = block_sum(gg, sh);
  __syncthreads();
  const double t2 = block_sum(dgg, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = t1;
    partials[kReduceBlocks + blockIdx.x] = t2;
  }
}

__global__ __launch_bounds__(256) void absmax_kernel(const float* __restrict__ x, long n,
                                                     double* __restrict__ partials) {
  __shared__ double sh[4];
  double m = 0.0;
  GRID_LOOP(k, n) m = fmax(m, (double)fabsf(x[k]));
  const double r = block_max(m, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__global__ __launch_bounds__(256) void grad_condition_kernel(const float* __restrict__ xi,
                                                             const float* __restrict__ p, double den,
                                                             long n, double* __restrict__ partials) {
  __shared__ double sh[4];
  double m = 0.0;
  GRID_LOOP(k, n) m = fmax(m, (double)fabsf(xi[k]) * fmax((double)fabsf(p[k]), 1.0) / den);
  const double r = block_max(m, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// Device-side scalars of the L-BFGS two-loop recursion.  Each double
// operation is the one the host form performs, rounded once (no contraction),
// so the recursion gives the same coefficients on either side.
__global__ void rho_times_kernel(const double* __restrict__ rho_den,
                                 const double* __restrict__ dot,
                                 const double* __restrict__ minus_from, double* __restrict__ out) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double rho = *rho_den != 0.0 ? 1.0 / *rho_den : 0.0;
  const double r = rho * *dot;
  *out = minus_from ? *minus_from - r : r;
}
__global__ void h0_scale_kernel(const double* __restrict__ sy, const double* __restrict__ yy,
                                int nimg, double* __restrict__ out) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double acc = 0.0;
  for (int i = 0; i < nimg; i++)
    if (yy[i] != 0.0) acc += sy[i] / yy[i];
  *out = acc;
}
__global__ void axpy_dev_kernel(float* __restrict__ y, const float* __restrict__ x,
                                const double* __restrict__ coef, int negate, long n) {
  const float a = (float)(negate ? -*coef : *coef);
  GRID_LOOP(k, n) y[k] += a * x[k];
}
__global__ void scale_dev_kernel(float* y, const float* x, const double* __restrict__ coef,
                                 long n) {
  const float a = (float)*coef;
  GRID_LOOP(k, n) y[k] = x[k] * a;
}

#undef GRID_LOOP

}  // namespace gvm

using namespace gvm;

#define LAUNCH_1D(kernel, n, ...)                                                        \
  do {                                                                                   \
    hipLaunchKernelGGL(kernel, dim3(reduce_grid(n)), dim3(256), 0, ctx->stream, __VA_ARGS__); \
    GVM_CHECK_LAUNCH();                                                                  \
  } while (0)

static long img_elems(const gvm_ctx* ctx) { return (long)ctx->g.M * ctx->g.N; }

extern "C" gvm_status gvm_entropy_value(gvm_ctx* ctx, const float* I, int index, float prior,
                                        float eta, double* out) {
  return prior_value(ctx, PRIOR_ENTROPY, I, index, prior, eta, out);
}
extern "C" gvm_status gvm_entropy_grad(gvm_ctx* ctx, const float* I, float* dgi, int index,
                                       float prior, float eta, float lambda) {
  return prior_grad(ctx, PRIOR_ENTROPY, I, dgi, index, prior, eta, lambda);
}
extern "C" gvm_status gvm_entropy_value_enqueue(gvm_ctx* ctx, const float* I, int index,
                                                float prior, float eta, int slot) {
  return prior_value_enqueue(ctx, PRIOR_ENTROPY, I, index, prior, eta, slot);
}

extern "C" gvm_status gvm_tsv_value_enqueue(gvm_ctx* ctx, const float* I, int index, int slot) {
  return prior_value_enqueue(ctx, PRIOR_TSV, I, index, 0.f, 0.f, slot);
}

extern "C" gvm_status gvm_laplacian_value_enqueue(gvm_ctx* ctx, const float* I, int index,
                                                  int slot) {
  return prior_value_enqueue(ctx, PRIOR_LAPLACIAN, I, index, 0.f, 0.f, slot);
}

extern "C" gvm_status gvm_tsv_value(gvm_ctx* ctx, const float* I, int index, double* out) {
  return prior_value(ctx, PRIOR_TSV, I, index, 0.f, 0.f, out);
}
extern "C" gvm_status gvm_tsv_grad(gvm_ctx* ctx, const float* I, float* dgi, int index,
                                   float lambda) {
  return prior_grad(ctx, PRIOR_TSV, I, dgi, index, 0.f, 0.f, lambda);
}
extern "C" gvm_status gvm_laplacian_value(gvm_ctx* ctx, const float* I, int index, double* out) {
  return prior_value(ctx, PRIOR_LAPLACIAN, I, index, 0.f, 0.f, out);
}
extern "C" gvm_status gvm_laplacian_grad(gvm_ctx* ctx, const float* I, float* dgi, int index,
                                         float lambda) {
  return prior_grad(ctx, PRIOR_LAPLACIAN, I, dgi, index, 0.f, 0.f, lambda);
}

namespace gvm {
// The block partials of n terms in partial rows row0 .. row0 + n - 1.
gvm_status priors_value_partials(gvm_ctx* ctx, const float* I, const gvm_prior_term* terms,
                                 int n, int row0, int* nb_out) {
  PriorTermsDev T;
  gvm_status st = check_terms(ctx, terms, n, &T);
  if (st) return st;
  const int N = ctx->g.N;
  const int nb = reduce_grid((long)N * N);
  *nb_out = nb;
  if (n == 0) return GVM_OK;
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  hipLaunchKernelGGL(priors_value_kernel, dim3(nb), dim3(256), 0, ctx->stream, T, I, ctx->noise,
                     ctx->g.noise_cut, N, ctx->partials + (long)row0 * kReduceBlocks);
  GVM_CHECK_LAUNCH();
  return GVM_OK;
}

gvm_status priors_value_partials_finish(gvm_ctx* ctx, const float* I,
                                        const gvm_prior_term* terms, int n, int first_slot) {
  int nb = 0;
  gvm_status st = priors_value_partials(ctx, I, terms, n, 0, &nb);
  if (st || n == 0) return st;
  // one block per term: finish_kernel's order, into consecutive value slots
  return finish_reductions_to(ctx, n, nb, ctx->dvalues + first_slot);
}
}  // namespace gvm

extern "C" gvm_status gvm_priors_value_enqueue(gvm_ctx* ctx, const float* I,
                                               const gvm_prior_term* terms, int n,
                                               int first_slot) {
  GVM_REQUIRE(ctx && I, "gvm_priors_value_enqueue: null argument");
  GVM_REQUIRE(first_slot >= 0 && n >= 0 && first_slot + n <= kValueSlots,
              "gvm_priors_value_enqueue: value slots out of range");
  return priors_value_partials_finish(ctx, I, terms, n, first_slot);
}

extern "C" gvm_status gvm_dphi_assemble(gvm_ctx* ctx, const float* I, const float* chi2_grad,
                                        float* dphi, float* xi, const gvm_prior_term* terms,
                                        int n) {
  GVM_REQUIRE(ctx && I && dphi && xi, "gvm_dphi_assemble: null argument");
  PriorTermsDev T;
  gvm_status st = check_terms(ctx, terms, n, &T);
  if (st) return st;
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  const int N = ctx->g.N;
  const long total = (long)N * N * ctx->g.image_count;
  LAUNCH_1D(dphi_assemble_kernel, total, T, I, ctx->noise, ctx->g.noise_cut, N,
            ctx->g.image_count, chi2_grad, dphi, xi);
  return GVM_OK;
}

extern "C" gvm_status gvm_add_to_dphi(gvm_ctx* ctx, float* dphi, const float* dgi, int index) {
  GVM_REQUIRE(ctx && dphi && dgi, "gvm_add_to_dphi: null argument");
  GVM_REQUIRE(index >= 0 && index < ctx->g.image_count, "gvm_add_to_dphi: bad image index");
  const long n = img_elems(ctx);
  LAUNCH_1D(add_kernel, n, dphi + n * index, dgi, n);
  return GVM_OK;
}

extern "C" gvm_status gvm_vec_zero(gvm_ctx* ctx, float* x, long n) {
  GVM_REQUIRE(ctx && x && n >= 0, "gvm_vec_zero: bad argument");
  if (n == 0) return GVM_OK;
  LAUNCH_1D(fill_kernel, n, x, 0.0f, n);
  return GVM_OK;
}

extern "C" gvm_status gvm_vec_copy(gvm_ctx* ctx, float* dst, const float* src, long n) {
  GVM_REQUIRE(ctx && dst && src && n >= 0, "gvm_vec_copy: bad argument");
  if (n == 0 || dst == src) return GVM_OK;
  LAUNCH_1D(copy_kernel, n, dst, src, n);
  return GVM_OK;
}

extern "C" gvm_status gvm_cg_search_direction(gvm_ctx* ctx, float* g, float* xi, float* h,
                                              int image) {
  GVM_REQUIRE(ctx && g && xi && h, "gvm_cg_search_direction: null argument");
  GVM_REQUIRE(image >= 0 && image < ctx->g.image_count, "gvm_cg_search_direction: bad image");
  const long n = img_elems(ctx), o = n * image;
  LAUNCH_1D(search_direction_kernel, n, g + o, xi + o, h + o, n);
  return GVM_OK;
}

extern "C" gvm_status gvm_cg_new_xi(gvm_ctx* ctx, float* g, float* xi, float* h, double gam,
                                    int image) {
  GVM_REQUIRE(ctx && g && xi && h, "gvm_cg_new_xi: null argument");
  GVM_REQUIRE(image >= 0 && image < ctx->g.image_count, "gvm_cg_new_xi: bad image");
  const long n = img_elems(ctx), o = n * image;
  LAUNCH_1D(new_xi_kernel, n, g + o, xi + o, h + o, (float)gam, n);
  return GVM_OK;
}

extern "C" gvm_status gvm_cg_grad_condition(gvm_ctx* ctx, const float* xi, const float* p,
                                            double den, int nimg, double* gmax) {
  GVM_REQUIRE(ctx && xi && p && gmax && nimg > 0, "gvm_cg_grad_condition: bad argument");
  const long n = img_elems(ctx) * nimg;
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(grad_condition_kernel, dim3(nb), dim3(256), 0, ctx->stream, xi, p, den, n,
                     ctx->partials);
  GVM_CHECK_LAUNCH();
  gvm_status st = finish_reductions_mode(ctx, 1, nb, 1);
  if (st) return st;
  return read_scalars(ctx, 1, gmax);
}

extern "C" gvm_status gvm_cg_gg_dgg(gvm_ctx* ctx, const float* xi, const float* g, int nimg,
                                    double* gg, double* dgg) {
  GVM_REQUIRE(ctx && xi && g && gg && dgg && nimg > 0, "gvm_cg_gg_dgg: bad argument");
  const long n = img_elems(ctx) * nimg;
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(gg_dgg_kernel, dim3(nb), dim3(256), 0, ctx->stream, xi, g, n, ctx->partials);
  GVM_CHECK_LAUNCH();
  gvm_status st = finish_reductions(ctx, 2, nb);
  if (st) return st;
  double v[2];
  st = read_scalars(ctx, 2, v);
  if (st) return st;
  *gg = v[0];
  *dgg = v[1];
  return GVM_OK;
}

extern "C" gvm_status gvm_evaluate_xt(gvm_ctx* ctx, float* xt, const float* pcom,
                                      const float* xicom, double x, int image, int positivity,
                                      float minpix, float eta) {
  GVM_REQUIRE(ctx && xt && pcom && xicom, "gvm_evaluate_xt: null argument");
  GVM_REQUIRE(image >= 0 && image < ctx->g.image_count, "gvm_evaluate_xt: bad image");
  const long n = img_elems(ctx), o = n * image;
  const float floor_v = (float)(-1.0 * (double)eta * (double)minpix);
  LAUNCH_1D(evaluate_xt_kernel, n, xt + o, pcom + o, xicom + o, (float)x, positivity, floor_v, n);
  return GVM_OK;
}

extern "C" gvm_status gvm_new_p(gvm_ctx* ctx, float* p, float* xi, double xmin, int image,
                                int positivity, float minpix, float eta) {
  GVM_REQUIRE(ctx && p && xi, "gvm_new_p: null argument");
  GVM_REQUIRE(image >= 0 && image < ctx->g.image_count, "gvm_new_p: bad image");
  const long n = img_elems(ctx), o = n * image;
  const float floor_v = (float)(-1.0 * (double)eta * (double)minpix);
  LAUNCH_1D(new_p_kernel, n, p + o, xi + o, (float)xmin, positivity, floor_v, n);
  return GVM_OK;
}

extern "C" gvm_status gvm_vec_absmax(gvm_ctx* ctx, const float* x, long n, double* out) {
  GVM_REQUIRE(ctx && x && out && n > 0, "gvm_vec_absmax: bad argument");
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(absmax_kernel, dim3(nb), dim3(256), 0, ctx->stream, x, n, ctx->partials);
  GVM_CHECK_LAUNCH();
  gvm_status st = finish_reductions_mode(ctx, 1, nb, 1);
  if (st) return st;
  return read_scalars(ctx, 1, out);
}

extern "C" gvm_status gvm_vec_scale(gvm_ctx* ctx, float* dst, const float* src, double scale,
                                    long n) {
  GVM_REQUIRE(ctx && dst && src && n >= 0, "gvm_vec_scale: bad argument");
  if (n == 0) return GVM_OK;
  LAUNCH_1D(scale_kernel, n, dst, src, (float)scale, n);
  return GVM_OK;
}

extern "C" gvm_status gvm_vec_dot(gvm_ctx* ctx, const float* a, const float* b, long n,
                                  double* out) {
  GVM_REQUIRE(ctx && a && b && out && n > 0, "gvm_vec_dot: bad argument");
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(dot_kernel, dim3(nb), dim3(256), 0, ctx->stream, a, b, n, ctx->partials);
  GVM_CHECK_LAUNCH();
  gvm_status st = finish_reductions(ctx, 1, nb);
  if (st) return st;
  return read_scalars(ctx, 1, out);
}

extern "C" gvm_status gvm_vec_axpy(gvm_ctx* ctx, float* y, const float* x, double alpha, long n) {
  GVM_REQUIRE(ctx && y && x && n >= 0, "gvm_vec_axpy: bad argument");
  if (n == 0) return GVM_OK;
  LAUNCH_1D(axpy_kernel, n, y, x, (float)alpha, n);
  return GVM_OK;
}

extern "C" gvm_status gvm_vec_dot_dev(gvm_ctx* ctx, const float* a, const float* b, long n,
                                      double* out_dev) {
  GVM_REQUIRE(ctx && a && b && out_dev && n > 0, "gvm_vec_dot_dev: bad argument");
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(dot_kernel, dim3(nb), dim3(256), 0, ctx->stream, a, b, n, ctx->partials);
  GVM_CHECK_LAUNCH();
  return finish_reductions_to(ctx, 1, nb, out_dev);
}

// ---------------------------------------------------------------------------
// gvm_vec_dot_dev without a context and without allocation.  The block
// partials live in `scratch`, a static device array of kReduceBlocks doubles
// (8 KiB) that every GPU running this library holds until the process ends.
// Host code never passes `scratch` itself to a kernel: on the host the name
// is only a stand-in, so its device address on the current GPU is looked up
// once per device (hipGetSymbolAddress) and kept in g_scratch_ptr.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[kReduceBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

gvm_status scratch_on_current_device(double** out) {
  int dev = 0;
  GVM_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("gvm_vec_dot_dev_noalloc: device index beyond the library's table");
    return GVM_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("gvm_vec_dot_dev_noalloc: hipGetSymbolAddress(scratch): ") +
                hipGetErrorString(e));
      return GVM_ERR_HIP;
    }
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return GVM_OK;
}
}  // namespace

extern "C" gvm_status gvm_vec_dot_dev_noalloc(const float* a, const float* b, long n,
                                              double* out_dev, void* stream) {
  GVM_REQUIRE(a && b && out_dev, "gvm_vec_dot_dev_noalloc: null argument");
  GVM_REQUIRE(n > 0 && n <= GVM_NOALLOC_DOT_MAX_N,
              "gvm_vec_dot_dev_noalloc: n must be in [1, GVM_NOALLOC_DOT_MAX_N]");
  double* part = nullptr;
  gvm_status st = scratch_on_current_device(&part);
  if (st) return st;
  const hipStream_t s = (hipStream_t)stream;
  // the kernels and launch shapes of gvm_vec_dot_dev: the same sum, bit for bit
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(dot_kernel, dim3(nb), dim3(256), 0, s, a, b, n, part);
  GVM_CHECK_LAUNCH();
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(kReduceThreads), 0, s, part, out_dev, nb, 0);
  GVM_CHECK_LAUNCH();
  return GVM_OK;
}

extern "C" gvm_status gvm_scalar_rho_times(gvm_ctx* ctx, const double* rho_den, const double* dot,
                                          const double* minus_from, double* out) {
  GVM_REQUIRE(ctx && rho_den && dot && out, "gvm_scalar_rho_times: null argument");
  hipLaunchKernelGGL(rho_times_kernel, dim3(1), dim3(64), 0, ctx->stream, rho_den, dot,
                     minus_from, out);
  GVM_CHECK_LAUNCH();
  return GVM_OK;
}

extern "C" gvm_status gvm_scalar_h0(gvm_ctx* ctx, const double* sy, const double* yy, int nimg,
                                   double* out) {
  GVM_REQUIRE(ctx && sy && yy && out && nimg >= 1, "gvm_scalar_h0: bad argument");
  hipLaunchKernelGGL(h0_scale_kernel, dim3(1), dim3(64), 0, ctx->stream, sy, yy, nimg, out);
  GVM_CHECK_LAUNCH();
  return GVM_OK;
}

extern "C" gvm_status gvm_vec_axpy_dev(gvm_ctx* ctx, float* y, const float* x, const double* coef,
                                      int negate, long n) {
  GVM_REQUIRE(ctx && y && x && coef && n >= 0, "gvm_vec_axpy_dev: bad argument");
  if (n == 0) return GVM_OK;
  LAUNCH_1D(axpy_dev_kernel, n, y, x, coef, negate, n);
  return GVM_OK;
}

extern "C" gvm_status gvm_vec_scale_dev(gvm_ctx* ctx, float* dst, const float* src,
                                       const double* coef, long n) {
  GVM_REQUIRE(ctx && dst && src && coef && n >= 0, "gvm_vec_scale_dev: bad argument");
  if (n == 0) return GVM_OK;
  LAUNCH_1D(scale_dev_kernel, n, dst, src, coef, n);
  return GVM_OK;
}

extern "C" gvm_status gvm_lbfgs_s_and_y(gvm_ctx* ctx, float* y, float* s, const float* p,
                                        const float* xi, const float* p_old, const float* xi_old,
                                        long n) {
  GVM_REQUIRE(ctx && y && s && p && xi && p_old && xi_old && n > 0,
              "gvm_lbfgs_s_and_y: bad argument");
  LAUNCH_1D(s_and_y_kernel, n, y, s, p, xi, p_old, xi_old, n);
  return GVM_OK;
}
