// Set-up of the hot path's inputs on the device (FP64), for observations
// whose visibilities are already in device memory:
//
//  * gvm_add_visibilities_dev: hermitianSymmetry (src/functions.cu:2256-2273:
//    rows with u > 0 get u, v and Im Vo negated; metres to wavelengths with
//    the float wavelength of src/MSFITSIO.cu:36-45), then the bilinear
//    stencil and off-grid weight rule of vis_mod (src/functions.cu:2566-2606)
//    and the exact-turn phase split of the adjoint, one thread per
//    visibility; the primary beam of the slice (attenuation,
//    src/functions.cu:2304-2333) one thread per pixel.  The operations are
//    the host set-up's, in the same order, so the slice arrays are identical
//    to the ones gvm_add_visibilities builds from the same rows.
//  * gvm_setup_beam_noise_dev: the weighted uv moments of
//    calculateNoiseAndBeam (src/functions.cu:1700-1814) as FP64 block sums.
//  * gvm_setup_noise_image_dev: the noise image, fg_scale and noise_cut of
//    src/mfs.cu:849-916.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "gvm_internal.h"

namespace gvm {

namespace {
constexpr double kPi = 3.14159265358979323846;
constexpr double kRad = kPi / 180.0;
constexpr double kRZ = 1.2196698912665045;  // first zero of J1 over pi
constexpr int kThreads = 256;

// attenuation(), as the host set-up computes it
__device__ double attenuation_dev(int kind, double diameter, double pb_factor, double pb_cutoff,
                                  double freq, int i, int j, float xobs, float yobs, double dx,
                                  double dy) {
  const int x0 = (int)xobs, y0 = (int)yobs;
  const double x = (j - x0) * dx * kRad;
  const double y = (i - y0) * dy * kRad;
  const double arc = sqrt(x * x + y * y);
  const double lambda = 2.99792458e8 / freq;
  double atten;
  if (kind == GVM_BEAM_GAUSSIAN) {
    const double fwhm = pb_factor * lambda / diameter;
    const double r = arc / fwhm;
    atten = exp(-4.0 * log(2.0) * r * r);
  } else {
    atten = 1.0;
    if (arc != 0.0) {
      const double z = kPi * arc * diameter / lambda * (kRZ / pb_factor);
      const double b = j1(z);
      atten = 4.0 * (b / z) * (b / z);
    }
  }
  return arc <= pb_cutoff ? atten : 0.0;
}

__global__ void beam_kernel(int kind, double diameter, double pb_factor, double pb_cutoff,
                            double freq, int M, int N, float xobs, float yobs, double dx,
                            double dy, float* __restrict__ out) {
  const long n = (long)M * N;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    const int i = (int)(k / M), j = (int)(k % M);
    out[k] = (float)attenuation_dev(kind, diameter, pb_factor, pb_cutoff, freq, i, j, xobs, yobs,
                                    dx, dy);
  }
}

__global__ void setup_vis_kernel(const double* __restrict__ uvw_m, const float* __restrict__ vo_in,
                                 const float* __restrict__ w_in, long nvis, double lambda,
                                 double deltau, double deltav, double dxr, double dyr,
                                 double hscale, long N, int* __restrict__ cell,
                                 float2* __restrict__ frac, float2* __restrict__ vo,
                                 float* __restrict__ weight, float4* __restrict__ phs,
                                 float* __restrict__ wcoord) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nvis;
       i += (long)gridDim.x * blockDim.x) {
    double u = uvw_m[3 * i], v = uvw_m[3 * i + 1], w = uvw_m[3 * i + 2];
    const float re = vo_in[2 * i];
    float im = vo_in[2 * i + 1];
    if (u > 0.0) {
      u *= -1.0;
      v *= -1.0;
      im *= -1.0f;
    }
    u /= lambda;
    v /= lambda;
    w /= lambda;
    double ux = u / deltau, uy = v / deltav;
    if (ux < 0.0) ux += N;
    if (uy < 0.0) uy += N;
    const long i1 = (long)floor(ux), j1 = (long)floor(uy);
    vo[i] = make_float2(re, im);
    if (i1 >= 0 && i1 < N && j1 >= 0 && j1 < N) {
      cell[i] = (int)(i1 | (j1 << 16));
      frac[i] = make_float2((float)(ux - i1), (float)(uy - j1));
      weight[i] = w_in[i];
      const double a = u * dxr, b = v * dyr;
      const double ah = rint(a * hscale) / hscale;
      const double bh = rint(b * hscale) / hscale;
      phs[i] = make_float4((float)ah, (float)(a - ah), (float)bh, (float)(b - bh));
    } else {
      cell[i] = -1;
      frac[i] = make_float2(0.f, 0.f);
      weight[i] = 0.0f;
      phs[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    wcoord[i] = (float)w;
  }
}

// per block: sum w u^2, w v^2, w u v, w (u, v in wavelengths)
__global__ void moments_kernel(const double* __restrict__ uvw_m, const float* __restrict__ w,
                               long nvis, double lambda, double* __restrict__ partials) {
  __shared__ double sh[kThreads / 64];
  double suu = 0.0, svv = 0.0, suv = 0.0, sw = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nvis;
       i += (long)gridDim.x * blockDim.x) {
    const double u = uvw_m[3 * i] / lambda, v = uvw_m[3 * i + 1] / lambda;
    const double wt = w[i];
    suu += u * u * wt;
    svv += v * v * wt;
    suv += u * v * wt;
    sw += wt;
  }
  double r[4] = {suu, svv, suv, sw};
  for (int q = 0; q < 4; q++) {
    const double t = block_sum(r[q], sh);
    __syncthreads();
    if (threadIdx.x == 0) partials[4 * blockIdx.x + q] = t;
  }
}

// per block: max of PB^2
__global__ void pb2_max_kernel(const float* __restrict__ pb, long n, double* __restrict__ partials) {
  __shared__ double sh[kThreads / 64];
  double m = 0.0;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x)
    m = fmax(m, (double)pb[k] * pb[k]);
  const double t = block_max(m, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// noise_image (src/functions.cu:2382-2422) and the per-block minimum
__global__ void noise_kernel(const float* __restrict__ pb, long n, double wmax, double noise_jypix,
                             float* __restrict__ noise, double* __restrict__ partials) {
  __shared__ double sh[kThreads / 64];
  double m = INFINITY;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    const double w = (double)pb[k] * pb[k];
    const double nw = (w / wmax) / (noise_jypix * noise_jypix);
    const double v = sqrt(1.0 / nw);
    noise[k] = (float)v;
    m = fmin(m, v);
  }
  const double t = -block_max(-m, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

int blocks_for(long n) { return (int)std::max(1L, std::min<long>((n + kThreads - 1) / kThreads, 1024)); }

}  // namespace

gvm_status primary_beam_dev(const gvm_geometry& g, double freq, float* out, hipStream_t stream) {
  const long n = (long)g.M * g.N;
  hipLaunchKernelGGL(beam_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, stream, g.beam_kind,
                     g.antenna_diameter, g.pb_factor, g.pb_cutoff, freq, g.M, g.N, g.x0_ref,
                     g.y0_ref, g.deltax, g.deltay, out);
  GVM_CHECK_LAUNCH();
  return GVM_OK;
}

}  // namespace gvm

using namespace gvm;

extern "C" int gvm_setup_beam_from_moments(double suu, double svv, double suv, double sw,
                                           double deltax_deg, double* bmaj_deg, double* bmin_deg,
                                           double* bpa_deg, double* sum_weights,
                                           double* vis_noise, double* noise_jypix) {
  GVM_REQUIRE(bmaj_deg && bmin_deg && bpa_deg && sum_weights && vis_noise && noise_jypix,
              "gvm_setup_beam_from_moments: null argument");
  if (!(sw > 0.0)) {
    set_error("gvm_setup_beam_from_moments: the sum of the visibility weights cannot be zero");
    return GVM_ERR_ARG;
  }
  suu /= sw;
  svv /= sw;
  suv /= sw;
  const double d = suu - svv, s = suu + svv;
  const double root = std::sqrt(d * d + 4.0 * suv * suv);
  const double bmaj = 1.0 / std::sqrt(2.0) / kPi / std::sqrt(s - root);
  const double bmin = 1.0 / std::sqrt(2.0) / kPi / std::sqrt(s + root);
  *bmaj_deg = bmaj / kRad;
  *bmin_deg = bmin / kRad;
  *bpa_deg = -0.5 * std::atan2(2.0 * suv, d) / kRad;
  *sum_weights = sw;
  *vis_noise = 0.5 * std::sqrt(1.0 / sw);
  const double bmaj_pix = *bmaj_deg / std::fabs(deltax_deg);
  const double bmin_pix = *bmin_deg / std::fabs(deltax_deg);
  *noise_jypix = *vis_noise / (kPi * bmaj_pix * bmin_pix / (4.0 * std::log(2.0)));
  return GVM_OK;
}

extern "C" gvm_status gvm_add_visibilities_dev(gvm_ctx* ctx, float freq, long nvis,
                                               const double* uvw_m_dev, const float* vo_dev,
                                               const float* weight_dev, int* slice_out) {
  GVM_REQUIRE(ctx && uvw_m_dev && vo_dev && weight_dev, "gvm_add_visibilities_dev: null argument");
  GVM_REQUIRE(nvis > 0 && nvis < (1L << 31), "gvm_add_visibilities_dev: nvis out of range");
  GVM_REQUIRE(freq > 0.f, "gvm_add_visibilities_dev: frequency must be positive");
  GVM_REQUIRE((int)ctx->slices.size() < kMaxSlots - 8, "gvm_add_visibilities_dev: too many slices");
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  const gvm_geometry& g = ctx->g;
  const long N = g.N;
  const double dxr = g.deltax * kRad, dyr = g.deltay * kRad;
  const double deltau = 1.0 / (g.M * dxr), deltav = 1.0 / (g.N * dyr);
  int lg = 0;
  while ((1L << lg) < N) lg++;
  const double hscale = std::ldexp(1.0, 23 - lg);
  const float lambda = 2.99792458E8f / freq;

  Slice s;
  s.freq = freq;
  s.nvis = nvis;
  auto alloc = [&](auto** p, size_t count) {
    return hipMalloc((void**)p, sizeof(**p) * count) == hipSuccess;
  };
  if (!alloc(&s.cell, nvis) || !alloc(&s.frac, nvis) || !alloc(&s.vo, nvis) ||
      !alloc(&s.weight, nvis) || !alloc(&s.phs, nvis) || !alloc(&s.wcoord, nvis) ||
      !alloc(&s.pb, (size_t)g.M * g.N) || !alloc(&s.vr, nvis) || !alloc(&s.cpol, nvis)) {
    free_slice(s);
    set_error("gvm_add_visibilities_dev: out of device memory");
    return GVM_ERR_OOM;
  }
  hipLaunchKernelGGL(setup_vis_kernel, dim3(blocks_for(nvis)), dim3(kThreads), 0, nullptr,
                     uvw_m_dev, vo_dev, weight_dev, nvis, (double)lambda, deltau, deltav, dxr, dyr,
                     hscale, N, s.cell, s.frac, s.vo,
                     s.weight, s.phs,
                     s.wcoord);
  if (hipGetLastError() != hipSuccess) {
    free_slice(s);
    set_error("gvm_add_visibilities_dev: set-up kernel launch failed");
    return GVM_ERR_HIP;
  }
  gvm_status st = primary_beam_dev(g, freq, s.pb, nullptr);
  if (st) {
    free_slice(s);
    return st;
  }
  st = register_slice(ctx, s);  // synchronises the device
  if (st) return st;
  if (slice_out) *slice_out = (int)ctx->slices.size() - 1;
  return GVM_OK;
}

extern "C" int gvm_setup_beam_noise_dev(int nch, const float* freq, const long* nvis,
                                        const double* const* uvw_m_dev,
                                        const float* const* weight_dev, double deltax_deg,
                                        double* bmaj_deg, double* bmin_deg, double* bpa_deg,
                                        double* sum_weights, double* vis_noise,
                                        double* noise_jypix) {
  GVM_REQUIRE(nch > 0 && freq && nvis && uvw_m_dev && weight_dev && bmaj_deg && bmin_deg &&
                  bpa_deg && sum_weights && vis_noise && noise_jypix,
              "gvm_setup_beam_noise_dev: null argument");
  double* part = nullptr;
  GVM_HIP_TRY(hipMalloc((void**)&part, sizeof(double) * 4 * 1024));
  std::vector<double> hp(4 * 1024);
  double suu = 0, svv = 0, suv = 0, sw = 0;
  for (int c = 0; c < nch; c++) {
    const float lambda = 2.99792458E8f / freq[c];
    const int nb = blocks_for(nvis[c]);
    hipLaunchKernelGGL(moments_kernel, dim3(nb), dim3(kThreads), 0, nullptr, uvw_m_dev[c],
                       weight_dev[c], nvis[c], (double)lambda, part);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(hp.data(), part, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost) != hipSuccess) {
      (void)hipFree(part);
      set_error("gvm_setup_beam_noise_dev: moment sums failed");
      return GVM_ERR_HIP;
    }
    for (int b = 0; b < nb; b++) {
      suu += hp[4 * b];
      svv += hp[4 * b + 1];
      suv += hp[4 * b + 2];
      sw += hp[4 * b + 3];
    }
  }
  (void)hipFree(part);
  if (!(sw > 0.0)) {
    set_error("gvm_setup_beam_noise_dev: the sum of the visibility weights cannot be zero");
    return GVM_ERR_ARG;
  }
  return gvm_setup_beam_from_moments(suu, svv, suv, sw, deltax_deg, bmaj_deg, bmin_deg, bpa_deg,
                                     sum_weights, vis_noise, noise_jypix);
}

extern "C" int gvm_setup_noise_image_dev(int beam_kind, double diameter, double pb_factor,
                                         double pb_cutoff, double nu_0, int M, int N,
                                         double deltax_deg, double deltay_deg, float x0_ref,
                                         float y0_ref, double noise_jypix,
                                         double noise_cut_factor, float* noise_out,
                                         float* fg_scale, float* noise_cut) {
  GVM_REQUIRE(noise_out && fg_scale && noise_cut && M > 0 && N > 0,
              "gvm_setup_noise_image_dev: bad argument");
  GVM_REQUIRE(beam_kind == GVM_BEAM_GAUSSIAN || beam_kind == GVM_BEAM_AIRYDISK,
              "gvm_setup_noise_image_dev: unknown beam kind");
  const long n = (long)M * N;
  gvm_geometry g{};
  g.M = M;
  g.N = N;
  g.beam_kind = beam_kind;
  g.antenna_diameter = diameter;
  g.pb_factor = pb_factor;
  g.pb_cutoff = pb_cutoff;
  g.x0_ref = x0_ref;
  g.y0_ref = y0_ref;
  g.deltax = deltax_deg;
  g.deltay = deltay_deg;
  float *pb = nullptr, *noise = nullptr;
  double* part = nullptr;
  auto cleanup = [&]() {
    (void)hipFree(pb);
    (void)hipFree(noise);
    (void)hipFree(part);
  };
  if (hipMalloc((void**)&pb, sizeof(float) * n) != hipSuccess ||
      hipMalloc((void**)&noise, sizeof(float) * n) != hipSuccess ||
      hipMalloc((void**)&part, sizeof(double) * 1024) != hipSuccess) {
    cleanup();
    set_error("gvm_setup_noise_image_dev: out of device memory");
    return GVM_ERR_OOM;
  }
  gvm_status st = primary_beam_dev(g, nu_0, pb, nullptr);
  if (st) {
    cleanup();
    return st;
  }
  const int nb = blocks_for(n);
  std::vector<double> hp(nb);
  hipLaunchKernelGGL(pb2_max_kernel, dim3(nb), dim3(kThreads), 0, nullptr, pb, n, part);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpy(hp.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost) != hipSuccess) {
    cleanup();
    set_error("gvm_setup_noise_image_dev: beam maximum failed");
    return GVM_ERR_HIP;
  }
  double wmax = 0.0;
  for (double v : hp) wmax = std::max(wmax, v);
  if (!(wmax > 0.0)) {
    cleanup();
    set_error("gvm_setup_noise_image_dev: the primary beam is zero everywhere");
    return GVM_ERR_ARG;
  }
  hipLaunchKernelGGL(noise_kernel, dim3(nb), dim3(kThreads), 0, nullptr, pb, n, wmax, noise_jypix,
                     noise, part);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpy(hp.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(noise_out, noise, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess) {
    cleanup();
    set_error("gvm_setup_noise_image_dev: noise image failed");
    return GVM_ERR_HIP;
  }
  cleanup();
  double nmin = std::numeric_limits<double>::infinity();
  for (double v : hp) nmin = std::min(nmin, v);
  *fg_scale = (float)nmin;
  *noise_cut = (float)(noise_cut_factor * nmin);
  return GVM_OK;
}
