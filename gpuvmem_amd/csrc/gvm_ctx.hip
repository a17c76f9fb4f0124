// Context, set-up of the visibility slices, device memory helpers and the
// reduction plumbing of libgvm_hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gvm_internal.h"

namespace gvm {

namespace {
thread_local std::string g_last_error;
constexpr double kPi = 3.14159265358979323846;
constexpr double kRad = kPi / 180.0;
constexpr double kRZ = 1.2196698912665045;  // first zero of J1 over pi

bool is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

template <typename T>
gvm_status dev_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  GVM_HIP_TRY(hipMalloc((void**)p, sizeof(T) * count));
  return GVM_OK;
}

template <typename T>
gvm_status upload(T** p, const std::vector<T>& v) {
  gvm_status st = dev_alloc(p, v.size());
  if (st != GVM_OK) return st;
  if (!v.empty()) GVM_HIP_TRY(hipMemcpy(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return GVM_OK;
}

// sin/cos of 2 pi t, argument reduced in turns first (FP64).
void sincos_turns(double t, double* s, double* c) {
  double r = t - std::nearbyint(t);
  *s = std::sin(2.0 * kPi * r);
  *c = std::cos(2.0 * kPi * r);
}

// attenuation(): reference src/functions.cu:2304-2333 (centre truncated to
// int, distance in radians, zero beyond pb_cutoff), GaussianBeam
// (:2293-2302), AiryDiskBeam (:2275-2291).  FP64.
double attenuation(int kind, double diameter, double pb_factor, double pb_cutoff,
                   double freq, int i, int j, float xobs, float yobs, double dx,
                   double dy) {
  const int x0 = (int)xobs, y0 = (int)yobs;
  const double x = (j - x0) * dx * kRad;
  const double y = (i - y0) * dy * kRad;
  const double arc = std::sqrt(x * x + y * y);
  const double lambda = 2.99792458e8 / freq;
  double atten;
  if (kind == GVM_BEAM_GAUSSIAN) {
    const double fwhm = pb_factor * lambda / diameter;
    const double r = arc / fwhm;
    atten = std::exp(-4.0 * std::log(2.0) * r * r);
  } else {
    atten = 1.0;
    if (arc != 0.0) {
      const double z = kPi * arc * diameter / lambda * (kRZ / pb_factor);
      const double j1 = ::jn(1, z);
      atten = 4.0 * (j1 / z) * (j1 / z);
    }
  }
  return arc <= pb_cutoff ? atten : 0.0;
}

}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error_cstr() { return g_last_error.c_str(); }

int reduce_grid(long n) {
  long b = (n + kReduceThreads - 1) / kReduceThreads;
  return (int)std::max(1L, std::min<long>(b, kReduceBlocks));
}

// Sums (or takes the max of) nblocks partials per slot, in order.
__global__ void finish_kernel(const double* __restrict__ partials,
                              double* __restrict__ out, int nblocks, int is_max) {
  __shared__ double sh[kReduceThreads / 64];
  const int slot = blockIdx.x;
  const double* p = partials + (long)slot * kReduceBlocks;
  double v = is_max ? 0.0 : 0.0;
  // fixed assignment of partials to threads -> deterministic
  for (int i = threadIdx.x; i < nblocks; i += blockDim.x) v = is_max ? fmax(v, p[i]) : v + p[i];
  double r = is_max ? block_max(v, sh) : block_sum(v, sh);
  if (threadIdx.x == 0) out[slot] = r;
}

gvm_status finish_reductions_mode(gvm_ctx* ctx, int nslots, int nblocks, int is_max) {
  hipLaunchKernelGGL(finish_kernel, dim3(nslots), dim3(kReduceThreads), 0, ctx->stream,
                     ctx->partials, ctx->dscalars, nblocks, is_max);
  GVM_CHECK_LAUNCH();
  return GVM_OK;
}

gvm_status finish_reductions_to(gvm_ctx* ctx, int nslots, int nblocks, double* out) {
  hipLaunchKernelGGL(finish_kernel, dim3(nslots), dim3(kReduceThreads), 0, ctx->stream,
                     ctx->partials, out, nblocks, 0);
  GVM_CHECK_LAUNCH();
  return GVM_OK;
}

// The kernels write the value slots in pinned host memory directly: nothing
// to copy, the values are there once the stream has reached this point.
gvm_status values_fetch_enqueue(gvm_ctx* ctx, int n) {
  (void)ctx;
  (void)n;
  return GVM_OK;
}

gvm_status values_host(gvm_ctx* ctx, int n, double* out) {
  GVM_REQUIRE(!ctx->capturing, "values: cannot wait for the stream while capturing");
  GVM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (int s = 0; s < n; s++) out[s] = ctx->hvalues[s];
  return GVM_OK;
}

gvm_status values_read(gvm_ctx* ctx, int n, double* out) {
  GVM_REQUIRE(!ctx->capturing, "values: cannot wait for the stream while capturing");
  gvm_status st = values_fetch_enqueue(ctx, n);
  if (st) return st;
  return values_host(ctx, n, out);
}

gvm_status finish_reductions(gvm_ctx* ctx, int nslots, int nblocks) {
  return finish_reductions_mode(ctx, nslots, nblocks, 0);
}

// Frees the device arrays of a slice.
void free_slice(Slice& s) {
  (void)hipFree(s.cell);
  (void)hipFree(s.frac);
  (void)hipFree(s.vo);
  (void)hipFree(s.weight);
  (void)hipFree(s.phs);
  (void)hipFree(s.wcoord);
  (void)hipFree(s.pb);
  (void)hipFree(s.vr);
  (void)hipFree(s.cpol);
  s = Slice{};
}

// Makes s (device arrays filled, vr / cpol allocated) the next slice of ctx:
// clears the per-evaluation arrays, grows the per-slice uv grids, writes the
// slice's device descriptor.  On failure the slice's arrays are freed.
gvm_status register_slice(gvm_ctx* ctx, Slice& s) {
  const gvm_geometry& g = ctx->g;
  auto fail = [&](gvm_status st) {
    free_slice(s);
    return st;
  };
  if (hipMemset(s.vr, 0, sizeof(float2) * s.nvis) != hipSuccess ||
      hipMemset(s.cpol, 0, sizeof(float2) * s.nvis) != hipSuccess) {
    set_error("register_slice: hipMemset failed");
    return fail(GVM_ERR_HIP);
  }
  // one uv grid per slice for the batched forward pass
  const size_t need = ctx->slices.size() + 1;
  if (need > ctx->grid_slices) {
    const size_t cap = std::max(need, 2 * ctx->grid_slices);
    float2* grid = nullptr;
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMalloc((void**)&grid, sizeof(float2) * (size_t)g.M * g.N * cap) != hipSuccess) {
      set_error("gvm_add_visibilities: out of device memory for the uv grids");
      return fail(GVM_ERR_OOM);
    }
    (void)hipFree(ctx->grid);
    ctx->grid = grid;
    ctx->grid_slices = cap;
  }
  SliceDev d;
  d.cell = s.cell;
  d.frac = s.frac;
  d.vo = s.vo;
  d.weight = s.weight;
  d.vr = s.vr;
  d.cpol = s.cpol;
  d.phs = s.phs;
  d.wcoord = s.wcoord;
  d.pb = s.pb;
  d.nvis = s.nvis;
  d.nudiv = s.freq / g.nu_0;
  d.pad = 0.f;
  if (hipMemcpy(ctx->slice_tab + ctx->slices.size(), &d, sizeof(d), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipDeviceSynchronize() != hipSuccess) {
    set_error("register_slice: descriptor upload failed");
    return fail(GVM_ERR_HIP);
  }
  ctx->slices.push_back(s);
  ctx->split_mode = -1;
  ctx->have_residuals = false;
  ctx->epoch++;
  return GVM_OK;
}

gvm_status read_scalars(gvm_ctx* ctx, int nslots, double* out) {
  GVM_REQUIRE(!ctx->capturing, "cannot wait for the stream while capturing");
  GVM_HIP_TRY(hipMemcpyAsync(ctx->hscalars, ctx->dscalars, sizeof(double) * nslots,
                             hipMemcpyDeviceToHost, ctx->stream));
  GVM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (int s = 0; s < nslots; s++) out[s] = ctx->hscalars[s];
  return GVM_OK;
}

}  // namespace gvm

using namespace gvm;

// ---------------------------------------------------------------------------
// runtime helpers
// ---------------------------------------------------------------------------
extern "C" const char* gvm_last_error(void) { return last_error_cstr(); }

extern "C" gvm_status gvm_device_info(char* name, int name_len, int* cu_count,
                                      int* clock_khz) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    set_error("no HIP device");
    return GVM_ERR_NODEV;
  }
  hipDeviceProp_t prop;
  GVM_HIP_TRY(hipGetDeviceProperties(&prop, 0));
  if (name && name_len > 0) {
    std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (clock_khz) *clock_khz = prop.clockRate;
  return GVM_OK;
}

extern "C" gvm_status gvm_malloc(void** ptr, size_t bytes) {
  GVM_REQUIRE(ptr != nullptr, "gvm_malloc: null ptr");
  GVM_HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1));
  return GVM_OK;
}

extern "C" gvm_status gvm_free(void* ptr) {
  if (ptr) GVM_HIP_TRY(hipFree(ptr));
  return GVM_OK;
}

// The helpers synchronise the whole device first: contexts run on their own
// non-blocking streams, which the null stream does not wait for.
extern "C" gvm_status gvm_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  GVM_HIP_TRY(hipDeviceSynchronize());
  GVM_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return GVM_OK;
}

extern "C" gvm_status gvm_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  GVM_HIP_TRY(hipDeviceSynchronize());
  GVM_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return GVM_OK;
}

extern "C" gvm_status gvm_sync(gvm_ctx* ctx) {
  GVM_REQUIRE(ctx, "gvm_sync: null context");
  GVM_REQUIRE(!ctx->capturing, "gvm_sync: cannot wait for the stream while capturing");
  GVM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return GVM_OK;
}

extern "C" gvm_status gvm_primary_beam(int beam_kind, double antenna_diameter,
                                       double pb_factor, double pb_cutoff,
                                       double freq, int M, int N, float xobs,
                                       float yobs, double deltax, double deltay,
                                       float* out) {
  GVM_REQUIRE(out && M > 0 && N > 0 && freq > 0, "gvm_primary_beam: bad argument");
  GVM_REQUIRE(beam_kind == GVM_BEAM_GAUSSIAN || beam_kind == GVM_BEAM_AIRYDISK,
              "gvm_primary_beam: unknown beam kind");
  for (int i = 0; i < N; i++)
    for (int j = 0; j < M; j++)
      out[(long)N * i + j] = (float)attenuation(beam_kind, antenna_diameter, pb_factor,
                                                pb_cutoff, freq, i, j, xobs, yobs, deltax,
                                                deltay);
  return GVM_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

extern "C" gvm_status gvm_destroy(gvm_ctx* ctx) {
  if (!ctx) return GVM_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& s : ctx->slices) free_slice(s);
  (void)hipFree(ctx->noise);
  (void)hipFree(ctx->mask_bits);
  if (ctx->gcf) (void)hipFree(ctx->gcf);
  (void)hipFree(ctx->grid);
  (void)hipFree(ctx->twiddle);
  (void)hipFree(ctx->rot_x);
  (void)hipFree(ctx->rot_y);
  (void)hipFree(ctx->xtab);
  (void)hipFree(ctx->ytab);
  (void)hipFree(ctx->partials);
  (void)hipFree(ctx->dscalars);
  (void)hipFree(ctx->cmax);
  (void)hipFree(ctx->adj_part);
  (void)hipFree(ctx->slice_tab);
  (void)hipFree(ctx->split_tab);
  (void)hipFree(ctx->slice_p0);
  if (ctx->hscalars) (void)hipHostFree(ctx->hscalars);
  if (ctx->dvalues) (void)hipHostFree(ctx->dvalues);
  (void)hipFree(ctx->nn_part);
  for (hipEvent_t& e : ctx->tev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return GVM_OK;
}

static gvm_status create_impl(const gvm_geometry* g, const float* noise_host, gvm_ctx* ctx) {
  const int M = g->M, N = g->N;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("gvm_create: no HIP device");
    return GVM_ERR_NODEV;
  }
  GVM_HIP_TRY(hipGetDevice(&ctx->device));
  hipDeviceProp_t prop;
  GVM_HIP_TRY(hipGetDeviceProperties(&prop, ctx->device));
  if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0) {
    set_error(std::string("gvm_create: device is ") + prop.gcnArchName + ", the library is built for gfx950");
    return GVM_ERR_NODEV;
  }
  ctx->cu_count = prop.multiProcessorCount;
  ctx->split_vis = default_split_vis(N);
  GVM_HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  for (hipEvent_t& e : ctx->tev) GVM_HIP_TRY(hipEventCreate(&e));

  const long n = (long)M * N;
  std::vector<float> noise(noise_host, noise_host + n);
  gvm_status st = upload(&ctx->noise, noise);
  if (st) return st;
  // the mask (noise < noise_cut) as bits, pixel k in bit k % 32 of word k / 32:
  // the adjoint writes its partial tiles on the mask only
  std::vector<unsigned int> bits((size_t)((n + 31) / 32), 0u);
  for (long k = 0; k < n; k++)
    if (noise[k] < g->noise_cut) bits[k >> 5] |= 1u << (k & 31);
  if ((st = upload(&ctx->mask_bits, bits))) return st;
  // the pixels the adjoint computes: the split-precision GEMM, the 128 x 64
  // sub-tiles (of its 256 x 256 workgroup tiles) with a pixel on the mask;
  // the direct form, its blocks of 256 consecutive pixels with one
  {
    const long wpr = N / 32;  // mask words per row
    long gemm_px = 0;
    for (int r0 = 0; r0 < N; r0 += 128)
      for (int c0 = 0; c0 < N; c0 += 64) {
        unsigned int any = 0u;
        for (int r = r0; r < std::min(N, r0 + 128); r++)
          for (int c = c0; c < std::min(N, c0 + 64); c += 32) any |= bits[(size_t)(r * wpr + c / 32)];
        if (any) gemm_px += (long)(std::min(N, r0 + 128) - r0) * (std::min(N, c0 + 64) - c0);
      }
    long direct_px = 0;
    for (long b = 0; b < n; b += 256) {
      unsigned int any = 0u;
      for (long k = b; k < std::min(n, b + 256); k += 32) any |= bits[(size_t)(k >> 5)];
      if (any) direct_px += std::min(n, b + 256) - b;
    }
    ctx->adj_px_gemm = gemm_px;
    ctx->adj_px_direct = direct_px;
  }
  if ((st = dev_alloc(&ctx->grid, n))) return st;
  ctx->grid_slices = 1;
  if ((st = dev_alloc(&ctx->slice_tab, kMaxSlots))) return st;
  if ((st = dev_alloc(&ctx->slice_p0, kMaxSlots + 1))) return st;

  // e^{+2 pi i k / N} (the CUFFT_INVERSE sign of the reference FFT)
  std::vector<float2> tw(N);
  for (int k = 0; k < N; k++) {
    double s, c;
    sincos_turns((double)k / N, &s, &c);
    tw[k] = make_float2((float)c, (float)s);
  }
  if ((st = upload(&ctx->twiddle, tw))) return st;

  // phase_rotate (src/functions.cu:2483-2518) factorised per axis:
  // e^{-2 pi i (x0/M) js} for column frequency js, likewise for rows.
  std::vector<float2> rx(M), ry(N);
  for (int j = 0; j < M; j++) {
    double js = j < M / 2 ? j : j - M;
    double s, c;
    sincos_turns(-((double)g->x0_phs / M) * js, &s, &c);
    rx[j] = make_float2((float)c, (float)s);
  }
  for (int i = 0; i < N; i++) {
    double is = i < N / 2 ? i : i - N;
    double s, c;
    sincos_turns(-((double)g->y0_phs / N) * is, &s, &c);
    ry[i] = make_float2((float)c, (float)s);
  }
  if ((st = upload(&ctx->rot_x, rx))) return st;
  if ((st = upload(&ctx->rot_y, ry))) return st;

  // DChi2 pixel coordinates (src/functions.cu:3728-3733): offset from the
  // int-truncated phase centre and f(t) = sqrt(1 - t^2) - 1 per axis.
  std::vector<float2> xt(M), yt(N);
  for (int j = 0; j < M; j++) {
    int o = j - (int)g->x0_phs;
    double x = o * g->deltax * kRad;
    xt[j] = make_float2((float)o, (float)(-(x * x) / (1.0 + std::sqrt(1.0 - x * x))));
  }
  for (int i = 0; i < N; i++) {
    int o = i - (int)g->y0_phs;
    double y = o * g->deltay * kRad;
    yt[i] = make_float2((float)o, (float)(-(y * y) / (1.0 + std::sqrt(1.0 - y * y))));
  }
  if ((st = upload(&ctx->xtab, xt))) return st;
  if ((st = upload(&ctx->ytab, yt))) return st;

  if ((st = dev_alloc(&ctx->partials, (size_t)kMaxSlots * kReduceBlocks))) return st;
  if ((st = dev_alloc(&ctx->dscalars, kMaxSlots))) return st;
  if ((st = dev_alloc(&ctx->cmax, kMaxSlots))) return st;
  GVM_HIP_TRY(hipHostMalloc((void**)&ctx->hscalars, sizeof(double) * kMaxSlots));
  // the value slots live in pinned host memory that the kernels write
  // directly, so reading values back takes no copy launch, only the wait
  GVM_HIP_TRY(hipHostMalloc((void**)&ctx->dvalues, sizeof(double) * kValueSlots));
  ctx->hvalues = ctx->dvalues;
  return GVM_OK;
}

extern "C" gvm_status gvm_create(const gvm_geometry* g, const float* noise_host,
                                 gvm_ctx** out) {
  GVM_REQUIRE(g && noise_host && out, "gvm_create: null argument");
  *out = nullptr;
  GVM_REQUIRE(g->M == g->N, "gvm_create: M must equal N (the reference assumes square images)");
  GVM_REQUIRE(is_pow2(g->N) && g->N >= 64 && g->N <= 4096,
              "gvm_create: N must be a power of two in [64, 4096]");
  GVM_REQUIRE(g->image_count == 2, "gvm_create: image_count must be 2 (I_nu_0, alpha)");
  GVM_REQUIRE(g->deltax != 0.0 && g->deltay != 0.0, "gvm_create: zero cell size");
  GVM_REQUIRE(g->beam_kind == GVM_BEAM_GAUSSIAN || g->beam_kind == GVM_BEAM_AIRYDISK,
              "gvm_create: unknown beam kind");
  GVM_REQUIRE(g->x0_ref >= 0 && g->x0_ref < g->M && g->y0_ref >= 0 && g->y0_ref < g->N,
              "gvm_create: pointing centre outside the image");
  GVM_REQUIRE(g->x0_phs >= 0 && g->x0_phs < g->M && g->y0_phs >= 0 && g->y0_phs < g->N,
              "gvm_create: phase centre outside the image");
  gvm_ctx* ctx = new gvm_ctx();
  ctx->g = *g;
  gvm_status st = create_impl(g, noise_host, ctx);
  if (st != GVM_OK) {
    gvm_destroy(ctx);
    return st;
  }
  *out = ctx;
  return GVM_OK;
}

extern "C" gvm_status gvm_num_slices(const gvm_ctx* ctx, int* n) {
  GVM_REQUIRE(ctx && n, "gvm_num_slices: null argument");
  *n = (int)ctx->slices.size();
  return GVM_OK;
}

extern "C" gvm_status gvm_get_geometry(const gvm_ctx* ctx, gvm_geometry* out) {
  GVM_REQUIRE(ctx && out, "gvm_get_geometry: null argument");
  *out = ctx->g;
  return GVM_OK;
}

extern "C" gvm_status gvm_set_adjoint(gvm_ctx* ctx, int mode) {
  GVM_REQUIRE(ctx, "gvm_set_adjoint: null context");
  GVM_REQUIRE(mode == GVM_ADJOINT_F32 || mode == GVM_ADJOINT_F16X3 ||
                  mode == GVM_ADJOINT_BF16X3 || mode == GVM_ADJOINT_DIRECT,
              "gvm_set_adjoint: unknown mode");
  if (ctx->adjoint_mode != mode) ctx->epoch++;
  ctx->adjoint_mode = mode;
  return GVM_OK;
}

extern "C" gvm_status gvm_add_visibilities(gvm_ctx* ctx, float freq, long nvis,
                                           const double* uvw, const float* vo,
                                           const float* weight, int* slice_out) {
  GVM_REQUIRE(ctx && uvw && vo && weight, "gvm_add_visibilities: null argument");
  GVM_REQUIRE(nvis > 0 && nvis < (1L << 31), "gvm_add_visibilities: nvis out of range");
  GVM_REQUIRE(freq > 0.f, "gvm_add_visibilities: frequency must be positive");
  GVM_REQUIRE((int)ctx->slices.size() < kMaxSlots - 8, "gvm_add_visibilities: too many slices");
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  const gvm_geometry& g = ctx->g;
  const long N = g.N;
  const double dxr = g.deltax * kRad, dyr = g.deltay * kRad;
  const double deltau = 1.0 / (g.M * dxr), deltav = 1.0 / (g.N * dyr);
  // fractional bits of the high part of u*dx: hi * n must stay exact in FP32
  int lg = 0;
  while ((1L << lg) < N) lg++;
  const double hscale = std::ldexp(1.0, 23 - lg);

  std::vector<int> cell(nvis);
  std::vector<float2> frac(nvis), vov(nvis);
  std::vector<float> wv(nvis), wc(nvis);
  std::vector<float4> phs(nvis);
  for (long i = 0; i < nvis; i++) {
    const double u = uvw[3 * i], v = uvw[3 * i + 1], w = uvw[3 * i + 2];
    // vis_mod (src/functions.cu:2566-2606): grid position, wrap, stencil
    double ux = u / deltau, uy = v / deltav;
    if (ux < 0.0) ux += N;
    if (uy < 0.0) uy += N;
    const long i1 = (long)std::floor(ux), j1 = (long)std::floor(uy);
    vov[i] = make_float2(vo[2 * i], vo[2 * i + 1]);
    if (i1 >= 0 && i1 < N && j1 >= 0 && j1 < N) {
      cell[i] = (int)(i1 | (j1 << 16));
      frac[i] = make_float2((float)(ux - i1), (float)(uy - j1));
      wv[i] = weight[i];
      const double a = u * dxr, b = v * dyr;
      const double ah = std::nearbyint(a * hscale) / hscale;
      const double bh = std::nearbyint(b * hscale) / hscale;
      phs[i] = make_float4((float)ah, (float)(a - ah), (float)bh, (float)(b - bh));
    } else {
      cell[i] = -1;
      frac[i] = make_float2(0.f, 0.f);
      wv[i] = 0.0f;
      phs[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    wc[i] = (float)w;
  }
  std::vector<float> pb((size_t)g.M * g.N);
  gvm_status st = gvm_primary_beam(g.beam_kind, g.antenna_diameter, g.pb_factor,
                                   g.pb_cutoff, freq, g.M, g.N, g.x0_ref, g.y0_ref,
                                   g.deltax, g.deltay, pb.data());
  if (st) return st;

  Slice s;
  s.freq = freq;
  s.nvis = nvis;
  if ((st = upload(&s.cell, cell)) || (st = upload(&s.frac, frac)) ||
      (st = upload(&s.vo, vov)) || (st = upload(&s.weight, wv)) ||
      (st = upload(&s.phs, phs)) || (st = upload(&s.wcoord, wc)) ||
      (st = upload(&s.pb, pb)) || (st = dev_alloc(&s.vr, nvis)) ||
      (st = dev_alloc(&s.cpol, nvis))) {
    free_slice(s);
    return st;
  }
  st = register_slice(ctx, s);
  if (st) return st;
  if (slice_out) *slice_out = (int)ctx->slices.size() - 1;
  return GVM_OK;
}

extern "C" gvm_status gvm_get_visibilities_out(gvm_ctx* ctx, int slice, const double* uvw_obs,
                                               float* model_host, float* residual_host) {
  GVM_REQUIRE(ctx && (model_host || residual_host), "gvm_get_visibilities_out: null argument");
  GVM_REQUIRE(slice >= 0 && slice < (int)ctx->slices.size(),
              "gvm_get_visibilities_out: bad slice");
  if (!ctx->have_residuals) {
    set_error("gvm_get_visibilities_out: no residuals (call gvm_chi2 first)");
    return GVM_ERR_STATE;
  }
  const Slice& s = ctx->slices[slice];
  std::vector<float2> vr(s.nvis), vo(s.nvis);
  GVM_HIP_TRY(hipMemcpyAsync(vr.data(), s.vr, sizeof(float2) * s.nvis, hipMemcpyDeviceToHost,
                             ctx->stream));
  GVM_HIP_TRY(hipMemcpyAsync(vo.data(), s.vo, sizeof(float2) * s.nvis, hipMemcpyDeviceToHost,
                             ctx->stream));
  GVM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (long i = 0; i < s.nvis; i++) {
    // Vm = Vo - Vr in the flipped frame; conjugate back where the observed u > 0
    const bool flipped = uvw_obs && uvw_obs[3 * i] > 0.0;
    const float mr = vo[i].x - vr[i].x, mi = vo[i].y - vr[i].y;
    if (model_host) {
      model_host[2 * i] = mr;
      model_host[2 * i + 1] = flipped ? -mi : mi;
    }
    if (residual_host) {
      residual_host[2 * i] = vr[i].x;
      residual_host[2 * i + 1] = flipped ? -vr[i].y : vr[i].y;
    }
  }
  return GVM_OK;
}

extern "C" gvm_status gvm_get_residuals(gvm_ctx* ctx, int slice, float* vr_host) {
  GVM_REQUIRE(ctx && vr_host, "gvm_get_residuals: null argument");
  GVM_REQUIRE(slice >= 0 && slice < (int)ctx->slices.size(), "gvm_get_residuals: bad slice");
  const Slice& s = ctx->slices[slice];
  GVM_HIP_TRY(hipMemcpyAsync(vr_host, s.vr, sizeof(float2) * s.nvis, hipMemcpyDeviceToHost,
                             ctx->stream));
  GVM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return GVM_OK;
}

extern "C" gvm_status gvm_chi2(gvm_ctx* ctx, float* I_dev, int flag_opt, int normalize,
                               double* chi2_out) {
  GVM_REQUIRE(ctx && I_dev && chi2_out, "gvm_chi2: null argument");
  GVM_REQUIRE(!ctx->slices.empty(), "gvm_chi2: no visibilities");
  GVM_REQUIRE(flag_opt >= 0, "gvm_chi2: negative flag_opt");
  return forward_chi2(ctx, I_dev, flag_opt, normalize, chi2_out);
}

extern "C" gvm_status gvm_chi2_enqueue(gvm_ctx* ctx, float* I_dev, int flag_opt, int normalize,
                                       int slot) {
  GVM_REQUIRE(ctx && I_dev, "gvm_chi2_enqueue: null argument");
  GVM_REQUIRE(!ctx->slices.empty(), "gvm_chi2_enqueue: no visibilities");
  GVM_REQUIRE(flag_opt >= 0, "gvm_chi2_enqueue: negative flag_opt");
  GVM_REQUIRE(slot >= 0 && slot < kValueSlots, "gvm_chi2_enqueue: value slot out of range");
  return forward_chi2_enqueue(ctx, I_dev, flag_opt, normalize, slot);
}

extern "C" gvm_status gvm_values_enqueue(gvm_ctx* ctx, float* I_dev, int flag_opt, int normalize,
                                         int chi2_slot, const gvm_prior_term* terms, int n,
                                         int first_slot) {
  GVM_REQUIRE(ctx && I_dev, "gvm_values_enqueue: null argument");
  GVM_REQUIRE(!ctx->slices.empty(), "gvm_values_enqueue: no visibilities");
  GVM_REQUIRE(flag_opt >= 0, "gvm_values_enqueue: negative flag_opt");
  GVM_REQUIRE(chi2_slot >= 0 && chi2_slot < kValueSlots && first_slot >= 0 && n >= 0 &&
                  first_slot + n <= kValueSlots &&
                  (n == 0 || chi2_slot < first_slot || chi2_slot >= first_slot + n),
              "gvm_values_enqueue: value slots out of range or overlapping");
  return values_enqueue(ctx, I_dev, flag_opt, normalize, chi2_slot, terms, n, first_slot);
}

extern "C" gvm_status gvm_values_read(gvm_ctx* ctx, int n, double* out) {
  GVM_REQUIRE(ctx && out, "gvm_values_read: null argument");
  GVM_REQUIRE(n >= 1 && n <= kValueSlots, "gvm_values_read: slot count out of range");
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  return values_read(ctx, n, out);
}

extern "C" gvm_status gvm_values_fetch_enqueue(gvm_ctx* ctx, int n) {
  GVM_REQUIRE(ctx, "gvm_values_fetch_enqueue: null context");
  GVM_REQUIRE(n >= 1 && n <= kValueSlots, "gvm_values_fetch_enqueue: slot count out of range");
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  return values_fetch_enqueue(ctx, n);
}

extern "C" gvm_status gvm_values_host(gvm_ctx* ctx, int n, double* out) {
  GVM_REQUIRE(ctx && out, "gvm_values_host: null argument");
  GVM_REQUIRE(n >= 1 && n <= kValueSlots, "gvm_values_host: slot count out of range");
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  return values_host(ctx, n, out);
}

// ---- HIP graphs ----------------------------------------------------------

struct gvm_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const gvm_ctx* ctx = nullptr;
  long epoch = 0;
};

extern "C" gvm_status gvm_capture_begin(gvm_ctx* ctx) {
  GVM_REQUIRE(ctx, "gvm_capture_begin: null context");
  GVM_REQUIRE(!ctx->capturing, "gvm_capture_begin: already capturing");
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  GVM_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = true;
  return GVM_OK;
}

extern "C" gvm_status gvm_capture_end(gvm_ctx* ctx, gvm_graph** out) {
  GVM_REQUIRE(ctx && out, "gvm_capture_end: null argument");
  GVM_REQUIRE(ctx->capturing, "gvm_capture_end: not capturing");
  *out = nullptr;
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  ctx->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    set_error(std::string("gvm_capture_end: the capture failed (a captured call waited or "
                          "failed): ") + hipGetErrorString(e));
    return GVM_ERR_STATE;
  }
  hipGraphExec_t x = nullptr;
  const hipError_t e2 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e2 != hipSuccess) {
    (void)hipGraphDestroy(g);
    set_error(std::string("gvm_capture_end: hipGraphInstantiate: ") + hipGetErrorString(e2));
    return GVM_ERR_HIP;
  }
  gvm_graph* gr = new gvm_graph;
  gr->graph = g;
  gr->exec = x;
  gr->ctx = ctx;
  gr->epoch = ctx->epoch;
  *out = gr;
  return GVM_OK;
}

extern "C" gvm_status gvm_graph_launch(gvm_ctx* ctx, gvm_graph* graph) {
  GVM_REQUIRE(ctx && graph, "gvm_graph_launch: null argument");
  GVM_REQUIRE(graph->ctx == ctx, "gvm_graph_launch: graph of another context");
  GVM_REQUIRE(!ctx->capturing, "gvm_graph_launch: capturing");
  if (graph->epoch != ctx->epoch) {
    set_error("gvm_graph_launch: the graph is stale (visibilities, adjoint form or GCF changed "
              "since it was captured)");
    return GVM_ERR_STATE;
  }
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  GVM_HIP_TRY(hipGraphLaunch(graph->exec, ctx->stream));
  return GVM_OK;
}

extern "C" gvm_status gvm_graph_destroy(gvm_graph* graph) {
  if (!graph) return GVM_OK;
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph) (void)hipGraphDestroy(graph->graph);
  delete graph;
  return GVM_OK;
}

extern "C" gvm_status gvm_epoch(const gvm_ctx* ctx, long* epoch) {
  GVM_REQUIRE(ctx && epoch, "gvm_epoch: null argument");
  *epoch = ctx->epoch;
  return GVM_OK;
}

extern "C" gvm_status gvm_dchi2(gvm_ctx* ctx, const float* I_dev, float* result_dev,
                                int flag_opt, int normalize) {
  GVM_REQUIRE(ctx && I_dev && result_dev, "gvm_dchi2: null argument");
  GVM_REQUIRE(flag_opt >= 0, "gvm_dchi2: negative flag_opt");
  if (!ctx->have_residuals) {
    set_error("gvm_dchi2: no residuals (call gvm_chi2 first)");
    return GVM_ERR_STATE;
  }
  return adjoint(ctx, I_dev, result_dev, flag_opt, normalize);
}

extern "C" gvm_status gvm_timing_reset(gvm_ctx* ctx) {
  GVM_REQUIRE(ctx, "gvm_timing_reset: null context");
  gvm_status st = gvm::timing_fold(ctx, -1);
  if (st) return st;
  ctx->adj_ms_total = 0.0;
  ctx->adj_count = 0;
  return GVM_OK;
}

extern "C" gvm_status gvm_timing_adjoint_ms(gvm_ctx* ctx, double* avg_ms, long* count) {
  GVM_REQUIRE(ctx && avg_ms, "gvm_timing_adjoint_ms: null argument");
  gvm_status st = gvm::timing_fold(ctx, -1);
  if (st) return st;
  *avg_ms = ctx->adj_count ? ctx->adj_ms_total / ctx->adj_count : 0.0;
  if (count) *count = ctx->adj_count;
  return GVM_OK;
}
