// Gridded-visibility mode on the device: convolutional gridding of the rows
// of one slice (do_gridding, src/functions.cu:1339-1653), the weighting
// schemes (src/naturalweightingscheme.cu, src/uniformweightingscheme.cu,
// src/briggsweightingscheme.cu, src/radialweightingscheme.cu) and the
// gridding correction image (apply_GCF, src/functions.cu:2468-2476).
//
// Both the gridder and the weighting schemes accumulate per-cell sums in
// float.  The reference adds rows to a cell one at a time in row order (with
// one OpenMP thread), so a parallel float sum would not reproduce its
// rounding.  Here every (row, tap) pair becomes a (cell, entry) key pair;
// a stable radix sort groups the pairs by cell while keeping the row order
// inside each cell, and one thread per cell then adds that cell's entries in
// sequence.  The result is the reference's single-thread sum, bit for bit,
// and does not depend on scheduling.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "gvm_internal.h"

namespace gvm {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kRad = kPi / 180.0;
constexpr int kThreads = 256;
constexpr int kSumBlocks = 1024;

int blocks_for(long n) {
  return (int)std::max(1L, std::min<long>((n + kThreads - 1) / kThreads, 65536));
}

// Device allocation released when it goes out of scope (set-up paths only).
struct DevMem {
  void* p = nullptr;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

gvm_status dev_alloc(DevMem& m, size_t bytes, const char* what) {
  if (m.p) {
    (void)hipFree(m.p);
    m.p = nullptr;
  }
  if (hipMalloc(&m.p, std::max<size_t>(bytes, 16)) != hipSuccess) {
    m.p = nullptr;
    set_error(std::string(what) + ": out of device memory");
    return GVM_ERR_OOM;
  }
  return GVM_OK;
}

// Number of bits a radix sort needs for keys in [0, sentinel].
int key_bits(uint32_t sentinel) {
  int b = 1;
  while (b < 32 && (sentinel >> b) != 0) b++;
  return b;
}

// Sorted (cell key, entry) pairs of one pass, with the buffers reused across
// slices.  After sort(), keys()/ids() are the sorted arrays.
struct CellSort {
  DevMem k0, k1, v0, v1, tmp;
  size_t cap = 0, tmp_bytes = 0;
  rocprim::double_buffer<uint32_t> kb, vb;

  gvm_status reserve(long n, const char* what) {
    if ((size_t)n <= cap) return GVM_OK;
    gvm_status st;
    if ((st = dev_alloc(k0, sizeof(uint32_t) * n, what)) ||
        (st = dev_alloc(k1, sizeof(uint32_t) * n, what)) ||
        (st = dev_alloc(v0, sizeof(uint32_t) * n, what)) ||
        (st = dev_alloc(v1, sizeof(uint32_t) * n, what)))
      return st;
    cap = (size_t)n;
    rocprim::double_buffer<uint32_t> k(k0.as<uint32_t>(), k1.as<uint32_t>());
    rocprim::double_buffer<uint32_t> v(v0.as<uint32_t>(), v1.as<uint32_t>());
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, k, v, (size_t)n, 0u, 32u) != hipSuccess) {
      set_error(std::string(what) + ": radix sort sizing failed");
      return GVM_ERR_HIP;
    }
    tmp_bytes = bytes;
    return dev_alloc(tmp, bytes, what);
  }
  uint32_t* in_keys() const { return k0.as<uint32_t>(); }
  uint32_t* in_ids() const { return v0.as<uint32_t>(); }
  gvm_status sort(long n, int bits, const char* what) {
    kb = rocprim::double_buffer<uint32_t>(k0.as<uint32_t>(), k1.as<uint32_t>());
    vb = rocprim::double_buffer<uint32_t>(v0.as<uint32_t>(), v1.as<uint32_t>());
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, kb, vb, (size_t)n, 0u, (unsigned)bits) !=
        hipSuccess) {
      set_error(std::string(what) + ": radix sort sizing failed");
      return GVM_ERR_HIP;
    }
    if (bytes > tmp_bytes) {
      gvm_status st = dev_alloc(tmp, bytes, what);
      if (st) return st;
      tmp_bytes = bytes;
    }
    bytes = tmp_bytes;
    if (rocprim::radix_sort_pairs(tmp.p, bytes, kb, vb, (size_t)n, 0u, (unsigned)bits) !=
        hipSuccess) {
      set_error(std::string(what) + ": radix sort failed");
      return GVM_ERR_HIP;
    }
    return GVM_OK;
  }
  const uint32_t* keys() { return kb.current(); }
  const uint32_t* ids() { return vb.current(); }
};

// ---- convolutional gridding -----------------------------------------------

struct GridGeom {
  double deltau, deltav;  // 1 / (M dx), 1 / (N dy), radians (dx < 0: deltau < 0)
  double cj, ck;          // floor(N / 2), floor(M / 2)
  double lambda;          // float wavelength of the slice, promoted
  int M, N;
  int km, kn;             // kernel rows / columns
  int supx, supy;         // supports
  int nx, taps;           // 2 supx + 1, (2 supy + 1) nx
  uint32_t off;           // key of entries that land off the grid (sorts last)
};

// One (row, tap) entry per thread: its cell key, or g.off.  Entry e is row
// z = e / taps (z >= nvis: the Hermitian twin of row z - nvis) and tap
// t = e % taps; keys are written in entry order, so a stable sort keeps the
// rows of a cell in row order (do_gridding, src/functions.cu:1424-1480).
__global__ __launch_bounds__(kThreads) void grid_keys_kernel(const double* __restrict__ uvw,
                                                             long nvis, GridGeom g,
                                                             uint32_t* __restrict__ keys,
                                                             uint32_t* __restrict__ ids) {
#pragma clang fp contract(off)
  const long n = 2 * nvis * g.taps;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const long z = e / g.taps;
    const int t = (int)(e % g.taps);
    const long vi = z < nvis ? z : z - nvis;
    double u = uvw[3 * vi], v = uvw[3 * vi + 1];
    if (z >= nvis) {
      u = -u;
      v = -v;
    }
    u = u / g.lambda;  // metres_to_lambda (src/MSFITSIO.cu:41-45)
    v = v / g.lambda;
    const double jf = u / g.deltau + g.cj + 0.5;
    const double kf = v / g.deltav + g.ck + 0.5;
    uint32_t key = g.off;
    if (fabs(jf) < 1073741824.0 && fabs(kf) < 1073741824.0) {
      const int j = (int)jf, k = (int)kf;  // C truncation, as int(j_fp)
      const int ki = t / g.nx, kj = t % g.nx;
      const int sj = j + kj - g.supx, sk = k + ki - g.supy;
      if (sk >= 0 && sk < g.M && sj >= 0 && sj < g.N && ki < g.km && kj < g.kn)
        key = (uint32_t)((long)g.N * sk + sj);
    }
    keys[e] = key;
    ids[e] = (uint32_t)e;
  }
}

// Run starts: flag[i] = 1 where a run of equal keys begins (the run of
// off-grid entries included, so every run ends where the next begins).
__global__ __launch_bounds__(kThreads) void run_heads_kernel(const uint32_t* __restrict__ keys,
                                                             long n, int* __restrict__ flag) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    flag[i] = (i == 0 || keys[i - 1] != keys[i]) ? 1 : 0;
}

// starts[rank[i]] = i at run starts, and starts[number of runs] = n.
__global__ __launch_bounds__(kThreads) void run_starts_kernel(const int* __restrict__ flag,
                                                              const int* __restrict__ rank,
                                                              long n, int* __restrict__ starts) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    if (flag[i]) starts[rank[i]] = (int)i;
    if (i == n - 1) starts[rank[i] + flag[i]] = (int)n;
  }
}

// One wave per cell (run r of the sorted entries): the float sums of
// do_gridding in entry order, then the normalisation
// (src/functions.cu:1424-1559).  The 64 lanes load a chunk of the run's
// entries and form the products w k, w k^2, w Vo.x k, w Vo.y k in parallel;
// lane 0 then adds them in entry order, so the rounding is the reference's
// single-thread sum.  Writes, at index r, whether the cell is kept
// (weight > 0) and its values.
__global__ __launch_bounds__(64) void grid_cells_kernel(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ ids,
    const int* __restrict__ starts, long nvis, GridGeom g, const float* __restrict__ taps,
    const float* __restrict__ weight, const float2* __restrict__ vo, int* __restrict__ keep,
    float2* __restrict__ cell_vo, float* __restrict__ cell_w) {
#pragma clang fp contract(off)
  __shared__ float prod[4][64];
  const int r = blockIdx.x;
  const int s = starts[r], e = starts[r + 1];
  const int lane = threadIdx.x;
  if (keys[s] == g.off) {
    if (lane == 0) keep[r] = 0;
    return;
  }
  float ws = 0.f, aux = 0.f, re = 0.f, im = 0.f;
  for (int c = s; c < e; c += 64) {
    const int p = c + lane;
    if (p < e) {
      const uint32_t id = ids[p];
      const long z = id / g.taps;
      const int t = (int)(id % g.taps);
      const long vi = z < nvis ? z : z - nvis;
      const float w = weight[vi];
      float2 v = vo[vi];
      if (z >= nvis) v.y = -v.y;
      const float kv = taps[g.kn * (t / g.nx) + (t % g.nx)];
      const float kv2 = kv * kv;
      prod[0][lane] = w * kv;
      prod[1][lane] = w * kv2;
      prod[2][lane] = w * v.x * kv;
      prod[3][lane] = w * v.y * kv;
    }
    __syncthreads();
    if (lane == 0) {
      const int cnt = min(64, e - c);
      for (int j = 0; j < cnt; j++) {
        ws = ws + prod[0][j];
        aux = aux + prod[1][j];
        re = re + prod[2][j];
        im = im + prod[3][j];
      }
    }
    __syncthreads();
  }
  if (lane == 0) {
    float weff = 0.f;
    float2 out = make_float2(0.f, 0.f);
    if (aux != 0.0f && ws != 0.0f) {
      weff = ws * ws / aux;
      out = make_float2(re / ws, im / ws);
    }
    keep[r] = weff > 0.0f ? 1 : 0;
    cell_vo[r] = out;
    cell_w[r] = weff;
  }
}

// Kept cells to rows at the cell centre, in row-major cell order
// (src/functions.cu:1527-1539, 1587-1607): uvw in metres, w = 0.  One
// thread per run.
__global__ __launch_bounds__(kThreads) void grid_emit_kernel(
    const uint32_t* __restrict__ keys, const int* __restrict__ starts,
    const int* __restrict__ keep, const int* __restrict__ pos, long nruns, GridGeom g,
    const float2* __restrict__ cell_vo, const float* __restrict__ cell_w, long capacity,
    double* __restrict__ uvw_out, float2* __restrict__ vo_out, float* __restrict__ w_out) {
#pragma clang fp contract(off)
  for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < nruns;
       r += (long)gridDim.x * blockDim.x) {
    if (!keep[r]) continue;
    const long o = pos[r];
    if (o >= capacity) continue;
    const uint32_t key = keys[starts[r]];
    const int gj = (int)(key % (uint32_t)g.N), gk = (int)(key / (uint32_t)g.N);
    const double ul = ((double)gj - g.cj) * g.deltau;
    const double vl = ((double)gk - g.ck) * g.deltav;
    uvw_out[3 * o] = ul * g.lambda;
    uvw_out[3 * o + 1] = vl * g.lambda;
    uvw_out[3 * o + 2] = 0.0;
    vo_out[o] = cell_vo[r];
    w_out[o] = cell_w[r];
  }
}

// ---- weighting schemes -----------------------------------------------------

struct WeightGeom {
  double absdu, absdv;  // |deltau|, |deltav|
  double cx, cy;        // int(floor(N / 2)), int(floor(M / 2))
  double lambda;
  int M, N;
  uint32_t off;
};

struct TaperCoef {
  float amplitude, a, b, c;
  double u0, v0;
  int on;
};

// Cell of one row for the weighting schemes: wavelengths, rows with u < 0
// mirrored, cells of |deltau| x |deltav| (src/uniformweightingscheme.cu:35-51).
__device__ inline uint32_t weight_cell(double u, double v, const WeightGeom& g) {
#pragma clang fp contract(off)
  const double xf = u / g.absdu + g.cx + 0.5;
  const double yf = v / g.absdv + g.cy + 0.5;
  if (!(fabs(xf) < 1073741824.0 && fabs(yf) < 1073741824.0)) return g.off;
  const int x = (int)xf, y = (int)yf;
  if (x >= 0 && y >= 0 && x < g.N && y < g.M) return (uint32_t)((long)g.N * y + x);
  return g.off;
}

__device__ inline void mirrored_lambda(const double* uvw, long z, double lambda, double* u,
                                       double* v) {
#pragma clang fp contract(off)
  *u = uvw[3 * z] / lambda;
  *v = uvw[3 * z + 1] / lambda;
  if (*u < 0.0) {
    *u = -*u;
    *v = -*v;
  }
}

__global__ __launch_bounds__(kThreads) void weight_keys_kernel(const double* __restrict__ uvw,
                                                               long nvis, WeightGeom g,
                                                               uint32_t* __restrict__ keys,
                                                               uint32_t* __restrict__ ids) {
  for (long z = blockIdx.x * (long)blockDim.x + threadIdx.x; z < nvis;
       z += (long)gridDim.x * blockDim.x) {
    double u, v;
    mirrored_lambda(uvw, z, g.lambda, &u, &v);
    keys[z] = weight_cell(u, v, g);
    ids[z] = (uint32_t)z;
  }
}

// Per cell: grid[cell] = (cumulative ? grid[cell] : 0) + the cell's weights
// added in row order.
__global__ __launch_bounds__(kThreads) void weight_cells_kernel(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ ids, long n, uint32_t off,
    const float* __restrict__ weight, int cumulative, float* __restrict__ grid) {
#pragma clang fp contract(off)
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const uint32_t key = keys[i];
    if (key == off || (i > 0 && keys[i - 1] == key)) continue;
    float s = cumulative ? grid[key] : 0.0f;
    for (long p = i; p < n && keys[p] == key; p++) s = s + weight[ids[p]];
    grid[key] = s;
  }
}

// Block partial sums (FP64) of sum_k w_k, or of G^2 over the columns
// n >= N/2 of every row (src/briggsweightingscheme.cu:47-53, 96-101).
__global__ __launch_bounds__(kThreads) void weight_sum_kernel(const float* __restrict__ x,
                                                              long n, double* __restrict__ part) {
  __shared__ double red[kThreads];
  double s = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    s += (double)x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kThreads) void grid_sq_sum_kernel(const float* __restrict__ grid,
                                                               int M, int N,
                                                               double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double red[kThreads];
  const int half = N - N / 2;
  const long n = (long)M * half;
  double s = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const float gv = grid[(i / half) * N + N / 2 + (i % half)];
    s += (double)(gv * gv);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// UVTaper::getValue (include/classes/uvtaper.cuh:101-114) with a, b, c
// computed on the host in float.
__device__ inline float taper_value(double u, double v, const TaperCoef& t) {
#pragma clang fp contract(off)
  const double x = u - t.u0, y = v - t.v0;
  const double e = -(double)t.a * x * x - (double)t.b * x * y - (double)t.c * y * y;
  return (float)((double)t.amplitude * exp(e));
}

// The per-row weight update of each scheme (second loops of
// src/uniformweightingscheme.cu:66-90, src/briggsweightingscheme.cu:151-175;
// src/naturalweightingscheme.cu:23-44; src/radialweightingscheme.cu:20-39).
__global__ __launch_bounds__(kThreads) void weight_apply_kernel(
    int scheme, const double* __restrict__ uvw, long nvis, WeightGeom g,
    const float* __restrict__ grid, float f2, TaperCoef taper, float* __restrict__ weight) {
#pragma clang fp contract(off)
  for (long z = blockIdx.x * (long)blockDim.x + threadIdx.x; z < nvis;
       z += (long)gridDim.x * blockDim.x) {
    float w = weight[z];
    double u, v;
    if (scheme == GVM_WEIGHTING_RADIAL) {
      u = uvw[3 * z] / g.lambda;
      v = uvw[3 * z + 1] / g.lambda;
      const float fu = (float)u, fv = (float)v;
      w = w * sqrtf(fu * fu + fv * fv);  // distance(u, v, 0, 0), src/MSFITSIO.cu:47-51
    } else {
      mirrored_lambda(uvw, z, g.lambda, &u, &v);
      if (scheme == GVM_WEIGHTING_UNIFORM || scheme == GVM_WEIGHTING_BRIGGS) {
        const uint32_t cell = weight_cell(u, v, g);
        if (cell == g.off)
          w = 0.0f;
        else if (scheme == GVM_WEIGHTING_UNIFORM)
          w = w / grid[cell];
        else
          w = (float)((double)w / (1.0 + (double)(grid[cell] * f2)));
      }
    }
    if (taper.on) w = w * taper_value(u, v, taper);
    weight[z] = w;
  }
}

gvm_status launch_ok(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return GVM_ERR_HIP;
  }
  return GVM_OK;
}

// FP64 sum of the block partials, in block order.
gvm_status read_partials(const double* part, int nb, double* out, const char* what) {
  std::vector<double> h(nb);
  if (hipMemcpy(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error(std::string(what) + ": partial sums copy failed");
    return GVM_ERR_HIP;
  }
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += h[b];
  *out = s;
  return GVM_OK;
}

}  // namespace
}  // namespace gvm

using namespace gvm;

extern "C" gvm_status gvm_grid_visibilities_dev(gvm_ctx* ctx, const gvm_ckernel* kernel,
                                                float freq, long nvis, const double* uvw_m_dev,
                                                const float* vo_dev, const float* weight_dev,
                                                double* uvw_out_dev, float* vo_out_dev,
                                                float* weight_out_dev, long capacity,
                                                long* nout) {
  static const char* what = "gvm_grid_visibilities_dev";
  GVM_REQUIRE(ctx && kernel && kernel->taps && uvw_m_dev && vo_dev && weight_dev && uvw_out_dev &&
                  vo_out_dev && weight_out_dev && nout,
              "gvm_grid_visibilities_dev: null argument");
  GVM_REQUIRE(freq > 0.f, "gvm_grid_visibilities_dev: frequency must be positive");
  GVM_REQUIRE(kernel->m >= 1 && kernel->n >= 1 && kernel->m <= 255 && kernel->n <= 255 &&
                  kernel->support_x >= 0 && kernel->support_y >= 0 &&
                  kernel->support_x <= 127 && kernel->support_y <= 127,
              "gvm_grid_visibilities_dev: kernel size out of range");
  GVM_REQUIRE(capacity >= 0, "gvm_grid_visibilities_dev: negative capacity");
  const gvm_geometry& geo = ctx->g;
  GridGeom g;
  g.M = geo.M;
  g.N = geo.N;
  const double dxr = geo.deltax * kRad, dyr = geo.deltay * kRad;
  g.deltau = 1.0 / (geo.M * dxr);
  g.deltav = 1.0 / (geo.N * dyr);
  g.cj = std::floor(geo.N / 2.0);
  g.ck = std::floor(geo.M / 2.0);
  g.lambda = (double)(2.99792458E8f / freq);
  g.km = kernel->m;
  g.kn = kernel->n;
  g.supx = kernel->support_x;
  g.supy = kernel->support_y;
  g.nx = 2 * g.supx + 1;
  g.taps = g.nx * (2 * g.supy + 1);
  g.off = (uint32_t)((long)geo.M * geo.N);
  *nout = 0;
  GVM_REQUIRE(nvis >= 0 && nvis <= (((1L << 31) - 1) / (2L * g.taps)),
              "gvm_grid_visibilities_dev: rows x taps exceed 2^31 entries");
  if (nvis == 0) return GVM_OK;
  const long n = 2 * nvis * g.taps;
  GVM_HIP_TRY(hipSetDevice(ctx->device));

  gvm_status st;
  DevMem taps_d, flag, rank, starts, keep, pos, cvo, cw;
  CellSort cs;
  if ((st = dev_alloc(taps_d, sizeof(float) * g.km * g.kn, what)) ||
      (st = cs.reserve(n, what)) || (st = dev_alloc(flag, sizeof(int) * n, what)) ||
      (st = dev_alloc(rank, sizeof(int) * n, what)) ||
      (st = dev_alloc(starts, sizeof(int) * (n + 1), what)))
    return st;
  GVM_HIP_TRY(hipMemcpy(taps_d.p, kernel->taps, sizeof(float) * g.km * g.kn,
                        hipMemcpyHostToDevice));
  hipLaunchKernelGGL(grid_keys_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, nullptr, uvw_m_dev,
                     nvis, g, cs.in_keys(), cs.in_ids());
  if ((st = launch_ok(what)) || (st = cs.sort(n, key_bits(g.off), what))) return st;
  // the runs of equal keys (cells, then the off-grid run): start of each
  hipLaunchKernelGGL(run_heads_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, nullptr, cs.keys(),
                     n, flag.as<int>());
  if ((st = launch_ok(what))) return st;
  size_t sbytes = 0;
  GVM_HIP_TRY(rocprim::exclusive_scan(nullptr, sbytes, flag.as<int>(), rank.as<int>(), 0,
                                      (size_t)n, rocprim::plus<int>()));
  DevMem stmp;
  if ((st = dev_alloc(stmp, sbytes, what))) return st;
  GVM_HIP_TRY(rocprim::exclusive_scan(stmp.p, sbytes, flag.as<int>(), rank.as<int>(), 0,
                                      (size_t)n, rocprim::plus<int>()));
  int last[2] = {0, 0};
  GVM_HIP_TRY(hipMemcpy(&last[0], rank.as<int>() + n - 1, sizeof(int), hipMemcpyDeviceToHost));
  GVM_HIP_TRY(hipMemcpy(&last[1], flag.as<int>() + n - 1, sizeof(int), hipMemcpyDeviceToHost));
  const long nruns = (long)last[0] + last[1];
  hipLaunchKernelGGL(run_starts_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, nullptr,
                     flag.as<int>(), rank.as<int>(), n, starts.as<int>());
  if ((st = launch_ok(what)) || (st = dev_alloc(keep, sizeof(int) * nruns, what)) ||
      (st = dev_alloc(pos, sizeof(int) * nruns, what)) ||
      (st = dev_alloc(cvo, sizeof(float2) * nruns, what)) ||
      (st = dev_alloc(cw, sizeof(float) * nruns, what)))
    return st;
  hipLaunchKernelGGL(grid_cells_kernel, dim3((unsigned)nruns), dim3(64), 0, nullptr, cs.keys(),
                     cs.ids(), starts.as<int>(), nvis, g, taps_d.as<float>(), weight_dev,
                     reinterpret_cast<const float2*>(vo_dev), keep.as<int>(), cvo.as<float2>(),
                     cw.as<float>());
  if ((st = launch_ok(what))) return st;
  // exclusive prefix sum of the keep flags: output row of each kept cell
  size_t kbytes = 0;
  GVM_HIP_TRY(rocprim::exclusive_scan(nullptr, kbytes, keep.as<int>(), pos.as<int>(), 0,
                                      (size_t)nruns, rocprim::plus<int>()));
  if (kbytes > sbytes && (st = dev_alloc(stmp, kbytes, what))) return st;
  kbytes = std::max(kbytes, sbytes);
  GVM_HIP_TRY(rocprim::exclusive_scan(stmp.p, kbytes, keep.as<int>(), pos.as<int>(), 0,
                                      (size_t)nruns, rocprim::plus<int>()));
  GVM_HIP_TRY(hipMemcpy(&last[0], pos.as<int>() + nruns - 1, sizeof(int), hipMemcpyDeviceToHost));
  GVM_HIP_TRY(hipMemcpy(&last[1], keep.as<int>() + nruns - 1, sizeof(int), hipMemcpyDeviceToHost));
  const long count = (long)last[0] + last[1];
  if (count > capacity) {
    set_error("gvm_grid_visibilities_dev: " + std::to_string(count) +
              " gridded rows exceed the output capacity " + std::to_string(capacity));
    return GVM_ERR_ARG;
  }
  hipLaunchKernelGGL(grid_emit_kernel, dim3(blocks_for(nruns)), dim3(kThreads), 0, nullptr,
                     cs.keys(), starts.as<int>(), keep.as<int>(), pos.as<int>(), nruns, g,
                     cvo.as<float2>(), cw.as<float>(), capacity, uvw_out_dev,
                     reinterpret_cast<float2*>(vo_out_dev), weight_out_dev);
  if ((st = launch_ok(what))) return st;
  GVM_HIP_TRY(hipDeviceSynchronize());
  *nout = count;
  return GVM_OK;
}

extern "C" gvm_status gvm_add_visibilities_gridded_dev(gvm_ctx* ctx, const gvm_ckernel* kernel,
                                                       float freq, long nvis,
                                                       const double* uvw_m_dev,
                                                       const float* vo_dev,
                                                       const float* weight_dev, long* ngridded,
                                                       int* slice_out) {
  static const char* what = "gvm_add_visibilities_gridded_dev";
  GVM_REQUIRE(ctx, "gvm_add_visibilities_gridded_dev: null context");
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  const long cap = (long)ctx->g.M * ctx->g.N;
  DevMem u, v, w;
  gvm_status st;
  if ((st = dev_alloc(u, sizeof(double) * 3 * cap, what)) ||
      (st = dev_alloc(v, sizeof(float) * 2 * cap, what)) ||
      (st = dev_alloc(w, sizeof(float) * cap, what)))
    return st;
  long count = 0;
  st = gvm_grid_visibilities_dev(ctx, kernel, freq, nvis, uvw_m_dev, vo_dev, weight_dev,
                                 u.as<double>(), v.as<float>(), w.as<float>(), cap, &count);
  if (st) return st;
  GVM_REQUIRE(count > 0, "gvm_add_visibilities_gridded_dev: no visibility falls on the grid");
  if (ngridded) *ngridded = count;
  return gvm_add_visibilities_dev(ctx, freq, count, u.as<double>(), v.as<float>(), w.as<float>(),
                                  slice_out);
}

extern "C" gvm_status gvm_set_gcf(gvm_ctx* ctx, const float* gcf_host) {
  GVM_REQUIRE(ctx, "gvm_set_gcf: null context");
  GVM_REQUIRE(!ctx->capturing, "gvm_set_gcf: capturing");
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  GVM_HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->epoch++;
  if (!gcf_host) {
    if (ctx->gcf) (void)hipFree(ctx->gcf);
    ctx->gcf = nullptr;
    return GVM_OK;
  }
  const size_t n = (size_t)ctx->g.M * ctx->g.N;
  if (!ctx->gcf) GVM_HIP_TRY(hipMalloc((void**)&ctx->gcf, sizeof(float) * n));
  GVM_HIP_TRY(hipMemcpy(ctx->gcf, gcf_host, sizeof(float) * n, hipMemcpyHostToDevice));
  return GVM_OK;
}

extern "C" gvm_status gvm_weighting_apply_dev(gvm_ctx* ctx, int scheme, float robust,
                                              const gvm_uvtaper* taper, int nsl,
                                              const float* freq, const long* nvis,
                                              const double* const* uvw_m_dev,
                                              float* const* weight_dev) {
  static const char* what = "gvm_weighting_apply_dev";
  GVM_REQUIRE(ctx && (nsl == 0 || (freq && nvis && uvw_m_dev && weight_dev)),
              "gvm_weighting_apply_dev: null argument");
  GVM_REQUIRE(scheme >= GVM_WEIGHTING_NATURAL && scheme <= GVM_WEIGHTING_RADIAL,
              "gvm_weighting_apply_dev: unknown weighting scheme");
  // BriggsWeightingScheme::setRobustParam (src/briggsweightingscheme.cu:13-21)
  GVM_REQUIRE(scheme != GVM_WEIGHTING_BRIGGS || (robust >= -2.0f && robust <= 2.0f),
              "gvm_weighting_apply_dev: robust parameter must be in [-2, 2]");
  GVM_HIP_TRY(hipSetDevice(ctx->device));
  const gvm_geometry& geo = ctx->g;
  const double dxr = geo.deltax * kRad, dyr = geo.deltay * kRad;
  WeightGeom g;
  g.absdu = std::fabs(1.0 / (geo.M * dxr));
  g.absdv = std::fabs(1.0 / (geo.N * dyr));
  g.cx = (double)(int)std::floor((double)(geo.N / 2));
  g.cy = (double)(int)std::floor((double)(geo.M / 2));
  g.M = geo.M;
  g.N = geo.N;
  g.off = (uint32_t)((long)geo.M * geo.N);
  TaperCoef tc{};
  if (taper) {
    // the float coefficients of UVTaper::getValue
    const float cb = std::cos(taper->bpa), sb = std::sin(taper->bpa);
    const float s2 = std::sin(2.0f * taper->bpa);
    const float smaj2 = 2.0f * taper->sigma_maj * taper->sigma_maj;
    const float smin2 = 2.0f * taper->sigma_min * taper->sigma_min;
    tc.a = (cb * cb) / smaj2 + (sb * sb) / smin2;
    tc.b = s2 / smaj2 - s2 / smin2;
    tc.c = (sb * sb) / smaj2 + (cb * cb) / smin2;
    tc.amplitude = taper->amplitude;
    tc.u0 = taper->u_0;
    tc.v0 = taper->v_0;
    tc.on = 1;
  }
  long nmax = 0;
  for (int s = 0; s < nsl; s++) {
    GVM_REQUIRE(nvis[s] >= 0 && nvis[s] < (1L << 31) && freq[s] > 0.f &&
                    (nvis[s] == 0 || (uvw_m_dev[s] && weight_dev[s])),
                "gvm_weighting_apply_dev: bad slice");
    nmax = std::max(nmax, nvis[s]);
  }
  if (nsl == 0 || nmax == 0) return GVM_OK;
  const bool gridded = scheme == GVM_WEIGHTING_UNIFORM || scheme == GVM_WEIGHTING_BRIGGS;
  const int bits = key_bits(g.off);
  gvm_status st;
  CellSort cs;
  DevMem grid, part;
  const size_t gbytes = sizeof(float) * (size_t)geo.M * geo.N;
  if (gridded && ((st = cs.reserve(nmax, what)) || (st = dev_alloc(grid, gbytes, what)) ||
                  (st = dev_alloc(part, sizeof(double) * kSumBlocks, what))))
    return st;

  float f2 = 0.0f;
  if (scheme == GVM_WEIGHTING_BRIGGS) {
    // Pass 1 (src/briggsweightingscheme.cu:44-105): the sum of the weights,
    // and per slice the sum of G^2 over the right half of a grid that is not
    // cleared between slices (the reference clears it only after the pass).
    double sum_w = 0.0, sum_g2 = 0.0;
    GVM_HIP_TRY(hipMemset(grid.p, 0, gbytes));
    for (int s = 0; s < nsl; s++) {
      if (nvis[s] == 0) continue;
      g.lambda = (double)(2.99792458E8f / freq[s]);
      const int nb = std::min(kSumBlocks, blocks_for(nvis[s]));
      double ps = 0.0;
      hipLaunchKernelGGL(weight_sum_kernel, dim3(nb), dim3(kThreads), 0, nullptr, weight_dev[s],
                         nvis[s], part.as<double>());
      if ((st = launch_ok(what)) || (st = read_partials(part.as<double>(), nb, &ps, what)))
        return st;
      sum_w += ps;
      hipLaunchKernelGGL(weight_keys_kernel, dim3(blocks_for(nvis[s])), dim3(kThreads), 0,
                         nullptr, uvw_m_dev[s], nvis[s], g, cs.in_keys(), cs.in_ids());
      if ((st = launch_ok(what)) || (st = cs.sort(nvis[s], bits, what))) return st;
      hipLaunchKernelGGL(weight_cells_kernel, dim3(blocks_for(nvis[s])), dim3(kThreads), 0,
                         nullptr, cs.keys(), cs.ids(), nvis[s], g.off, weight_dev[s], 1,
                         grid.as<float>());
      const int nbg = std::min(kSumBlocks, blocks_for((long)geo.M * (geo.N - geo.N / 2)));
      hipLaunchKernelGGL(grid_sq_sum_kernel, dim3(nbg), dim3(kThreads), 0, nullptr,
                         grid.as<float>(), geo.M, geo.N, part.as<double>());
      if ((st = launch_ok(what)) || (st = read_partials(part.as<double>(), nbg, &ps, what)))
        return st;
      sum_g2 += ps;
    }
    GVM_REQUIRE(sum_w != 0.0, "gvm_weighting_apply_dev: the sum of the weights is zero");
    const float average_weights = (float)sum_g2 / (float)sum_w;
    const float f = 5.0f * std::pow(10.0f, -robust);
    f2 = f * f / average_weights;
  }

  // Pass 2: per slice, the slice's own weight grid, then the row update.
  for (int s = 0; s < nsl; s++) {
    if (nvis[s] == 0) continue;
    g.lambda = (double)(2.99792458E8f / freq[s]);
    if (gridded) {
      GVM_HIP_TRY(hipMemset(grid.p, 0, gbytes));
      hipLaunchKernelGGL(weight_keys_kernel, dim3(blocks_for(nvis[s])), dim3(kThreads), 0,
                         nullptr, uvw_m_dev[s], nvis[s], g, cs.in_keys(), cs.in_ids());
      if ((st = launch_ok(what)) || (st = cs.sort(nvis[s], bits, what))) return st;
      hipLaunchKernelGGL(weight_cells_kernel, dim3(blocks_for(nvis[s])), dim3(kThreads), 0,
                         nullptr, cs.keys(), cs.ids(), nvis[s], g.off, weight_dev[s], 0,
                         grid.as<float>());
      if ((st = launch_ok(what))) return st;
    }
    hipLaunchKernelGGL(weight_apply_kernel, dim3(blocks_for(nvis[s])), dim3(kThreads), 0, nullptr,
                       scheme, uvw_m_dev[s], nvis[s], g,
                       gridded ? grid.as<float>() : nullptr, f2, tc, weight_dev[s]);
    if ((st = launch_ok(what))) return st;
  }
  GVM_HIP_TRY(hipDeviceSynchronize());
  return GVM_OK;
}
