// Internal declarations shared by the HIP translation units of libgvm_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "gvm_hip.h"

namespace gvm {

// Thread-local text of the last error (gvm_last_error).
void set_error(const std::string& msg);
const char* last_error_cstr();

#define GVM_HIP_TRY(expr)                                                   \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) {                                                 \
      ::gvm::set_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
      return e_ == hipErrorOutOfMemory ? GVM_ERR_OOM : GVM_ERR_HIP;         \
    }                                                                       \
  } while (0)

#define GVM_CHECK_LAUNCH() GVM_HIP_TRY(hipGetLastError())

#define GVM_REQUIRE(cond, msg)     \
  do {                             \
    if (!(cond)) {                 \
      ::gvm::set_error(msg);       \
      return GVM_ERR_ARG;          \
    }                              \
  } while (0)

// One channel x stokes slice of visibilities, resident in HBM.
struct Slice {
  float freq = 0.f;
  long nvis = 0;
  // static per-visibility data (set-up)
  int* cell = nullptr;       // i1 | (j1 << 16) of the bilinear stencil, -1 off-grid
  float2* frac = nullptr;    // (du, dv) bilinear fractions
  float2* vo = nullptr;      // observed visibilities
  float* weight = nullptr;   // weights (0 off-grid)
  float4* phs = nullptr;     // (a_hi, a_lo, b_hi, b_lo): u dx and v dy split for exact turns
  float* wcoord = nullptr;   // w in wavelengths
  float* pb = nullptr;       // primary beam at this frequency, float[M*N]
  // per-evaluation data
  float2* vr = nullptr;      // residuals Vo - Vm
  float2* cpol = nullptr;    // (w |Vr|, -arg(Vr) / 2 pi): c = w conj(Vr) in polar form, turns
};

// What the batched kernels read about one slice (a copy of the pointers of
// a Slice plus per-slice scalars), kept in a device table indexed by slice.
struct SliceDev {
  const int* cell;
  const float2* frac;
  const float2* vo;
  const float* weight;
  float2* vr;
  float2* cpol;
  const float4* phs;
  const float* wcoord;
  const float* pb;
  long nvis;
  float nudiv;  // freq / nu_0
  float pad;
};

constexpr int kReduceBlocks = 1024;
constexpr int kReduceThreads = 256;
constexpr int kMaxSlots = 4096;  // slices per context (and reduction slots)
constexpr int kValueSlots = 16;  // term values enqueued before one read-back

}  // namespace gvm

struct gvm_ctx {
  gvm_geometry g{};
  int device = 0;
  int cu_count = 0;
  hipStream_t stream = nullptr;
  // image-space tables
  float* noise = nullptr;       // float[M*N]
  unsigned int* mask_bits = nullptr;  // noise < noise_cut, one bit per pixel
  long adj_px_gemm = 0;    // pixels the split-precision adjoint computes
  long adj_px_direct = 0;  // pixels the direct-sum adjoint computes
  float* gcf = nullptr;         // gridding correction image float[M*N], or none
  float2* grid = nullptr;       // uv grid, complex float[N*M]
  float2* twiddle = nullptr;    // e^{+2 pi i k / N}, k < N
  float2* rot_x = nullptr;      // phase rotation factor per column frequency
  float2* rot_y = nullptr;      // phase rotation factor per row frequency
  float2* xtab = nullptr;       // (j - int(x0_phs), sqrt(1 - x^2) - 1) per column
  float2* ytab = nullptr;       // (i - int(y0_phs), sqrt(1 - y^2) - 1) per row
  // reductions
  double* partials = nullptr;   // kMaxSlots * kReduceBlocks doubles
  double* dscalars = nullptr;   // kMaxSlots doubles (device)
  double* hscalars = nullptr;   // kMaxSlots doubles (pinned host)
  double* dvalues = nullptr;    // kValueSlots term values (pinned host, written by kernels)
  double* hvalues = nullptr;    // the same memory, read on the host after a wait
  unsigned int* cmax = nullptr; // per-slice max |w Vr| (float bits), device
  // batched slices: device descriptor table, uv grids (one per slice)
  gvm::SliceDev* slice_tab = nullptr;  // kMaxSlots entries
  size_t grid_slices = 0;              // slices the grid buffer holds
  // adjoint workspace and its work split: entry p = (slice, kbeg, kend, 0)
  // of the visibility ranges; slice s owns partial tiles [p0[s], p0[s+1])
  float* adj_part = nullptr;    // split-K partial tiles
  size_t adj_part_elems = 0;
  int4* split_tab = nullptr;
  int* slice_p0 = nullptr;      // kMaxSlots + 1 entries
  int split_cap = 0;
  int nsplit_total = 0;
  int split_mode = -1;          // adjoint mode the table was built for (-1: stale)
  int adjoint_mode = GVM_ADJOINT_BF16X3;
  long split_vis = 6144;        // target visibilities per adjoint range (K tile), default_split_vis
  int fft_mode = 0;             // FFT batching: 0 automatic, 1 wide, 2 narrow
  std::vector<gvm::Slice> slices;
  bool have_residuals = false;
  // graph capture: true between gvm_capture_begin and gvm_capture_end; the
  // epoch changes whenever recorded device pointers or launches go stale
  bool capturing = false;
  long epoch = 0;
  // network workspace: the weight-gradient partials of a convolution
  float* nn_part = nullptr;
  size_t nn_part_elems = 0;
  // timing of the adjoint GEMM: a ring of event pairs recorded around each
  // eager launch and read back lazily (when the ring is full or the timing
  // is queried), so an evaluation never waits on the host for its timing
  static constexpr int kTimingPairs = 32;
  hipEvent_t tev[2 * kTimingPairs] = {};
  int tev_head = 0, tev_pending = 0;
  bool timing_paused = false;
  double adj_ms_total = 0.0;
  long adj_count = 0;
};

namespace gvm {

// The adjoint's default range length and the longest the autotuner tries, by
// image size.  The FP32 accumulation error of a range grows with the square
// root of its length; at 2048^2 and 4.35 M visibilities ranges of 6144 put
// the gradient at 8e-6 (bf16x3) and 1.0e-5 (f16x3) relative L2 against FP64
// sums on sampled pixels, ranges of 3072 at 5-6e-6.
inline long default_split_vis(int N) { return N >= 2048 ? 3072 : 6144; }
inline long max_tuned_split_vis(int N) { return N >= 2048 ? 3072 : 8192; }

// Adds the durations of up to n of the oldest pending timed adjoint launches
// to the totals (waiting for each); n < 0 for all.
gvm_status timing_fold(gvm_ctx* ctx, int n);

// Launchers implemented in the .hip files.
gvm_status forward_chi2(gvm_ctx* ctx, float* I, int flag_opt, int normalize,
                        double* out);
gvm_status adjoint(gvm_ctx* ctx, const float* I, float* result, int flag_opt,
                   int normalize);
// Rebuilds the adjoint's range table and workspace for the current form.
gvm_status rebuild_splits(gvm_ctx* ctx);
// Block-partial reduction helpers: reduce `nblocks` partials of slot s into
// dscalars[s] (device), in a fixed order.
gvm_status finish_reductions(gvm_ctx* ctx, int nslots, int nblocks);
gvm_status finish_reductions_mode(gvm_ctx* ctx, int nslots, int nblocks, int is_max);
// The same, writing the nslots results to `out` (device) instead of dscalars.
gvm_status finish_reductions_to(gvm_ctx* ctx, int nslots, int nblocks, double* out);
// Block partials of n fused prior terms in partial rows row0.. (blocks per
// row in *nb_out), and the same with its finish into dvalues[first_slot..].
gvm_status priors_value_partials(gvm_ctx* ctx, const float* I, const gvm_prior_term* terms,
                                 int n, int row0, int* nb_out);
gvm_status priors_value_partials_finish(gvm_ctx* ctx, const float* I,
                                        const gvm_prior_term* terms, int n, int first_slot);
// chi2 into dvalues[chi2_slot] and n prior terms into dvalues[first_slot..]:
// for up to 4 slices the FFT passes, one launch for the partials and one to
// finish them.
gvm_status values_enqueue(gvm_ctx* ctx, float* I, int flag_opt, int normalize, int chi2_slot,
                          const gvm_prior_term* terms, int n, int first_slot);
// Enqueues the chi2 evaluation; 0.5 * sum over slices lands in dvalues[slot].
gvm_status forward_chi2_enqueue(gvm_ctx* ctx, float* I, int flag_opt, int normalize, int slot);
// Reads the value slots [0, n) once the stream has reached this point (one
// synchronisation; the kernels write the slots in pinned host memory).
gvm_status values_read(gvm_ctx* ctx, int n, double* out);
gvm_status values_fetch_enqueue(gvm_ctx* ctx, int n);
gvm_status values_host(gvm_ctx* ctx, int n, double* out);
// Slices: frees one slice's device arrays; makes a filled slice the next
// slice of ctx (grids, descriptor), freeing it on failure.
void free_slice(Slice& s);
gvm_status register_slice(gvm_ctx* ctx, Slice& s);
// Primary beam image on the device (FP64 formula, stored as float).
gvm_status primary_beam_dev(const gvm_geometry& g, double freq, float* out, hipStream_t stream);
gvm_status read_scalars(gvm_ctx* ctx, int nslots, double* out);
int reduce_grid(long n);
// Sums (or takes the max of) nblocks FP64 partials per slot, slot s reading
// partials[s * kReduceBlocks ...], in a fixed order (gvm_ctx.hip).
__global__ void finish_kernel(const double* __restrict__ partials, double* __restrict__ out,
                              int nblocks, int is_max);

// Loads and stores through global-address-space pointers.  Pointers read
// from device tables (the slice descriptors) are generic, and generic (flat)
// accesses also count against the LDS wait counter, so every LDS wait would
// wait for them too.  HIP's float2 / float4 are classes whose copy goes
// through a generic reference, so they are moved as native vectors.
typedef float gf2_t __attribute__((ext_vector_type(2)));
typedef float gf4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float ldg(const float* p) {
  return *(const __attribute__((address_space(1))) float*)p;
}
__device__ __forceinline__ int ldg(const int* p) {
  return *(const __attribute__((address_space(1))) int*)p;
}
__device__ __forceinline__ float2 ldg(const float2* p) {
  const gf2_t v = *(const __attribute__((address_space(1))) gf2_t*)p;
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 ldg(const float4* p) {
  const gf4_t v = *(const __attribute__((address_space(1))) gf4_t*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void stg(float2* p, float2 v) {
  gf2_t w;
  w.x = v.x;
  w.y = v.y;
  *(__attribute__((address_space(1))) gf2_t*)p = w;
}

// |r|^2 of a residual with each operation rounded (no multiply-add
// contraction), so the kernels that recompute it get the same bits.
__device__ __forceinline__ float resid_abs2(float2 r) {
#pragma clang fp contract(off)
  return r.x * r.x + r.y * r.y;
}

// Wave-level (64 lanes) and block-level sums in double.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sums v over a block of kReduceThreads threads; the result is valid in
// thread 0.  `sh` must hold kReduceThreads / 64 doubles.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += sh[w];
  }
  return r;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

__device__ __forceinline__ double block_max(double v, double* sh) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    r = sh[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = fmax(r, sh[w]);
  }
  return r;
}

}  // namespace gvm
