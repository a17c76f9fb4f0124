// Optimisers and line search over device vectors (libgvm_hip C ABI).
//
// ConjugateGradient follows src/frprmn.cu:85-195, LBFGS src/lbfgs.cu:99-341
// (with its storage addressing, see LBFGS::optimize), linmin
// src/linmin.cu:52-110, f1dim src/f1dim.cu:49-80.  mnbrak and brent implement
// golden-ratio bracketing with parabolic extrapolation and Brent's
// parabolic/golden-section minimisation (Brent 1973) with the reference's
// constants and decision order (src/mnbrak.cu, src/brent.cu).  All scalars
// are FP64 (the reference uses float); vectors never leave the device.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "gvm/factory.h"
#include "gvm/framework.h"
#include "gvm/optimizers.h"

using gvm::check;
using gvm::globals;

namespace {
gvm_ctx* ctx() {
  gvm_ctx* c = globals().ctx;
  if (!c) throw gvm::Error("no current gvm context", GVM_ERR_STATE);
  return c;
}
long img_n() { return globals().M * globals().N; }
float* dev_zeros(long n) {
  void* p = nullptr;
  check(gvm_malloc(&p, sizeof(float) * (size_t)n), "gvm_malloc");
  check(gvm_vec_zero(ctx(), (float*)p, n), "gvm_vec_zero");
  return (float*)p;
}

// State shared by linmin and f1dim (the reference's file-scope globals
// device_pcom / device_xicom / testof / I).
ObjectiveFunction* testof = nullptr;
Image* I = nullptr;
float* device_pcom = nullptr;
float* device_xicom = nullptr;
float* device_xt = nullptr;
}  // namespace

// ---------------------------------------------------------------------------
// image function mapping: src/functions.cu:4566-4592, src/mfs.cu:794-811
// ---------------------------------------------------------------------------
void defaultNewP(float* p, float* xi, double xmin, int image) {
  check(gvm_new_p(ctx(), p, xi, xmin, image, 0, 0.f, globals().eta), "newPNoPositivity");
}
void particularNewP(float* p, float* xi, double xmin, int image) {
  check(gvm_new_p(ctx(), p, xi, xmin, image, 1, globals().initial_values[image], globals().eta),
        "newP");
}
void defaultEvaluateXt(float* xt, float* pcom, float* xicom, double x, int image) {
  check(gvm_evaluate_xt(ctx(), xt, pcom, xicom, x, image, 0, 0.f, globals().eta),
        "evaluateXtNoPositivity");
}
void particularEvaluateXt(float* xt, float* pcom, float* xicom, double x, int image) {
  check(gvm_evaluate_xt(ctx(), xt, pcom, xicom, x, image, 1, globals().initial_values[image],
                        globals().eta),
        "evaluateXt");
}
void defaultNewP(float* p, float* xi, float xmin, int image) {
  defaultNewP(p, xi, (double)xmin, image);
}
void particularNewP(float* p, float* xi, float xmin, int image) {
  particularNewP(p, xi, (double)xmin, image);
}
void defaultEvaluateXt(float* xt, float* pcom, float* xicom, float x, int image) {
  defaultEvaluateXt(xt, pcom, xicom, (double)x, image);
}
void particularEvaluateXt(float* xt, float* pcom, float* xicom, float x, int image) {
  particularEvaluateXt(xt, pcom, xicom, (double)x, image);
}
imageMap* makeFunctionMapping(int image_count, bool nopositivity) {
  imageMap* m = new imageMap[image_count];
  for (int i = 0; i < image_count; i++) {
    if (!nopositivity && i == 0) {
      m[i].evaluateXt = particularEvaluateXt;
      m[i].newP = particularNewP;
      m[i].evaluateXtD = particularEvaluateXt;
      m[i].newPD = particularNewP;
    } else {
      m[i].evaluateXt = defaultEvaluateXt;
      m[i].newP = defaultNewP;
      m[i].evaluateXtD = defaultEvaluateXt;
      m[i].newPD = defaultNewP;
    }
  }
  return m;
}

namespace {
// The double-step form of a mapping entry when it has one, else its float
// form with the step rounded to float (the reference's arithmetic).
void applyNewP(const imageMap& m, float* p, float* xi, double xmin, int image) {
  if (m.newPD)
    m.newPD(p, xi, xmin, image);
  else
    m.newP(p, xi, (float)xmin, image);
}
void applyEvaluateXt(const imageMap& m, float* xt, float* pcom, float* xicom, double x,
                     int image) {
  if (m.evaluateXtD)
    m.evaluateXtD(xt, pcom, xicom, x, image);
  else
    m.evaluateXt(xt, pcom, xicom, (float)x, image);
}
}  // namespace

// ---------------------------------------------------------------------------
// line search
// ---------------------------------------------------------------------------
namespace {
constexpr double kGold = 1.618034;
constexpr double kGlimit = 100.0;
constexpr double kTiny = 1.0e-20;
constexpr int kItmax = 500;
constexpr double kCgold = 0.3819660;
constexpr double kZeps = 1.0e-10;
inline double sign_of(double a, double b) { return b >= 0.0 ? std::fabs(a) : -std::fabs(a); }
}  // namespace

void mnbrak(double* ax, double* bx, double* cx, double* fa, double* fb, double* fc,
            f1dim_fn func) {
  *fa = func(*ax);
  *fb = func(*bx);
  if (*fb > *fa) {  // go downhill from a to b
    std::swap(*ax, *bx);
    std::swap(*fa, *fb);
  }
  *cx = *bx + kGold * (*bx - *ax);
  *fc = func(*cx);
  while (*fb > *fc) {
    const double r = (*bx - *ax) * (*fb - *fc);
    const double q = (*bx - *cx) * (*fb - *fa);
    const double u0 = *bx - ((*bx - *cx) * q - (*bx - *ax) * r) /
                                (2.0 * sign_of(std::max(std::fabs(q - r), kTiny), q - r));
    const double ulim = *bx + kGlimit * (*cx - *bx);
    double u = u0, fu;
    if ((*bx - u) * (u - *cx) > 0.0) {  // parabolic minimum between b and c
      fu = func(u);
      if (fu < *fc) {
        *ax = *bx;
        *bx = u;
        *fa = *fb;
        *fb = fu;
        return;
      } else if (fu > *fb) {
        *cx = u;
        *fc = fu;
        return;
      }
      u = *cx + kGold * (*cx - *bx);
      fu = func(u);
    } else if ((*cx - u) * (u - ulim) > 0.0) {  // between c and the limit
      fu = func(u);
      if (fu < *fc) {
        *bx = *cx;
        *cx = u;
        u = *cx + kGold * (*cx - *bx);
        *fb = *fc;
        *fc = fu;
        fu = func(u);
      }
    } else if ((u - ulim) * (ulim - *cx) >= 0.0) {  // clamp to the limit
      u = ulim;
      fu = func(u);
    } else {  // default golden step
      u = *cx + kGold * (*cx - *bx);
      fu = func(u);
    }
    *ax = *bx;
    *bx = *cx;
    *cx = u;
    *fa = *fb;
    *fb = *fc;
    *fc = fu;
  }
}

double brent(double ax, double bx, double cx, double tol, double* xmin, f1dim_fn f) {
  double a = std::min(ax, cx), b = std::max(ax, cx);
  double x = bx, w = bx, v = bx;
  double fx = f(x), fw = fx, fv = fx;
  double d = 0.0, e = 0.0;
  for (int it = 1; it <= kItmax; it++) {
    const double xm = 0.5 * (a + b);
    const double tol1 = tol * std::fabs(x) + kZeps;
    const double tol2 = 2.0 * tol1;
    if (std::fabs(x - xm) <= (tol2 - 0.5 * (b - a))) {
      *xmin = x;
      return fx;
    }
    if (std::fabs(e) > tol1) {  // trial parabolic fit
      const double r = (x - w) * (fx - fv);
      double q = (x - v) * (fx - fw);
      double p = (x - v) * q - (x - w) * r;
      q = 2.0 * (q - r);
      if (q > 0.0) p = -p;
      q = std::fabs(q);
      const double etemp = e;
      e = d;
      if (std::fabs(p) >= std::fabs(0.5 * q * etemp) || p <= q * (a - x) || p >= q * (b - x)) {
        e = (x >= xm ? a - x : b - x);
        d = kCgold * e;
      } else {
        d = p / q;
        const double u = x + d;
        if (u - a < tol2 || b - u < tol2) d = sign_of(tol1, xm - x);
      }
    } else {
      e = (x >= xm ? a - x : b - x);
      d = kCgold * e;
    }
    const double u = (std::fabs(d) >= tol1 ? x + d : x + sign_of(tol1, d));
    const double fu = f(u);
    if (fu <= fx) {
      if (u >= x)
        a = x;
      else
        b = x;
      v = w;
      w = x;
      x = u;
      fv = fw;
      fw = fx;
      fx = fu;
    } else {
      if (u < x)
        a = u;
      else
        b = u;
      if (fu <= fw || w == x) {
        v = w;
        w = u;
        fv = fw;
        fw = fu;
      } else if (fu <= fv || v == x || v == w) {
        v = u;
        fv = fu;
      }
    }
  }
  if (globals().verbose) std::printf("Too many iterations in brent\n");
  *xmin = x;
  return fx;
}

// f(x) = Phi(p + x xi) with the per-image positivity of evaluateXt.
double f1dim(double x) {
  for (int i = 0; i < I->getImageCount(); i++)
    applyEvaluateXt(I->getFunctionMapping()[i], device_xt, device_pcom, device_xicom, x, i);
  return testof->calcFunctionD(device_xt);
}

// Line-search replay (not in the reference): when a table of step lengths
// is set, linmin takes the next one as xmin instead of bracketing and
// minimising, and returns f1dim at that step (what brent returns at its
// minimum).  Replaying another implementation's accepted steps separates
// arithmetic differences from the line search's branch decisions.
namespace {
std::vector<double> g_replay;
size_t g_replay_pos = 0;
std::vector<double> g_steps;  // the step lengths applied since the last reset
}  // namespace

void setLineSearchReplay(const double* xmins, int n) {
  g_replay.assign(xmins, xmins + std::max(0, n));
  g_replay_pos = 0;
  g_steps.clear();
}

const std::vector<double>& lineSearchSteps() { return g_steps; }

void linmin(float* p, float* xi, double* fret, f1dim_fn func) {
  const long tot = img_n() * globals().image_count;
  (void)func;
  check(gvm_vec_copy(ctx(), device_pcom, p, tot), "gvm_vec_copy(pcom)");
  check(gvm_vec_copy(ctx(), device_xicom, xi, tot), "gvm_vec_copy(xicom)");
  double ax = 0.0, xx = 1.0, bx, fa, fx, fb, xmin = 0.0;
  if (!g_replay.empty()) {
    if (g_replay_pos >= g_replay.size())
      throw gvm::Error("linmin: the line-search replay table is exhausted", GVM_ERR_STATE);
    xmin = g_replay[g_replay_pos++];
    *fret = f1dim(xmin);
  } else {
    mnbrak(&ax, &xx, &bx, &fa, &fx, &fb, f1dim);
    *fret = brent(ax, xx, bx, 1.0e-7, &xmin, f1dim);
  }
  g_steps.push_back(xmin);
  if (globals().verbose) std::printf("Alpha for linear minimization = %f\n\n", xmin);
  for (int i = 0; i < I->getImageCount(); i++)
    applyNewP(I->getFunctionMapping()[i], p, xi, xmin, i);
}

// Line-search buffers are allocated once per optimize() (the reference
// allocates pcom/xicom per linmin call and xt per f1dim call).
static void linmin_alloc(long tot) {
  device_pcom = dev_zeros(tot);
  device_xicom = dev_zeros(tot);
  device_xt = dev_zeros(tot);
}
static void linmin_free() {
  gvm_free(device_pcom);
  gvm_free(device_xicom);
  gvm_free(device_xt);
  device_pcom = device_xicom = device_xt = nullptr;
}

// ---------------------------------------------------------------------------
// ConjugateGradient ("CG-FRPRMN"): src/frprmn.cu:54-195
// ---------------------------------------------------------------------------
void ConjugateGradient::allocateMemoryGpu() {
  const long tot = img_n() * image->getImageCount();
  device_g = dev_zeros(tot);
  device_h = dev_zeros(tot);
  xi = dev_zeros(tot);
  linmin_alloc(tot);
}

void ConjugateGradient::deallocateMemoryGpu() {
  gvm_free(device_g);
  gvm_free(device_h);
  gvm_free(xi);
  device_g = device_h = xi = nullptr;
  linmin_free();
}

void ConjugateGradient::optimize() {
  constexpr double EPS = 1.0e-10;
  I = image;
  globals().flag_opt = this->flag;
  allocateMemoryGpu();
  testof = of;
  startTrajectory();
  exit_reason = 0;
  if (configured) {
    of->configure(globals().N, globals().M, image->getImageCount());
    configured = 0;
  }
  const int nimg = image->getImageCount();
  float* p = image->getImage();
  auto finish = [&](int reason) {
    exit_reason = reason;
    final_value = of->calcFunctionD(p);
    deallocateMemoryGpu();
  };

  fp = of->calcFunctionD(p);
  of->calcGradient(p, xi, 0);
  for (int k = 0; k < nimg; k++) check(gvm_cg_search_direction(ctx(), device_g, xi, device_h, k), "searchDirection");

  for (int i = 1; i <= this->total_iterations; i++) {
    this->current_iteration = i;
    linmin(p, xi, &fret, nullptr);
    recordTrajectory(fret);
    if (2.0 * std::fabs(fret - fp) <= this->ftol * (std::fabs(fret) + std::fabs(fp) + EPS)) {
      finish(1);
      return;
    }
    fp = of->calcFunctionD(p);
    of->calcGradient(p, xi, i);
    const double den = std::max(fp, 1.0);
    double gmax = 0.0;
    check(gvm_cg_grad_condition(ctx(), xi, p, den, nimg, &gmax), "CGGradCondition");
    if (gmax < this->gtol) {
      finish(2);
      return;
    }
    double gg = 0.0, dgg = 0.0;
    check(gvm_cg_gg_dgg(ctx(), xi, device_g, nimg, &gg, &dgg), "getGGandDGG");
    if (gg == 0.0) {
      finish(3);
      return;
    }
    const double gam = std::max(0.0, dgg / gg);
    for (int k = 0; k < nimg; k++) check(gvm_cg_new_xi(ctx(), device_g, xi, device_h, gam, k), "newXi");
  }
  finish(0);
}

// ---------------------------------------------------------------------------
// LBFGS ("CG-LBFGS"): src/lbfgs.cu:64-341
//
// Storage addressing is the reference's: d_y / d_s hold K*image_count
// image-sized blocks; for image `im`, slot `it` is written at block
// im*it + im (calculateSandY), dot products read block im*k + im
// (getDot_LBFGS_ff) and the q / r updates read block im + k (updateQ).  For
// image 0 every write and dot product therefore uses block 0.  y is the new
// gradient plus the previous search direction (xi_old).
// ---------------------------------------------------------------------------
void LBFGS::allocateMemoryGpu() {
  const long tot = img_n() * image->getImageCount();
  d_y = dev_zeros(tot * K);
  d_s = dev_zeros(tot * K);
  p_old = dev_zeros(tot);
  xi = dev_zeros(tot);
  xi_old = dev_zeros(tot);
  d_q = dev_zeros(tot);
  d_r = dev_zeros(tot);
  // alpha for every (image, memory slot), then the scratch scalars
  void* sc = nullptr;
  check(gvm_malloc(&sc, sizeof(double) * ((size_t)image->getImageCount() * K + 16)),
        "gvm_malloc(lbfgs scalars)");
  d_scal = (double*)sc;
  linmin_alloc(tot);
}

void LBFGS::deallocateMemoryGpu() {
  gvm_free(d_y);
  gvm_free(d_s);
  gvm_free(p_old);
  gvm_free(xi);
  gvm_free(xi_old);
  gvm_free(d_q);
  gvm_free(d_r);
  gvm_free(d_scal);
  d_y = d_s = p_old = xi = xi_old = d_q = d_r = nullptr;
  d_scal = nullptr;
  linmin_free();
}

void LBFGS::optimize() {
  I = image;
  globals().flag_opt = this->flag;
  allocateMemoryGpu();
  testof = of;
  startTrajectory();
  exit_reason = 0;
  if (configured) {
    of->configure(globals().N, globals().M, image->getImageCount());
    configured = 0;
  }
  const int nimg = image->getImageCount();
  const long n = img_n(), tot = n * nimg;
  float* p = image->getImage();
  auto finish = [&](int reason) {
    exit_reason = reason;
    final_value = of->calcFunctionD(p);
    deallocateMemoryGpu();
  };

  fp = of->calcFunctionD(p);
  of->calcGradient(p, xi, 0);
  for (int k = 0; k < nimg; k++) check(gvm_vec_scale(ctx(), xi + n * k, xi + n * k, -1.0, n), "searchDirection_LBFGS");

  for (int i = 1; i <= this->total_iterations; i++) {
    this->current_iteration = i;
    check(gvm_vec_copy(ctx(), p_old, p, tot), "copy p_old");
    check(gvm_vec_copy(ctx(), xi_old, xi, tot), "copy xi_old");
    linmin(p, xi, &fret, nullptr);
    recordTrajectory(fret);
    if ((fp - fret) / std::max({std::fabs(fret), std::fabs(fp), 1.0}) <= this->ftol) {
      finish(1);
      return;
    }
    double max_per_it = 0.0;
    check(gvm_vec_absmax(ctx(), xi, tot, &max_per_it), "normArray");
    if (max_per_it <= this->gtol) {
      finish(2);
      return;
    }
    fp = of->calcFunctionD(p);
    of->calcGradient(p, xi, i);
    const int slot = (this->current_iteration - 1) % this->K;
    for (int im = 0; im < nimg; im++) {
      const long blk = (long)im * slot + im;
      check(gvm_lbfgs_s_and_y(ctx(), d_y + n * blk, d_s + n * blk, p + n * im, xi + n * im,
                              p_old + n * im, xi_old + n * im, n),
            "calculateSandY");
    }
    LBFGS_recursion(d_y, d_s, xi, std::min(this->K, this->current_iteration), slot,
                    (int)globals().M, (int)globals().N);
  }
  finish(0);
}

// Two-loop recursion over the stored pairs (src/lbfgs.cu:252-341), with the
// reference's slot addressing: dot products read pair k of image im at
// offset n (im k + im), the updates at n (im + k).
void LBFGS::LBFGS_recursion(float* d_y, float* d_s, float* xi, int par_M, int lbfgs_it, int M,
                            int N) {
  (void)d_y;
  (void)d_s;
  (void)xi;
  const long n = (long)M * N;
  const int nimg = image->getImageCount();
  check(gvm_vec_copy(ctx(), d_q, this->xi, n * nimg), "copy q");
  check(gvm_vec_zero(ctx(), d_r, n * nimg), "zero r");
  if (host_scalars)
    recursion_host(par_M, lbfgs_it, n, nimg);
  else
    recursion_device(par_M, lbfgs_it, n, nimg);
  check(gvm_vec_scale(ctx(), this->xi, d_r, -1.0, n * nimg), "searchDirection_LBFGS(r)");
}

// Every scalar read back to the host, as the reference does.
void LBFGS::recursion_host(int par_M, int lbfgs_it, long n, int nimg) {
  std::vector<double> alpha((size_t)nimg * par_M, 0.0);
  auto blk_dot = [&](float* buf, int im, int k) { return buf + n * ((long)im * k + im); };
  auto blk_upd = [&](float* buf, int im, int k) { return buf + n * ((long)im + k); };
  for (int im = 0; im < nimg; im++) {
    for (int k = par_M - 1; k >= 0; k--) {
      double rho_den = 0.0, sq = 0.0;
      check(gvm_vec_dot(ctx(), blk_dot(d_y, im, k), blk_dot(d_s, im, k), n, &rho_den), "y.s");
      const double rho = rho_den != 0.0 ? 1.0 / rho_den : 0.0;
      check(gvm_vec_dot(ctx(), blk_dot(d_s, im, k), d_q + n * im, n, &sq), "s.q");
      alpha[(size_t)im * par_M + k] = rho * sq;
      check(gvm_vec_axpy(ctx(), d_q + n * im, blk_upd(d_y, im, k), -alpha[(size_t)im * par_M + k], n),
            "updateQ");
    }
  }
  double sy_yy = 0.0;
  for (int im = 0; im < nimg; im++) {
    double sy = 0.0, yy = 0.0;
    check(gvm_vec_dot(ctx(), blk_dot(d_y, im, lbfgs_it), blk_dot(d_s, im, lbfgs_it), n, &sy), "s.y");
    check(gvm_vec_dot(ctx(), blk_dot(d_y, im, lbfgs_it), blk_dot(d_y, im, lbfgs_it), n, &yy), "y.y");
    if (yy != 0.0) sy_yy += sy / yy;
  }
  for (int im = 0; im < nimg; im++) {
    check(gvm_vec_scale(ctx(), d_r + n * im, d_q + n * im, sy_yy, n), "getR");
    for (int k = 0; k < par_M; k++) {
      double rho_den = 0.0, yr = 0.0;
      check(gvm_vec_dot(ctx(), blk_dot(d_y, im, k), blk_dot(d_s, im, k), n, &rho_den), "y.s");
      const double rho = rho_den != 0.0 ? 1.0 / rho_den : 0.0;
      check(gvm_vec_dot(ctx(), blk_dot(d_y, im, k), d_r + n * im, n, &yr), "y.r");
      const double beta = rho * yr;
      check(gvm_vec_axpy(ctx(), d_r + n * im, blk_upd(d_s, im, k), alpha[(size_t)im * par_M + k] - beta, n),
            "updateQ(r)");
    }
  }
}

// The same operations with the scalars kept in device doubles: the whole
// recursion is enqueued without a host round trip.
void LBFGS::recursion_device(int par_M, int lbfgs_it, long n, int nimg) {
  auto blk_dot = [&](float* buf, int im, int k) { return buf + n * ((long)im * k + im); };
  auto blk_upd = [&](float* buf, int im, int k) { return buf + n * ((long)im + k); };
  double* alpha = d_scal;                     // [nimg * par_M]
  double* t = d_scal + (size_t)nimg * K;      // scratch
  double *rho_den = t, *dot = t + 1, *coef = t + 2, *sy = t + 3, *yy = t + 3 + nimg,
         *h0 = t + 3 + 2 * nimg;
  for (int im = 0; im < nimg; im++) {
    for (int k = par_M - 1; k >= 0; k--) {
      double* a = alpha + (size_t)im * par_M + k;
      check(gvm_vec_dot_dev(ctx(), blk_dot(d_y, im, k), blk_dot(d_s, im, k), n, rho_den), "y.s");
      check(gvm_vec_dot_dev(ctx(), blk_dot(d_s, im, k), d_q + n * im, n, dot), "s.q");
      check(gvm_scalar_rho_times(ctx(), rho_den, dot, nullptr, a), "alpha");
      check(gvm_vec_axpy_dev(ctx(), d_q + n * im, blk_upd(d_y, im, k), a, 1, n), "updateQ");
    }
  }
  for (int im = 0; im < nimg; im++) {
    check(gvm_vec_dot_dev(ctx(), blk_dot(d_y, im, lbfgs_it), blk_dot(d_s, im, lbfgs_it), n, sy + im),
          "s.y");
    check(gvm_vec_dot_dev(ctx(), blk_dot(d_y, im, lbfgs_it), blk_dot(d_y, im, lbfgs_it), n, yy + im),
          "y.y");
  }
  check(gvm_scalar_h0(ctx(), sy, yy, nimg, h0), "H0 scale");
  for (int im = 0; im < nimg; im++) {
    check(gvm_vec_scale_dev(ctx(), d_r + n * im, d_q + n * im, h0, n), "getR");
    for (int k = 0; k < par_M; k++) {
      check(gvm_vec_dot_dev(ctx(), blk_dot(d_y, im, k), blk_dot(d_s, im, k), n, rho_den), "y.s");
      check(gvm_vec_dot_dev(ctx(), blk_dot(d_y, im, k), d_r + n * im, n, dot), "y.r");
      check(gvm_scalar_rho_times(ctx(), rho_den, dot, alpha + (size_t)im * par_M + k, coef),
            "alpha - beta");
      check(gvm_vec_axpy_dev(ctx(), d_r + n * im, blk_upd(d_s, im, k), coef, 0, n), "updateQ(r)");
    }
  }
}

namespace {
Optimizer* CreateFrprmn() { return new ConjugateGradient; }
Optimizer* CreateLbfgs() { return new LBFGS; }
const bool RegisteredFrprmn = registerCreationFunction<Optimizer, std::string>("CG-FRPRMN", CreateFrprmn);
const bool RegisteredLbfgs = registerCreationFunction<Optimizer, std::string>("CG-LBFGS", CreateLbfgs);
}  // namespace
