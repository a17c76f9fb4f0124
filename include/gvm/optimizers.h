// Optimisers and line search of the accelerated path:
// ConjugateGradient ("CG-FRPRMN", src/frprmn.cu), LBFGS ("CG-LBFGS",
// src/lbfgs.cu), linmin / mnbrak / brent / f1dim (src/linmin.cu, mnbrak.cu,
// brent.cu, f1dim.cu).  Scalars are FP64; vectors stay on the device.
#pragma once

#include <vector>

#include "gvm/optimizer.h"

class ConjugateGradient : public Optimizer {
 public:
  void allocateMemoryGpu() override;
  void deallocateMemoryGpu() override;
  void optimize() override;

 private:
  double fret = 0, fp = 0;
  float *device_g = nullptr, *device_h = nullptr, *xi = nullptr;
  int configured = 1;
};

class LBFGS : public Optimizer {
 public:
  int getK() override { return K; }
  void setK(int K) override { this->K = K; }
  void allocateMemoryGpu() override;
  void deallocateMemoryGpu() override;
  void optimize() override;
  void LBFGS_recursion(float* d_y, float* d_s, float* xi, int par_M, int lbfgs_it, int M, int N);
  // Not in the reference: read every dot product of the recursion back to
  // the host, as the reference does, instead of keeping the recursion's
  // scalars on the device (the default; both give the same direction).
  void setHostScalars(bool on) { host_scalars = on; }

 private:
  void recursion_host(int par_M, int lbfgs_it, long n, int nimg);
  void recursion_device(int par_M, int lbfgs_it, long n, int nimg);
  double fret = 0, fp = 0;
  float *d_s = nullptr, *d_y = nullptr, *xi = nullptr, *xi_old = nullptr, *p_old = nullptr;
  float *d_q = nullptr, *d_r = nullptr;
  double* d_scal = nullptr;  // device scalars of the recursion
  bool host_scalars = false;
  int configured = 1;
  int K = 100;
};

// Line search (one-dimensional minimisation along xi from p).
typedef double (*f1dim_fn)(double);
void mnbrak(double* ax, double* bx, double* cx, double* fa, double* fb, double* fc,
            f1dim_fn func);
double brent(double ax, double bx, double cx, double tol, double* xmin, f1dim_fn f);
double f1dim(double x);
void linmin(float* p, float* xi, double* fret, f1dim_fn func);
// Not in the reference: linmin takes its step lengths from xmins[0..n) in
// turn instead of searching (n = 0 restores the search).
void setLineSearchReplay(const double* xmins, int n);
// Not in the reference: the step lengths linmin has applied since the last
// setLineSearchReplay call, searched or replayed, in order.
const std::vector<double>& lineSearchSteps();
