// Sum of objective terms: the reference's
// include/classes/objectivefunction.cuh:6-103.  Terms with lambda == 0 are
// dropped at addFi; calcGradient restarts every term, then per term (in
// insertion order) sets the iteration, computes its gradient and adds it to
// dphi, then copies dphi to xi.  Registered as "ObjectiveFunction".
#pragma once

#include <vector>

#include "gvm/fi.h"

class ObjectiveFunction {
 public:
  ObjectiveFunction() {}
  ~ObjectiveFunction();
  void addFi(Fi* fi);
  float calcFunction(float* p);
  double calcFunctionD(float* p);  // FP64 sum of the terms' FP64 values
  void calcGradient(float* p, float* xi, int iter);
  void restartDPhi();
  void copyDphiToXi(float* xi);
  std::vector<Fi*> getFi() { return this->fis; }
  Fi* getFiByName(std::string fi_name);
  void setN(long N) { this->N = N; }
  void setM(long M) { this->M = M; }
  void setImageCount(int I) { this->image_count = I; }
  void configure(long N, long M, int I);
  std::vector<float> get_fi_values() { return this->fi_values; }
  long getFunctionCalls() const { return n_fcalls; }
  long getGradientCalls() const { return n_gcalls; }

 private:
  std::vector<Fi*> fis;
  std::vector<float> fi_values;
  float* dphi = nullptr;
  long N = 0;
  long M = 0;
  int image_count = 1;
  long n_fcalls = 0, n_gcalls = 0;
};
