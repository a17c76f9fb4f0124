// Objective term interface: the reference's include/classes/fi.cuh:13-104
// (same members and virtuals).  All float* crossing this interface are device
// pointers to float[image_count*M*N].  calcFiD is an addition: the FP64 value
// of the term (the default forwards to calcFi), which the optimisers use for
// their line-search scalars.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "gvm/framework.h"

class CKernel;  // gridding kernels are not part of this path

// The reference's cufftComplex (8 bytes, {x, y} = {re, im}).
typedef struct gvm_complex {
  float x, y;
} gvm_complex;

class Fi {
 public:
  Fi();
  virtual ~Fi();

  virtual float calcFi(float* p) = 0;
  virtual void calcGi(float* p, float* xi) = 0;
  virtual void restartDGi() = 0;
  virtual void addToDphi(float* device_dphi) = 0;
  virtual double calcFiD(float* p) { return (double)calcFi(p); }
  // Asynchronous value (not in the reference): when every term of an
  // ObjectiveFunction has one, it enqueues all term values and waits once.
  // enqueueValue returns false when the value is zero without evaluation
  // (an inactive prior); finishValue stores the raw value like calcFi and
  // returns the weighted one.
  virtual bool hasAsyncValue() const { return false; }
  virtual bool enqueueValue(float* p, int slot) { return false; }
  // Appends every value that decides which kernels this term launches and
  // with which scalar arguments (the key of a captured evaluation graph).
  virtual void graphKey(std::vector<int64_t>& key) const;
  double finishValue(double raw) {
    set_fivalue(raw);
    return (double)penalization_factor * raw;
  }
  // A prior evaluated with others in one pass (not in the reference): fills
  // t with this term's kind, images, constants and activity (the value's
  // gating, and for `gradient` also the flag_opt gating of calcGi) and
  // returns true; false for a term that has no such form.
  virtual bool priorTerm(gvm_prior_term* t, bool gradient) const { return false; }
  virtual void setPrior(float prior) {}
  virtual void setPrior(float* prior) {}
  virtual float getEta() { return 0.0f; }
  virtual void setEta(float eta) {}
  virtual void setCKernel(CKernel* ckernel) {}
  virtual void setFgScale(float fg_scale) {}
  virtual float getFgScale() { return 0.0f; }

  std::string getName() { return this->name; }
  void setName(std::string name) { this->name = name; }

  float get_fivalue() { return (float)this->fi_value; }
  double get_fivalue_d() { return this->fi_value; }
  bool getNormalize() { return this->normalize; }
  float getPenalizationFactor() { return this->penalization_factor; }
  void set_fivalue(double fi) { this->fi_value = fi; }
  void setPenalizationFactor(float p) { this->penalization_factor = p; }
  // include/classes/fi.cuh:44-53.  Inu is a device pointer the term keeps
  // for its caller (the terms here do not read it).  setS / setDS hand the
  // term a device buffer of M*N floats for its values / gradient: the term
  // frees the one it held (gvm_free) and owns the new one.
  void setInu(gvm_complex* Inu) { this->Inu = Inu; }
  gvm_complex* getInu() { return this->Inu; }
  void setS(float* S);
  void setDS(float* DS);
  void setIteration(int iteration) { this->iteration = iteration; }
  int getIteration() { return this->iteration; }
  void setNormalize(bool normalize) { this->normalize = normalize; }

  virtual float calculateSecondDerivate() { return 0.0f; }
  // include/classes/fi.cuh:58-89: image indices, -Z lambda lookup, device_S /
  // device_DS allocation.
  virtual void configure(int penalizatorIndex, int imageIndex, int imageToAdd, bool normalize);

 protected:
  double fi_value = 0.0;
  float* device_S = nullptr;
  float* device_DS = nullptr;
  float penalization_factor = 1.0f;
  int imageIndex = 0;
  int iteration = 0;
  std::string name;
  gvm_complex* Inu = nullptr;
  int imageToAdd = 0;
  bool normalize = false;
};
