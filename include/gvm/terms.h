// Objective terms on the accelerated path, registered under the reference's
// factory ids: "Chi2" (src/chi2.cu), "Entropy" (src/entropy.cu),
// "TotalSquaredVariation" (src/totalsquaredvariation.cu), "Laplacian"
// (src/laplacian.cu).
#pragma once

#include "gvm/fi.h"

class Chi2 : public Fi {
 public:
  Chi2();
  ~Chi2() override;
  float calcFi(float* p) override;
  double calcFiD(float* p) override;
  bool hasAsyncValue() const override { return true; }
  bool enqueueValue(float* p, int slot) override;
  void graphKey(std::vector<int64_t>& key) const override;
  void calcGi(float* p, float* xi) override;
  void restartDGi() override;
  void addToDphi(float* device_dphi) override;
  void configure(int penalizatorIndex, int imageIndex, int imageToAdd, bool normalize) override;
  void setFgScale(float fg_scale) override { this->fg_scale = fg_scale; }
  float getFgScale() override { return this->fg_scale; }
  // the device gradient buffer calcGi accumulates into
  float* getResult() const { return result_dchi2; }

 private:
  float* result_dchi2 = nullptr;
  float fg_scale = 1.0f;
};

class Entropy : public Fi {
 public:
  Entropy();
  explicit Entropy(float prior_value);
  Entropy(float prior_value, float eta);
  float calcFi(float* p) override;
  double calcFiD(float* p) override;
  bool hasAsyncValue() const override { return true; }
  bool enqueueValue(float* p, int slot) override;
  void graphKey(std::vector<int64_t>& key) const override;
  void calcGi(float* p, float* xi) override;
  void restartDGi() override;
  void addToDphi(float* device_dphi) override;
  bool priorTerm(gvm_prior_term* t, bool gradient) const override;
  float getPrior() { return prior_value; }
  void setPrior(float prior_value) override { this->prior_value = prior_value; }
  float getEta() override { return eta; }
  void setEta(float eta) override { this->eta = eta; }

 private:
  float prior_value;
  float eta;
};

class TotalSquaredVariationP : public Fi {
 public:
  TotalSquaredVariationP();
  bool priorTerm(gvm_prior_term* t, bool gradient) const override;
  float calcFi(float* p) override;
  double calcFiD(float* p) override;
  bool hasAsyncValue() const override { return true; }
  bool enqueueValue(float* p, int slot) override;
  void calcGi(float* p, float* xi) override;
  void restartDGi() override;
  void addToDphi(float* device_dphi) override;
};

class Laplacian : public Fi {
 public:
  Laplacian();
  bool priorTerm(gvm_prior_term* t, bool gradient) const override;
  float calcFi(float* p) override;
  double calcFiD(float* p) override;
  bool hasAsyncValue() const override { return true; }
  bool enqueueValue(float* p, int slot) override;
  void calcGi(float* p, float* xi) override;
  void restartDGi() override;
  void addToDphi(float* device_dphi) override;
};
