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
  // Replays captured HIP graphs of the value and gradient evaluations: an
  // evaluation whose key (pointers, epoch, flags, each term's graphKey) has
  // been run once is captured the next time and replayed after that.  Off by
  // default.  Results are identical to the uncaptured evaluation.
  void setGraphs(bool on);
  bool getGraphs() const { return graphs; }
  long getGraphLaunches() const { return graph_launches; }
  // With Chi2 first and up to GVM_MAX_FUSED_PRIORS priors after it, their
  // values are enqueued in one pass and the gradient's restart / AddToDPhi /
  // copy sequence is one kernel (gvm_priors_value_enqueue,
  // gvm_dphi_assemble), bit-identical to the term-by-term sequence.  On by
  // default.
  void setFusedTerms(bool on) { fused = on; }
  bool getFusedTerms() const { return fused; }

 private:
  struct GraphEntry {
    std::vector<int64_t> key;
    gvm_graph* graph = nullptr;
    std::vector<char> queued;
    int seen = 0;
    bool failed = false;
    long used = 0;
  };
  GraphEntry* graphEntry(std::vector<GraphEntry>& cache, const std::vector<int64_t>& key);
  std::vector<int64_t> baseKey() const;
  void clearGraphs();
  void gradientSequence(float* p, float* xi, int iter);
  bool fusable() const;
  void enqueueValues(float* p, std::vector<char>& queued);
  bool fused = true;
  bool graphs = false;
  std::vector<GraphEntry> value_graphs, grad_graphs;
  long graph_clock = 0, graph_launches = 0;

  std::vector<Fi*> fis;
  std::vector<float> fi_values;
  float* dphi = nullptr;
  long N = 0;
  long M = 0;
  int image_count = 1;
  long n_fcalls = 0, n_gcalls = 0;
};
