// Fi base, ObjectiveFunction and the objective terms (Chi2, Entropy,
// TotalSquaredVariation, Laplacian) over the libgvm_hip C ABI.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "gvm/factory.h"
#include "gvm/framework.h"
#include "gvm/objectivefunction.h"
#include "gvm/terms.h"

namespace gvm {

Globals& globals() {
  static Globals g;
  return g;
}

void check(gvm_status st, const char* what) {
  if (st != GVM_OK)
    throw Error(std::string(what) + " failed (status " + std::to_string(st) + "): " +
                    gvm_last_error(),
                st);
}

static float* dev_alloc_zero(long count) {
  void* p = nullptr;
  check(gvm_malloc(&p, sizeof(float) * (size_t)count), "gvm_malloc");
  check(gvm_vec_zero(globals().ctx, (float*)p, count), "gvm_vec_zero");
  return (float*)p;
}

static gvm_ctx* ctx() {
  gvm_ctx* c = globals().ctx;
  if (!c) throw Error("no current gvm context (gvm::globals().ctx)", GVM_ERR_STATE);
  return c;
}

}  // namespace gvm

using gvm::check;
using gvm::globals;

// ---------------------------------------------------------------------------
// Fi: include/classes/fi.cuh:13-104
// ---------------------------------------------------------------------------
Fi::Fi() : name("default") {}

Fi::~Fi() {
  gvm_free(device_S);
  gvm_free(device_DS);
}

namespace {
int64_t key_bits(float f) {
  int32_t b;
  std::memcpy(&b, &f, sizeof b);
  return b;
}
int64_t key_addr(const void* p) { return (int64_t)(uintptr_t)p; }
constexpr size_t kGraphCacheEntries = 8;
}  // namespace

void Fi::graphKey(std::vector<int64_t>& key) const {
  key.push_back(iteration > 0 ? 1 : 0);
  key.push_back(key_bits(penalization_factor));
  key.push_back(imageIndex);
  key.push_back(imageToAdd);
  key.push_back(normalize ? 1 : 0);
  key.push_back(key_addr(device_S));
  key.push_back(key_addr(device_DS));
}

void Fi::setS(float* S) {
  if (S != device_S) gvm_free(device_S);
  device_S = S;
}

void Fi::setDS(float* DS) {
  if (DS != device_DS) gvm_free(device_DS);
  device_DS = DS;
}

void Fi::configure(int penalizatorIndex, int imageIndex, int imageToAdd, bool normalize) {
  const auto& g = globals();
  this->imageIndex = imageIndex;
  this->imageToAdd = imageToAdd;
  this->normalize = normalize;
  if (imageIndex > g.image_count - 1 || imageToAdd > g.image_count - 1)
    throw gvm::Error("There is no image for the provided index " + name, GVM_ERR_ARG);
  if (penalizatorIndex != -1) {
    if (penalizatorIndex < 0)
      throw gvm::Error("invalid index for penalizator (" + name + ")", GVM_ERR_ARG);
    else if (penalizatorIndex > (int)g.penalizators.size() - 1)
      this->penalization_factor = 0.0f;
    else
      this->penalization_factor = g.penalizators[penalizatorIndex];
  }
  gvm_free(device_S);
  gvm_free(device_DS);
  device_S = gvm::dev_alloc_zero(g.M * g.N);
  device_DS = gvm::dev_alloc_zero(g.M * g.N);
}

// ---------------------------------------------------------------------------
// ObjectiveFunction: include/classes/objectivefunction.cuh:6-103
// ---------------------------------------------------------------------------
ObjectiveFunction::~ObjectiveFunction() {
  clearGraphs();
  gvm_free(dphi);
}

void ObjectiveFunction::clearGraphs() {
  for (auto* c : {&value_graphs, &grad_graphs}) {
    for (auto& e : *c) gvm_graph_destroy(e.graph);
    c->clear();
  }
}

void ObjectiveFunction::setGraphs(bool on) {
  if (!on) clearGraphs();
  graphs = on;
}

std::vector<int64_t> ObjectiveFunction::baseKey() const {
  long epoch = 0;
  check(gvm_epoch(gvm::ctx(), &epoch), "gvm_epoch");
  const auto& g = globals();
  return {key_addr(gvm::ctx()), epoch, g.flag_opt, g.M, g.N, g.image_count, key_addr(dphi),
          (int64_t)fis.size(), fusable() ? 1 : 0};
}

// The cache entry of `key` (created if new; the least recently used entry is
// dropped beyond kGraphCacheEntries).
ObjectiveFunction::GraphEntry* ObjectiveFunction::graphEntry(std::vector<GraphEntry>& cache,
                                                             const std::vector<int64_t>& key) {
  graph_clock++;
  for (auto& e : cache)
    if (e.key == key) {
      e.used = graph_clock;
      return &e;
    }
  if (cache.size() >= kGraphCacheEntries) {
    size_t old = 0;
    for (size_t i = 1; i < cache.size(); i++)
      if (cache[i].used < cache[old].used) old = i;
    gvm_graph_destroy(cache[old].graph);
    cache.erase(cache.begin() + (long)old);
  }
  GraphEntry e;
  e.key = key;
  e.used = graph_clock;
  cache.push_back(e);
  return &cache.back();
}

void ObjectiveFunction::addFi(Fi* fi) {
  if (fi->getPenalizationFactor()) {
    fis.push_back(fi);
    fi_values.push_back(0.0f);
  }
}

// Chi2 first, then priors that have a one-pass form.
bool ObjectiveFunction::fusable() const {
  if (!fused || fis.empty() || fis.size() - 1 > GVM_MAX_FUSED_PRIORS) return false;
  if (!dynamic_cast<Chi2*>(fis[0])) return false;
  gvm_prior_term t;
  for (size_t k = 1; k < fis.size(); k++)
    if (!fis[k]->priorTerm(&t, false)) return false;
  return true;
}

// Enqueues every term's value into slot k; queued[k] says whether term k
// produced one (an inactive prior is 0 without evaluation).
void ObjectiveFunction::enqueueValues(float* p, std::vector<char>& queued) {
  if (!fusable()) {
    for (size_t k = 0; k < fis.size(); k++) queued[k] = fis[k]->enqueueValue(p, (int)k) ? 1 : 0;
    return;
  }
  // Chi2 and the priors with one finishing launch (what Chi2::enqueueValue
  // and the priors' enqueueValue would give, in one pass over the image)
  queued[0] = 1;
  std::vector<gvm_prior_term> terms(fis.size() - 1);
  for (size_t k = 1; k < fis.size(); k++) {
    fis[k]->priorTerm(&terms[k - 1], false);
    queued[k] = terms[k - 1].active ? 1 : 0;
  }
  check(gvm_values_enqueue(gvm::ctx(), p, globals().flag_opt, fis[0]->getNormalize() ? 1 : 0, 0,
                           terms.data(), (int)terms.size(), 1),
        "gvm_values_enqueue");
}

double ObjectiveFunction::calcFunctionD(float* p) {
  double value = 0.0;
  n_fcalls++;
  bool async = !fis.empty() && fis.size() <= GVM_VALUE_SLOTS;
  for (Fi* fi : fis) async = async && fi->hasAsyncValue();
  if (async) {
    // the same evaluations in the same stream order, one synchronisation
    const int nv = (int)fis.size();
    std::vector<char> queued(fis.size());
    std::vector<double> raw(fis.size(), 0.0);
    GraphEntry* e = nullptr;
    if (graphs) {
      std::vector<int64_t> key = baseKey();
      key.push_back(1);
      key.push_back(key_addr(p));
      for (Fi* fi : fis) fi->graphKey(key);
      e = graphEntry(value_graphs, key);
      if (!e->graph && !e->failed && e->seen >= 1) {
        check(gvm_capture_begin(gvm::ctx()), "gvm_capture_begin");
        bool ok = true;
        try {
          enqueueValues(p, queued);
          check(gvm_values_fetch_enqueue(gvm::ctx(), nv), "gvm_values_fetch_enqueue");
        } catch (const gvm::Error&) {
          ok = false;
        }
        gvm_graph* gr = nullptr;
        if (gvm_capture_end(gvm::ctx(), &gr) == GVM_OK && ok) {
          e->graph = gr;
          e->queued = queued;
        } else {
          gvm_graph_destroy(gr);
          e->failed = true;
        }
      }
    }
    if (e && e->graph) {
      check(gvm_graph_launch(gvm::ctx(), e->graph), "gvm_graph_launch");
      check(gvm_values_host(gvm::ctx(), nv, raw.data()), "gvm_values_host");
      queued = e->queued;
      graph_launches++;
    } else {
      enqueueValues(p, queued);
      check(gvm_values_read(gvm::ctx(), nv, raw.data()), "gvm_values_read");
      if (e) e->seen++;
    }
    for (size_t k = 0; k < fis.size(); k++) {
      value += fis[k]->finishValue(queued[k] ? raw[k] : 0.0);
      fi_values[k] = fis[k]->get_fivalue();
    }
    return value;
  }
  for (size_t k = 0; k < fis.size(); k++) {
    value += fis[k]->calcFiD(p);
    fi_values[k] = fis[k]->get_fivalue();
  }
  return value;
}

float ObjectiveFunction::calcFunction(float* p) { return (float)calcFunctionD(p); }

void ObjectiveFunction::gradientSequence(float* p, float* xi, int iter) {
  if (fusable()) {
    // Chi2's restart and gradient, then the priors' gradients, the
    // AddToDPhi sequence and the copy to xi in one kernel
    for (Fi* fi : fis) fi->setIteration(iter);
    Chi2* chi2 = static_cast<Chi2*>(fis[0]);
    chi2->restartDGi();
    chi2->calcGi(p, xi);
    std::vector<gvm_prior_term> terms(fis.size() - 1);
    for (size_t k = 1; k < fis.size(); k++) fis[k]->priorTerm(&terms[k - 1], true);
    check(gvm_dphi_assemble(gvm::ctx(), p, chi2->getResult(), dphi, xi, terms.data(),
                            (int)terms.size()),
          "gvm_dphi_assemble");
    return;
  }
  restartDPhi();
  for (Fi* fi : fis) {
    fi->setIteration(iter);
    fi->calcGi(p, xi);
    fi->addToDphi(dphi);
  }
  copyDphiToXi(xi);
}

void ObjectiveFunction::calcGradient(float* p, float* xi, int iter) {
  n_gcalls++;
  GraphEntry* e = nullptr;
  if (graphs) {
    for (Fi* fi : fis) fi->setIteration(iter);
    std::vector<int64_t> key = baseKey();
    key.push_back(2);
    key.push_back(key_addr(p));
    key.push_back(key_addr(xi));
    for (Fi* fi : fis) fi->graphKey(key);
    e = graphEntry(grad_graphs, key);
    if (!e->graph && !e->failed && e->seen >= 1) {
      check(gvm_capture_begin(gvm::ctx()), "gvm_capture_begin");
      bool ok = true;
      try {
        gradientSequence(p, xi, iter);
      } catch (const gvm::Error&) {
        ok = false;
      }
      gvm_graph* gr = nullptr;
      if (gvm_capture_end(gvm::ctx(), &gr) == GVM_OK && ok) {
        e->graph = gr;
      } else {
        gvm_graph_destroy(gr);
        e->failed = true;
      }
    }
    if (e->graph) {
      check(gvm_graph_launch(gvm::ctx(), e->graph), "gvm_graph_launch");
      graph_launches++;
      return;
    }
  }
  gradientSequence(p, xi, iter);
  if (e) e->seen++;
}

void ObjectiveFunction::restartDPhi() {
  for (Fi* fi : fis) fi->restartDGi();
  check(gvm_vec_zero(gvm::ctx(), dphi, M * N * image_count), "gvm_vec_zero(dphi)");
}

void ObjectiveFunction::copyDphiToXi(float* xi) {
  check(gvm_vec_copy(gvm::ctx(), xi, dphi, M * N * image_count), "gvm_vec_copy(xi)");
}

Fi* ObjectiveFunction::getFiByName(std::string fi_name) {
  for (Fi* fi : fis)
    if (fi->getName() == fi_name) return fi;
  return nullptr;
}

void ObjectiveFunction::configure(long N, long M, int I) {
  setN(N);
  setM(M);
  setImageCount(I);
  gvm_free(dphi);
  dphi = gvm::dev_alloc_zero(M * N * I);
}

namespace {
ObjectiveFunction* CreateObjectiveFunction() { return new ObjectiveFunction; }
const bool RegisteredObjectiveFunction =
    registerCreationFunction<ObjectiveFunction, std::string>("ObjectiveFunction",
                                                             CreateObjectiveFunction);
}  // namespace

// ---------------------------------------------------------------------------
// Chi2: src/chi2.cu:13-92
// ---------------------------------------------------------------------------
Chi2::Chi2() { this->name = "Chi2"; }

Chi2::~Chi2() { gvm_free(result_dchi2); }

void Chi2::configure(int penalizatorIndex, int imageIndex, int imageToAdd, bool normalize) {
  const auto& g = globals();
  this->imageIndex = imageIndex;
  this->imageToAdd = imageToAdd;
  this->normalize = normalize;
  if (penalizatorIndex != -1) {
    if (penalizatorIndex > (int)g.penalizators.size() - 1 || penalizatorIndex < 0)
      throw gvm::Error("invalid index for penalizator (" + name + ")", GVM_ERR_ARG);
    this->penalization_factor = g.penalizators[penalizatorIndex];
  }
  gvm_free(result_dchi2);
  result_dchi2 = gvm::dev_alloc_zero(g.M * g.N * g.image_count);
}

double Chi2::calcFiD(float* p) {
  double v = 0.0;
  check(gvm_chi2(gvm::ctx(), p, globals().flag_opt, normalize ? 1 : 0, &v), "gvm_chi2");
  set_fivalue(v);
  return (double)penalization_factor * v;
}

float Chi2::calcFi(float* p) { return (float)calcFiD(p); }

void Chi2::graphKey(std::vector<int64_t>& key) const {
  Fi::graphKey(key);
  key.push_back(key_addr(result_dchi2));
}

bool Chi2::enqueueValue(float* p, int slot) {
  check(gvm_chi2_enqueue(gvm::ctx(), p, globals().flag_opt, normalize ? 1 : 0, slot),
        "gvm_chi2_enqueue");
  return true;
}

void Chi2::calcGi(float* p, float* xi) {
  check(gvm_dchi2(gvm::ctx(), p, result_dchi2, globals().flag_opt, normalize ? 1 : 0),
        "gvm_dchi2");
}

void Chi2::restartDGi() {
  const auto& g = globals();
  check(gvm_vec_zero(gvm::ctx(), result_dchi2, g.M * g.N * g.image_count), "gvm_vec_zero");
}

// src/chi2.cu:60-70: with more than one image dphi is cleared and replaced by
// the chi2 gradient, so Chi2 has to be the first term.
void Chi2::addToDphi(float* device_dphi) {
  const auto& g = globals();
  if (g.image_count == 1) check(gvm_add_to_dphi(gvm::ctx(), device_dphi, result_dchi2, 0), "AddToDPhi");
  if (g.image_count > 1) {
    check(gvm_vec_zero(gvm::ctx(), device_dphi, g.M * g.N * g.image_count), "gvm_vec_zero");
    check(gvm_vec_copy(gvm::ctx(), device_dphi, result_dchi2, g.M * g.N * g.image_count),
          "gvm_vec_copy");
  }
}

// ---------------------------------------------------------------------------
// Priors.  Host launchers src/functions.cu:4722-4964: value and gradient are
// zero at iteration 0 or with lambda 0; the gradient is computed only when
// flag_opt % 2 == imageIndex.
// ---------------------------------------------------------------------------
static bool prior_active(int iteration, float lambda) { return iteration > 0 && lambda != 0.0f; }

// The one-pass descriptor of a prior (value gating; for a gradient also the
// flag_opt gating of calcGi).
static void fill_prior_term(gvm_prior_term* t, int kind, int imageIndex, int imageToAdd,
                            float prior, float eta, float lambda, int iteration, bool gradient) {
  t->kind = kind;
  t->index = imageIndex;
  t->image_to_add = imageToAdd;
  t->prior = prior;
  t->eta = eta;
  t->lambda = lambda;
  t->pad = 0.0f;
  bool on = prior_active(iteration, lambda);
  if (gradient) on = on && globals().flag_opt % 2 == imageIndex;
  t->active = on ? 1 : 0;
}

bool Entropy::priorTerm(gvm_prior_term* t, bool gradient) const {
  fill_prior_term(t, GVM_PRIOR_ENTROPY, imageIndex, imageToAdd, prior_value, eta,
                  penalization_factor, iteration, gradient);
  return true;
}
bool TotalSquaredVariationP::priorTerm(gvm_prior_term* t, bool gradient) const {
  fill_prior_term(t, GVM_PRIOR_TSV, imageIndex, imageToAdd, 0.f, 0.f, penalization_factor,
                  iteration, gradient);
  return true;
}
bool Laplacian::priorTerm(gvm_prior_term* t, bool gradient) const {
  fill_prior_term(t, GVM_PRIOR_LAPLACIAN, imageIndex, imageToAdd, 0.f, 0.f, penalization_factor,
                  iteration, gradient);
  return true;
}

Entropy::Entropy() : prior_value(1.0f), eta(-1.0f) { this->name = "Entropy"; }
Entropy::Entropy(float prior_value) : prior_value(prior_value), eta(-1.0f) { this->name = "Entropy"; }
Entropy::Entropy(float prior_value, float eta) : prior_value(prior_value), eta(eta) {
  this->name = "Entropy";
}

double Entropy::calcFiD(float* p) {
  double v = 0.0;
  if (prior_active(iteration, penalization_factor))
    check(gvm_entropy_value(gvm::ctx(), p, imageIndex, prior_value, eta, &v), "gvm_entropy_value");
  set_fivalue(v);
  return (double)penalization_factor * v;
}
float Entropy::calcFi(float* p) { return (float)calcFiD(p); }
void Entropy::graphKey(std::vector<int64_t>& key) const {
  Fi::graphKey(key);
  key.push_back(key_bits(prior_value));
  key.push_back(key_bits(eta));
}

bool Entropy::enqueueValue(float* p, int slot) {
  if (!prior_active(iteration, penalization_factor)) return false;
  check(gvm_entropy_value_enqueue(gvm::ctx(), p, imageIndex, prior_value, eta, slot),
        "gvm_entropy_value_enqueue");
  return true;
}

void Entropy::calcGi(float* p, float* xi) {
  if (prior_active(iteration, penalization_factor) && globals().flag_opt % 2 == imageIndex)
    check(gvm_entropy_grad(gvm::ctx(), p, device_DS, imageIndex, prior_value, eta,
                           penalization_factor),
          "gvm_entropy_grad");
}
void Entropy::restartDGi() {
  check(gvm_vec_zero(gvm::ctx(), device_DS, globals().M * globals().N), "gvm_vec_zero");
}
void Entropy::addToDphi(float* device_dphi) {
  check(gvm_add_to_dphi(gvm::ctx(), device_dphi, device_DS, imageToAdd), "AddToDPhi");
}

TotalSquaredVariationP::TotalSquaredVariationP() { this->name = "Total Squared Variation"; }
double TotalSquaredVariationP::calcFiD(float* p) {
  double v = 0.0;
  if (prior_active(iteration, penalization_factor))
    check(gvm_tsv_value(gvm::ctx(), p, imageIndex, &v), "gvm_tsv_value");
  set_fivalue(v);
  return (double)penalization_factor * v;
}
float TotalSquaredVariationP::calcFi(float* p) { return (float)calcFiD(p); }
bool TotalSquaredVariationP::enqueueValue(float* p, int slot) {
  if (!prior_active(iteration, penalization_factor)) return false;
  check(gvm_tsv_value_enqueue(gvm::ctx(), p, imageIndex, slot), "gvm_tsv_value_enqueue");
  return true;
}
void TotalSquaredVariationP::calcGi(float* p, float* xi) {
  if (prior_active(iteration, penalization_factor) && globals().flag_opt % 2 == imageIndex)
    check(gvm_tsv_grad(gvm::ctx(), p, device_DS, imageIndex, penalization_factor), "gvm_tsv_grad");
}
void TotalSquaredVariationP::restartDGi() {
  check(gvm_vec_zero(gvm::ctx(), device_DS, globals().M * globals().N), "gvm_vec_zero");
}
void TotalSquaredVariationP::addToDphi(float* device_dphi) {
  check(gvm_add_to_dphi(gvm::ctx(), device_dphi, device_DS, imageToAdd), "AddToDPhi");
}

Laplacian::Laplacian() { this->name = "Laplacian"; }
double Laplacian::calcFiD(float* p) {
  double v = 0.0;
  if (prior_active(iteration, penalization_factor))
    check(gvm_laplacian_value(gvm::ctx(), p, imageIndex, &v), "gvm_laplacian_value");
  set_fivalue(v);
  return (double)penalization_factor * v;
}
float Laplacian::calcFi(float* p) { return (float)calcFiD(p); }
bool Laplacian::enqueueValue(float* p, int slot) {
  if (!prior_active(iteration, penalization_factor)) return false;
  check(gvm_laplacian_value_enqueue(gvm::ctx(), p, imageIndex, slot),
        "gvm_laplacian_value_enqueue");
  return true;
}
void Laplacian::calcGi(float* p, float* xi) {
  if (prior_active(iteration, penalization_factor) && globals().flag_opt % 2 == imageIndex)
    check(gvm_laplacian_grad(gvm::ctx(), p, device_DS, imageIndex, penalization_factor),
          "gvm_laplacian_grad");
}
void Laplacian::restartDGi() {
  check(gvm_vec_zero(gvm::ctx(), device_DS, globals().M * globals().N), "gvm_vec_zero");
}
void Laplacian::addToDphi(float* device_dphi) {
  check(gvm_add_to_dphi(gvm::ctx(), device_dphi, device_DS, imageToAdd), "AddToDPhi");
}

namespace {
Fi* CreateChi2() { return new Chi2; }
Fi* CreateEntropy() { return new Entropy; }
Fi* CreateTSV() { return new TotalSquaredVariationP; }
Fi* CreateLaplacian() { return new Laplacian; }
const bool RegisteredChi2 = registerCreationFunction<Fi, std::string>("Chi2", CreateChi2);
const bool RegisteredEntropy = registerCreationFunction<Fi, std::string>("Entropy", CreateEntropy);
const bool RegisteredEntropyInt = registerCreationFunction<Fi, int>(0, CreateEntropy);
const bool RegisteredTSV =
    registerCreationFunction<Fi, std::string>("TotalSquaredVariation", CreateTSV);
const bool RegisteredLaplacian =
    registerCreationFunction<Fi, std::string>("Laplacian", CreateLaplacian);
}  // namespace
