// C entry points of the framework (include/gvm/fw.h): composition of the
// objective by factory id, evaluation, and optimisation, as src/main.cu
// :147-203 and src/mfs.cu:978-1012 do it.
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "gvm/factory.h"
#include "gvm/framework.h"
#include "gvm/fw.h"
#include "gvm/optimizers.h"
#include "gvm/terms.h"

struct gvm_fw_objective {
  gvm_ctx* ctx = nullptr;
  gvm_fw_config cfg{};
  std::vector<std::unique_ptr<Fi>> terms;
  std::unique_ptr<ObjectiveFunction> of;
  imageMap* mapping = nullptr;
  bool lbfgs_host_scalars = false;
  std::vector<double> replay;  // line-search step lengths to replay (empty: search)
  std::vector<double> steps;   // the step lengths the last optimisation applied
  ~gvm_fw_objective() { delete[] mapping; }
};

namespace {
thread_local std::string g_fw_error;

gvm_status fail(const std::exception& e) {
  g_fw_error = e.what();
  if (auto* ge = dynamic_cast<const gvm::Error*>(&e)) return ge->status;
  return GVM_ERR_ARG;
}

// Globals of the reference that the composition sets
// (src/mfs.cu:158-180, src/functions.cu:181-301).
void make_current(gvm_fw_objective* obj) {
  auto& g = gvm::globals();
  gvm_geometry geo;
  gvm::check(gvm_get_geometry(obj->ctx, &geo), "gvm_get_geometry");
  g.ctx = obj->ctx;
  g.M = geo.M;
  g.N = geo.N;
  g.image_count = geo.image_count;
  g.penalizators.assign(obj->cfg.lambdas, obj->cfg.lambdas + obj->cfg.nlambdas);
  g.eta = obj->cfg.eta;
  g.initial_values = {obj->cfg.initial_values[0], obj->cfg.initial_values[1]};
  g.nopositivity = obj->cfg.nopositivity != 0;
  g.flag_opt = obj->cfg.flag_opt;
}
}  // namespace

extern "C" const char* gvm_fw_last_error(void) { return g_fw_error.c_str(); }

extern "C" gvm_status gvm_fw_objective_create(gvm_ctx* ctx, const gvm_fw_config* cfg,
                                              gvm_fw_objective** out) {
  if (!ctx || !cfg || !out || cfg->nterms <= 0 || cfg->nterms > GVM_FW_MAX_TERMS ||
      cfg->nlambdas < 0 || cfg->nlambdas > GVM_FW_MAX_TERMS) {
    g_fw_error = "gvm_fw_objective_create: bad argument";
    return GVM_ERR_ARG;
  }
  *out = nullptr;
  auto obj = std::make_unique<gvm_fw_objective>();
  obj->ctx = ctx;
  obj->cfg = *cfg;
  try {
    make_current(obj.get());
    obj->of.reset(createObject<ObjectiveFunction, std::string>("ObjectiveFunction"));
    for (int t = 0; t < cfg->nterms; t++) {
      const gvm_fw_term& tc = cfg->terms[t];
      if (!tc.name) throw gvm::Error("term without a name", GVM_ERR_ARG);
      std::unique_ptr<Fi> fi(createObject<Fi, std::string>(std::string(tc.name)));
      fi->configure(tc.penalizator, tc.image_index, tc.image_to_add, tc.normalize != 0);
      if (std::string(tc.name) == "Entropy") fi->setPrior(tc.prior);
      obj->of->addFi(fi.get());
      obj->terms.push_back(std::move(fi));
    }
    auto& g = gvm::globals();
    obj->of->configure(g.N, g.M, g.image_count);
    obj->mapping = makeFunctionMapping(g.image_count, g.nopositivity);
  } catch (const std::exception& e) {
    return fail(e);
  }
  *out = obj.release();
  return GVM_OK;
}

extern "C" gvm_status gvm_fw_objective_destroy(gvm_fw_objective* obj) {
  delete obj;
  return GVM_OK;
}

extern "C" gvm_status gvm_fw_value(gvm_fw_objective* obj, float* p_dev, double* value) {
  if (!obj || !p_dev || !value) return GVM_ERR_ARG;
  try {
    make_current(obj);
    *value = obj->of->calcFunctionD(p_dev);
  } catch (const std::exception& e) {
    return fail(e);
  }
  return GVM_OK;
}

extern "C" gvm_status gvm_fw_gradient(gvm_fw_objective* obj, float* p_dev, float* xi_dev,
                                      int iter) {
  if (!obj || !p_dev || !xi_dev) return GVM_ERR_ARG;
  try {
    make_current(obj);
    obj->of->calcGradient(p_dev, xi_dev, iter);
  } catch (const std::exception& e) {
    return fail(e);
  }
  return GVM_OK;
}

extern "C" gvm_status gvm_fw_term_values(gvm_fw_objective* obj, double* values, int cap, int* n) {
  if (!obj || !n) return GVM_ERR_ARG;
  auto fis = obj->of->getFi();
  *n = (int)fis.size();
  for (int k = 0; k < (int)fis.size() && k < cap; k++) values[k] = fis[k]->get_fivalue_d();
  return GVM_OK;
}

extern "C" gvm_status gvm_fw_optimize(gvm_fw_objective* obj, const char* optimizer,
                                      int total_iterations, double ftol, double gtol, int lbfgs_k,
                                      float* p_dev, gvm_fw_result* res, double* traj, int cap) {
  return gvm_fw_optimize_timed(obj, optimizer, total_iterations, ftol, gtol, lbfgs_k, p_dev, res,
                               traj, nullptr, cap);
}

extern "C" gvm_status gvm_fw_optimize_timed(gvm_fw_objective* obj, const char* optimizer,
                                            int total_iterations, double ftol, double gtol,
                                            int lbfgs_k, float* p_dev, gvm_fw_result* res,
                                            double* traj, double* traj_seconds, int cap) {
  if (!obj || !optimizer || !p_dev || !res || total_iterations < 0) return GVM_ERR_ARG;
  try {
    make_current(obj);
    std::unique_ptr<Optimizer> opt(createObject<Optimizer, std::string>(std::string(optimizer)));
    Image image(p_dev, gvm::globals().image_count);
    image.setFunctionMapping(obj->mapping);
    opt->setObjectiveFunction(obj->of.get());
    opt->setTotalIterations(total_iterations);
    opt->setFTol((float)ftol);
    opt->setGTol((float)gtol);
    if (lbfgs_k > 0) opt->setK(lbfgs_k);
    if (auto* lb = dynamic_cast<LBFGS*>(opt.get())) lb->setHostScalars(obj->lbfgs_host_scalars);
    opt->setImage(&image);
    opt->setFlag(obj->cfg.flag_opt);
    const long f0 = obj->of->getFunctionCalls(), g0 = obj->of->getGradientCalls();
    const auto t0 = std::chrono::steady_clock::now();
    setLineSearchReplay(obj->replay.data(), (int)obj->replay.size());
    struct ReplayReset {
      ~ReplayReset() { setLineSearchReplay(nullptr, 0); }
    } replay_reset;
    obj->steps.clear();
    opt->optimize();
    obj->steps = lineSearchSteps();
    gvm::check(gvm_sync(obj->ctx), "gvm_sync");
    const auto t1 = std::chrono::steady_clock::now();
    res->iterations = opt->getCurrentIteration();
    res->exit_reason = opt->getExitReason();
    res->final_value = opt->getFinalValue();
    res->function_calls = obj->of->getFunctionCalls() - f0;
    res->gradient_calls = obj->of->getGradientCalls() - g0;
    res->seconds = std::chrono::duration<double>(t1 - t0).count();
    if (traj) {
      const auto& tr = opt->getTrajectory();
      for (int k = 0; k < (int)tr.size() && k < cap; k++) traj[k] = tr[k];
    }
    if (traj_seconds) {
      const auto& ts = opt->getTrajectorySeconds();
      for (int k = 0; k < (int)ts.size() && k < cap; k++) traj_seconds[k] = ts[k];
    }
  } catch (const std::exception& e) {
    return fail(e);
  }
  return GVM_OK;
}

extern "C" gvm_status gvm_fw_set_line_search_replay(gvm_fw_objective* obj, const double* xmins,
                                                    int n) {
  if (!obj || n < 0 || (n > 0 && !xmins)) {
    g_fw_error = "gvm_fw_set_line_search_replay: bad argument";
    return GVM_ERR_ARG;
  }
  obj->replay.assign(xmins, xmins + n);
  return GVM_OK;
}

extern "C" gvm_status gvm_fw_line_search_steps(const gvm_fw_objective* obj, double* xmins,
                                                int cap, int* n) {
  if (!obj || !n || cap < 0 || (cap > 0 && !xmins)) {
    g_fw_error = "gvm_fw_line_search_steps: bad argument";
    return GVM_ERR_ARG;
  }
  *n = (int)obj->steps.size();
  for (int k = 0; k < *n && k < cap; k++) xmins[k] = obj->steps[k];
  return GVM_OK;
}

extern "C" gvm_status gvm_fw_set_flag(gvm_fw_objective* obj, int flag_opt) {
  if (!obj || flag_opt < 0) {
    g_fw_error = "gvm_fw_set_flag: bad argument";
    return GVM_ERR_ARG;
  }
  obj->cfg.flag_opt = flag_opt;
  return GVM_OK;
}

extern "C" gvm_status gvm_fw_set_graphs(gvm_fw_objective* obj, int on) {
  if (!obj || !obj->of) {
    g_fw_error = "gvm_fw_set_graphs: bad argument";
    return GVM_ERR_ARG;
  }
  obj->of->setGraphs(on != 0);
  return GVM_OK;
}

extern "C" gvm_status gvm_fw_set_fused_terms(gvm_fw_objective* obj, int on) {
  if (!obj || !obj->of) {
    g_fw_error = "gvm_fw_set_fused_terms: bad argument";
    return GVM_ERR_ARG;
  }
  obj->of->setFusedTerms(on != 0);
  return GVM_OK;
}

extern "C" gvm_status gvm_fw_set_lbfgs_host_scalars(gvm_fw_objective* obj, int on) {
  if (!obj) {
    g_fw_error = "gvm_fw_set_lbfgs_host_scalars: null objective";
    return GVM_ERR_ARG;
  }
  obj->lbfgs_host_scalars = on != 0;
  return GVM_OK;
}

extern "C" gvm_status gvm_fw_graph_launches(gvm_fw_objective* obj, long* n) {
  if (!obj || !obj->of || !n) {
    g_fw_error = "gvm_fw_graph_launches: bad argument";
    return GVM_ERR_ARG;
  }
  *n = obj->of->getGraphLaunches();
  return GVM_OK;
}

extern "C" int gvm_fw_registered(const char* kind, const char* name) {
  if (!kind || !name) return 0;
  const std::string k(kind), n(name);
  try {
    if (k == "Fi") {
      std::unique_ptr<Fi> p(createObject<Fi, std::string>(n));
      return p ? 1 : 0;
    }
    if (k == "Optimizer") {
      std::unique_ptr<Optimizer> p(createObject<Optimizer, std::string>(n));
      return p ? 1 : 0;
    }
    if (k == "ObjectiveFunction") {
      std::unique_ptr<ObjectiveFunction> p(createObject<ObjectiveFunction, std::string>(n));
      return p ? 1 : 0;
    }
  } catch (const gvm::FactoryError&) {
    return 0;
  }
  return 0;
}
