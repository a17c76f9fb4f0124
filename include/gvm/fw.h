/* C entry points of the C++ framework (libgvm): compose the objective from
 * factory ids the way src/main.cu:147-203 does, evaluate it, and run an
 * optimiser, all on a gvm_ctx.  For harnesses and bindings (Python ctypes,
 * cgo, ...) that cannot instantiate the C++ classes themselves. */
#ifndef GVM_FW_H
#define GVM_FW_H

#include "gvm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GVM_FW_MAX_TERMS 8

typedef struct {
  const char* name;   /* factory id: "Chi2", "Entropy", "TotalSquaredVariation", "Laplacian" */
  int penalizator;    /* index into lambdas, or -1 (keep the term's default lambda 1) */
  int image_index;
  int image_to_add;
  float prior;        /* Entropy prior G (setPrior) */
  int normalize;
} gvm_fw_term;

typedef struct {
  int nterms;
  gvm_fw_term terms[GVM_FW_MAX_TERMS];
  int nlambdas;
  float lambdas[GVM_FW_MAX_TERMS]; /* the -Z list */
  float initial_values[2];         /* initial_values after src/mfs.cu:167-179 */
  float eta;                       /* -e (default -1) */
  int nopositivity;                /* -x */
  int flag_opt;
} gvm_fw_config;

typedef struct gvm_fw_objective gvm_fw_objective;

/* Makes ctx current, creates the terms via createObject<Fi,std::string>,
 * configures them and adds them to an ObjectiveFunction ("ObjectiveFunction"). */
gvm_status gvm_fw_objective_create(gvm_ctx* ctx, const gvm_fw_config* cfg,
                                   gvm_fw_objective** out);
gvm_status gvm_fw_objective_destroy(gvm_fw_objective* obj);
/* ObjectiveFunction::calcFunction (FP64 sum of the terms). p_dev is clipped
 * in place by the chi2 term, as in the reference. */
gvm_status gvm_fw_value(gvm_fw_objective* obj, float* p_dev, double* value);
/* ObjectiveFunction::calcGradient(p, xi, iter). */
gvm_status gvm_fw_gradient(gvm_fw_objective* obj, float* p_dev, float* xi_dev, int iter);
/* Values of the individual terms from the last gvm_fw_value (without lambda). */
gvm_status gvm_fw_term_values(gvm_fw_objective* obj, double* values, int cap, int* n);

typedef struct {
  int iterations;
  int exit_reason;
  double final_value;
  long function_calls;
  long gradient_calls;
  double seconds;
} gvm_fw_result;

/* Runs createObject<Optimizer,std::string>(optimizer) ("CG-FRPRMN" or
 * "CG-LBFGS") on the image p_dev (float[2*M*N], device, updated in place),
 * like optimizationOrder (src/main.cu:88-98).  lbfgs_k <= 0 keeps the
 * default memory.  traj (optional, capacity cap) receives the value after
 * each line search. */
gvm_status gvm_fw_optimize(gvm_fw_objective* obj, const char* optimizer, int total_iterations,
                           double ftol, double gtol, int lbfgs_k, float* p_dev,
                           gvm_fw_result* res, double* traj, int cap);
/* gvm_fw_optimize that also writes, into traj_seconds (optional, capacity
 * cap), the wall-clock seconds from the start of the run to each trajectory
 * entry (for time-to-image measurements). */
gvm_status gvm_fw_optimize_timed(gvm_fw_objective* obj, const char* optimizer,
                                 int total_iterations, double ftol, double gtol, int lbfgs_k,
                                 float* p_dev, gvm_fw_result* res, double* traj,
                                 double* traj_seconds, int cap);

/* Optimizer::setFlag for the next gvm_fw_optimize / value / gradient: the
 * schedule flag_opt (image flag_opt % 2 is optimised; flag_opt > 0 zeroes
 * alpha where I_nu_0 < threshold).  MFS::run (src/mfs.cu:995-1009) runs the
 * optimiser with flags 0, 1, 2, 3 in turn on the same image. */
gvm_status gvm_fw_set_flag(gvm_fw_objective* obj, int flag_opt);

/* ObjectiveFunction::setGraphs: replay captured HIP graphs of repeated value
 * and gradient evaluations (off by default; results are identical). */
gvm_status gvm_fw_set_graphs(gvm_fw_objective* obj, int on);
/* Graph replays so far (value and gradient evaluations served by a graph). */
gvm_status gvm_fw_graph_launches(gvm_fw_objective* obj, long* n);
/* CG-LBFGS's two-loop recursion (src/lbfgs.cu:252-341) reads every dot product
 * back to the host when on != 0, as the reference does; by default (0) its
 * scalars stay on the device and the recursion runs without host round
 * trips.  Both give the same search direction, bit for bit. */
gvm_status gvm_fw_set_lbfgs_host_scalars(gvm_fw_objective* obj, int on);
/* ObjectiveFunction::setFusedTerms: the priors after Chi2 in one pass (on by
 * default; off gives the term-by-term launches, with identical results). */
gvm_status gvm_fw_set_fused_terms(gvm_fw_objective* obj, int on);
/* Line-search replay (no counterpart in the reference): the next
 * gvm_fw_optimize takes xmins[0..n) in turn as linmin's step lengths instead of
 * bracketing and minimising (src/linmin.cu:52-110), and fails with
 * GVM_ERR_STATE if it needs more.  n = 0 restores the search.  Replaying the
 * steps another implementation accepted (the FP64 oracle's) compares two
 * optimiser paths without the line search's branch decisions. */
gvm_status gvm_fw_set_line_search_replay(gvm_fw_objective* obj, const double* xmins, int n);
/* The step lengths the last gvm_fw_optimize applied, searched or replayed (no
 * counterpart in the reference): *n receives their count, xmins[0..min(n,
 * cap)) the values.  Feeding them to gvm_fw_set_line_search_replay of another
 * objective (another adjoint form) runs the same path there. */
gvm_status gvm_fw_line_search_steps(const gvm_fw_objective* obj, double* xmins, int cap, int* n);
const char* gvm_fw_last_error(void);

/* 1 if `name` is registered with the factory of `kind` ("Fi", "Optimizer",
 * "ObjectiveFunction"), else 0. */
int gvm_fw_registered(const char* kind, const char* name);

#ifdef __cplusplus
}
#endif
#endif
