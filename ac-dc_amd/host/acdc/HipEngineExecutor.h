// HipEngineExecutor.h -- the gfx950 executor behind the reference's engine
// interface.
//
// AbstractEngineExecutor<L,D> restates cpp/src/solver/AbstractEngineExecutor.h:13-39;
// HipEngineExecutor<L,D> takes the place of MulticoreEngineExecutor<L,D>
// (cpp/src/solver/MulticoreEngineExecutor.h:25-97): same constructor
// arguments, the same side effects on the caller's ProblemData (x zeroed and
// kept current, Li filled), the same four operations, but every operation is
// a call into the C ABI of include/acdc_hip.h.  Errors, which the reference
// never reports, surface as std::runtime_error carrying acdc_last_error().
#pragma once
#include <algorithm>
#include <cstdint>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/acdc_hip.h"
#include "losses.h"
#include "structures.h"

template <typename L, typename D>
class AbstractEngineExecutor {
 public:
  OptimizationSettings *settings = nullptr;
  virtual ~AbstractEngineExecutor() {}
  virtual std::string getResultLogHeaders() { return "Function is not implemented"; }
  virtual std::string getResultLogRow(OptimizationStatistics &) { return "Function is not implemented"; }
  virtual void executeBulkOfIterations() {}
  virtual void recomputeResiduals() {}
  virtual D getObjectiveValue() { return 0; }
};

namespace acdc_host {

inline void check(acdc_status st, const char *what) {
  if (st != ACDC_OK) throw std::runtime_error(std::string(what) + ": " + acdc_last_error());
}

template <typename L>
std::vector<int32_t> to_i32(const std::vector<L> &v, const char *what) {
  std::vector<int32_t> out(v.size());
  for (size_t i = 0; i < v.size(); ++i) {
    const long long e = (long long)v[i];
    if (e < std::numeric_limits<int32_t>::min() || e > std::numeric_limits<int32_t>::max())
      throw std::runtime_error(std::string(what) + ": index exceeds int32");
    out[i] = (int32_t)e;
  }
  return out;
}

template <typename D>
std::vector<double> to_f64(const std::vector<D> &v) {
  return std::vector<double>(v.begin(), v.end());
}

// The ESO scaling a solver gets from settings.hipBetaMode (structures.h):
// forcedSigma > 0 first (ClusterSolver.cpp:203-205); the reference rule --
// beta-less Lasso/NNLS (MulticoreSquareLoss.h:27-40 ignores sigma), SVM with
// the caller's sigma or the omega rule (MulticoreSVMDualLoss.h:48-61,
// MultiCoreSolver.cpp:108-110); else the acdc_beta_mode asked for.
inline void resolve_beta(int loss, const OptimizationSettings &set, double sigma,
                         acdc_settings *s) {
  s->forced_beta = 1.0;
  if (set.forcedSigma > 0) {
    s->beta_mode = ACDC_BETA_FORCED;
    s->forced_beta = set.forcedSigma;
  } else if (set.hipBetaMode == kHipBetaReference) {
    if (loss != ACDC_LOSS_SVM_DUAL_HINGE) {
      s->beta_mode = ACDC_BETA_NONE;
    } else if (sigma > 0) {
      s->beta_mode = ACDC_BETA_FORCED;
      s->forced_beta = sigma;
    } else {
      s->beta_mode = ACDC_BETA_OMEGA;
    }
  } else if (set.hipBetaMode == ACDC_BETA_FORCED) {
    if (!(sigma > 0)) throw std::runtime_error("hipBetaMode = ACDC_BETA_FORCED needs instance.sigma > 0");
    s->beta_mode = ACDC_BETA_FORCED;
    s->forced_beta = sigma;
  } else {
    s->beta_mode = set.hipBetaMode;
  }
}

}  // namespace acdc_host

template <typename L, typename D>
class HipEngineExecutor : public AbstractEngineExecutor<L, D> {
 public:
  ProblemData<L, D> &instance;

  HipEngineExecutor(ProblemData<L, D> &_instance, OptimizationSettings *_settings)
      : instance(_instance) {
    using namespace acdc_host;
    this->settings = _settings;
    const int loss = _settings->lossFunction;
    acdc_status st;
    if (loss == ACDC_LOSS_SVM_DUAL_HINGE) {
      const std::vector<int32_t> rp = to_i32(instance.A_csr_row_ptr, "A_csr_row_ptr");
      const std::vector<int32_t> ci = to_i32(instance.A_csr_col_idx, "A_csr_col_idx");
      const std::vector<double> v = to_f64(instance.A_csr_values), y = to_f64(instance.b);
      st = acdc_problem_create_csr((int64_t)instance.n, (int64_t)instance.m, (int64_t)ci.size(),
                                   rp.data(), ci.data(), v.data(), y.data(), &problem_);
    } else {
      const std::vector<int32_t> cp = to_i32(instance.A_csc_col_ptr, "A_csc_col_ptr");
      const std::vector<int32_t> ri = to_i32(instance.A_csc_row_idx, "A_csc_row_idx");
      const std::vector<double> v = to_f64(instance.A_csc_values), b = to_f64(instance.b);
      st = acdc_problem_create_csc((int64_t)instance.m, (int64_t)instance.n, (int64_t)ri.size(),
                                   cp.data(), ri.data(), v.data(), b.data(), &problem_);
    }
    check(st, "acdc_problem_create");
    acdc_settings s{};
    s.lambda = (double)instance.lambda;
    s.tau = std::max(1, _settings->totalThreads);
    s.seed = _settings->hipSeed;
    resolve_beta(loss, *_settings, (double)instance.sigma, &s);
    s.precision = _settings->hipPrecision;
    s.step_mode = _settings->hipStepMode;
    s.lanes_per_coord = _settings->hipLanesPerCoord;
    s.refresh_period = _settings->hipRefreshPeriod;
    s.launch_mode = _settings->hipLaunchMode;
    st = acdc_solver_create(problem_, loss, &s, &solver_);
    if (st != ACDC_OK) {
      acdc_problem_destroy(problem_);
      check(st, "acdc_solver_create");
    }
    double beta = 1;
    acdc_get_beta(solver_, &beta);
    instance.sigma = (D)beta;
    instance.x.assign((size_t)instance.n, (D)0);
    std::vector<double> li((size_t)instance.n);
    check(acdc_get_lipschitz(solver_, li.data(), (int64_t)li.size()), "acdc_get_lipschitz");
    instance.Li.assign(li.begin(), li.end());
  }

  ~HipEngineExecutor() override {
    // the caller's x ends current, as the reference's in-place x does
    if (!this->settings->hipPullXEachBulk) {
      try {
        pull_x();
      } catch (const std::exception &) {
      }
    }
    acdc_solver_destroy(solver_);
    acdc_problem_destroy(problem_);
  }
  HipEngineExecutor(const HipEngineExecutor &) = delete;
  HipEngineExecutor &operator=(const HipEngineExecutor &) = delete;

  // returns when the bulk has run (the caller times it as pure computation,
  // Solver.h:45-50); x is copied back when hipPullXEachBulk asks for it
  void executeBulkOfIterations() override {
    acdc_host::check(acdc_run_updates(solver_, (uint64_t)this->settings->innerIterations),
                     "acdc_run_updates");
    if (this->settings->hipPullXEachBulk)
      pull_x();
    else
      acdc_host::check(acdc_synchronize(solver_), "acdc_synchronize");
  }

  // copies the device's x into instance.x
  void syncX() { pull_x(); }

  void recomputeResiduals() override {
    acdc_host::check(acdc_recompute_residuals(solver_), "acdc_recompute_residuals");
  }

  // Lasso/NNLS: the objective (MulticoreL1RegSquareLoss.h:30-39); SVM: the
  // duality gap P - D (MulticoreSVMDualLoss.h:114-116)
  D getObjectiveValue() override {
    acdc_host::check(acdc_report_get(solver_, &last_), "acdc_report_get");
    instance.primalObjective = (D)last_.primal;
    instance.dualObjective = (D)last_.dual;
    instance.oneZeroAccuracy = (D)last_.accuracy;
    if (this->settings->lossFunction == ACDC_LOSS_SVM_DUAL_HINGE && !instance.test_b.empty())
      test_accuracy_ = test_accuracy();
    return (D)(this->settings->lossFunction == ACDC_LOSS_SVM_DUAL_HINGE ? last_.gap : last_.primal);
  }

  std::string getResultLogHeaders() override {
    return acdc_losses::log_headers(this->settings->lossFunction);
  }
  std::string getResultLogRow(OptimizationStatistics &statistics) override {
    return acdc_losses::log_row(this->settings->lossFunction, last_, statistics, test_accuracy_);
  }

  const acdc_report &lastReport() const { return last_; }
  acdc_solver *handle() { return solver_; }

 private:
  // the test-set 0-1 accuracy of w, on the host as in
  // MulticoreSVMDualLoss.h:93-110: a test row counts when y <x, w> > 0
  double test_accuracy() {
    std::vector<double> w((size_t)instance.m);
    acdc_host::check(acdc_get_residuals(solver_, w.data(), (int64_t)w.size()),
                     "acdc_get_residuals");
    const size_t rows = instance.A_test_csr_row_ptr.empty() ? 0 : instance.A_test_csr_row_ptr.size() - 1;
    if (rows == 0) return 0.0;
    size_t good = 0;
    for (size_t j = 0; j < rows; ++j) {
      double e = 0;
      for (L i = instance.A_test_csr_row_ptr[j]; i < instance.A_test_csr_row_ptr[j + 1]; ++i) {
        const L f = instance.A_test_csr_col_idx[i];
        if (f >= 0 && (size_t)f < w.size()) e += (double)instance.A_test_csr_values[i] * w[(size_t)f];
      }
      if ((double)instance.test_b[j] * e > 0) ++good;
    }
    return (double)good / (double)rows;
  }

  void pull_x() {
    std::vector<double> x((size_t)instance.n);
    acdc_host::check(acdc_get_x(solver_, x.data(), (int64_t)x.size()), "acdc_get_x");
    instance.x.assign(x.begin(), x.end());
  }

  acdc_problem *problem_ = nullptr;
  acdc_solver *solver_ = nullptr;
  acdc_report last_{};
  double test_accuracy_ = 0.0;
};
