// HipEngineExecutor's Lipschitz constants under each step-scaling rule:
//   reference rule, Lasso at 8 threads: L_j = 1/||a_j||^2, sigma ignored
//     (MulticoreSquareLoss.h:27-40), instance.sigma left at 1;
//   reference rule, SVM with sigma = 1.7: L_i = 1/(1.7 ||x_i||^2)
//     (MulticoreSVMDualLoss.h:48-61);
//   hipBetaMode = ACDC_BETA_OMEGA, Lasso at 8 threads: beta = 1 + (omega-1)
//     (tau-1)/(n-1) (MultiCoreSolver.cpp:108-110);
//   forcedSigma = 2.5 overrides the rule.
// Prints "ok" when every constant matches within 1e-14 relative.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "acdc/HipEngineExecutor.h"

static int fails = 0;

static void expect(const char *what, const std::vector<double> &li,
                   const std::vector<double> &sq, double beta) {
  for (size_t j = 0; j < li.size(); ++j) {
    const double want = sq[j] > 0 ? 1.0 / (beta * sq[j]) : 0.0;
    if (std::fabs(li[j] - want) > 1e-14 * std::fabs(want)) {
      std::printf("%s: L[%zu] = %.17g, want %.17g\n", what, j, li[j], want);
      ++fails;
      return;
    }
  }
}

int main() {
  std::mt19937 rng(3);
  std::normal_distribution<double> nrm(0.0, 1.0);
  const int m = 200, n = 700, per = 7;
  std::uniform_int_distribution<int> row(0, m - 1);
  ProblemData<int, double> lasso;
  lasso.m = m;
  lasso.n = n;
  lasso.lambda = 0.05;
  lasso.A_csc_col_ptr.push_back(0);
  std::vector<double> sq(n, 0.0);
  std::vector<int> rowcount(m, 0);
  for (int j = 0; j < n; ++j) {
    for (int k = 0; k < per; ++k) {
      const int r = row(rng);
      const double v = nrm(rng);
      lasso.A_csc_row_idx.push_back(r);
      lasso.A_csc_values.push_back(v);
      sq[j] += v * v;
      ++rowcount[r];
    }
    lasso.A_csc_col_ptr.push_back((int)lasso.A_csc_row_idx.size());
  }
  int omega = 1;
  for (int c : rowcount) omega = std::max(omega, c);
  for (int i = 0; i < m; ++i) lasso.b.push_back(nrm(rng));

  OptimizationSettings st;
  st.lossFunction = ACDC_LOSS_L1_SQUARE;
  st.totalThreads = 8;
  {
    ProblemData<int, double> inst = lasso;
    inst.sigma = 3.0;  // set by MultiCoreSolver.cpp:108-110, ignored by the multicore Lasso
    HipEngineExecutor<int, double> ex(inst, &st);
    expect("reference rule, Lasso", inst.Li, sq, 1.0);
    if (inst.sigma != 1.0) {
      std::printf("reference rule, Lasso: sigma %.17g\n", inst.sigma);
      ++fails;
    }
  }
  {
    ProblemData<int, double> inst = lasso;
    OptimizationSettings so = st;
    so.hipBetaMode = ACDC_BETA_OMEGA;
    HipEngineExecutor<int, double> ex(inst, &so);
    expect("omega rule, Lasso", inst.Li, sq, 1.0 + (omega - 1) * (8 - 1) / (n - 1.0));
  }
  {
    ProblemData<int, double> inst = lasso;
    OptimizationSettings so = st;
    so.forcedSigma = 2.5;
    HipEngineExecutor<int, double> ex(inst, &so);
    expect("forcedSigma, Lasso", inst.Li, sq, 2.5);
  }
  {
    // the same data as SVM sample rows (rows = samples): L_i from the rows
    ProblemData<int, double> svm;
    svm.n = n;
    svm.m = m;
    svm.lambda = 1.0 / n;
    svm.A_csr_row_ptr = lasso.A_csc_col_ptr;
    svm.A_csr_col_idx = lasso.A_csc_row_idx;
    svm.A_csr_values = lasso.A_csc_values;
    for (int i = 0; i < n; ++i) svm.b.push_back(i % 3 ? 1.0 : -1.0);
    svm.sigma = 1.7;
    OptimizationSettings so = st;
    so.lossFunction = ACDC_LOSS_SVM_DUAL_HINGE;
    HipEngineExecutor<int, double> ex(svm, &so);
    expect("reference rule, SVM", svm.Li, sq, 1.7);
  }
  if (fails) return 1;
  std::printf("ok\n");
  return 0;
}
