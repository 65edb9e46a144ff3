// The SVM executor's test-set 0-1 accuracy (MulticoreSVMDualLoss.h:93-110),
// computed on the host from the device's w: with the training rows given as
// the test set it must equal the training 0-1 accuracy the device reports.
// Prints "ok <accuracy>" on success.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "acdc/HipEngineExecutor.h"

int main() {
  const int n = 400, m = 600, per_row = 25;  // samples, features
  std::mt19937 rng(5);
  std::uniform_int_distribution<int> feat(0, m - 1);
  std::normal_distribution<double> nrm(0.0, 1.0);
  std::vector<double> wstar(m);
  for (double &v : wstar) v = nrm(rng);
  ProblemData<int, double> inst;
  inst.n = n;
  inst.m = m;
  inst.lambda = 1.0 / n;
  inst.A_csr_row_ptr.push_back(0);
  for (int i = 0; i < n; ++i) {
    std::vector<int> cols;
    while ((int)cols.size() < per_row) {
      const int f = feat(rng);
      bool seen = false;
      for (int c : cols) seen = seen || c == f;
      if (!seen) cols.push_back(f);
    }
    std::sort(cols.begin(), cols.end());
    double dot = 0;
    for (int c : cols) {
      const double v = std::fabs(nrm(rng));
      inst.A_csr_col_idx.push_back(c);
      inst.A_csr_values.push_back(v);
      dot += v * wstar[c];
    }
    inst.A_csr_row_ptr.push_back((int)inst.A_csr_col_idx.size());
    inst.b.push_back(dot > 0 ? 1.0 : -1.0);
  }
  inst.A_test_csr_row_ptr = inst.A_csr_row_ptr;
  inst.A_test_csr_col_idx = inst.A_csr_col_idx;
  inst.A_test_csr_values = inst.A_csr_values;
  inst.test_b = inst.b;
  OptimizationSettings st;
  st.lossFunction = ACDC_LOSS_SVM_DUAL_HINGE;
  st.totalThreads = 16;
  st.innerIterations = 4 * n;
  HipEngineExecutor<int, double> ex(inst, &st);
  OptimizationStatistics stats;
  for (int bulk = 0; bulk < 3; ++bulk) {
    ex.executeBulkOfIterations();
    ex.getObjectiveValue();
    // columns: iterations, gap, primal, dual, 0-1, pure time, total time, test 0-1
    std::stringstream row(ex.getResultLogRow(stats));
    std::vector<std::string> f;
    std::string tok;
    while (std::getline(row, tok, ',')) f.push_back(tok);
    if (f.size() != 8) {
      std::printf("bad row\n");
      return 1;
    }
    const double train = std::stod(f[4]), test = std::stod(f[7]);
    if (!(train > 0.5) || test != train) {
      std::printf("train %.17g test %.17g\n", train, test);
      return 1;
    }
    if (bulk == 2) std::printf("ok %.6f\n", test);
  }
  return 0;
}
