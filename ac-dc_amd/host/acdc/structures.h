// structures.h -- problem data and settings of the host interface.
//
// Same field names and meaning as the reference's ProblemData
// (cpp/src/solver/structures.h:30-97), OptimizationSettings and
// OptimizationStatistics (cpp/src/solver/settingsAndStatistics.h:11-149), so
// code written against the reference fills them unchanged.  Fields the PCDM
// path does not read are kept out; the HIP-specific knobs are grouped at the
// end of OptimizationSettings with their defaults.
#pragma once
#include <fstream>
#include <string>
#include <vector>

#include "../../../include/acdc_hip.h"

constexpr int kHipBetaReference = -1;

template <typename L, typename D>
class ProblemData {
 public:
  // column-major A for losses 0/1/10 (A_csc_col_ptr[n+1], row_idx/values[nnz])
  std::vector<D> A_csc_values;
  std::vector<L> A_csc_row_idx;
  std::vector<L> A_csc_col_ptr;
  // sample-major data for the SVM dual (rows = samples)
  std::vector<D> A_csr_values;
  std::vector<L> A_csr_col_idx;
  std::vector<L> A_csr_row_ptr;
  // optional test set of the SVM (0-1 accuracy only)
  std::vector<D> A_test_csr_values;
  std::vector<L> A_test_csr_col_idx;
  std::vector<L> A_test_csr_row_ptr;
  std::vector<D> test_b;

  std::vector<D> Li;  // reciprocal ESO-scaled Lipschitz constants (filled by the executor)
  std::vector<D> b;   // right-hand side, or labels +-1 for the SVM
  std::vector<D> x;   // solution (alpha for the SVM), zeroed by the executor

  std::string experimentName;
  L m = 0;  // rows (features for the SVM)
  L n = 0;  // columns (samples for the SVM)
  D lambda = 0;
  D sigma = 0;  // ESO beta; 0 = computed from omega and tau by the executor
  D omegaMin = 0;
  D omegaAvg = 0;
  D oneZeroAccuracy = 0;
  D dualObjective = 0;
  D primalObjective = 0;
};

class OptimizationSettings {
 public:
  int lossFunction = 0;  // 0 L1-square (Lasso), 2 SVM dual hinge, 10 NNLS
  bool showInitialObjectiveValue = false;
  bool showIntermediateObjectiveValue = true;
  bool showLastObjectiveValue = true;
  int totalThreads = 8;  // tau: coordinates per synchronous step
  bool recomputeResidualAfterEachBulkIteration = false;
  bool bulkIterations = false;
  bool verbose = false;
  bool logToFile = false;
  std::ofstream *logFile = nullptr;
  double forcedSigma = -1;  // > 0: use this ESO beta
  unsigned long iters_bulkIterations_count = 10;
  unsigned long innerIterations = 100;
  // seconds of pure computation after which runSolver stops early (0: the
  // bulk count alone); the reference's field (settingsAndStatistics.h:35),
  // which its GPU solver loops on (randomL2L1Solver.h:168)
  double total_execution_time = 0;

  // HIP path (no counterpart in the reference)
  unsigned long long hipSeed = 1;             // Philox key of the coordinate stream
  int hipPrecision = ACDC_PREC_FP64;          // ACDC_PREC_FP64 / ACDC_PREC_FP32_REFRESH
  int hipRefreshPeriod = 5;                   // FP32 path: FP64 refresh every k epochs
  int hipLanesPerCoord = 0;                   // 0 = auto
  int hipLaunchMode = ACDC_LAUNCH_AUTO;
  int hipStepMode = ACDC_STEP_ASYNC;          // the multicore executor's unsynchronised
                                              // workers, tau = totalThreads
                                              // (MulticoreEngineExecutor.h:83-89);
                                              // ACDC_STEP_SYNC: tau-nice synchronous steps
                                              // (cdn_common.h:235-282)
  int hipBetaMode = kHipBetaReference;        // step scaling, see below
  bool hipPullXEachBulk = true;               // copy x into instance.x after every bulk (the
                                              // reference executor updates the caller's x
                                              // in place); false: only syncX() and the
                                              // executor's destructor copy it, which saves
                                              // an n-vector device-to-host copy per bulk
};

// hipBetaMode: kHipBetaReference (the default) applies the reference multicore
// executor's rule -- the Lasso and NNLS take L_j = 1/||a_j||^2 and ignore
// instance.sigma (MulticoreSquareLoss.h:27-40), the SVM dual takes
// L_i = 1/(sigma ||x_i||^2) with the caller's instance.sigma
// (MulticoreSVMDualLoss.h:48-61), or the omega rule of MultiCoreSolver.cpp:108-110
// when the caller left sigma at 0.  Any acdc_beta_mode value applies that rule
// to every loss instead (ACDC_BETA_FORCED takes instance.sigma).
// settings.forcedSigma > 0 overrides both (ClusterSolver.cpp:203-205).

class OptimizationStatistics {
 public:
  unsigned long elapsedIterations = 0;
  double elapsedPureComputationTime = 0;
  double elapsed_time = 0;
};
