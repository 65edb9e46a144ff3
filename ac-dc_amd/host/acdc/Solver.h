// Solver.h -- the bulk driver loop (cpp/src/solver/Solver.h:13-79): run
// iters_bulkIterations_count bulks of innerIterations updates, accumulate the
// pure computation time around each bulk, optionally refresh the residuals,
// and log one CSV row per bulk.  Works with any AbstractEngineExecutor.
// With total_execution_time > 0 the loop also stops once the pure
// computation time reaches it (the time budget of the reference's GPU
// solver loop, randomL2L1Solver.h:168).
#pragma once
#include <iomanip>
#include <iostream>
#include <sys/time.h>

#include "HipEngineExecutor.h"

inline double acdc_gettime() {
  struct timeval t;
  if (gettimeofday(&t, nullptr)) return -1.0;
  return t.tv_sec + 1.0e-6 * t.tv_usec;
}

template <typename L, typename D>
class Solver {
 public:
  AbstractEngineExecutor<L, D> &executor;
  OptimizationSettings *settings;
  OptimizationStatistics statistics;

  explicit Solver(AbstractEngineExecutor<L, D> &e) : executor(e), settings(e.settings) {}

  void log(const std::string &row) {
    std::cout << row << std::endl;
    if (settings->logToFile && settings->logFile) (*settings->logFile) << row << std::endl;
  }

  D runSolver() {
    if (settings->verbose)
      std::cout << "Number of large iterations: " << settings->iters_bulkIterations_count
                << std::endl;
    D objVal = 0;
    if (settings->showInitialObjectiveValue) {
      objVal = executor.getObjectiveValue();
      if (settings->verbose) {
        std::cout << executor.getResultLogHeaders() << std::endl;
        log(executor.getResultLogRow(statistics));
      }
    }
    const double totalStart = acdc_gettime();
    for (unsigned long i = 0; i < settings->iters_bulkIterations_count; ++i) {
      const double bulkStart = acdc_gettime();
      executor.executeBulkOfIterations();
      const double bulkEnd = acdc_gettime();
      statistics.elapsedPureComputationTime += bulkEnd - bulkStart;
      statistics.elapsed_time = bulkEnd - totalStart;
      statistics.elapsedIterations++;
      if (settings->recomputeResidualAfterEachBulkIteration) executor.recomputeResiduals();
      if (settings->showIntermediateObjectiveValue) {
        objVal = executor.getObjectiveValue();
        if (settings->verbose) log(executor.getResultLogRow(statistics));
      }
      if (settings->total_execution_time > 0 &&
          statistics.elapsedPureComputationTime >= settings->total_execution_time)
        break;
    }
    if (settings->showLastObjectiveValue) {
      objVal = executor.getObjectiveValue();
      if (settings->verbose)
        std::cout << "Last objective value: " << std::setprecision(16) << objVal << std::endl;
    }
    return objVal;
  }
};
