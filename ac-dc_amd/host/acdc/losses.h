// losses.h -- CSV log headers and rows of the loss objects.
//
// SquareLoss::getResultLogHeaders/Row (squareLoss/SquareLoss.h:28-38) and
// SVMDualLoss::getResultLogHeaders/Row (SVMDualLoss/SVMDualLoss.h:31-42):
// the same columns in the same order, so scripts that parse the reference's
// logs parse these, followed by columns the reference does not log (the Lasso
// duality gap, the NNLS KKT residual).
#pragma once
#include <iomanip>
#include <sstream>
#include <string>

#include "../../../include/acdc_hip.h"
#include "structures.h"

namespace acdc_losses {

inline std::string log_headers(int loss) {
  if (loss == ACDC_LOSS_SVM_DUAL_HINGE)
    return "Elapsed iterations, Current duality gap, Primal objective, Dual objective, 0-1 loss, "
           "Pure computation time, Total time, Test 0-1";
  if (loss == ACDC_LOSS_NNLS)
    return "Elapsed iterations, Current objective value, Pure computation time, Total time, "
           "KKT residual";
  return "Elapsed iterations, Current objective value, Pure computation time, Total time, "
         "Duality gap";
}

// test_accuracy: the SVM's test-set 0-1 accuracy (0 without a test set, as
// in MulticoreSVMDualLoss.h:93-110)
inline std::string log_row(int loss, const acdc_report &r, const OptimizationStatistics &st,
                           double test_accuracy = 0.0) {
  std::stringstream ss;
  ss << st.elapsedIterations << "," << std::setprecision(16);
  if (loss == ACDC_LOSS_SVM_DUAL_HINGE) {
    ss << r.gap << "," << r.primal << "," << r.dual << "," << r.accuracy << ","
       << st.elapsedPureComputationTime << "," << st.elapsed_time << "," << test_accuracy;
  } else {
    ss << r.primal << "," << st.elapsedPureComputationTime << "," << st.elapsed_time << ","
       << (loss == ACDC_LOSS_NNLS ? r.kkt : r.gap);
  }
  return ss.str();
}

}  // namespace acdc_losses
