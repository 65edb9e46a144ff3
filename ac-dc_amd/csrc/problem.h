// problem.h -- host-side layout of an acdc_problem handle (shared by the
// single-instance and the batched solvers).
#pragma once
#include <cstdint>
#include <vector>

struct acdc_problem {
  int kind;        // 0: column-major input (Lasso/NNLS), 1: sample rows (SVM)
  int64_t ncoord;  // coordinates: columns n, or samples
  int64_t nres;    // residual entries: rows m, or features
  int64_t nnz;
  std::vector<int32_t> cptr, cidx;  // coordinate-major segments
  std::vector<double> cval;
  std::vector<double> rhs;  // b (m) or labels y (samples)
  int32_t omega;           // max entries per residual index (ESO omega)
};
