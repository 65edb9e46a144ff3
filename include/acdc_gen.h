/*
 * acdc_gen.h -- seeded synthetic problem generators of the PCDM benchmarks.
 *
 * No counterpart in the reference's solver interface.  The reference ships
 * its own instance generators (Nesterov-style Lasso with known optimum,
 * cpp/src/problem_generator/generator_nesterov.h:24-186; 3D truss ground
 * structures, cpp/src/problem_generator/ttd/TTDGenerator.h:65-234); these
 * produce the instances the benchmarks run on.  All randomness comes from
 * Philox4x32-10 streams keyed by the seed, one counter range per purpose and
 * per column/row, so a dataset is a pure function of its configuration.
 */
#ifndef ACDC_GEN_H
#define ACDC_GEN_H
#include <stdint.h>
#include "acdc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum acdc_nnz_law {
  ACDC_NNZ_FIXED = 0,       /* every column has nnz_fixed entries */
  ACDC_NNZ_ZIPF_TRUNC = 1,  /* P(k) ~ k^-s renormalised on [nnz_lo, nnz_hi] */
  ACDC_NNZ_ZIPF_CLIP = 2    /* Zipf(s) on [1, inf) clipped to [nnz_lo, nnz_hi] */
} acdc_nnz_law;

/* Lasso instance: A (m x n) with distinct uniform rows per column, N(0,1)
 * values, unit-norm columns; x* with xstar_nnz nonzeros (+-1 or N(0,1));
 * b = A x* + noise N(0,1); lambda = lambda_frac ||A^T b||_inf. */
typedef struct acdc_gen_lasso_cfg {
  int64_t m, n;
  int32_t nnz_law;   /* acdc_nnz_law */
  int32_t nnz_fixed;
  int32_t nnz_lo, nnz_hi;
  double zipf_s;
  int64_t xstar_nnz;
  int32_t xstar_normal; /* 0: +-1 entries, 1: N(0,1) entries */
  double noise;
  double lambda_frac;
  uint64_t seed;
} acdc_gen_lasso_cfg;

/* Sparse binary classification: n samples over d features, row nnz =
 * round(LogNormal(ln(median), sigma)) clipped to [1, d], distinct uniform
 * features, LogNormal(0,1) values, rows L2-normalised; labels sign(<w*,x>)
 * (0 -> +1) from a wstar_nnz-sparse N(0,1) w*, a flip_frac share flipped. */
typedef struct acdc_gen_svm_cfg {
  int64_t n_samples, n_features;
  double row_median, row_sigma;
  int64_t wstar_nnz;
  double flip_frac;
  uint64_t seed;
} acdc_gen_svm_cfg;

typedef struct acdc_dataset acdc_dataset;

typedef struct acdc_dataset_info {
  int32_t kind;   /* 0: column-major (A, b), 1: sample-major (X, labels) */
  int64_t m, n;   /* kind 0: rows, columns; kind 1: features, samples */
  int64_t nnz;
  double lambda;  /* lambda the configuration asks for (0 for TTD) */
  double lambda_max; /* ||A^T b||_inf (kind 0) */
  int64_t xstar_nnz;
} acdc_dataset_info;

acdc_status acdc_gen_lasso(const acdc_gen_lasso_cfg *cfg, acdc_dataset **out);
acdc_status acdc_gen_svm(const acdc_gen_svm_cfg *cfg, acdc_dataset **out);
/* 3D truss ground structure on a cells^3 grid of unit cells: bars between
 * all node pairs at Chebyshev distance 1; the z = 0 face is fixed and its
 * DOFs removed; f = -1 on the z DOF of every top-face node.  A bar from node
 * a to node b = a + d (d lexicographically positive, so b is not below a)
 * holds +d/||d||^2 on a's free DOFs and -d/||d||^2 on b's: x_j >= 0 is the
 * bar's compressive force.  (TTDGenerator.h:186-216 orients bars the other
 * way, which suits its sign-free L1 formulation; with x >= 0 that
 * orientation makes x = 0 optimal.)  Zero components are not stored; rows
 * are dof-major: x block, y block, z block of the free nodes. */
acdc_status acdc_gen_ttd(int32_t cells, acdc_dataset **out);

acdc_status acdc_dataset_info_get(const acdc_dataset *d, acdc_dataset_info *out);
/* Copies the arrays out: ptr has (kind 0: n+1, kind 1: n+1 over samples)
 * entries, idx/vals nnz, rhs (b or labels) m or n entries, xstar n (kind 0)
 * or n_features (kind 1) entries; any pointer may be null to skip it. */
acdc_status acdc_dataset_copy(const acdc_dataset *d, int32_t *ptr, int32_t *idx,
                              double *vals, double *rhs, double *xstar);
/* Builds a solver problem straight from the dataset (no host round trip). */
acdc_status acdc_dataset_problem(const acdc_dataset *d, acdc_problem **out);
acdc_status acdc_dataset_destroy(acdc_dataset *d);

#ifdef __cplusplus
}
#endif
#endif /* ACDC_GEN_H */
