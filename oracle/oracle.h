/*
 * CPU FP64 oracle for the PCDM hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker, never the product: only tests/, the smoke()
 * entry and bench.py's cpu_baseline leg may load it.  The shipped path
 * (libacdc_hip.so) never links or calls anything here.
 *
 * Every function restates an algorithm of optml/ac-dc in plain C (no code
 * copied); the reference file:line it follows is cited next to each
 * declaration.  How each function is pinned (known-answer vectors, closed
 * forms, brute force, scipy solvers) is listed in the module header of
 * oracle.c.
 */
#ifndef ACDC_ORACLE_H
#define ACDC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based sampling (no reference counterpart: the reference uses
 *      GSL mt19937 per thread, randomNumbersUtil.h:83-125; this build's
 *      coordinate stream is Philox4x32-10, Salmon et al. SC'11) ---------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);
/* round keys of the step-s permutation */
void orc_step_keys(uint64_t seed, uint64_t step, uint32_t keys[4]);
/* slot i of the keyed permutation of [0,n) */
uint32_t orc_perm(uint32_t i, uint32_t n, const uint32_t keys[4]);
/* coordinates of steps [step0, step0+nsteps), tau per step, row-major */
void orc_sample_steps(uint64_t seed, uint64_t step0, uint64_t nsteps,
                      uint32_t n, uint32_t tau, int32_t *out);

/* ---- scalar rules ------------------------------------------------------ */
/* sgn(v)(|v|-t)_+ : treshhold_functions.h:27-33, helpers/utils.h:29-32 */
double orc_soft_threshold(double t, double v);

/* ---- square loss over CSC (Lasso loss 0, NNLS loss 10) ----------------- */
/* r = A x - b, column-scatter order: MulticoreSquareLoss.h:42-53 */
void orc_csc_residual(int64_t m, int64_t n, const int32_t *col_ptr,
                      const int32_t *row_idx, const double *vals,
                      const double *x, const double *b, double *r);
/* L_j = 1/(beta ||a_j||^2), 0 for empty columns: loss_square.h:183-194 */
void orc_csc_lipschitz(int64_t n, const int32_t *col_ptr, const double *vals,
                       double beta, double *L);
/* omega = max number of entries per index value (max nnz per row of a CSC
 * matrix, max samples per feature of a CSR one): MultiCoreSolver.cpp:32-38,
 * :55-62 */
int32_t orc_omega(int64_t nidx, int64_t nnz, const int32_t *idx);
/* ESO beta = 1 + (omega-1)(tau-1)/(n-1): MultiCoreSolver.cpp:108-110 */
double orc_eso_beta(int32_t omega, int64_t tau, int64_t n);

/* the spectral ESO (mSDCA.cpp:188-265, :286): beta = 1 + (tau-1)(lmax-1)/(n-1)
 * with lmax the largest eigenvalue of D^-1/2 A^T A D^-1/2 by `iters` power
 * steps from the seeded start orc_pm_start */
double orc_pm_start(uint64_t seed, int64_t c);
double orc_spectral_lmax(int64_t nseg, int64_t nidx, const int32_t *ptr, const int32_t *idx,
                         const double *vals, uint64_t seed, int iters);

/* nsteps synchronous tau-nice steps (cdn_common.h:235-282 semantics: all
 * deltas of a step from the same state, then all applied).  tau=1 is the
 * serial method do_single_iteration_serial, loss_square.h:72-83 with
 * compute_update :137-148.  loss 0 = soft threshold; loss 10 = clip at 0.
 * x (n) and r (m) are updated in place. */
void orc_cd_run(int loss, int64_t m, int64_t n, const int32_t *col_ptr,
                const int32_t *row_idx, const double *vals, const double *L,
                double lambda, uint64_t seed, int64_t tau, uint64_t step0,
                uint64_t nsteps, double *x, double *r);

/* FP64 evaluation from x alone (r recomputed).
 * Lasso out[0..7] = {primal, dual, gap, rel_gap, ||A^T r||_inf, ||x||_1,
 *                    ||r||^2, nnz(x)}
 * primal: MulticoreL1RegSquareLoss.h:30-39.  The dual is the scaled-residual
 * point theta = -r min(1, lambda/||A^T r||_inf) (no reference counterpart:
 * the reference logs no Lasso gap).
 * NNLS out[0..7] = {primal, kkt=max|min(x_j,g_j)|, x^T g, min_j g_j,
 *                   ||g||_inf, ||x||_1, ||r||^2, nnz(x)}, g = A^T r. */
void orc_lasso_eval(int64_t m, int64_t n, const int32_t *col_ptr,
                    const int32_t *row_idx, const double *vals,
                    const double *b, const double *x, double lambda,
                    double out[8]);
void orc_nnls_eval(int64_t m, int64_t n, const int32_t *col_ptr,
                   const int32_t *row_idx, const double *vals,
                   const double *b, const double *x, double out[8]);

/* ---- hinge-SVM dual over CSR rows = samples (loss 2) ------------------- */
/* L_i = 1/(beta ||x_i||^2): MulticoreSVMDualLoss.h:48-61 */
void orc_svm_lipschitz(int64_t ns, const int32_t *row_ptr, const double *vals,
                       double beta, double *L);
/* w = 1/(lambda n) sum alpha_i y_i x_i after clipping alpha to [0,1]:
 * MulticoreSVMDualLoss.h:31-46 */
void orc_svm_w(int64_t ns, int64_t nf, const int32_t *row_ptr,
               const int32_t *col_idx, const double *vals, const double *y,
               double lambda, double *alpha, double *w);
/* synchronous tau-nice dual SDCA steps; tau=1 is
 * MulticoreSVMDualLoss::performOneCoordinateUpdate, :119-141 */
void orc_svm_run(int64_t ns, int64_t nf, const int32_t *row_ptr,
                 const int32_t *col_idx, const double *vals, const double *y,
                 const double *L, double lambda, uint64_t seed, int64_t tau,
                 uint64_t step0, uint64_t nsteps, double *alpha, double *w);
/* out[0..7] = {gap, primal, dual, train 0-1 accuracy, rel_gap, ||w||^2,
 *              sum alpha, sum hinge}: MulticoreSVMDualLoss.h:63-117 with w
 * recomputed from alpha in FP64 */
/* serial SDCA steps on coords[0..count) in order */
void orc_svm_list(int64_t ns, int64_t nf, const int32_t *row_ptr, const int32_t *col_idx,
                  const double *vals, const double *y, const double *L, double lambda,
                  const int32_t *coords, uint64_t count, double *alpha, double *w);
void orc_svm_eval(int64_t ns, int64_t nf, const int32_t *row_ptr,
                  const int32_t *col_idx, const double *vals, const double *y,
                  double lambda, const double *alpha, double out[8]);

/* ---- asynchronous workers -------------------------------------------- */
/* Coordinates of updates [k0, k0+count) of asynchronous worker `worker`:
 * with replacement, one Philox4x32-10 stream per worker (the reference's
 * per-thread generators drawn by gsl_rng_uniform_int,
 * MulticoreEngineExecutor.h:86, randomNumbersUtil.h:83-125). */
void orc_async_coords(uint64_t seed, uint32_t worker, uint64_t k0, uint64_t count,
                      uint32_t n, int32_t *out);
/* serial updates over the given coordinates, in order (the tau = 1 method,
 * loss_square.h:72-83; loss 0 soft threshold, 10 clip at 0) */
void orc_cd_list(int loss, int64_t m, int64_t n, const int32_t *col_ptr,
                 const int32_t *row_idx, const double *vals, const double *L,
                 double lambda, const int32_t *coords, uint64_t count,
                 double *x, double *r);
/* MulticoreEngineExecutor::executeBulkOfIterations, :83-89: `threads` OpenMP
 * workers, each running updates [k0, k0+per_worker) of its own stream with
 * unsynchronised reads of x and r and CAS-loop atomic adds
 * (parallel_essentials_posix.h:62-68).  Returns the number of threads. */
int orc_cd_async_omp(int loss, int64_t m, int64_t n, const int32_t *col_ptr,
                     const int32_t *row_idx, const double *vals,
                     const double *L, double lambda, uint64_t seed,
                     uint64_t k0, uint64_t per_worker, int threads,
                     double *x, double *r);

#ifdef __cplusplus
}
#endif
#endif
