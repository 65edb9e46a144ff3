/*
 * CPU FP64 oracle for the PCDM hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker.  The product library never links it.
 *
 * How each part is pinned (tests/test_oracle.py):
 *   orc_philox4x32_10   Random123 known-answer vectors for Philox4x32-10
 *                       (tests/golden/philox4x32_10_kat.json).
 *   orc_perm            bijection on [0,n) for many n (brute force), chi^2
 *                       uniformity of slot 0 over steps, distinct slots.
 *   orc_soft_threshold  closed-form table incl. ties |v|=t and signed zeros.
 *   orc_cd_run (Lasso)  one-column closed form x = S(a^T b, lambda)/||a||^2;
 *                       tau=1 step is an exact coordinate minimiser (each step
 *                       never increases the objective); converged x meets the
 *                       Lasso KKT conditions and gap -> 0; planted
 *                       Nesterov-style instance with known optimum.
 *   orc_cd_run (NNLS)   scipy.optimize.nnls golden vectors (tests/golden).
 *   orc_lasso_eval      gap >= 0, gap = 0 at x = 0 when lambda >= lambda_max,
 *                       closed form gap(0) = ||b||^2 (1-s)^2/2.
 *   orc_svm_*           scipy L-BFGS-B on the box-constrained dual (golden),
 *                       brute force for n <= 8, P >= D, w identity;
 *                       orc_svm_list equals orc_svm_run(tau=1) on the
 *                       samples orc_sample_steps gives it (bit for bit).
 *   orc_async_coords    multiply-shift map checked against exact integer
 *                       arithmetic; chi^2 uniformity; per-worker streams
 *                       distinct.
 *   orc_cd_list         equals orc_cd_run(tau=1) on the coordinates that
 *                       orc_sample_steps gives it (bit for bit).
 *   orc_cd_async_omp    1 thread == orc_cd_list over orc_async_coords, bit
 *                       for bit; many threads converge to the serial optimum.
 *   orc_spectral_lmax   closed forms: orthonormal segments give 1, n equal
 *                       segments give n, two disjoint groups of p and q equal
 *                       segments give max(p, q); numpy eigvalsh on a random
 *                       matrix with a separated top eigenvalue.
 *
 * Compiled with -ffp-contract=off so that every a*b+c rounds twice, like the
 * reference's plain C++ loops.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------
 * Philox4x32-10 (J. Salmon, M. Moraes, R. Dror, D. Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11).  Written from the paper: two
 * 32x32->64 multiplies per round, Weyl key schedule, 10 rounds.
 * ---------------------------------------------------------------------- */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    uint64_t p0 = (uint64_t)PHILOX_M0 * (uint64_t)c0;
    uint64_t p1 = (uint64_t)PHILOX_M1 * (uint64_t)c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += PHILOX_W0;
    k1 += PHILOX_W1;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* stream tag of the coordinate sampler ("PCDS") */
#define SAMPLER_TAG 0x50434453u

void orc_step_keys(uint64_t seed, uint64_t step, uint32_t keys[4]) {
  uint32_t ctr[4] = {(uint32_t)step, (uint32_t)(step >> 32), SAMPLER_TAG, 0u};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  orc_philox4x32_10(ctr, key, keys);
}

/* round function: 32-bit integer mixer (two multiply/xor-shift rounds) */
static uint32_t orc_mix(uint32_t x, uint32_t k) {
  uint32_t h = x ^ k;
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}

/* Four-round alternating-xor Feistel network on [0, 2^b) with 2^b >= n,
 * restricted to [0,n) by cycle walking.  Each round is invertible given the
 * other half, so the network is a bijection and so is the walk. */
uint32_t orc_perm(uint32_t i, uint32_t n, const uint32_t keys[4]) {
  uint32_t b = 1;
  while (b < 32 && ((uint64_t)1 << b) < (uint64_t)n) ++b;
  uint32_t hb = b >> 1, ha = b - hb;
  uint32_t mb = (hb == 0) ? 0u : (uint32_t)(((uint64_t)1 << hb) - 1);
  uint32_t ma = (uint32_t)(((uint64_t)1 << ha) - 1);
  uint32_t v = i;
  do {
    uint32_t L = (hb == 32) ? 0 : (v >> hb);
    uint32_t R = v & mb;
    L ^= orc_mix(R, keys[0]) & ma;
    R ^= orc_mix(L, keys[1]) & mb;
    L ^= orc_mix(R, keys[2]) & ma;
    R ^= orc_mix(L, keys[3]) & mb;
    v = (L << hb) | R;
  } while (v >= n);
  return v;
}

void orc_sample_steps(uint64_t seed, uint64_t step0, uint64_t nsteps,
                      uint32_t n, uint32_t tau, int32_t *out) {
  for (uint64_t s = 0; s < nsteps; ++s) {
    uint32_t k[4];
    orc_step_keys(seed, step0 + s, k);
    for (uint32_t i = 0; i < tau; ++i)
      out[s * tau + i] = (int32_t)orc_perm(i, n, k);
  }
}

/* ------------------------------------------------------------------------ */
double orc_soft_threshold(double t, double v) {
  double a = fabs(v) - t;
  if (a < 0) a = 0;
  int sg = (v > 0) - (v < 0);
  return sg * a;
}

/* Neumaier-compensated running sum */
typedef struct {
  double s, c;
} ksum;
static void ks_add(ksum *k, double v) {
  double t = k->s + v;
  if (fabs(k->s) >= fabs(v))
    k->c += (k->s - t) + v;
  else
    k->c += (v - t) + k->s;
  k->s = t;
}
static double ks_val(const ksum *k) { return k->s + k->c; }

void orc_csc_residual(int64_t m, int64_t n, const int32_t *col_ptr,
                      const int32_t *row_idx, const double *vals,
                      const double *x, const double *b, double *r) {
  for (int64_t i = 0; i < m; ++i) r[i] = -b[i];
  for (int64_t j = 0; j < n; ++j)
    for (int32_t p = col_ptr[j]; p < col_ptr[j + 1]; ++p)
      r[row_idx[p]] += vals[p] * x[j];
}

void orc_csc_lipschitz(int64_t n, const int32_t *col_ptr, const double *vals,
                       double beta, double *L) {
  for (int64_t j = 0; j < n; ++j) {
    double s = 0;
    for (int32_t p = col_ptr[j]; p < col_ptr[j + 1]; ++p) s += vals[p] * vals[p];
    L[j] = (s > 0) ? 1.0 / (beta * s) : 0.0;
  }
}

int32_t orc_omega(int64_t nidx, int64_t nnz, const int32_t *idx) {
  int32_t *cnt = (int32_t *)calloc((size_t)(nidx > 0 ? nidx : 1), sizeof(int32_t));
  int32_t om = 1; /* the reference starts omega at 1 */
  for (int64_t p = 0; p < nnz; ++p) {
    int32_t c = ++cnt[idx[p]];
    if (c > om) om = c;
  }
  free(cnt);
  return om;
}

double orc_eso_beta(int32_t omega, int64_t tau, int64_t n) {
  if (n <= 1) return 1.0;
  return 1.0 + (double)(omega - 1) * (double)(tau - 1) / ((double)n - 1.0);
}

/* start vector of the spectral power steps: coordinate c takes the first
 * word of Philox4x32-10 at counter (c, c >> 32, "PCPM", 0) under the seed,
 * over 2^32 */
#define ORC_PM_TAG 0x5043504Du
double orc_pm_start(uint64_t seed, int64_t c) {
  const uint32_t ctr[4] = {(uint32_t)c, (uint32_t)((uint64_t)c >> 32), ORC_PM_TAG, 0u};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  orc_philox4x32_10(ctr, key, o);
  return (double)o[0] * (1.0 / 4294967296.0);
}

/* Largest eigenvalue of D^-1/2 A^T A D^-1/2 (A's segments as its columns,
 * D = diag ||a_c||^2) by power steps, as mSDCA.cpp:188-265: x normalised,
 * then per step y = sum_c a_c x_c / ||a_c||, x_c = <a_c, y> / ||a_c||,
 * estimate ||x||, x /= estimate.  Empty segments take 0. */
double orc_spectral_lmax(int64_t nseg, int64_t nidx, const int32_t *ptr, const int32_t *idx,
                         const double *vals, uint64_t seed, int iters) {
  double *x = (double *)malloc(sizeof(double) * (size_t)(nseg > 0 ? nseg : 1));
  double *w = (double *)malloc(sizeof(double) * (size_t)(nseg > 0 ? nseg : 1));
  double *y = (double *)malloc(sizeof(double) * (size_t)(nidx > 0 ? nidx : 1));
  double nrm = 0;
  for (int64_t c = 0; c < nseg; ++c) {
    double s2 = 0;
    for (int32_t p = ptr[c]; p < ptr[c + 1]; ++p) s2 += vals[p] * vals[p];
    w[c] = s2 > 0 ? 1.0 / sqrt(s2) : 0.0;
    x[c] = orc_pm_start(seed, c);
    nrm += x[c] * x[c];
  }
  nrm = sqrt(nrm);
  for (int64_t c = 0; c < nseg; ++c) x[c] /= nrm;
  double est = 1.0;
  for (int it = 0; it < iters; ++it) {
    for (int64_t k = 0; k < nidx; ++k) y[k] = 0;
    for (int64_t c = 0; c < nseg; ++c)
      for (int32_t p = ptr[c]; p < ptr[c + 1]; ++p) y[idx[p]] += vals[p] * w[c] * x[c];
    double s2 = 0;
    for (int64_t c = 0; c < nseg; ++c) {
      double g = 0;
      for (int32_t p = ptr[c]; p < ptr[c + 1]; ++p) g += vals[p] * y[idx[p]];
      x[c] = g * w[c];
      s2 += x[c] * x[c];
    }
    est = sqrt(s2);
    if (est > 0)
      for (int64_t c = 0; c < nseg; ++c) x[c] /= est;
  }
  free(x);
  free(w);
  free(y);
  return est;
}

/* one coordinate's delta from the current state (compute_update) */
static double cd_delta(int loss, const int32_t *col_ptr, const int32_t *row_idx,
                       const double *vals, const double *L, double lambda,
                       const double *x, const double *r, int32_t j) {
  double g = 0;
  for (int32_t p = col_ptr[j]; p < col_ptr[j + 1]; ++p)
    g += vals[p] * r[row_idx[p]];
  double xj = x[j];
  if (loss == 10) {
    double v = xj - L[j] * g;
    return (v > 0 ? v : 0.0) - xj;
  }
  return orc_soft_threshold(L[j] * lambda, xj - L[j] * g) - xj;
}

void orc_cd_run(int loss, int64_t m, int64_t n, const int32_t *col_ptr,
                const int32_t *row_idx, const double *vals, const double *L,
                double lambda, uint64_t seed, int64_t tau, uint64_t step0,
                uint64_t nsteps, double *x, double *r) {
  (void)m;
  int32_t *S = (int32_t *)malloc(sizeof(int32_t) * (size_t)tau);
  double *d = (double *)malloc(sizeof(double) * (size_t)tau);
  for (uint64_t s = 0; s < nsteps; ++s) {
    uint32_t k[4];
    orc_step_keys(seed, step0 + s, k);
    for (int64_t i = 0; i < tau; ++i) S[i] = (int32_t)orc_perm((uint32_t)i, (uint32_t)n, k);
    for (int64_t i = 0; i < tau; ++i)
      d[i] = cd_delta(loss, col_ptr, row_idx, vals, L, lambda, x, r, S[i]);
    for (int64_t i = 0; i < tau; ++i) {
      int32_t j = S[i];
      x[j] += d[i];
      for (int32_t p = col_ptr[j]; p < col_ptr[j + 1]; ++p)
        r[row_idx[p]] += d[i] * vals[p];
    }
  }
  free(S);
  free(d);
}

/* g = A^T r, r = A x - b recomputed */
static void csc_grad(int64_t m, int64_t n, const int32_t *col_ptr,
                     const int32_t *row_idx, const double *vals,
                     const double *b, const double *x, double *r, double *g) {
  orc_csc_residual(m, n, col_ptr, row_idx, vals, x, b, r);
  for (int64_t j = 0; j < n; ++j) {
    ksum k = {0, 0};
    for (int32_t p = col_ptr[j]; p < col_ptr[j + 1]; ++p)
      ks_add(&k, vals[p] * r[row_idx[p]]);
    g[j] = ks_val(&k);
  }
}

void orc_lasso_eval(int64_t m, int64_t n, const int32_t *col_ptr,
                    const int32_t *row_idx, const double *vals,
                    const double *b, const double *x, double lambda,
                    double out[8]) {
  double *r = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  double *g = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  csc_grad(m, n, col_ptr, row_idx, vals, b, x, r, g);
  ksum rr = {0, 0}, br = {0, 0}, l1 = {0, 0};
  double ginf = 0, nz = 0;
  for (int64_t i = 0; i < m; ++i) {
    ks_add(&rr, r[i] * r[i]);
    ks_add(&br, b[i] * r[i]);
  }
  for (int64_t j = 0; j < n; ++j) {
    ks_add(&l1, fabs(x[j]));
    if (fabs(g[j]) > ginf) ginf = fabs(g[j]);
    if (x[j] != 0) nz += 1;
  }
  double R2 = ks_val(&rr), bTr = ks_val(&br), X1 = ks_val(&l1);
  double P = 0.5 * R2 + lambda * X1;
  double s = (ginf > lambda) ? lambda / ginf : 1.0;
  double D = -s * bTr - 0.5 * s * s * R2;
  out[0] = P;
  out[1] = D;
  out[2] = P - D;
  out[3] = (P != 0) ? (P - D) / fabs(P) : 0.0;
  out[4] = ginf;
  out[5] = X1;
  out[6] = R2;
  out[7] = nz;
  free(r);
  free(g);
}

void orc_nnls_eval(int64_t m, int64_t n, const int32_t *col_ptr,
                   const int32_t *row_idx, const double *vals,
                   const double *b, const double *x, double out[8]) {
  double *r = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  double *g = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  csc_grad(m, n, col_ptr, row_idx, vals, b, x, r, g);
  ksum rr = {0, 0}, xg = {0, 0}, l1 = {0, 0};
  double kkt = 0, gmin = INFINITY, ginf = 0, nz = 0;
  for (int64_t i = 0; i < m; ++i) ks_add(&rr, r[i] * r[i]);
  for (int64_t j = 0; j < n; ++j) {
    double mn = fabs(fmin(x[j], g[j]));
    if (mn > kkt) kkt = mn;
    if (g[j] < gmin) gmin = g[j];
    if (fabs(g[j]) > ginf) ginf = fabs(g[j]);
    ks_add(&xg, x[j] * g[j]);
    ks_add(&l1, fabs(x[j]));
    if (x[j] != 0) nz += 1;
  }
  out[0] = 0.5 * ks_val(&rr);
  out[1] = kkt;
  out[2] = ks_val(&xg);
  out[3] = (n > 0) ? gmin : 0.0;
  out[4] = ginf;
  out[5] = ks_val(&l1);
  out[6] = ks_val(&rr);
  out[7] = nz;
  free(r);
  free(g);
}

/* ------------------------------------------------------------------------
 * hinge-SVM dual
 * ---------------------------------------------------------------------- */
void orc_svm_lipschitz(int64_t ns, const int32_t *row_ptr, const double *vals,
                       double beta, double *L) {
  orc_csc_lipschitz(ns, row_ptr, vals, beta, L); /* same rule, rows = samples */
}

void orc_svm_w(int64_t ns, int64_t nf, const int32_t *row_ptr,
               const int32_t *col_idx, const double *vals, const double *y,
               double lambda, double *alpha, double *w) {
  const double delta = 1.0 / (lambda * (double)ns);
  for (int64_t f = 0; f < nf; ++f) w[f] = 0;
  for (int64_t i = 0; i < ns; ++i) {
    if (alpha[i] > 1) alpha[i] = 1;
    if (alpha[i] < 0) alpha[i] = 0;
    for (int32_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p)
      w[col_idx[p]] += delta * y[i] * vals[p] * alpha[i];
  }
}

static double svm_delta(const int32_t *row_ptr, const int32_t *col_idx,
                        const double *vals, const double *y, const double *L,
                        double lambda, int64_t ns, const double *alpha,
                        const double *w, int32_t i) {
  double t = 0;
  for (int32_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p) t += vals[p] * w[col_idx[p]];
  t = (1 - t * y[i]) * lambda * (double)ns * L[i];
  if (t < -alpha[i])
    t = -alpha[i];
  else if (t > 1 - alpha[i])
    t = 1 - alpha[i];
  return t;
}

void orc_svm_run(int64_t ns, int64_t nf, const int32_t *row_ptr,
                 const int32_t *col_idx, const double *vals, const double *y,
                 const double *L, double lambda, uint64_t seed, int64_t tau,
                 uint64_t step0, uint64_t nsteps, double *alpha, double *w) {
  (void)nf;
  int32_t *S = (int32_t *)malloc(sizeof(int32_t) * (size_t)tau);
  double *d = (double *)malloc(sizeof(double) * (size_t)tau);
  const double delta = 1.0 / (lambda * (double)ns);
  for (uint64_t s = 0; s < nsteps; ++s) {
    uint32_t k[4];
    orc_step_keys(seed, step0 + s, k);
    for (int64_t i = 0; i < tau; ++i) S[i] = (int32_t)orc_perm((uint32_t)i, (uint32_t)ns, k);
    for (int64_t i = 0; i < tau; ++i)
      d[i] = svm_delta(row_ptr, col_idx, vals, y, L, lambda, ns, alpha, w, S[i]);
    for (int64_t i = 0; i < tau; ++i) {
      int32_t j = S[i];
      alpha[j] += d[i];
      for (int32_t p = row_ptr[j]; p < row_ptr[j + 1]; ++p)
        w[col_idx[p]] += d[i] * vals[p] * y[j] * delta;
    }
  }
  free(S);
  free(d);
}

/* serial SDCA steps on the given samples, in order (the one-worker case of
 * the asynchronous executor, MulticoreEngineExecutor.h:83-89, with
 * MulticoreSVMDualLoss.h:119-141 as the step) */
void orc_svm_list(int64_t ns, int64_t nf, const int32_t *row_ptr, const int32_t *col_idx,
                  const double *vals, const double *y, const double *L, double lambda,
                  const int32_t *coords, uint64_t count, double *alpha, double *w) {
  (void)nf;
  const double delta = 1.0 / (lambda * (double)ns);
  for (uint64_t k = 0; k < count; ++k) {
    const int32_t j = coords[k];
    const double d = svm_delta(row_ptr, col_idx, vals, y, L, lambda, ns, alpha, w, j);
    alpha[j] += d;
    for (int32_t p = row_ptr[j]; p < row_ptr[j + 1]; ++p)
      w[col_idx[p]] += d * vals[p] * y[j] * delta;
  }
}

void orc_svm_eval(int64_t ns, int64_t nf, const int32_t *row_ptr,
                  const int32_t *col_idx, const double *vals, const double *y,
                  double lambda, const double *alpha, double out[8]) {
  double *a = (double *)malloc(sizeof(double) * (size_t)(ns > 0 ? ns : 1));
  double *w = (double *)malloc(sizeof(double) * (size_t)(nf > 0 ? nf : 1));
  memcpy(a, alpha, sizeof(double) * (size_t)ns);
  orc_svm_w(ns, nf, row_ptr, col_idx, vals, y, lambda, a, w);
  ksum ww = {0, 0}, hl = {0, 0}, sa = {0, 0};
  int64_t good = 0;
  for (int64_t f = 0; f < nf; ++f) ks_add(&ww, w[f] * w[f]);
  for (int64_t i = 0; i < ns; ++i) {
    ksum e = {0, 0};
    for (int32_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p) ks_add(&e, vals[p] * w[col_idx[p]]);
    double err = ks_val(&e);
    if (y[i] * err > 0) ++good;
    err = 1 - y[i] * err;
    if (err < 0) err = 0;
    ks_add(&hl, err);
    ks_add(&sa, a[i]);
  }
  double W2 = ks_val(&ww);
  double P = lambda * 0.5 * W2 + ks_val(&hl) / (double)ns;
  double D = -lambda * 0.5 * W2 + ks_val(&sa) / (double)ns;
  out[0] = P - D;
  out[1] = P;
  out[2] = D;
  out[3] = (double)good / (double)ns;
  out[4] = (P != 0) ? (P - D) / fabs(P) : 0.0;
  out[5] = W2;
  out[6] = ks_val(&sa);
  out[7] = ks_val(&hl);
  free(a);
  free(w);
}

/* ------------------------------------------------------------------------
 * asynchronous OpenMP restatement of MulticoreEngineExecutor + atomics
 * ---------------------------------------------------------------------- */
static void cas_add(double *a, double v) {
  uint64_t *p = (uint64_t *)a;
  uint64_t oldb = __atomic_load_n(p, __ATOMIC_RELAXED);
  for (;;) {
    double o, nv;
    memcpy(&o, &oldb, sizeof o);
    nv = o + v;
    uint64_t newb;
    memcpy(&newb, &nv, sizeof newb);
    if (__atomic_compare_exchange_n(p, &oldb, newb, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
      return;
  }
}
static double racy_load(const double *a) {
  uint64_t b = __atomic_load_n((const uint64_t *)a, __ATOMIC_RELAXED);
  double d;
  memcpy(&d, &b, sizeof d);
  return d;
}

/* update k of asynchronous worker w: 64 bits of Philox at counter
 * (k/2, w, "PCDA") mapped to [0,n) by the high word of the product */
#define ASYNC_TAG 0x50434441u

static uint32_t async_coord(uint64_t seed, uint32_t w, uint64_t k, uint32_t n) {
  uint64_t blk = k >> 1;
  uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), w, ASYNC_TAG};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  orc_philox4x32_10(ctr, key, o);
  uint32_t lo = (k & 1) ? o[2] : o[0], hi = (k & 1) ? o[3] : o[1];
  /* floor((hi 2^32 + lo) n / 2^64) with a 128-bit product */
  unsigned __int128 u = ((unsigned __int128)(((uint64_t)hi << 32) | lo)) * n;
  return (uint32_t)(u >> 64);
}

void orc_async_coords(uint64_t seed, uint32_t worker, uint64_t k0, uint64_t count,
                      uint32_t n, int32_t *out) {
  for (uint64_t k = 0; k < count; ++k)
    out[k] = (int32_t)async_coord(seed, worker, k0 + k, n);
}

void orc_cd_list(int loss, int64_t m, int64_t n, const int32_t *col_ptr,
                 const int32_t *row_idx, const double *vals, const double *L,
                 double lambda, const int32_t *coords, uint64_t count,
                 double *x, double *r) {
  (void)m;
  (void)n;
  for (uint64_t k = 0; k < count; ++k) {
    int32_t j = coords[k];
    double d = cd_delta(loss, col_ptr, row_idx, vals, L, lambda, x, r, j);
    x[j] += d;
    for (int32_t p = col_ptr[j]; p < col_ptr[j + 1]; ++p)
      r[row_idx[p]] += d * vals[p];
  }
}

int orc_cd_async_omp(int loss, int64_t m, int64_t n, const int32_t *col_ptr,
                     const int32_t *row_idx, const double *vals,
                     const double *L, double lambda, uint64_t seed,
                     uint64_t k0, uint64_t per_worker, int threads,
                     double *x, double *r) {
  (void)m;
  if (threads < 1) threads = 1;
#ifdef _OPENMP
  omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(static, 1)
  for (int w = 0; w < threads; ++w) {
    for (uint64_t k = 0; k < per_worker; ++k) {
      int32_t j = (int32_t)async_coord(seed, (uint32_t)w, k0 + k, (uint32_t)n);
      double g = 0;
      for (int32_t p = col_ptr[j]; p < col_ptr[j + 1]; ++p)
        g += vals[p] * racy_load(&r[row_idx[p]]);
      double xj = racy_load(&x[j]);
      double d;
      if (loss == 10) {
        double v = xj - L[j] * g;
        d = (v > 0 ? v : 0.0) - xj;
      } else {
        d = orc_soft_threshold(L[j] * lambda, xj - L[j] * g) - xj;
      }
      cas_add(&x[j], d);
      for (int32_t p = col_ptr[j]; p < col_ptr[j + 1]; ++p)
        cas_add(&r[row_idx[p]], d * vals[p]);
    }
  }
  return threads;
}
