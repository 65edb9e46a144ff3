// solver.hip -- the C ABI of include/acdc_hip.h: problem and solver handles,
// device data layout, kernel dispatch and the FP64 evaluation.
//
// Layout in HBM (one solver):
//   A  coordinate-major segments  ptr int32[n+1], idx int32[nnz], val T[nnz]
//   R  residual-major segments    ptr int32[m+1], idx int32[nnz], val T[nnz]
//   x double[n], L double[n], b double[m] (Lasso/NNLS) or y double[n] (SVM)
//   res T[m] (running residual r or w), res64 double[m] (FP64 refresh; the
//   same buffer as res on the FP64 path), coord int32[tau], coef double[tau]
// T = double on the FP64 path, float on the FP32+refresh path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/acdc_hip.h"
#include "kernels.h"

using namespace acdc;

// ----------------------------------------------------------------- errors
namespace {
thread_local std::string g_err;

acdc_status fail(acdc_status st, const std::string &msg) {
  g_err = msg;
  return st;
}
}  // namespace

namespace acdc {
void set_error(const char *msg) { g_err = msg; }
}  // namespace acdc

namespace {

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(ACDC_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }
}  // namespace

// ----------------------------------------------------------------- problem
struct acdc_problem {
  int kind;          // 0: column-major input (Lasso/NNLS), 1: sample rows (SVM)
  int64_t ncoord;    // coordinates: columns n, or samples
  int64_t nres;      // residual entries: rows m, or features
  int64_t nnz;
  std::vector<int32_t> cptr, cidx;  // coordinate-major
  std::vector<double> cval;
  std::vector<int32_t> rptr, ridx;  // residual-major
  std::vector<double> rval;
  std::vector<double> rhs;  // b (m) or labels y (samples)
  int32_t omega;
};

namespace {
// stable counting-sort transpose of a segment matrix (getCSR_from_CSC,
// helpers/matrix_conversions.h:65-104)
void transpose(int64_t nseg, int64_t nout, const std::vector<int32_t> &ptr,
               const std::vector<int32_t> &idx, const std::vector<double> &val,
               std::vector<int32_t> &optr, std::vector<int32_t> &oidx,
               std::vector<double> &oval) {
  const int64_t nnz = (int64_t)idx.size();
  optr.assign(nout + 1, 0);
  for (int64_t p = 0; p < nnz; ++p) optr[idx[p] + 1]++;
  for (int64_t k = 0; k < nout; ++k) optr[k + 1] += optr[k];
  std::vector<int32_t> fill(optr.begin(), optr.end() - 1);
  oidx.resize(nnz);
  oval.resize(nnz);
  for (int64_t s = 0; s < nseg; ++s)
    for (int32_t p = ptr[s]; p < ptr[s + 1]; ++p) {
      const int32_t o = fill[idx[p]]++;
      oidx[o] = (int32_t)s;
      oval[o] = val[p];
    }
}

acdc_status check_segments(int64_t nseg, int64_t nidx, int64_t nnz,
                           const int32_t *ptr, const int32_t *idx,
                           const char *what) {
  if (nseg < 1 || nidx < 1 || nnz < 0)
    return fail(ACDC_ERR_INVALID, std::string(what) + ": sizes must be positive");
  if (nnz > INT32_MAX || nseg >= INT32_MAX || nidx >= INT32_MAX)
    return fail(ACDC_ERR_INVALID, std::string(what) + ": sizes exceed int32 indexing");
  if (ptr[0] != 0 || ptr[nseg] != nnz)
    return fail(ACDC_ERR_INVALID, std::string(what) + ": ptr[0] must be 0 and ptr[n] = nnz");
  for (int64_t s = 0; s < nseg; ++s)
    if (ptr[s + 1] < ptr[s])
      return fail(ACDC_ERR_INVALID, std::string(what) + ": ptr must be non-decreasing");
  for (int64_t p = 0; p < nnz; ++p)
    if (idx[p] < 0 || idx[p] >= nidx)
      return fail(ACDC_ERR_INVALID, std::string(what) + ": index out of range");
  return ACDC_OK;
}

int32_t max_segment(const std::vector<int32_t> &ptr) {
  int32_t om = 1;  // the reference starts omega at 1 (MultiCoreSolver.cpp:24)
  for (size_t k = 0; k + 1 < ptr.size(); ++k) om = std::max(om, ptr[k + 1] - ptr[k]);
  return om;
}
}  // namespace

extern "C" const char *acdc_last_error(void) { return g_err.c_str(); }

extern "C" acdc_status acdc_device_count(int32_t *count) {
  if (!count) return fail(ACDC_ERR_INVALID, "count is null");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  int good = 0;
  for (int d = 0; d < n; ++d) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) == hipSuccess &&
        std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
      ++good;
  }
  *count = good;
  return ACDC_OK;
}

extern "C" acdc_status acdc_problem_create_csc(int64_t m, int64_t n, int64_t nnz,
                                               const int32_t *col_ptr,
                                               const int32_t *row_idx,
                                               const double *vals,
                                               const double *b,
                                               acdc_problem **out) {
  if (!out || !col_ptr || (nnz > 0 && (!row_idx || !vals)) || !b)
    return fail(ACDC_ERR_INVALID, "acdc_problem_create_csc: null argument");
  acdc_status st = check_segments(n, m, nnz, col_ptr, row_idx, "acdc_problem_create_csc");
  if (st != ACDC_OK) return st;
  acdc_problem *p = new (std::nothrow) acdc_problem();
  if (!p) return fail(ACDC_ERR_NOMEM, "acdc_problem_create_csc: out of host memory");
  try {
    p->kind = 0;
    p->ncoord = n;
    p->nres = m;
    p->nnz = nnz;
    p->cptr.assign(col_ptr, col_ptr + n + 1);
    p->cidx.assign(row_idx, row_idx + nnz);
    p->cval.assign(vals, vals + nnz);
    p->rhs.assign(b, b + m);
    transpose(n, m, p->cptr, p->cidx, p->cval, p->rptr, p->ridx, p->rval);
    p->omega = max_segment(p->rptr);
  } catch (const std::bad_alloc &) {
    delete p;
    return fail(ACDC_ERR_NOMEM, "acdc_problem_create_csc: out of host memory");
  }
  *out = p;
  return ACDC_OK;
}

extern "C" acdc_status acdc_problem_create_csr(int64_t n_samples, int64_t n_features,
                                               int64_t nnz, const int32_t *row_ptr,
                                               const int32_t *col_idx,
                                               const double *vals,
                                               const double *labels,
                                               acdc_problem **out) {
  if (!out || !row_ptr || (nnz > 0 && (!col_idx || !vals)) || !labels)
    return fail(ACDC_ERR_INVALID, "acdc_problem_create_csr: null argument");
  acdc_status st = check_segments(n_samples, n_features, nnz, row_ptr, col_idx,
                                  "acdc_problem_create_csr");
  if (st != ACDC_OK) return st;
  acdc_problem *p = new (std::nothrow) acdc_problem();
  if (!p) return fail(ACDC_ERR_NOMEM, "acdc_problem_create_csr: out of host memory");
  try {
    p->kind = 1;
    p->ncoord = n_samples;
    p->nres = n_features;
    p->nnz = nnz;
    p->cptr.assign(row_ptr, row_ptr + n_samples + 1);
    p->cidx.assign(col_idx, col_idx + nnz);
    p->cval.assign(vals, vals + nnz);
    p->rhs.assign(labels, labels + n_samples);
    transpose(n_samples, n_features, p->cptr, p->cidx, p->cval, p->rptr, p->ridx, p->rval);
    p->omega = max_segment(p->rptr);
  } catch (const std::bad_alloc &) {
    delete p;
    return fail(ACDC_ERR_NOMEM, "acdc_problem_create_csr: out of host memory");
  }
  *out = p;
  return ACDC_OK;
}

extern "C" acdc_status acdc_problem_destroy(acdc_problem *p) {
  delete p;
  return ACDC_OK;
}

extern "C" acdc_status acdc_problem_omega(const acdc_problem *p, int32_t *omega) {
  if (!p || !omega) return fail(ACDC_ERR_INVALID, "acdc_problem_omega: null argument");
  *omega = p->omega;
  return ACDC_OK;
}

// ------------------------------------------------------------------ solver
struct DevSeg {
  int32_t *ptr = nullptr, *idx = nullptr;
  void *val = nullptr;
};

struct acdc_solver {
  int loss = 0;
  acdc_settings set{};
  int64_t ncoord = 0, nres = 0, nnz = 0;
  double beta = 1.0;
  int lanes = 64;
  bool fp32 = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevSeg A, R;
  double *x = nullptr, *L = nullptr, *rhs = nullptr;
  void *res = nullptr;
  double *res64 = nullptr;
  int32_t *coord = nullptr;
  double *coef = nullptr;
  double *partial = nullptr;
  int32_t *samples = nullptr;
  int64_t nblk_eval = 0;
  uint64_t step = 0;
  uint64_t updates = 0;
  uint64_t updates_since_refresh = 0;
  double kernel_ms = 0, eval_ms = 0;
  std::vector<void *> allocs;
};

namespace {
template <typename T>
Seg<T> seg(const DevSeg &d) {
  return Seg<T>{d.ptr, d.idx, static_cast<const T *>(d.val)};
}

acdc_status dalloc(acdc_solver *s, void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess)
    return fail(ACDC_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  s->allocs.push_back(*p);
  return ACDC_OK;
}

template <typename P>
acdc_status dalloc_t(acdc_solver *s, P **p, size_t count) {
  void *v = nullptr;
  acdc_status st = dalloc(s, &v, count * sizeof(P));
  *p = static_cast<P *>(v);
  return st;
}

acdc_status upload_seg(acdc_solver *s, DevSeg &d, const std::vector<int32_t> &ptr,
                       const std::vector<int32_t> &idx,
                       const std::vector<double> &val) {
  acdc_status st;
  if ((st = dalloc_t(s, &d.ptr, ptr.size())) != ACDC_OK) return st;
  if ((st = dalloc_t(s, &d.idx, idx.size())) != ACDC_OK) return st;
  HIP_TRY(hipMemcpy(d.ptr, ptr.data(), ptr.size() * 4, hipMemcpyHostToDevice));
  if (!idx.empty())
    HIP_TRY(hipMemcpy(d.idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice));
  if (s->fp32) {
    std::vector<float> f(val.begin(), val.end());
    if ((st = dalloc(s, &d.val, f.size() * 4)) != ACDC_OK) return st;
    if (!f.empty())
      HIP_TRY(hipMemcpy(d.val, f.data(), f.size() * 4, hipMemcpyHostToDevice));
  } else {
    if ((st = dalloc(s, &d.val, val.size() * 8)) != ACDC_OK) return st;
    if (!val.empty())
      HIP_TRY(hipMemcpy(d.val, val.data(), val.size() * 8, hipMemcpyHostToDevice));
  }
  return ACDC_OK;
}

StepScalars scalars(const acdc_solver *s) {
  StepScalars sc;
  sc.lambda = s->set.lambda;
  sc.lambda_n = s->set.lambda * (double)s->ncoord;
  sc.inv_lambda_n = 1.0 / sc.lambda_n;
  return sc;
}

int auto_lanes(int64_t nnz, int64_t nseg) {
  const double avg = nseg > 0 ? (double)nnz / (double)nseg : 0.0;
  if (avg <= 8) return 8;
  if (avg <= 16) return 16;
  if (avg <= 32) return 32;
  return 64;
}

// ---- dispatch helpers: value type x lanes x loss
template <template <typename, int, int> class F, typename... Args>
void dispatch3(const acdc_solver *s, Args &&...args) {
  auto go = [&](auto tag_t, auto tag_g) {
    using T = typename decltype(tag_t)::type;
    constexpr int G = decltype(tag_g)::value;
    switch (s->loss) {
      case kLasso: F<T, G, kLasso>::run(s, args...); break;
      case kNnls: F<T, G, kNnls>::run(s, args...); break;
      default: F<T, G, kSvm>::run(s, args...); break;
    }
  };
  auto by_g = [&](auto tag_t) {
    switch (s->lanes) {
      case 8: go(tag_t, std::integral_constant<int, 8>{}); break;
      case 16: go(tag_t, std::integral_constant<int, 16>{}); break;
      case 32: go(tag_t, std::integral_constant<int, 32>{}); break;
      default: go(tag_t, std::integral_constant<int, 64>{}); break;
    }
  };
  if (s->fp32)
    by_g(std::type_identity<float>{});
  else
    by_g(std::type_identity<double>{});
}

// part: 1 = gather (phase A), 2 = scatter (phase B), 3 = both
template <typename T, int G, int LOSS>
struct LaunchStep {
  static void run(const acdc_solver *s, uint64_t step, int part) {
    const int64_t tau = s->set.tau;
    const unsigned blocks = (unsigned)div_up(tau * G, kBlock);
    if (part & 1)
      k_step_gather<T, G, LOSS><<<blocks, kBlock, 0, s->stream>>>(
          seg<T>(s->A), static_cast<const T *>(s->res), s->x, s->L, s->rhs,
          scalars(s), Perm::make((uint32_t)s->ncoord), s->set.seed, step, tau,
          s->coord, s->coef);
    if (part & 2)
      k_step_scatter<T, G><<<blocks, kBlock, 0, s->stream>>>(
          seg<T>(s->A), static_cast<T *>(s->res), s->coord, s->coef, tau);
  }
};

template <typename T, int G, int LOSS>
struct LaunchSerial {
  static void run(const acdc_solver *s, uint64_t step0, uint64_t nsteps) {
    const Perm perm = Perm::make((uint32_t)s->ncoord);
    k_serial<T, LOSS><<<1, 64, 0, s->stream>>>(
        seg<T>(s->A), static_cast<T *>(s->res), s->x, s->L, s->rhs, scalars(s),
        perm, s->set.seed, step0, nsteps);
  }
};

template <typename T, int G, int LOSS>
struct LaunchLipschitz {
  static void run(const acdc_solver *s) {
    const unsigned blocks = (unsigned)div_up(s->ncoord * G, kBlock);
    k_lipschitz<T, G><<<blocks, kBlock, 0, s->stream>>>(seg<T>(s->A), s->ncoord,
                                                        s->beta, s->L);
  }
};

template <typename T, int G, int LOSS>
struct LaunchRefresh {
  static void run(const acdc_solver *s, bool write_back) {
    const unsigned blocks = (unsigned)div_up(s->nres * G, kBlock);
    T *resT = write_back && static_cast<void *>(s->res64) != s->res
                  ? static_cast<T *>(s->res)
                  : nullptr;
    k_refresh<T, G, LOSS><<<blocks, kBlock, 0, s->stream>>>(
        seg<T>(s->R), s->nres, s->x, s->rhs, s->rhs, scalars(s).inv_lambda_n,
        s->res64, resT);
  }
};

template <typename T, int G, int LOSS>
struct LaunchEvalCoord {
  static void run(const acdc_solver *s) {
    k_eval_coord<T, G, LOSS><<<(unsigned)s->nblk_eval, kBlock, 0, s->stream>>>(
        seg<T>(s->A), s->ncoord, s->res64, s->x, s->rhs, s->partial);
  }
};

// residual-major lanes: chosen from the residual segments' mean length
int res_lanes(const acdc_solver *s) { return auto_lanes(s->nnz, s->nres); }

acdc_status launch_refresh(acdc_solver *s, bool write_back) {
  const int keep = s->lanes;
  s->lanes = res_lanes(s);
  dispatch3<LaunchRefresh>(s, write_back);
  s->lanes = keep;
  HIP_TRY(hipGetLastError());
  return ACDC_OK;
}

acdc_status timed_begin(acdc_solver *s) {
  HIP_TRY(hipEventRecord(s->ev0, s->stream));
  return ACDC_OK;
}
acdc_status timed_end(acdc_solver *s, double *acc) {
  HIP_TRY(hipEventRecord(s->ev1, s->stream));
  HIP_TRY(hipEventSynchronize(s->ev1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  *acc += ms;
  return ACDC_OK;
}

void destroy_solver(acdc_solver *s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (void *p : s->allocs) (void)hipFree(p);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}
}  // namespace

extern "C" acdc_status acdc_solver_create(const acdc_problem *p, int32_t loss,
                                          const acdc_settings *set,
                                          acdc_solver **out) {
  if (!p || !set || !out) return fail(ACDC_ERR_INVALID, "acdc_solver_create: null argument");
  if (loss != kLasso && loss != kNnls && loss != kSvm)
    return fail(ACDC_ERR_INVALID, "acdc_solver_create: loss must be 0 (L1-square), 2 (SVM dual hinge) or 10 (NNLS)");
  if ((loss == kSvm) != (p->kind == 1))
    return fail(ACDC_ERR_STATE, "acdc_solver_create: SVM needs a sample-major (CSR) problem, losses 0/10 a column-major (CSC) one");
  if (set->tau < 1 || set->tau > p->ncoord)
    return fail(ACDC_ERR_INVALID, "acdc_solver_create: tau must lie in [1, n]");
  if (set->precision != ACDC_PREC_FP64 && set->precision != ACDC_PREC_FP32_REFRESH)
    return fail(ACDC_ERR_INVALID, "acdc_solver_create: unknown precision");
  if (set->step_mode != ACDC_STEP_SYNC)
    return fail(ACDC_ERR_INVALID, "acdc_solver_create: unknown step mode");
  if (set->lanes_per_coord != 0 && set->lanes_per_coord != 8 && set->lanes_per_coord != 16 &&
      set->lanes_per_coord != 32 && set->lanes_per_coord != 64)
    return fail(ACDC_ERR_INVALID, "acdc_solver_create: lanes_per_coord must be 0, 8, 16, 32 or 64");
  if (!(set->lambda >= 0) || (loss == kSvm && !(set->lambda > 0)))
    return fail(ACDC_ERR_INVALID, "acdc_solver_create: lambda must be >= 0 (> 0 for the SVM)");
  if (set->refresh_period < 0)
    return fail(ACDC_ERR_INVALID, "acdc_solver_create: refresh_period must be >= 0");
  int32_t ndev = 0;
  acdc_device_count(&ndev);
  if (ndev < 1) return fail(ACDC_ERR_NO_DEVICE, "acdc_solver_create: no gfx950 device visible");

  acdc_solver *s = new (std::nothrow) acdc_solver();
  if (!s) return fail(ACDC_ERR_NOMEM, "acdc_solver_create: out of host memory");
  s->loss = loss;
  s->set = *set;
  s->ncoord = p->ncoord;
  s->nres = p->nres;
  s->nnz = p->nnz;
  s->fp32 = set->precision == ACDC_PREC_FP32_REFRESH;
  s->lanes = set->lanes_per_coord ? set->lanes_per_coord : auto_lanes(p->nnz, p->ncoord);
  switch (set->beta_mode) {
    case ACDC_BETA_OMEGA:
      s->beta = p->ncoord > 1 ? 1.0 + (double)(p->omega - 1) * (double)(set->tau - 1) /
                                          ((double)p->ncoord - 1.0)
                              : 1.0;
      break;
    case ACDC_BETA_FORCED:
      if (!(set->forced_beta > 0)) {
        delete s;
        return fail(ACDC_ERR_INVALID, "acdc_solver_create: forced_beta must be > 0");
      }
      s->beta = set->forced_beta;
      break;
    case ACDC_BETA_NONE: s->beta = 1.0; break;
    default: delete s; return fail(ACDC_ERR_INVALID, "acdc_solver_create: unknown beta_mode");
  }

  acdc_status st = ACDC_OK;
  auto bail = [&](acdc_status e) {
    std::string msg = g_err;
    destroy_solver(s);
    g_err = msg;
    return e;
  };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
    return bail(fail(ACDC_ERR_HIP, "acdc_solver_create: stream/event creation failed"));
  if ((st = upload_seg(s, s->A, p->cptr, p->cidx, p->cval)) != ACDC_OK) return bail(st);
  if ((st = upload_seg(s, s->R, p->rptr, p->ridx, p->rval)) != ACDC_OK) return bail(st);
  const size_t tsz = s->fp32 ? 4 : 8;
  if ((st = dalloc_t(s, &s->x, s->ncoord)) != ACDC_OK) return bail(st);
  if ((st = dalloc_t(s, &s->L, s->ncoord)) != ACDC_OK) return bail(st);
  if ((st = dalloc_t(s, &s->rhs, p->rhs.size())) != ACDC_OK) return bail(st);
  if ((st = dalloc(s, &s->res, s->nres * tsz)) != ACDC_OK) return bail(st);
  if (s->fp32) {
    if ((st = dalloc_t(s, &s->res64, s->nres)) != ACDC_OK) return bail(st);
  } else {
    s->res64 = static_cast<double *>(s->res);
  }
  if ((st = dalloc_t(s, &s->coord, set->tau)) != ACDC_OK) return bail(st);
  if ((st = dalloc_t(s, &s->coef, set->tau)) != ACDC_OK) return bail(st);
  s->nblk_eval = 1024;
  if ((st = dalloc_t(s, &s->partial, (size_t)s->nblk_eval * kStats)) != ACDC_OK) return bail(st);
  if (hipMemcpy(s->rhs, p->rhs.data(), p->rhs.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(s->x, 0, s->ncoord * 8) != hipSuccess)
    return bail(fail(ACDC_ERR_HIP, "acdc_solver_create: upload failed"));
  dispatch3<LaunchLipschitz>(s);
  if ((st = launch_refresh(s, true)) != ACDC_OK) return bail(st);
  if (hipStreamSynchronize(s->stream) != hipSuccess)
    return bail(fail(ACDC_ERR_HIP, "acdc_solver_create: setup kernels failed"));
  *out = s;
  return ACDC_OK;
}

extern "C" acdc_status acdc_solver_destroy(acdc_solver *s) {
  destroy_solver(s);
  return ACDC_OK;
}

extern "C" acdc_status acdc_run_updates(acdc_solver *s, uint64_t n_updates) {
  if (!s) return fail(ACDC_ERR_INVALID, "acdc_run_updates: null solver");
  const uint64_t tau = (uint64_t)s->set.tau;
  const uint64_t nsteps = (n_updates + tau - 1) / tau;
  if (nsteps == 0) return ACDC_OK;
  const uint64_t period =
      s->fp32 && s->set.refresh_period > 0 ? (uint64_t)s->set.refresh_period * (uint64_t)s->ncoord : 0;
  acdc_status st;
  if ((st = timed_begin(s)) != ACDC_OK) return st;
  uint64_t done = 0;
  while (done < nsteps) {
    uint64_t chunk = nsteps - done;
    if (period) {
      const uint64_t left = period > s->updates_since_refresh ? period - s->updates_since_refresh : 1;
      chunk = std::min(chunk, (left + tau - 1) / tau);
    }
    if (tau == 1) {
      dispatch3<LaunchSerial>(s, s->step, chunk);
    } else {
      for (uint64_t k = 0; k < chunk; ++k) dispatch3<LaunchStep>(s, s->step + k, 3);
    }
    s->step += chunk;
    s->updates += chunk * tau;
    s->updates_since_refresh += chunk * tau;
    done += chunk;
    if (period && s->updates_since_refresh >= period) {
      if ((st = launch_refresh(s, true)) != ACDC_OK) return st;
      s->updates_since_refresh = 0;
    }
  }
  HIP_TRY(hipGetLastError());
  return timed_end(s, &s->kernel_ms);
}

extern "C" acdc_status acdc_recompute_residuals(acdc_solver *s) {
  if (!s) return fail(ACDC_ERR_INVALID, "acdc_recompute_residuals: null solver");
  acdc_status st;
  if ((st = timed_begin(s)) != ACDC_OK) return st;
  if ((st = launch_refresh(s, true)) != ACDC_OK) return st;
  s->updates_since_refresh = 0;
  return timed_end(s, &s->eval_ms);
}

extern "C" acdc_status acdc_report_get(acdc_solver *s, acdc_report *out) {
  if (!s || !out) return fail(ACDC_ERR_INVALID, "acdc_report_get: null argument");
  acdc_status st;
  if ((st = timed_begin(s)) != ACDC_OK) return st;
  if (s->fp32) {
    if ((st = launch_refresh(s, true)) != ACDC_OK) return st;
    s->updates_since_refresh = 0;
  }
  dispatch3<LaunchEvalCoord>(s);
  HIP_TRY(hipGetLastError());
  std::vector<double> pc((size_t)s->nblk_eval * kStats), pr((size_t)s->nblk_eval * kStats);
  HIP_TRY(hipMemcpyAsync(pc.data(), s->partial, pc.size() * 8, hipMemcpyDeviceToHost, s->stream));
  k_eval_res<<<(unsigned)s->nblk_eval, kBlock, 0, s->stream>>>(
      s->nres, s->res64, s->loss == kSvm ? nullptr : s->rhs, s->partial);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(pr.data(), s->partial, pr.size() * 8, hipMemcpyDeviceToHost, s->stream));
  if ((st = timed_end(s, &s->eval_ms)) != ACDC_OK) return st;

  // host: fixed-order compensated sums of the block rows
  auto reduce = [&](const std::vector<double> &v, int k, int how) {
    double sum = 0, c = 0, ext = how == 1 ? -INFINITY : INFINITY;
    for (int64_t b = 0; b < s->nblk_eval; ++b) {
      const double a = v[(size_t)b * kStats + k];
      if (how == 1) {
        ext = std::max(ext, a);
      } else if (how == 2) {
        ext = std::min(ext, a);
      } else {
        const double t = sum + a;
        c += std::fabs(sum) >= std::fabs(a) ? (sum - t) + a : (a - t) + sum;
        sum = t;
      }
    }
    return how == 0 ? sum + c : ext;
  };
  acdc_report r{};
  const double R2 = reduce(pr, 0, 0), bTr = reduce(pr, 1, 0);
  const double lam = s->set.lambda;
  if (s->loss == kLasso) {
    const double ginf = reduce(pc, 0, 1), x1 = reduce(pc, 1, 0);
    r.primal = 0.5 * R2 + lam * x1;
    const double sc = ginf > lam ? lam / ginf : 1.0;
    r.dual = -sc * bTr - 0.5 * sc * sc * R2;
    r.grad_inf = ginf;
    r.x_l1 = x1;
    r.res_sq = R2;
    r.nnz_x = reduce(pc, 2, 0);
  } else if (s->loss == kNnls) {
    const double xg = reduce(pc, 1, 0);
    r.primal = 0.5 * R2;
    r.dual = r.primal - xg;
    r.kkt = reduce(pc, 0, 1);
    r.grad_inf = reduce(pc, 3, 1);
    r.x_l1 = reduce(pc, 4, 0);
    r.res_sq = R2;
    r.nnz_x = reduce(pc, 5, 0);
  } else {
    const double n = (double)s->ncoord;
    const double hinge = reduce(pc, 0, 0), good = reduce(pc, 1, 0), sa = reduce(pc, 2, 0);
    r.primal = lam * 0.5 * R2 + hinge / n;
    r.dual = -lam * 0.5 * R2 + sa / n;
    r.accuracy = good / n;
    r.x_l1 = sa;
    r.res_sq = R2;
    r.nnz_x = reduce(pc, 3, 0);
  }
  r.gap = r.primal - r.dual;
  r.rel_gap = r.primal != 0 ? r.gap / std::fabs(r.primal) : 0.0;
  r.updates = s->updates;
  r.epochs = (double)s->updates / (double)s->ncoord;
  r.kernel_ms = s->kernel_ms;
  r.eval_ms = s->eval_ms;
  *out = r;
  return ACDC_OK;
}

extern "C" acdc_status acdc_get_x(acdc_solver *s, double *x, int64_t n) {
  if (!s || !x) return fail(ACDC_ERR_INVALID, "acdc_get_x: null argument");
  if (n != s->ncoord) return fail(ACDC_ERR_INVALID, "acdc_get_x: length must equal n");
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(x, s->x, n * 8, hipMemcpyDeviceToHost));
  return ACDC_OK;
}

extern "C" acdc_status acdc_set_x(acdc_solver *s, const double *x, int64_t n) {
  if (!s || !x) return fail(ACDC_ERR_INVALID, "acdc_set_x: null argument");
  if (n != s->ncoord) return fail(ACDC_ERR_INVALID, "acdc_set_x: length must equal n");
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(s->x, x, n * 8, hipMemcpyHostToDevice));
  acdc_status st = launch_refresh(s, true);
  if (st != ACDC_OK) return st;
  s->updates_since_refresh = 0;
  HIP_TRY(hipStreamSynchronize(s->stream));
  return ACDC_OK;
}

extern "C" acdc_status acdc_get_residuals(acdc_solver *s, double *r, int64_t len) {
  if (!s || !r) return fail(ACDC_ERR_INVALID, "acdc_get_residuals: null argument");
  if (len != s->nres) return fail(ACDC_ERR_INVALID, "acdc_get_residuals: wrong length");
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (s->fp32) {
    std::vector<float> f((size_t)len);
    HIP_TRY(hipMemcpy(f.data(), s->res, len * 4, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < len; ++k) r[k] = f[k];
  } else {
    HIP_TRY(hipMemcpy(r, s->res, len * 8, hipMemcpyDeviceToHost));
  }
  return ACDC_OK;
}

extern "C" acdc_status acdc_get_lipschitz(acdc_solver *s, double *Li, int64_t n) {
  if (!s || !Li) return fail(ACDC_ERR_INVALID, "acdc_get_lipschitz: null argument");
  if (n != s->ncoord) return fail(ACDC_ERR_INVALID, "acdc_get_lipschitz: length must equal n");
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(Li, s->L, n * 8, hipMemcpyDeviceToHost));
  return ACDC_OK;
}

extern "C" acdc_status acdc_set_lambda(acdc_solver *s, double lambda) {
  if (!s) return fail(ACDC_ERR_INVALID, "acdc_set_lambda: null solver");
  if (!(lambda >= 0) || (s->loss == kSvm && !(lambda > 0)))
    return fail(ACDC_ERR_INVALID, "acdc_set_lambda: lambda must be >= 0 (> 0 for the SVM)");
  if (s->loss == kSvm && lambda != s->set.lambda) {
    const double f = s->set.lambda / lambda;
    const unsigned blocks = (unsigned)div_up(s->nres, kBlock);
    if (s->fp32) {
      k_scale<float><<<blocks, kBlock, 0, s->stream>>>(static_cast<float *>(s->res), s->nres, f);
      k_scale<double><<<blocks, kBlock, 0, s->stream>>>(s->res64, s->nres, f);
    } else {
      k_scale<double><<<blocks, kBlock, 0, s->stream>>>(s->res64, s->nres, f);
    }
    HIP_TRY(hipGetLastError());
  }
  s->set.lambda = lambda;
  return ACDC_OK;
}

extern "C" acdc_status acdc_get_beta(const acdc_solver *s, double *beta) {
  if (!s || !beta) return fail(ACDC_ERR_INVALID, "acdc_get_beta: null argument");
  *beta = s->beta;
  return ACDC_OK;
}

extern "C" acdc_status acdc_sample_coordinates(acdc_solver *s, uint64_t step0,
                                               uint64_t nsteps, int32_t *out) {
  if (!s || !out) return fail(ACDC_ERR_INVALID, "acdc_sample_coordinates: null argument");
  const int64_t total = (int64_t)nsteps * s->set.tau;
  if (total <= 0) return ACDC_OK;
  if (total > ((int64_t)1 << 30))
    return fail(ACDC_ERR_INVALID, "acdc_sample_coordinates: at most 2^30 samples per call");
  int32_t *d = nullptr;
  HIP_TRY(hipMalloc(&d, (size_t)total * 4));
  k_sample<<<(unsigned)div_up(total, kBlock), kBlock, 0, s->stream>>>(
      Perm::make((uint32_t)s->ncoord), s->set.seed, step0, nsteps, s->set.tau, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, (size_t)total * 4, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(ACDC_ERR_HIP, std::string("acdc_sample_coordinates: ") + hipGetErrorString(e));
  return ACDC_OK;
}

extern "C" acdc_status acdc_synchronize(acdc_solver *s) {
  if (!s) return fail(ACDC_ERR_INVALID, "acdc_synchronize: null solver");
  HIP_TRY(hipStreamSynchronize(s->stream));
  return ACDC_OK;
}

extern "C" acdc_status acdc_profile_step_kernels(acdc_solver *s, uint64_t nsteps,
                                                 double *gather_ms, double *scatter_ms) {
  if (!s || !gather_ms || !scatter_ms)
    return fail(ACDC_ERR_INVALID, "acdc_profile_step_kernels: null argument");
  if (s->set.tau < 2)
    return fail(ACDC_ERR_STATE, "acdc_profile_step_kernels: needs tau >= 2 (two-phase steps)");
  if (nsteps < 1 || nsteps > 4096)
    return fail(ACDC_ERR_INVALID, "acdc_profile_step_kernels: nsteps must lie in [1, 4096]");
  std::vector<hipEvent_t> ev(3 * nsteps, nullptr);
  acdc_status st = ACDC_OK;
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) st = fail(ACDC_ERR_HIP, "acdc_profile_step_kernels: event");
  double ga = 0, sc = 0;
  if (st == ACDC_OK) {
    for (uint64_t k = 0; k < nsteps; ++k) {
      (void)hipEventRecord(ev[3 * k], s->stream);
      dispatch3<LaunchStep>(s, s->step, 1);
      (void)hipEventRecord(ev[3 * k + 1], s->stream);
      dispatch3<LaunchStep>(s, s->step, 2);
      (void)hipEventRecord(ev[3 * k + 2], s->stream);
      s->step += 1;
      s->updates += (uint64_t)s->set.tau;
      s->updates_since_refresh += (uint64_t)s->set.tau;
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
      st = fail(ACDC_ERR_HIP, std::string("acdc_profile_step_kernels: ") + hipGetErrorString(e));
    } else {
      for (uint64_t k = 0; k < nsteps; ++k) {
        float a = 0, b = 0;
        (void)hipEventElapsedTime(&a, ev[3 * k], ev[3 * k + 1]);
        (void)hipEventElapsedTime(&b, ev[3 * k + 1], ev[3 * k + 2]);
        ga += a;
        sc += b;
      }
      s->kernel_ms += ga + sc;
    }
  }
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  *gather_ms = ga / (double)nsteps;
  *scatter_ms = sc / (double)nsteps;
  return st;
}
