// solver.hip -- the C ABI of include/acdc_hip.h: problem and solver handles,
// device data layout, kernel dispatch and the FP64 evaluation.
//
// Layout in HBM (one solver):
//   A   coordinate-major segments  ptr int32[n+1], idx int32[nnz], val T[nnz]
//       (FP32 path: plus val64 double[nnz], the FP64 values read by the
//       refresh, evaluation and Lipschitz passes)
//   x   double[n] (alpha for the SVM), L double[n] (1/(beta ||a_j||^2))
//   rhs double[m] (b) or double[n] (SVM labels)
//   res T[m]      FP64 path: running residual r (SVM: w)
//                 FP32 path: correction dr since the last refresh
//   r64 double[m] FP64 path: same buffer as res; FP32 path: anchor Ax - b
//   gref double[n] FP32 path: A^T r64 at the last refresh
//   coord int32[tau], coef double[tau]  staged between the two step phases
//   streamed scatter only: tent TEntry<T>[nnz] (the transposed copy, by
//   residual index) and cvec T[2n] (the step coefficients by coordinate)
// T = double on the FP64 path, float on the FP32+refresh path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/acdc_hip.h"
#include "problem.h"
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

#define HIP_TRY(expr)                                                               \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail(ACDC_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_));   \
  } while (0)

int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

// bounded wait of the grid barrier: ~2^22 polls of ~50 ns
constexpr unsigned int kSpinLimit = 1u << 22;
// blocks of the evaluation grid (grid-stride beyond)
constexpr int64_t kEvalBlocks = 8192;
// steps per persistent launch
constexpr uint64_t kStepsPerLaunch = 4096;
// updates per asynchronous worker and launch (bounds one kernel's duration)
constexpr uint64_t kAsyncPerLaunch = 1024;
// segment entries an async worker caches per lane
constexpr int kAsyncChunks = 4;
// segment entries per lane kept in registers by the persistent kernel
constexpr int kRegChunks = 2;
// dynamic LDS of the streamed scatter: the step's n coefficients (at most
// kStreamCoefBytes) and a block's accumulator
constexpr size_t kStreamLdsBytes = 160 * 1024 - 1024;
constexpr size_t kStreamCoefBytes = 96 * 1024;
}  // namespace


namespace {
acdc_status check_segments(int64_t nseg, int64_t nidx, int64_t nnz, const int32_t *ptr,
                           const int32_t *idx, const char *what) {
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

// omega = max number of entries per index value; the reference starts its
// running maximum at 1 (MultiCoreSolver.cpp:24, :32-38)
int32_t max_count(int64_t nidx, const std::vector<int32_t> &idx) {
  std::vector<int32_t> cnt(nidx, 0);
  int32_t om = 1;
  for (int32_t k : idx) om = std::max(om, ++cnt[k]);
  return om;
}

acdc_status make_problem(int kind, int64_t nseg, int64_t nidx, int64_t nnz,
                         const int32_t *ptr, const int32_t *idx, const double *vals,
                         const double *rhs, int64_t nrhs, acdc_problem **out,
                         const char *what) {
  acdc_status st = check_segments(nseg, nidx, nnz, ptr, idx, what);
  if (st != ACDC_OK) return st;
  acdc_problem *p = new (std::nothrow) acdc_problem();
  if (!p) return fail(ACDC_ERR_NOMEM, std::string(what) + ": out of host memory");
  try {
    p->kind = kind;
    p->ncoord = nseg;
    p->nres = nidx;
    p->nnz = nnz;
    p->cptr.assign(ptr, ptr + nseg + 1);
    p->cidx.assign(idx, idx + nnz);
    p->cval.assign(vals, vals + nnz);
    p->rhs.assign(rhs, rhs + nrhs);
    p->omega = max_count(nidx, p->cidx);
  } catch (const std::bad_alloc &) {
    delete p;
    return fail(ACDC_ERR_NOMEM, std::string(what) + ": out of host memory");
  }
  *out = p;
  return ACDC_OK;
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
                                               const int32_t *col_ptr, const int32_t *row_idx,
                                               const double *vals, const double *b,
                                               acdc_problem **out) {
  if (!out || !col_ptr || (nnz > 0 && (!row_idx || !vals)) || !b)
    return fail(ACDC_ERR_INVALID, "acdc_problem_create_csc: null argument");
  return make_problem(0, n, m, nnz, col_ptr, row_idx, vals, b, m, out,
                      "acdc_problem_create_csc");
}

extern "C" acdc_status acdc_problem_create_csr(int64_t n_samples, int64_t n_features,
                                               int64_t nnz, const int32_t *row_ptr,
                                               const int32_t *col_idx, const double *vals,
                                               const double *labels, acdc_problem **out) {
  if (!out || !row_ptr || (nnz > 0 && (!col_idx || !vals)) || !labels)
    return fail(ACDC_ERR_INVALID, "acdc_problem_create_csr: null argument");
  for (int64_t i = 0; i < n_samples; ++i)
    if (labels[i] != 1.0 && labels[i] != -1.0)
      return fail(ACDC_ERR_INVALID, "acdc_problem_create_csr: labels must be +1 or -1");
  return make_problem(1, n_samples, n_features, nnz, row_ptr, col_idx, vals, labels,
                      n_samples, out, "acdc_problem_create_csr");
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
struct acdc_solver {
  int loss = 0;
  acdc_settings set{};
  int64_t ncoord = 0, nres = 0, nnz = 0;
  double beta = 1.0;
  int lanes = 64;       // lanes per coordinate in the step kernels
  int eval_lanes = 64;  // lanes per coordinate in refresh/evaluation
  bool fp32 = false;
  bool persistent = false;
  bool async = false;  // ACDC_STEP_ASYNC
  bool staged = false;  // asynchronous workers fed through LDS (k_async_staged)
  // per-step launches whose scatter streams the transposed matrix
  // (k_step_scatter_t): its entries, the two coefficient halves, the grid
  bool streamed = false;
  void *tent = nullptr, *cvec = nullptr;
  int32_t *teb = nullptr, *tfb = nullptr;  // block ranges: entries, residual indices
  // SVM: FP64 transposed copy for the refresh's gather of w (k_seg_dot)
  int32_t *wptr = nullptr, *widx = nullptr;
  double *wval = nullptr, *wcoef = nullptr;
  int wlanes = 8;
  unsigned tgrid = 0;
  size_t tlds = 0;  // dynamic LDS of the streamed scatter
  // asynchronous mode: per-coordinate records and interleaved entries; the
  // records hold x while workers run, the plain x array between runs
  void *rec = nullptr, *ent = nullptr;
  bool rec_cur = false;  // records hold the current x (and anchor)
  bool x_cur = true;     // the plain x array is current
  bool stats_cur = false;  // FP32: stats of the last refresh describe the current x
  int64_t maxlen = 0;  // longest segment
  int64_t coresident_blocks = 0;
  hipStream_t stream = nullptr;
  int device = 0;  // the device current at create (the stream's)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int32_t *ptr = nullptr, *idx = nullptr;
  void *val = nullptr;
  double *val64 = nullptr;  // FP64 values (same buffer as val on the FP64 path)
  double *x = nullptr, *L = nullptr, *rhs = nullptr;
  void *res = nullptr;
  double *r64 = nullptr, *gref = nullptr;
  int32_t *coord = nullptr;
  double *coef = nullptr;
  double *partial = nullptr, *stats = nullptr;
  GridSync *gsync = nullptr;
  uint64_t step = 0;
  uint64_t updates = 0;
  uint64_t steps_since_refresh = 0;  // steps (async: updates per worker)
  double kernel_ms = 0, eval_ms = 0, step_ms = 0;
  uint64_t step_launches = 0;
  // event pairs around the update-kernel launches of one acdc_run_updates
  std::vector<hipEvent_t> step_ev;
  size_t step_ev_used = 0;
  // pairs bracketing whole acdc_run_updates calls, read (with the step
  // pairs) by settle_timing, so that a call returns once its work is queued
  std::vector<hipEvent_t> run_ev;
  size_t run_ev_used = 0;
  std::vector<void *> allocs;
};

namespace {
acdc_status dalloc(acdc_solver *s, void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(ACDC_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
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

StepScalars scalars(const acdc_solver *s) {
  StepScalars sc;
  sc.lambda = s->set.lambda;
  sc.lambda_n = s->set.lambda * (double)s->ncoord;
  sc.inv_lambda_n = 1.0 / sc.lambda_n;
  return sc;
}

template <typename T>
StepArgs<T> step_args(const acdc_solver *s) {
  StepArgs<T> a;
  a.A = Seg<T>{s->ptr, s->idx, static_cast<const T *>(s->val)};
  a.res = static_cast<T *>(s->res);
  a.gref = s->fp32 ? s->gref : nullptr;
  a.x = s->x;
  a.L = s->L;
  a.y = s->loss == kSvm ? s->rhs : nullptr;
  a.sc = scalars(s);
  a.perm = Perm::make((uint32_t)s->ncoord);
  a.seed = s->set.seed;
  a.tau = s->set.tau;
  a.coord = s->coord;
  a.coef = s->coef;
  a.cvec = s->streamed ? static_cast<T *>(s->cvec) : nullptr;
  a.ncoord = s->ncoord;
  return a;
}

// the FP64 view of A used by the refresh, evaluation and setup passes
Seg<double> seg64(const acdc_solver *s) { return Seg<double>{s->ptr, s->idx, s->val64}; }

int auto_lanes(int64_t nnz, int64_t nseg) {
  const double avg = nseg > 0 ? (double)nnz / (double)nseg : 0.0;
  if (avg <= 8) return 8;
  if (avg <= 16) return 16;
  if (avg <= 32) return 32;
  return 64;
}

// refresh, evaluation and setup passes (one pass over every segment): half
// waves up to 64 entries on average, two segments in flight per wave
// (measured on the 1M-column Lasso, 46.7 entries per column: 399 against
// 426 us for the gradient pass, 38 against 64 us for the scatter)
int auto_eval_lanes(int64_t nnz, int64_t nseg) {
  const double avg = nseg > 0 ? (double)nnz / (double)nseg : 0.0;
  return avg <= 64 ? std::min(32, auto_lanes(nnz, nseg)) : 64;
}

// synchronous steps: a whole block per coordinate once segments average more
// than 192 entries (the SVM's lognormal sample rows: a step waits for its
// longest row, and 256 lanes walk it four times faster than one wave)
int auto_step_lanes(int64_t nnz, int64_t nseg) {
  const double avg = nseg > 0 ? (double)nnz / (double)nseg : 0.0;
  return avg > 192 ? kBlock : auto_lanes(nnz, nseg);
}

// ---- dispatch: value type x lanes x loss (WIDE: also a block per coordinate,
// for the synchronous per-step and persistent kernels)
template <template <typename, int, int> class F, bool WIDE = false, typename... Args>
void dispatch3(const acdc_solver *s, int lanes, Args &&...args) {
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
    switch (lanes) {
      case 8: go(tag_t, std::integral_constant<int, 8>{}); break;
      case 16: go(tag_t, std::integral_constant<int, 16>{}); break;
      case 32: go(tag_t, std::integral_constant<int, 32>{}); break;
      case kBlock:
        if constexpr (WIDE) {
          go(tag_t, std::integral_constant<int, kBlock>{});
          break;
        }
        [[fallthrough]];
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
    const unsigned blocks = (unsigned)div_up(s->set.tau * G, kBlock);
    const StepArgs<T> a = step_args<T>(s);
    if (part & 1) k_step_gather<T, G, LOSS><<<blocks, kBlock, 0, s->stream>>>(a, step);
    if ((part & 2) && s->streamed) {
      TScatterArgs<T> t;
      t.ent = static_cast<const TEntry<T> *>(s->tent);
      t.eb = s->teb;
      t.fb = s->tfb;
      t.cvec = a.cvec + (int64_t)(step & 1) * s->ncoord;
      t.ncoord = (int32_t)s->ncoord;
      t.res = a.res;
      k_step_scatter_t<T><<<s->tgrid, kTBlock, s->tlds, s->stream>>>(t);
    } else if (part & 2) {
      k_step_scatter<T, G><<<blocks, kBlock, 0, s->stream>>>(a);
    }
  }
};

// async workers: count updates each, from per-worker update k0 on
template <typename T, int G, int LOSS>
struct LaunchAsync {
  static void run(const acdc_solver *s, uint64_t k0, uint64_t count) {
    AsyncArgs<T> a;
    a.rec = static_cast<AsyncRec<T> *>(s->rec);
    a.ent = static_cast<const Entry<T> *>(s->ent);
    a.nnz = s->nnz;
    a.res = static_cast<T *>(s->res);
    a.sc = scalars(s);
    a.seed = s->set.seed;
    a.tau = s->set.tau;
    a.n = (uint32_t)s->ncoord;
    if constexpr (G == 64) {
      if (s->staged) {
        const unsigned sb = (unsigned)div_up(s->set.tau, kStagedWorkers);
        if (s->maxlen <= kStagedCached<T>)
          k_async_staged<T, LOSS, false><<<sb, kStagedBlock, 0, s->stream>>>(a, k0, count);
        else
          k_async_staged<T, LOSS, true><<<sb, kStagedBlock, 0, s->stream>>>(a, k0, count);
        return;
      }
    }
    const unsigned blocks = (unsigned)div_up(s->set.tau * G, kAsyncBlock);
    if (s->maxlen <= kAsyncChunks * G)
      k_async<T, G, LOSS, kAsyncChunks, false><<<blocks, kAsyncBlock, 0, s->stream>>>(a, k0, count);
    else
      k_async<T, G, LOSS, kAsyncChunks, true><<<blocks, kAsyncBlock, 0, s->stream>>>(a, k0, count);
  }
};

// x / anchor exchange between the plain arrays and the async records
acdc_status ensure_rec(acdc_solver *s) {
  if (!s->async || s->rec_cur) return ACDC_OK;
  const unsigned nb = (unsigned)div_up(s->ncoord, kBlock);
  if (s->fp32)
    k_rec_put<float><<<nb, kBlock, 0, s->stream>>>(static_cast<AsyncRec<float> *>(s->rec),
                                                   s->ncoord, s->x, s->gref);
  else
    k_rec_put<double><<<nb, kBlock, 0, s->stream>>>(static_cast<AsyncRec<double> *>(s->rec),
                                                    s->ncoord, s->x, nullptr);
  HIP_TRY(hipGetLastError());
  s->rec_cur = true;
  return ACDC_OK;
}
acdc_status ensure_x(acdc_solver *s) {
  if (!s->async || s->x_cur) return ACDC_OK;
  const unsigned nb = (unsigned)div_up(s->ncoord, kBlock);
  if (s->fp32)
    k_rec_get_x<float><<<nb, kBlock, 0, s->stream>>>(static_cast<AsyncRec<float> *>(s->rec),
                                                     s->ncoord, s->x);
  else
    k_rec_get_x<double><<<nb, kBlock, 0, s->stream>>>(static_cast<AsyncRec<double> *>(s->rec),
                                                      s->ncoord, s->x);
  HIP_TRY(hipGetLastError());
  s->x_cur = true;
  return ACDC_OK;
}

template <typename T, int G, int LOSS>
struct LaunchSerial {
  static void run(const acdc_solver *s, uint64_t step0, uint64_t nsteps) {
    k_serial<T, LOSS><<<1, 64, 0, s->stream>>>(step_args<T>(s), step0, nsteps);
  }
};

template <typename T, int G, int LOSS>
struct PersistentKernel {
  static const void *fn() {
    return reinterpret_cast<const void *>(&k_sync_persistent<T, G, LOSS, kRegChunks>);
  }
  static void run(const acdc_solver *s, uint64_t step0, uint64_t nsteps, hipError_t *err) {
    StepArgs<T> a = step_args<T>(s);
    GridSync *gs = s->gsync;
    unsigned int spin = kSpinLimit;
    void *args[] = {&a, &step0, &nsteps, &gs, &spin};
    const unsigned blocks = (unsigned)div_up(s->set.tau * G, kPersistentBlock);
    *err = hipLaunchCooperativeKernel(fn(), dim3(blocks), dim3(kPersistentBlock), args, 0,
                                      s->stream);
  }
};

template <typename T, int G, int LOSS>
struct Occupancy {
  static void run(const acdc_solver *s, int *blocks_per_cu) {
    (void)s;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, PersistentKernel<T, G, LOSS>::fn(),
                                                     kPersistentBlock, 0) != hipSuccess)
      *blocks_per_cu = 0;
  }
};

template <typename T, int G, int LOSS>
struct LaunchLipschitz {
  static void run(const acdc_solver *s) {
    const unsigned blocks = (unsigned)div_up(s->ncoord * G, kBlock);
    k_lipschitz<double, G><<<blocks, kBlock, 0, s->stream>>>(seg64(s), s->ncoord, s->beta, s->L);
  }
};

// one power step of the spectral ESO: y = sum_c a_c w_c v_c scale (y zeroed
// before), v_c = w_c <a_c, y>, block partials of ||v||^2
template <typename T, int G, int LOSS>
struct LaunchPowerStep {
  static void run(const acdc_solver *s, double *v, const double *w, double scale, double *y,
                  double *partial, unsigned gather_blocks) {
    const unsigned blocks = (unsigned)div_up(s->ncoord * G, kBlock);
    k_pm_scatter<G><<<blocks, kBlock, 0, s->stream>>>(seg64(s), s->ncoord, v, w, scale, y);
    k_pm_gather<G><<<gather_blocks, kBlock, 0, s->stream>>>(seg64(s), s->ncoord, y, w, v, partial);
  }
};

template <typename T, int G, int LOSS>
struct LaunchRefreshScatter {
  static void run(const acdc_solver *s, double *target) {
    const unsigned blocks = (unsigned)div_up(s->ncoord * G, kBlock);
    k_refresh_scatter<double, G, LOSS><<<blocks, kBlock, 0, s->stream>>>(
        seg64(s), s->ncoord, s->x, s->rhs, scalars(s).inv_lambda_n, target);
  }
};

template <typename T, int G, int LOSS>
struct LaunchEvalCoord {
  static void run(const acdc_solver *s, double *gref) {
    const int64_t blocks = std::min(kEvalBlocks, div_up(s->ncoord * G, kBlock));
    k_eval_coord<double, G, LOSS><<<(unsigned)blocks, kBlock, 0, s->stream>>>(
        seg64(s), s->ncoord, s->r64, s->x, s->rhs, gref, s->partial);
    k_reduce_rows<<<1, kReduceBlock, 0, s->stream>>>(s->partial, blocks, EvalMasks<LOSS>::max,
                                               EvalMasks<LOSS>::min, s->stats);
  }
};

// r64 = A x - b (SVM: w) from the nonzero coordinates; FP32 path: then
// gref = A^T r64 with the evaluation statistics in stats[0..7], dr = 0.
acdc_status refresh(acdc_solver *s) {
  acdc_status st = ensure_x(s);
  if (st != ACDC_OK) return st;
  if (s->fp32) s->rec_cur = false;  // the anchor changes
  const unsigned rb = (unsigned)div_up(s->nres, kBlock);
  if (s->wptr) {  // SVM: w gathered per feature from the transposed copy
    k_svm_coef<<<(unsigned)div_up(s->ncoord, kBlock), kBlock, 0, s->stream>>>(
        s->ncoord, s->x, s->rhs, scalars(s).inv_lambda_n, s->wcoef);
    const Seg<double> wt{s->wptr, s->widx, s->wval};
    auto go = [&](auto tag) {
      constexpr int G = decltype(tag)::value;
      const int64_t blocks = std::min(kEvalBlocks, div_up(s->nres * G, kBlock));
      k_seg_dot<G><<<(unsigned)blocks, kBlock, 0, s->stream>>>(wt, s->nres, s->wcoef, s->r64);
    };
    switch (s->wlanes) {
      case 8: go(std::integral_constant<int, 8>{}); break;
      case 16: go(std::integral_constant<int, 16>{}); break;
      case 32: go(std::integral_constant<int, 32>{}); break;
      default: go(std::integral_constant<int, 64>{}); break;
    }
  } else {
    k_res_init<<<rb, kBlock, 0, s->stream>>>(s->nres, s->loss == kSvm ? nullptr : s->rhs, s->r64);
    dispatch3<LaunchRefreshScatter>(s, s->eval_lanes, s->r64);
  }
  if (s->fp32) {
    dispatch3<LaunchEvalCoord>(s, s->eval_lanes, s->gref);
    k_fill<float><<<rb, kBlock, 0, s->stream>>>(static_cast<float *>(s->res), s->nres, 0.0f);
  }
  HIP_TRY(hipGetLastError());
  s->steps_since_refresh = 0;
  s->stats_cur = s->fp32;
  return ACDC_OK;
}

acdc_status timed_begin(acdc_solver *s) {
  HIP_TRY(hipEventRecord(s->ev0, s->stream));
  return ACDC_OK;
}
// records the next event of an event pool on the solver's stream
acdc_status pool_mark(acdc_solver *s, std::vector<hipEvent_t> &pool, size_t &used) {
  if (used == pool.size()) {
    hipEvent_t e = nullptr;
    HIP_TRY(hipEventCreate(&e));
    pool.push_back(e);
  }
  HIP_TRY(hipEventRecord(pool[used++], s->stream));
  return ACDC_OK;
}
// pairs of step marks bracket the update-kernel launches of acdc_run_updates
acdc_status step_mark(acdc_solver *s) { return pool_mark(s, s->step_ev, s->step_ev_used); }
acdc_status run_mark(acdc_solver *s) { return pool_mark(s, s->run_ev, s->run_ev_used); }

// waits for the last pending run mark and adds the pending pairs to
// kernel_ms (whole calls) and step_ms (update kernels); every mark is on the
// solver's stream, so the last one completes after all the others
// (sums are committed, and the marks released, only when every pair reads
// back, so that a failed call can be retried without counting pairs twice)
acdc_status settle_timing(acdc_solver *s) {
  if (s->run_ev_used == 0) return ACDC_OK;
  HIP_TRY(hipEventSynchronize(s->run_ev[s->run_ev_used - 1]));
  double run_ms = 0, steps_ms = 0;
  for (size_t e = 0; e + 1 < s->run_ev_used; e += 2) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, s->run_ev[e], s->run_ev[e + 1]));
    run_ms += ms;
  }
  for (size_t e = 0; e + 1 < s->step_ev_used; e += 2) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, s->step_ev[e], s->step_ev[e + 1]));
    steps_ms += ms;
  }
  s->kernel_ms += run_ms;
  s->step_ms += steps_ms;
  s->run_ev_used = 0;
  s->step_ev_used = 0;
  return ACDC_OK;
}
// a failed acdc_run_updates drops the marks it made, so pairs stay aligned
struct PendingMarks {
  acdc_solver *s;
  size_t run0, step0;
  bool keep = false;
  ~PendingMarks() {
    if (!keep) {
      s->run_ev_used = run0;
      s->step_ev_used = step0;
    }
  }
};
constexpr size_t kMaxPendingMarks = 512;

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
  for (hipEvent_t e : s->step_ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : s->run_ev) (void)hipEventDestroy(e);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

// steps [step0, step0 + nsteps) in persistent launches of kStepsPerLaunch.
// A refused cooperative launch runs nothing; in ACDC_LAUNCH_AUTO the solver
// then switches to per-step launches (same arithmetic) for good.
// The spectral ESO's eigenvalue (mSDCA.cpp:188-265): the largest eigenvalue
// of D^-1/2 A^T A D^-1/2 by kPowerSteps power steps from a seeded start
// (coordinate c: the first word of Philox4x32-10 at counter (c, c >> 32,
// "PCPM", 0) under the solver's seed, over 2^32), normalised.  Set-up time
// only: the temporaries are freed before returning.
constexpr int kPowerSteps = 20;
constexpr uint32_t kPowerTag = 0x5043504Du;  // "PCPM"

acdc_status spectral_lmax(acdc_solver *s, const acdc_problem *p, double *lmax) {
  const int64_t n = s->ncoord, m = std::max<int64_t>(s->nres, 1);
  std::vector<double> w(std::max<int64_t>(n, 1)), v(std::max<int64_t>(n, 1));
  double nrm = 0;
  for (int64_t c = 0; c < n; ++c) {
    double s2 = 0;
    for (int32_t q = p->cptr[c]; q < p->cptr[c + 1]; ++q) s2 += p->cval[q] * p->cval[q];
    w[c] = s2 > 0 ? 1.0 / std::sqrt(s2) : 0.0;
    uint32_t ctr[4] = {(uint32_t)c, (uint32_t)((uint64_t)c >> 32), kPowerTag, 0u};
    acdc::Philox::gen(ctr, (uint32_t)s->set.seed, (uint32_t)(s->set.seed >> 32));
    v[c] = (double)ctr[0] * (1.0 / 4294967296.0);
    nrm += v[c] * v[c];
  }
  nrm = std::sqrt(nrm);
  for (int64_t c = 0; c < n; ++c) v[c] /= nrm;
  const unsigned gblocks =
      (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, div_up(n * s->eval_lanes, kBlock)));
  double *dv = nullptr, *dw = nullptr, *dy = nullptr, *dpart = nullptr;
  auto cleanup = [&]() {
    (void)hipFree(dv);
    (void)hipFree(dw);
    (void)hipFree(dy);
    (void)hipFree(dpart);
  };
  if (hipMalloc(&dv, v.size() * 8) != hipSuccess || hipMalloc(&dw, w.size() * 8) != hipSuccess ||
      hipMalloc(&dy, (size_t)m * 8) != hipSuccess || hipMalloc(&dpart, (size_t)gblocks * 8) != hipSuccess) {
    cleanup();
    return fail(ACDC_ERR_NOMEM, "acdc_solver_create: spectral beta buffers");
  }
  hipError_t e = hipMemcpy(dv, v.data(), v.size() * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dw, w.data(), w.size() * 8, hipMemcpyHostToDevice);
  std::vector<double> part(gblocks);
  double est = 1.0, scale = 1.0;
  for (int it = 0; it < kPowerSteps && e == hipSuccess; ++it) {
    e = hipMemsetAsync(dy, 0, (size_t)m * 8, s->stream);
    if (e != hipSuccess) break;
    dispatch3<LaunchPowerStep>(s, s->eval_lanes, dv, dw, scale, dy, dpart, gblocks);
    e = hipMemcpyAsync(part.data(), dpart, (size_t)gblocks * 8, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    double sq = 0;
    for (double q : part) sq += q;
    est = std::sqrt(sq);
    scale = est > 0 ? 1.0 / est : 0.0;
  }
  cleanup();
  if (e != hipSuccess)
    return fail(ACDC_ERR_HIP, std::string("acdc_solver_create: spectral beta: ") + hipGetErrorString(e));
  *lmax = est;
  return ACDC_OK;
}

acdc_status run_persistent(acdc_solver *s, uint64_t step0, uint64_t nsteps) {
  for (uint64_t done = 0; done < nsteps;) {
    const uint64_t chunk = std::min(kStepsPerLaunch, nsteps - done);
    HIP_TRY(hipMemsetAsync(s->gsync, 0, sizeof(GridSync), s->stream));
    hipError_t err = hipSuccess;
    dispatch3<PersistentKernel, true>(s, s->lanes, step0 + done, chunk, &err);
    if (err != hipSuccess) {
      if (s->set.launch_mode != ACDC_LAUNCH_AUTO)
        return fail(ACDC_ERR_HIP, std::string("cooperative launch: ") + hipGetErrorString(err));
      (void)hipGetLastError();
      s->persistent = false;
      for (uint64_t k = done; k < nsteps; ++k) dispatch3<LaunchStep, true>(s, s->lanes, step0 + k, 3);
      return ACDC_OK;
    }
    done += chunk;
  }
  return ACDC_OK;
}
}  // namespace

extern "C" acdc_status acdc_solver_create(const acdc_problem *p, int32_t loss,
                                          const acdc_settings *set, acdc_solver **out) {
  if (!p || !set || !out) return fail(ACDC_ERR_INVALID, "acdc_solver_create: null argument");
  if (loss != kLasso && loss != kNnls && loss != kSvm)
    return fail(ACDC_ERR_INVALID,
                "acdc_solver_create: loss must be 0 (L1-square), 2 (SVM dual hinge) or 10 (NNLS)");
  if ((loss == kSvm) != (p->kind == 1))
    return fail(ACDC_ERR_STATE,
                "acdc_solver_create: SVM needs a sample-major (CSR) problem, losses 0/10 a "
                "column-major (CSC) one");
  if (set->tau < 1 || set->tau > p->ncoord)
    return fail(ACDC_ERR_INVALID, "acdc_solver_create: tau must lie in [1, n]");
  if (set->precision != ACDC_PREC_FP64 && set->precision != ACDC_PREC_FP32_REFRESH)
    return fail(ACDC_ERR_INVALID, "acdc_solver_create: unknown precision");
  if (set->step_mode != ACDC_STEP_SYNC && set->step_mode != ACDC_STEP_ASYNC)
    return fail(ACDC_ERR_INVALID, "acdc_solver_create: unknown step mode");
  if (set->launch_mode < ACDC_LAUNCH_AUTO || set->launch_mode > ACDC_LAUNCH_ASYNC_STAGED)
    return fail(ACDC_ERR_INVALID, "acdc_solver_create: unknown launch mode");
  if (set->launch_mode >= ACDC_LAUNCH_ASYNC_DIRECT && set->step_mode != ACDC_STEP_ASYNC)
    return fail(ACDC_ERR_INVALID,
                "acdc_solver_create: launch modes 4 and 5 are for asynchronous workers");
  if (set->launch_mode == ACDC_LAUNCH_ASYNC_STAGED && set->lanes_per_coord != 0 &&
      set->lanes_per_coord != 64)
    return fail(ACDC_ERR_INVALID,
                "acdc_solver_create: staged asynchronous workers take 64 lanes per coordinate");
  const size_t coef_bytes = (size_t)p->ncoord * (set->precision == ACDC_PREC_FP32_REFRESH ? 4 : 8);
  const bool stream_fits = coef_bytes <= kStreamCoefBytes && p->ncoord <= 65535;
  if (set->launch_mode == ACDC_LAUNCH_PER_STEP_STREAMED &&
      (set->step_mode != ACDC_STEP_SYNC || set->tau < 2 || !stream_fits))
    return fail(ACDC_ERR_INVALID,
                "acdc_solver_create: the streamed scatter needs synchronous steps with tau >= 2 "
                "and n <= 65535 coefficients within " + std::to_string(kStreamCoefBytes) +
                    " bytes of LDS");
  if (set->lanes_per_coord != 0 && set->lanes_per_coord != 8 && set->lanes_per_coord != 16 &&
      set->lanes_per_coord != 32 && set->lanes_per_coord != 64 && set->lanes_per_coord != kBlock)
    return fail(ACDC_ERR_INVALID,
                "acdc_solver_create: lanes_per_coord must be 0, 8, 16, 32, 64 or 256");
  if (set->lanes_per_coord == kBlock && set->step_mode == ACDC_STEP_ASYNC)
    return fail(ACDC_ERR_INVALID,
                "acdc_solver_create: 256 lanes per coordinate (a block per coordinate) is for "
                "synchronous steps");
  if (!(set->lambda >= 0) || (loss == kSvm && !(set->lambda > 0)))
    return fail(ACDC_ERR_INVALID, "acdc_solver_create: lambda must be >= 0 (> 0 for the SVM)");
  if (set->refresh_period < 0)
    return fail(ACDC_ERR_INVALID, "acdc_solver_create: refresh_period must be >= 0");
  double beta = 1.0;
  switch (set->beta_mode) {
    case ACDC_BETA_OMEGA:
      beta = p->ncoord > 1 ? 1.0 + (double)(p->omega - 1) * (double)(set->tau - 1) /
                                       ((double)p->ncoord - 1.0)
                           : 1.0;
      break;
    case ACDC_BETA_FORCED:
      if (!(set->forced_beta > 0))
        return fail(ACDC_ERR_INVALID, "acdc_solver_create: forced_beta must be > 0");
      beta = set->forced_beta;
      break;
    case ACDC_BETA_NONE: beta = 1.0; break;
    case ACDC_BETA_SPECTRAL: beta = 1.0; break;  // after the upload: spectral_lmax
    default: return fail(ACDC_ERR_INVALID, "acdc_solver_create: unknown beta_mode");
  }
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
  s->beta = beta;
  s->fp32 = set->precision == ACDC_PREC_FP32_REFRESH;
  s->staged = set->step_mode == ACDC_STEP_ASYNC && set->launch_mode == ACDC_LAUNCH_ASYNC_STAGED;
  s->lanes = s->staged ? 64
             : set->lanes_per_coord ? set->lanes_per_coord
                                   : (set->step_mode == ACDC_STEP_SYNC &&
                                              set->launch_mode != ACDC_LAUNCH_PERSISTENT
                                          ? auto_step_lanes(p->nnz, p->ncoord)
                                          : auto_lanes(p->nnz, p->ncoord));
  s->eval_lanes = auto_eval_lanes(p->nnz, p->ncoord);

  auto bail = [&](acdc_status e) {
    std::string msg = g_err;
    destroy_solver(s);
    g_err = msg;
    return e;
  };
  acdc_status st = ACDC_OK;
  if (hipGetDevice(&s->device) != hipSuccess)
    return bail(fail(ACDC_ERR_HIP, "acdc_solver_create: no current device"));
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
    return bail(fail(ACDC_ERR_HIP, "acdc_solver_create: stream/event creation failed"));
  const size_t tsz = s->fp32 ? 4 : 8;
  if ((st = dalloc_t(s, &s->ptr, p->cptr.size())) != ACDC_OK) return bail(st);
  if ((st = dalloc_t(s, &s->idx, p->cidx.size())) != ACDC_OK) return bail(st);
  if ((st = dalloc(s, &s->val, p->cval.size() * tsz)) != ACDC_OK) return bail(st);
  if (s->fp32) {
    if ((st = dalloc_t(s, &s->val64, p->cval.size())) != ACDC_OK) return bail(st);
  } else {
    s->val64 = static_cast<double *>(s->val);
  }
  if ((st = dalloc_t(s, &s->x, s->ncoord)) != ACDC_OK) return bail(st);
  if ((st = dalloc_t(s, &s->L, s->ncoord)) != ACDC_OK) return bail(st);
  if ((st = dalloc_t(s, &s->rhs, p->rhs.size())) != ACDC_OK) return bail(st);
  if ((st = dalloc(s, &s->res, s->nres * tsz)) != ACDC_OK) return bail(st);
  if (s->fp32) {
    if ((st = dalloc_t(s, &s->r64, s->nres)) != ACDC_OK) return bail(st);
    if ((st = dalloc_t(s, &s->gref, s->ncoord)) != ACDC_OK) return bail(st);
  } else {
    s->r64 = static_cast<double *>(s->res);
  }
  if ((st = dalloc_t(s, &s->coord, set->tau)) != ACDC_OK) return bail(st);
  if ((st = dalloc_t(s, &s->coef, set->tau)) != ACDC_OK) return bail(st);
  if ((st = dalloc_t(s, &s->partial, (size_t)kEvalBlocks * kStats)) != ACDC_OK) return bail(st);
  if ((st = dalloc_t(s, &s->stats, 2 * kStats)) != ACDC_OK) return bail(st);
  if ((st = dalloc_t(s, &s->gsync, 1)) != ACDC_OK) return bail(st);
  if (set->step_mode == ACDC_STEP_ASYNC) {
    const size_t rsz = s->fp32 ? sizeof(AsyncRec<float>) : sizeof(AsyncRec<double>);
    const size_t esz = s->fp32 ? sizeof(Entry<float>) : sizeof(Entry<double>);
    if ((st = dalloc(s, &s->rec, (size_t)s->ncoord * rsz)) != ACDC_OK) return bail(st);
    // two entries of zero padding: the staged loader copies whole 16-byte
    // pairs and may touch the pair after the last segment
    if ((st = dalloc(s, &s->ent, ((size_t)p->nnz + 2) * esz)) != ACDC_OK) return bail(st);
    if (hipMemset(static_cast<char *>(s->ent) + (size_t)p->nnz * esz, 0, 2 * esz) != hipSuccess)
      return bail(fail(ACDC_ERR_HIP, "acdc_solver_create: entry padding"));
  }

  hipError_t e = hipMemcpy(s->ptr, p->cptr.data(), p->cptr.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && p->nnz > 0)
    e = hipMemcpy(s->idx, p->cidx.data(), p->cidx.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && p->nnz > 0) {
    if (s->fp32) {
      std::vector<float> f(p->cval.begin(), p->cval.end());
      e = hipMemcpy(s->val, f.data(), f.size() * 4, hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = hipMemcpy(s->val64, p->cval.data(), p->cval.size() * 8, hipMemcpyHostToDevice);
    } else {
      e = hipMemcpy(s->val, p->cval.data(), p->cval.size() * 8, hipMemcpyHostToDevice);
    }
  }
  if (e == hipSuccess) e = hipMemcpy(s->rhs, p->rhs.data(), p->rhs.size() * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(s->x, 0, s->ncoord * 8);
  if (e != hipSuccess)
    return bail(fail(ACDC_ERR_HIP, std::string("acdc_solver_create: upload: ") + hipGetErrorString(e)));

  // persistent grid: every slot's group must be resident at once
  s->async = set->step_mode == ACDC_STEP_ASYNC;
  for (size_t c = 0; c + 1 < p->cptr.size(); ++c)
    s->maxlen = std::max<int64_t>(s->maxlen, (int64_t)p->cptr[c + 1] - p->cptr[c]);
  // (a block per coordinate takes the persistent grid only when asked: on
  // the SVM's sample rows a step waits for its longest row either way, and
  // per-step launches measured 5% faster)
  if (!s->async && set->tau >= 2 &&
      ((set->launch_mode == ACDC_LAUNCH_AUTO && s->lanes != kBlock) ||
       set->launch_mode == ACDC_LAUNCH_PERSISTENT)) {
    int per_cu = 0, dev = 0;
    hipDeviceProp_t prop;
    (void)hipGetDevice(&dev);
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
      dispatch3<Occupancy, true>(s, s->lanes, &per_cu);
      // at most one 1024-thread block per CU: a 512-block grid at two blocks
      // per CU (which the occupancy query allows) did not make every block
      // arrive at the grid barrier within its bound on the MI355X
      s->coresident_blocks = (int64_t)std::min(per_cu, 1) * prop.multiProcessorCount;
      // with automatic lanes, narrow the groups until tau slots fit the grid
      // (a block per coordinate keeps its width: per-step launches then)
      if (!set->lanes_per_coord && s->lanes <= 64)
        while (s->lanes > 8 && div_up(set->tau * s->lanes, kPersistentBlock) > s->coresident_blocks)
          s->lanes /= 2;
    }
    const int64_t need = div_up(set->tau * s->lanes, kPersistentBlock);
    s->persistent = need <= s->coresident_blocks;
    if (!s->persistent && set->launch_mode == ACDC_LAUNCH_PERSISTENT)
      return bail(fail(ACDC_ERR_INVALID,
                       "acdc_solver_create: tau * lanes_per_coord exceeds the co-resident grid "
                       "of the persistent kernel"));
  }
  // FP32 SVM: the FP64 transposed copy of the periodic refresh (w by
  // feature, no atomics); every alpha_j is usually nonzero, so the scatter
  // form would add all entries with FP64 atomics.  (The FP64 path refreshes
  // only at creation and acdc_set_x, with the scatter form.)
  if (loss == kSvm && s->fp32) {
    std::vector<int32_t> wp((size_t)s->nres + 1, 0);
    for (int32_t i : p->cidx) ++wp[(size_t)i + 1];
    for (int64_t i = 0; i < s->nres; ++i) wp[i + 1] += wp[i];
    std::vector<int32_t> wi((size_t)p->nnz);
    std::vector<double> wv((size_t)p->nnz);
    {
      std::vector<int32_t> at(wp.begin(), wp.end() - 1);
      for (int64_t j = 0; j < s->ncoord; ++j)
        for (int32_t q = p->cptr[j]; q < p->cptr[j + 1]; ++q) {
          const int32_t k = at[p->cidx[q]]++;
          wi[k] = (int32_t)j;
          wv[k] = p->cval[q];
        }
    }
    if ((st = dalloc_t(s, &s->wptr, wp.size())) != ACDC_OK) return bail(st);
    if ((st = dalloc_t(s, &s->widx, wi.size())) != ACDC_OK) return bail(st);
    if ((st = dalloc_t(s, &s->wval, wv.size())) != ACDC_OK) return bail(st);
    if ((st = dalloc_t(s, &s->wcoef, (size_t)s->ncoord)) != ACDC_OK) return bail(st);
    hipError_t we = hipMemcpy(s->wptr, wp.data(), wp.size() * 4, hipMemcpyHostToDevice);
    if (we == hipSuccess && p->nnz)
      we = hipMemcpy(s->widx, wi.data(), wi.size() * 4, hipMemcpyHostToDevice);
    if (we == hipSuccess && p->nnz)
      we = hipMemcpy(s->wval, wv.data(), wv.size() * 8, hipMemcpyHostToDevice);
    if (we != hipSuccess)
      return bail(fail(ACDC_ERR_HIP, std::string("acdc_solver_create: transposed copy: ") +
                                         hipGetErrorString(we)));
    s->wlanes = auto_eval_lanes(p->nnz, s->nres);
  }
  // streamed scatter: forced, or automatic when a step holds at least a
  // sixteenth of the coordinates (measured crossover on the 20,000-sample
  // SVM: tau ~1,100; one coalesced pass over every entry against one atomic
  // per sampled entry)
  s->streamed = !s->async && !s->persistent && set->tau >= 2 && stream_fits &&
                (set->launch_mode == ACDC_LAUNCH_PER_STEP_STREAMED ||
                 (set->launch_mode == ACDC_LAUNCH_AUTO && 16 * set->tau >= s->ncoord));
  if (s->streamed) {
    int dev = 0, cus = 1;
    hipError_t te = hipGetDevice(&dev);
    if (te == hipSuccess) te = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // block ranges at residual-index boundaries: about nnz / CUs entries
    // each, at most `cap` indices (the accumulator's LDS, and 16-bit local
    // indices); one index with more entries is a block of its own
    const size_t tsz_ = s->fp32 ? 4 : 8;
    const int64_t cap = std::min<int64_t>(65536, (int64_t)((kStreamLdsBytes - coef_bytes) / tsz_));
    const int64_t target = std::max<int64_t>(1024, div_up(p->nnz, std::max(1, cus)));
    std::vector<int64_t> at(s->nres + 1, 0);  // entries per index, then starts
    for (int32_t i : p->cidx) ++at[i + 1];
    std::vector<int32_t> eb{0}, fb{0};
    int64_t span = 0, ents = 0, maxspan = 1;
    for (int64_t i = 0; i < s->nres; ++i) {
      const int64_t c = at[i + 1];
      if (span > 0 && (ents + c > target || span + 1 > cap)) {
        eb.push_back(eb.back() + (int32_t)ents);
        fb.push_back((int32_t)i);
        maxspan = std::max(maxspan, span);
        span = ents = 0;
      }
      ++span;
      ents += c;
    }
    eb.push_back(eb.back() + (int32_t)ents);
    fb.push_back((int32_t)s->nres);
    maxspan = std::max(maxspan, span);
    for (int64_t i = 0; i < s->nres; ++i) at[i + 1] += at[i];
    const size_t nb = eb.size() - 1;
    const size_t tesz = s->fp32 ? sizeof(TEntry<float>) : sizeof(TEntry<double>);
    if ((st = dalloc(s, &s->tent, (size_t)p->nnz * tesz)) != ACDC_OK) return bail(st);
    if ((st = dalloc(s, &s->cvec, 2 * coef_bytes)) != ACDC_OK) return bail(st);
    if ((st = dalloc_t(s, &s->teb, nb + 1)) != ACDC_OK) return bail(st);
    if ((st = dalloc_t(s, &s->tfb, nb + 1)) != ACDC_OK) return bail(st);
    // entries by residual index (coordinates ascending within one index),
    // the index stored relative to its block's first
    std::vector<int32_t> blk_of((size_t)s->nres);
    for (size_t b = 0; b < nb; ++b)
      for (int32_t i = fb[b]; i < fb[b + 1]; ++i) blk_of[i] = (int32_t)b;
    auto fill = [&](auto tag) {
      using T = decltype(tag);
      std::vector<TEntry<T>> tv((size_t)p->nnz);
      for (int64_t j = 0; j < s->ncoord; ++j)
        for (int32_t q = p->cptr[j]; q < p->cptr[j + 1]; ++q) {
          const int32_t i = p->cidx[q];
          tv[at[i]++] = TEntry<T>{(uint16_t)(i - fb[blk_of[i]]), (uint16_t)j, (T)p->cval[q]};
        }
      return hipMemcpy(s->tent, tv.data(), tv.size() * sizeof(TEntry<T>), hipMemcpyHostToDevice);
    };
    if (te == hipSuccess && p->nnz) te = s->fp32 ? fill(0.0f) : fill(0.0);
    if (te == hipSuccess) te = hipMemcpy(s->teb, eb.data(), eb.size() * 4, hipMemcpyHostToDevice);
    if (te == hipSuccess) te = hipMemcpy(s->tfb, fb.data(), fb.size() * 4, hipMemcpyHostToDevice);
    // both halves start at zero; the first step clears "its previous
    // coordinates" at slot coordinates 0
    if (te == hipSuccess) te = hipMemset(s->cvec, 0, 2 * coef_bytes);
    if (te == hipSuccess) te = hipMemset(s->coord, 0, (size_t)set->tau * 4);
    s->tlds = coef_bytes + (size_t)maxspan * tsz_;
    if (te == hipSuccess)
      te = hipFuncSetAttribute(s->fp32 ? reinterpret_cast<const void *>(&k_step_scatter_t<float>)
                                       : reinterpret_cast<const void *>(&k_step_scatter_t<double>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)s->tlds);
    if (te != hipSuccess)
      return bail(fail(ACDC_ERR_HIP, std::string("acdc_solver_create: streamed scatter setup: ") +
                                         hipGetErrorString(te)));
    s->tgrid = (unsigned)nb;
  }
  if (set->beta_mode == ACDC_BETA_SPECTRAL) {
    double lmax = 1.0;
    if ((st = spectral_lmax(s, p, &lmax)) != ACDC_OK) return bail(st);
    s->beta = s->ncoord > 1 ? 1.0 + (double)(set->tau - 1) * (lmax - 1.0) /
                                        ((double)s->ncoord - 1.0)
                            : 1.0;
  }
  dispatch3<LaunchLipschitz>(s, s->eval_lanes);
  if (s->async) {
    const unsigned nb = (unsigned)div_up(s->ncoord, kBlock);
    const unsigned eb = (unsigned)std::max<int64_t>(1, div_up(s->nnz, kBlock));
    const double *y = loss == kSvm ? s->rhs : nullptr;
    if (s->fp32) {
      k_rec_init<float><<<nb, kBlock, 0, s->stream>>>(static_cast<AsyncRec<float> *>(s->rec),
                                                      s->ncoord, s->ptr, s->L, y);
      k_pack_entries<float><<<eb, kBlock, 0, s->stream>>>(
          s->idx, static_cast<const float *>(s->val), s->nnz, static_cast<Entry<float> *>(s->ent));
    } else {
      k_rec_init<double><<<nb, kBlock, 0, s->stream>>>(static_cast<AsyncRec<double> *>(s->rec),
                                                       s->ncoord, s->ptr, s->L, y);
      k_pack_entries<double><<<eb, kBlock, 0, s->stream>>>(
          s->idx, static_cast<const double *>(s->val), s->nnz,
          static_cast<Entry<double> *>(s->ent));
    }
    s->rec_cur = false;
  }
  if ((st = refresh(s)) != ACDC_OK) return bail(st);
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
  // refresh every refresh_period epochs of ceil(n / tau) steps each, so that
  // runs of whole epochs end exactly on a refresh
  const uint64_t period = s->fp32 && s->set.refresh_period > 0
                              ? (uint64_t)s->set.refresh_period * (((uint64_t)s->ncoord + tau - 1) / tau)
                              : 0;
  acdc_status st;
  PendingMarks pending{s, s->run_ev_used, s->step_ev_used};
  if ((st = run_mark(s)) != ACDC_OK) return st;
  uint64_t done = 0;
  while (done < nsteps) {
    uint64_t chunk = nsteps - done;
    if (period) {
      const uint64_t left = period > s->steps_since_refresh ? period - s->steps_since_refresh : 1;
      chunk = std::min(chunk, left);
    }
    if ((st = ensure_rec(s)) != ACDC_OK) return st;
    if ((st = step_mark(s)) != ACDC_OK) return st;
    if (s->async) {
      s->x_cur = false;
      // equal launches of at most kAsyncPerLaunch updates per worker
      const uint64_t nl = (chunk + kAsyncPerLaunch - 1) / kAsyncPerLaunch;
      const uint64_t per = (chunk + nl - 1) / nl;
      for (uint64_t k = 0; k < chunk; k += per)
        dispatch3<LaunchAsync>(s, s->lanes, s->step + k, std::min(per, chunk - k));
      s->step_launches += nl;
    } else if (tau == 1) {
      dispatch3<LaunchSerial>(s, s->lanes, s->step, chunk);
      s->step_launches += 1;
    } else if (s->persistent) {
      if ((st = run_persistent(s, s->step, chunk)) != ACDC_OK) return st;
      s->step_launches += s->persistent ? (chunk + kStepsPerLaunch - 1) / kStepsPerLaunch : 2 * chunk;
    } else {
      for (uint64_t k = 0; k < chunk; ++k) dispatch3<LaunchStep, true>(s, s->lanes, s->step + k, 3);
      s->step_launches += 2 * chunk;
    }
    if ((st = step_mark(s)) != ACDC_OK) return st;
    s->stats_cur = false;
    s->step += chunk;
    s->updates += chunk * tau;
    s->steps_since_refresh += chunk;
    done += chunk;
    if (period && s->steps_since_refresh >= period)
      if ((st = refresh(s)) != ACDC_OK) return st;
  }
  HIP_TRY(hipGetLastError());
  if ((st = run_mark(s)) != ACDC_OK) return st;
  pending.keep = true;
  // the call returns once its work is queued (the next call's launches
  // follow on the stream without a gap); the persistent kernel's barrier
  // status is read after it has finished
  if (s->persistent || s->run_ev_used >= kMaxPendingMarks || s->step_ev_used >= kMaxPendingMarks)
    if ((st = settle_timing(s)) != ACDC_OK) return st;
  if (s->persistent) {
    GridSync g{};
    HIP_TRY(hipMemcpy(&g, s->gsync, sizeof g, hipMemcpyDeviceToHost));
    if (g.error.v) return fail(ACDC_ERR_HIP, "acdc_run_updates: grid barrier wait exceeded its bound");
  }
  return ACDC_OK;
}

extern "C" acdc_status acdc_recompute_residuals(acdc_solver *s) {
  if (!s) return fail(ACDC_ERR_INVALID, "acdc_recompute_residuals: null solver");
  acdc_status st;
  if ((st = timed_begin(s)) != ACDC_OK) return st;
  if ((st = refresh(s)) != ACDC_OK) return st;
  return timed_end(s, &s->eval_ms);
}

extern "C" acdc_status acdc_report_get(acdc_solver *s, acdc_report *out) {
  if (!s || !out) return fail(ACDC_ERR_INVALID, "acdc_report_get: null argument");
  acdc_status st;
  if ((st = settle_timing(s)) != ACDC_OK) return st;
  if ((st = timed_begin(s)) != ACDC_OK) return st;
  if (s->fp32) {
    // stats come with the anchor; a refresh that no update has followed yet
    // (the periodic one at the end of a run) already holds them
    if (!s->stats_cur && (st = refresh(s)) != ACDC_OK) return st;
  } else {
    if ((st = ensure_x(s)) != ACDC_OK) return st;
    dispatch3<LaunchEvalCoord>(s, s->eval_lanes, nullptr);
  }
  const int64_t rblocks = std::min(kEvalBlocks, div_up(s->nres, kBlock));
  k_eval_res<<<(unsigned)rblocks, kBlock, 0, s->stream>>>(
      s->nres, s->r64, s->loss == kSvm ? nullptr : s->rhs, s->partial);
  k_reduce_rows<<<1, kReduceBlock, 0, s->stream>>>(s->partial, rblocks, 0u, 0u, s->stats + kStats);
  HIP_TRY(hipGetLastError());
  double h[2 * kStats];
  HIP_TRY(hipMemcpyAsync(h, s->stats, sizeof h, hipMemcpyDeviceToHost, s->stream));
  if ((st = timed_end(s, &s->eval_ms)) != ACDC_OK) return st;
  const double *pc = h, *pr = h + kStats;
  acdc_report r{};
  const double R2 = pr[0], bTr = pr[1];
  const double lam = s->set.lambda;
  if (s->loss == kLasso) {
    const double ginf = pc[0], x1 = pc[1];
    r.primal = 0.5 * R2 + lam * x1;
    const double sc = ginf > lam ? lam / ginf : 1.0;
    r.dual = -sc * bTr - 0.5 * sc * sc * R2;
    r.grad_inf = ginf;
    r.x_l1 = x1;
    r.res_sq = R2;
    r.nnz_x = pc[2];
  } else if (s->loss == kNnls) {
    r.primal = 0.5 * R2;
    r.dual = r.primal - pc[1];
    r.kkt = pc[0];
    r.grad_inf = pc[3];
    r.x_l1 = pc[4];
    r.res_sq = R2;
    r.nnz_x = pc[5];
  } else {
    const double n = (double)s->ncoord;
    r.primal = lam * 0.5 * R2 + pc[0] / n;
    r.dual = -lam * 0.5 * R2 + pc[2] / n;
    r.accuracy = pc[1] / n;
    r.x_l1 = pc[2];
    r.res_sq = R2;
    r.nnz_x = pc[3];
  }
  r.gap = r.primal - r.dual;
  r.rel_gap = r.primal != 0 ? r.gap / std::fabs(r.primal) : 0.0;
  r.updates = s->updates;
  r.epochs = (double)s->updates / (double)s->ncoord;
  r.kernel_ms = s->kernel_ms;
  r.eval_ms = s->eval_ms;
  r.step_ms = s->step_ms;
  r.step_launches = s->step_launches;
  *out = r;
  return ACDC_OK;
}

extern "C" acdc_status acdc_get_x(acdc_solver *s, double *x, int64_t n) {
  if (!s || !x) return fail(ACDC_ERR_INVALID, "acdc_get_x: null argument");
  if (n != s->ncoord) return fail(ACDC_ERR_INVALID, "acdc_get_x: length must equal n");
  acdc_status st = ensure_x(s);
  if (st != ACDC_OK) return st;
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(x, s->x, n * 8, hipMemcpyDeviceToHost));
  return ACDC_OK;
}

extern "C" acdc_status acdc_set_x(acdc_solver *s, const double *x, int64_t n) {
  if (!s || !x) return fail(ACDC_ERR_INVALID, "acdc_set_x: null argument");
  if (n != s->ncoord) return fail(ACDC_ERR_INVALID, "acdc_set_x: length must equal n");
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(s->x, x, n * 8, hipMemcpyHostToDevice));
  s->x_cur = true;
  s->rec_cur = false;
  acdc_status st = refresh(s);
  if (st != ACDC_OK) return st;
  HIP_TRY(hipStreamSynchronize(s->stream));
  return ACDC_OK;
}

extern "C" acdc_status acdc_get_residuals(acdc_solver *s, double *r, int64_t len) {
  if (!s || !r) return fail(ACDC_ERR_INVALID, "acdc_get_residuals: null argument");
  if (len != s->nres) return fail(ACDC_ERR_INVALID, "acdc_get_residuals: wrong length");
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (s->fp32) {  // r = r64 + dr
    std::vector<float> f((size_t)len);
    HIP_TRY(hipMemcpy(f.data(), s->res, len * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(r, s->r64, len * 8, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < len; ++k) r[k] += f[k];
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
    // w = (1/(lambda n)) sum alpha_i y_i x_i: w, its anchor and the anchored
    // margins all scale by lambda_old / lambda_new
    const double f = s->set.lambda / lambda;
    const unsigned rb = (unsigned)div_up(s->nres, kBlock);
    if (s->fp32) {
      // asynchronous workers keep the live alpha in the records: bring it to
      // the plain array first, since the records are rebuilt from that array
      // (with the rescaled anchor) before the next updates
      acdc_status st = ensure_x(s);
      if (st != ACDC_OK) return st;
      k_scale<float><<<rb, kBlock, 0, s->stream>>>(static_cast<float *>(s->res), s->nres, f);
      k_scale<double><<<rb, kBlock, 0, s->stream>>>(s->r64, s->nres, f);
      k_scale<double><<<(unsigned)div_up(s->ncoord, kBlock), kBlock, 0, s->stream>>>(
          s->gref, s->ncoord, f);
      s->rec_cur = false;
    } else {
      k_scale<double><<<rb, kBlock, 0, s->stream>>>(s->r64, s->nres, f);
    }
    HIP_TRY(hipGetLastError());
  }
  s->set.lambda = lambda;
  s->stats_cur = false;
  return ACDC_OK;
}

extern "C" acdc_status acdc_get_beta(const acdc_solver *s, double *beta) {
  if (!s || !beta) return fail(ACDC_ERR_INVALID, "acdc_get_beta: null argument");
  *beta = s->beta;
  return ACDC_OK;
}

extern "C" acdc_status acdc_sample_coordinates(acdc_solver *s, uint64_t step0, uint64_t nsteps,
                                               int32_t *out) {
  if (!s || !out) return fail(ACDC_ERR_INVALID, "acdc_sample_coordinates: null argument");
  const int64_t total = (int64_t)nsteps * s->set.tau;
  if (total <= 0) return ACDC_OK;
  if (total > ((int64_t)1 << 30))
    return fail(ACDC_ERR_INVALID, "acdc_sample_coordinates: at most 2^30 samples per call");
  int32_t *d = nullptr;
  HIP_TRY(hipMalloc(&d, (size_t)total * 4));
  if (s->async)
    k_sample_async<<<(unsigned)div_up(total, kBlock), kBlock, 0, s->stream>>>(
        (uint32_t)s->ncoord, s->set.seed, step0, nsteps, s->set.tau, d);
  else
    k_sample<<<(unsigned)div_up(total, kBlock), kBlock, 0, s->stream>>>(
        Perm::make((uint32_t)s->ncoord), s->set.seed, step0, nsteps, s->set.tau, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, (size_t)total * 4, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  if (e != hipSuccess)
    return fail(ACDC_ERR_HIP, std::string("acdc_sample_coordinates: ") + hipGetErrorString(e));
  return ACDC_OK;
}

// ---- allocation-free sampling: the staging array is a file-scope device
// array (one copy per device), whose address on each device is looked up
// once with hipGetSymbolAddress while that device is current; host code uses
// only that address, never the array's name.
static __device__ __align__(256) int32_t sample_scratch[ACDC_NOALLOC_MAX_SAMPLES];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t *g_scratch[kMaxDevices] = {};  // per-device address of sample_scratch

// the address of sample_scratch on `device` (made current for the lookup)
acdc_status scratch_on(int device, int32_t **out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(ACDC_ERR_INVALID, "acdc_sample_coordinates_noalloc: device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    int prev = 0;
    HIP_TRY(hipGetDevice(&prev));
    HIP_TRY(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(sample_scratch));
    (void)hipSetDevice(prev);
    if (e != hipSuccess || !p)
      return fail(ACDC_ERR_HIP, std::string("acdc_sample_coordinates_noalloc: hipGetSymbolAddress: ") +
                                    hipGetErrorString(e));
    g_scratch[device] = static_cast<int32_t *>(p);
  }
  *out = g_scratch[device];
  return ACDC_OK;
}
}  // namespace

extern "C" acdc_status acdc_sample_coordinates_noalloc(acdc_solver *s, uint64_t step0,
                                                       uint64_t nsteps, int32_t *out) {
  if (!s || !out) return fail(ACDC_ERR_INVALID, "acdc_sample_coordinates_noalloc: null argument");
  const uint64_t tau = (uint64_t)s->set.tau;
  if (nsteps > (uint64_t)ACDC_NOALLOC_MAX_SAMPLES / tau)
    return fail(ACDC_ERR_INVALID, "acdc_sample_coordinates_noalloc: nsteps * tau exceeds " +
                                      std::to_string(ACDC_NOALLOC_MAX_SAMPLES) + " samples");
  const int64_t total = (int64_t)(nsteps * tau);
  if (total == 0) return ACDC_OK;
  int32_t *d = nullptr;
  acdc_status st = scratch_on(s->device, &d);
  if (st != ACDC_OK) return st;
  // the kernels write every entry of [0, total) before the copy reads it
  if (s->async)
    k_sample_async<<<(unsigned)div_up(total, kBlock), kBlock, 0, s->stream>>>(
        (uint32_t)s->ncoord, s->set.seed, step0, nsteps, s->set.tau, d);
  else
    k_sample<<<(unsigned)div_up(total, kBlock), kBlock, 0, s->stream>>>(
        Perm::make((uint32_t)s->ncoord), s->set.seed, step0, nsteps, s->set.tau, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, (size_t)total * 4, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess)
    return fail(ACDC_ERR_HIP, std::string("acdc_sample_coordinates_noalloc: ") + hipGetErrorString(e));
  return ACDC_OK;
}

extern "C" acdc_status acdc_synchronize(acdc_solver *s) {
  if (!s) return fail(ACDC_ERR_INVALID, "acdc_synchronize: null solver");
  HIP_TRY(hipStreamSynchronize(s->stream));
  return settle_timing(s);
}

extern "C" acdc_status acdc_profile_step_kernels(acdc_solver *s, uint64_t nsteps, double *gather_ms,
                                                 double *scatter_ms) {
  if (!s || !gather_ms || !scatter_ms)
    return fail(ACDC_ERR_INVALID, "acdc_profile_step_kernels: null argument");
  if (s->set.tau < 2 || s->async)
    return fail(ACDC_ERR_STATE,
                "acdc_profile_step_kernels: needs synchronous steps with tau >= 2 (two phases)");
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
      dispatch3<LaunchStep, true>(s, s->lanes, s->step, 1);
      (void)hipEventRecord(ev[3 * k + 1], s->stream);
      dispatch3<LaunchStep, true>(s, s->lanes, s->step, 2);
      (void)hipEventRecord(ev[3 * k + 2], s->stream);
      s->step += 1;
      s->updates += (uint64_t)s->set.tau;
      s->steps_since_refresh += 1;
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
