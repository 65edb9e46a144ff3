// batch.hip -- C ABI of the batched multi-instance solver: Lasso, NNLS or
// hinge-SVM dual instances (acdc_batch_* in include/acdc_hip.h).
//
// Layout: all instances' arrays are packed into one allocation per kind
// (coordinate records, (row, value) entries, FP64 values, b, r64, global dr),
// and a device table of BatchInstance records points into them.  Instances whose rows fit
// in 160 KB of LDS form one work list, the rest a second one; each list is a
// persistent launch of k_batch_lasso over a work-queue counter, largest
// instances first.  The reset kernel and both launches are captured once in a
// HIP graph and replayed by acdc_batch_run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/acdc_hip.h"
#include "problem.h"
#include "batch_kernels.h"

using namespace acdc;


namespace {
acdc_status bfail(acdc_status st, const std::string &msg) {
  acdc::set_error(msg.c_str());
  return st;
}
#define BHIP(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return bfail(ACDC_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_));     \
  } while (0)

// inside acdc_batch_create: a failing call also frees the partial batch
#define CREATE_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return bail(bfail(ACDC_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_))); \
  } while (0)

constexpr size_t kLdsBytes = 160 * 1024 - 1024;  // dynamic LDS per block (static scratch aside)
}  // namespace

struct acdc_batch {
  acdc_batch_settings set{};
  bool fp32 = true;
  bool async = false;  // ACDC_STEP_ASYNC inside instances
  bool svm = false;    // hinge-SVM dual instances (kernels of their own)
  bool ws = false;     // screened Lasso instances with asynchronous workers (kernels of their own)
  int lanes = 16;
  int32_t count = 0, path_len = 0;
  std::vector<int64_t> n, m, xoff;
  std::vector<int32_t> tau;
  std::vector<double> beta;
  int64_t total_n = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  BatchInstance *dinst = nullptr;
  CoordRec *dcs = nullptr;    // records of all instances, contiguous
  int32_t *dorder = nullptr;  // instances in processing order (estimated cost, largest first)
  double *dlambda = nullptr;
  unsigned int *dqueue = nullptr;
  BatchResult *dres = nullptr;
  int blocks = 0;  // resident blocks of the batch kernel
  int64_t maxlen = 0;  // longest column of the batch
  double last_ms = 0;
  double clock_khz = 1e5;  // device wall clock (hipDeviceAttributeWallClockRate)
  std::vector<void *> allocs;
};

namespace {
template <typename P>
acdc_status balloc(acdc_batch *b, P **p, size_t count) {
  void *v = nullptr;
  hipError_t e = hipMalloc(&v, std::max<size_t>(count * sizeof(P), 16));
  if (e != hipSuccess) return bfail(ACDC_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  b->allocs.push_back(v);
  *p = static_cast<P *>(v);
  return ACDC_OK;
}

void destroy_batch(acdc_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (void *p : b->allocs) (void)hipFree(p);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const void *batch_kernel(const acdc_batch *b) {
  if (b->svm) return batch_svm_kernel(b->fp32, b->lanes, b->async, b->maxlen);
  if (b->ws) return batch_ws_kernel(b->fp32, b->lanes, b->maxlen);
  return batch_kernel_for<false>(b->fp32, b->lanes, b->async, b->maxlen);
}

// enqueue the reset and the persistent launch on b->stream
acdc_status enqueue_run(acdc_batch *b) {
  const int64_t nreset = std::max<int64_t>(b->total_n, 1);
  k_batch_reset<<<(unsigned)((nreset + kBlock - 1) / kBlock), kBlock, 0, b->stream>>>(
      b->dcs, b->total_n, b->dqueue, 1);
  BHIP(hipGetLastError());
  const size_t tsz = b->fp32 ? 4 : 8;
  BatchArgs a;
  a.inst = b->dinst;
  a.order = b->dorder;
  a.count = b->count;
  a.path_len = b->path_len;
  a.lambdas = b->dlambda;
  a.refresh_period = b->set.refresh_period;
  a.max_epochs = b->set.max_epochs;
  a.gap_tol = b->set.gap_tol;
  a.queue = b->dqueue;
  a.chunk = std::max(1, b->set.queue_chunk);
  a.results = b->dres;
  a.lds_doubles = (int32_t)(kLdsBytes / 8);
  a.screening = b->set.screening;
  a.spectral = b->set.beta_mode == ACDC_BETA_SPECTRAL;
  a.ws_newton = b->set.ws_newton_iters;
  void *args[] = {&a};
  const size_t shmem = (kLdsBytes / tsz) * tsz;
  BHIP(hipLaunchKernel(batch_kernel(b), dim3(std::min(b->blocks, b->count)), dim3(kBatchBlock),
                       args, shmem, b->stream));
  return ACDC_OK;
}

template <typename T>
void pack_entries(const acdc_problem *p, std::vector<Entry<T>> &out) {
  out.resize(p->nnz);
  for (int64_t q = 0; q < p->nnz; ++q) {
    Entry<T> e{};
    e.row = p->cidx[q];
    e.val = (T)p->cval[q];
    out[q] = e;
  }
}
}  // namespace

extern "C" acdc_status acdc_batch_create(const acdc_problem *const *problems, int32_t count,
                                         const double *lambdas, int32_t path_len,
                                         const acdc_batch_settings *set, acdc_batch **out) {
  if (!problems || !lambdas || !set || !out || count < 1 || path_len < 1)
    return bfail(ACDC_ERR_INVALID, "acdc_batch_create: null argument or empty batch");
  if (set->precision != ACDC_PREC_FP64 && set->precision != ACDC_PREC_FP32_REFRESH)
    return bfail(ACDC_ERR_INVALID, "acdc_batch_create: unknown precision");
  if (set->step_mode != ACDC_STEP_SYNC && set->step_mode != ACDC_STEP_ASYNC)
    return bfail(ACDC_ERR_INVALID, "acdc_batch_create: unknown step mode");
  if (set->refresh_period < 1 || set->max_epochs < 0)
    return bfail(ACDC_ERR_INVALID, "acdc_batch_create: refresh_period >= 1 and max_epochs >= 0");
  if (set->lanes_per_coord != 0 && set->lanes_per_coord != 8 && set->lanes_per_coord != 16 &&
      set->lanes_per_coord != 32 && set->lanes_per_coord != 64)
    return bfail(ACDC_ERR_INVALID, "acdc_batch_create: lanes_per_coord must be 0, 8, 16, 32 or 64");
  if (set->beta_mode != ACDC_BETA_OMEGA && set->beta_mode != ACDC_BETA_SPECTRAL)
    return bfail(ACDC_ERR_INVALID, "acdc_batch_create: beta_mode must be omega or spectral");
  if (set->screening != 0 && set->screening != 1)
    return bfail(ACDC_ERR_INVALID, "acdc_batch_create: screening must be 0 or 1");
  if (set->ws_newton_iters < 0 || set->ws_newton_iters > ACDC_WS_NEWTON_MAX_ITERS)
    return bfail(ACDC_ERR_INVALID, "acdc_batch_create: ws_newton_iters must be in [0, 10000]");
  if (set->loss != ACDC_LOSS_L1_SQUARE && set->loss != ACDC_LOSS_NNLS &&
      set->loss != ACDC_LOSS_SVM_DUAL_HINGE)
    return bfail(ACDC_ERR_INVALID, "acdc_batch_create: loss must be L1-square, NNLS or SVM dual");
  const bool svm = set->loss == ACDC_LOSS_SVM_DUAL_HINGE;
  for (int32_t i = 0; i < count; ++i) {
    if (!problems[i] || problems[i]->kind != (svm ? 1 : 0))
      return bfail(ACDC_ERR_INVALID, svm ? "acdc_batch_create: SVM problems must be sample rows (CSR)"
                                         : "acdc_batch_create: Lasso and NNLS problems must be column-major");
    for (int32_t l = 0; l < path_len; ++l) {
      const double lam = lambdas[(int64_t)i * path_len + l];
      if (svm ? !(lam > 0) : !(lam >= 0))
        return bfail(ACDC_ERR_INVALID, svm ? "acdc_batch_create: SVM lambdas must be > 0"
                                           : "acdc_batch_create: lambdas must be >= 0");
    }
  }
  int32_t ndev = 0;
  acdc_device_count(&ndev);
  if (ndev < 1) return bfail(ACDC_ERR_NO_DEVICE, "acdc_batch_create: no gfx950 device visible");

  acdc_batch *b = new (std::nothrow) acdc_batch();
  if (!b) return bfail(ACDC_ERR_NOMEM, "acdc_batch_create: out of host memory");
  auto bail = [&](acdc_status e) {
    std::string msg = acdc_last_error();
    destroy_batch(b);
    acdc::set_error(msg.c_str());
    return e;
  };
  b->set = *set;
  b->fp32 = set->precision == ACDC_PREC_FP32_REFRESH;
  b->async = set->step_mode == ACDC_STEP_ASYNC;
  b->svm = svm;
  b->count = count;
  b->path_len = path_len;
  int64_t tot_nnz = 0, tot_n = 0, tot_m = 0;
  for (int32_t i = 0; i < count; ++i) {
    const std::vector<int32_t> &cp = problems[i]->cptr;
    for (size_t c = 0; c + 1 < cp.size(); ++c) b->maxlen = std::max<int64_t>(b->maxlen, cp[c + 1] - cp[c]);
    tot_nnz += problems[i]->nnz;
    tot_n += problems[i]->ncoord;
    tot_m += problems[i]->nres;
  }
  // lanes per coordinate (measured on the 200k-column instances of the batched
  // Lasso workload: asynchronous workers peak at 16 lanes = 64 workers per
  // block up to ~48 entries per column on average)
  if (set->lanes_per_coord)
    b->lanes = set->lanes_per_coord;
  else if (b->async)
    b->lanes = tot_nnz <= 48 * tot_n ? 16 : 32;
  else
    b->lanes = tot_nnz <= 8 * tot_n ? 8 : (tot_nnz <= 24 * tot_n ? 16 : 32);
  // tau per instance: the block's coordinate slots per synchronous step, or
  // its lane groups as asynchronous workers
  const int32_t cap = (b->async ? 1 : kSlotsPerGroup) * (kBatchBlock / b->lanes);
  if (set->queue_chunk < 0)
    return bail(bfail(ACDC_ERR_INVALID, "acdc_batch_create: queue_chunk must be >= 0"));
  if (set->tau < 0 || set->tau > cap)
    return bail(bfail(ACDC_ERR_INVALID, "acdc_batch_create: tau must lie in [0, " +
                                            std::to_string(cap) + "] for these lanes and mode"));
  const int32_t slots = set->tau ? set->tau : cap;
  const size_t tsz = b->fp32 ? 4 : 8;
  const size_t esz = b->fp32 ? sizeof(Entry<float>) : sizeof(Entry<double>);
  const int64_t lds_rows = (int64_t)(kLdsBytes / tsz);
  b->total_n = tot_n;

  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
    return bail(bfail(ACDC_ERR_HIP, "acdc_batch_create: stream/event creation failed"));
  unsigned char *dent, *ddr = nullptr;
  RefEntry *drent;
  double *db, *dr64;
  acdc_status st;
  if ((st = balloc(b, &b->dcs, tot_n)) != ACDC_OK) return bail(st);
  // working sets (screened Lasso instances with asynchronous workers only)
  const bool ws = set->screening && !svm && set->loss == ACDC_LOSS_L1_SQUARE && b->async;
  b->ws = ws;
  CoordRec *dws = nullptr;
  int32_t *dwsi = nullptr;
  double *dan = nullptr;  // Anderson snapshots and Newton-step vectors of the working sets
  if (ws && ((st = balloc(b, &dws, tot_n)) != ACDC_OK || (st = balloc(b, &dwsi, tot_n)) != ACDC_OK ||
             (st = balloc(b, &dan, tot_n * kWsScratchSlots)) != ACDC_OK))
    return bail(st);
  // one pad entry after the last segment: workers load a padding lane's
  // entry unconditionally (from the segment start, which for an empty last
  // segment is the pad)
  if ((st = balloc(b, &dent, (tot_nnz + 1) * esz)) != ACDC_OK) return bail(st);
  if (hipMemset(dent + tot_nnz * esz, 0, esz) != hipSuccess)  // the pad entry: row 0, value 0
    return bail(bfail(ACDC_ERR_HIP, "acdc_batch_create: pad entry"));
  if ((st = balloc(b, &drent, tot_nnz)) != ACDC_OK) return bail(st);
  RefEntry *dsent = nullptr;  // the support entries of the working sets' Newton steps
  if (ws && (st = balloc(b, &dsent, tot_nnz)) != ACDC_OK) return bail(st);
  if ((st = balloc(b, &db, tot_m)) != ACDC_OK) return bail(st);
  if ((st = balloc(b, &dr64, tot_m)) != ACDC_OK) return bail(st);
  int64_t glob_m = 0;
  for (int32_t i = 0; i < count; ++i)
    if (problems[i]->nres > lds_rows) glob_m += problems[i]->nres - lds_rows;
  if (glob_m > 0 && (st = balloc(b, &ddr, glob_m * tsz)) != ACDC_OK) return bail(st);
  if ((st = balloc(b, &b->dinst, count)) != ACDC_OK) return bail(st);
  if ((st = balloc(b, &b->dorder, count)) != ACDC_OK) return bail(st);
  if ((st = balloc(b, &b->dlambda, (size_t)count * path_len)) != ACDC_OK) return bail(st);
  if ((st = balloc(b, &b->dqueue, 1)) != ACDC_OK) return bail(st);
  if ((st = balloc(b, &b->dres, (size_t)count * path_len)) != ACDC_OK) return bail(st);

  std::vector<BatchInstance> hinst(count);
  int64_t onnz = 0, on = 0, om = 0, og = 0;
  hipError_t e = hipSuccess;
  std::vector<CoordRec> recs;
  std::vector<Entry<float>> ef;
  std::vector<Entry<double>> ed;
  std::vector<RefEntry> re;
  for (int32_t i = 0; i < count && e == hipSuccess; ++i) {
    const acdc_problem *p = problems[i];
    BatchInstance &I = hinst[i];
    I.cs = b->dcs + on;
    I.ws = ws ? dws + on : nullptr;
    I.wsi = ws ? dwsi + on : nullptr;
    I.anx = ws ? dan + on * kWsScratchSlots : nullptr;
    I.ent = dent + onnz * esz;
    I.rent = drent + onnz;
    I.sent = ws ? dsent + onnz : nullptr;
    I.nnz = (int32_t)p->nnz;
    I.loss = set->loss;  // the LossId values are the ABI's
    I.scale = 0;
    I.colsq = 0;
    for (int64_t j = 0; !svm && j < p->ncoord; ++j) {  // ||A^T b||_inf and mean ||a_j||^2
      double g = 0, c = 0;
      for (int32_t q = p->cptr[j]; q < p->cptr[j + 1]; ++q) {
        g += p->cval[q] * p->rhs[p->cidx[q]];
        c += p->cval[q] * p->cval[q];
      }
      I.scale = std::max(I.scale, std::fabs(g));
      I.colsq += c;
    }
    if (p->ncoord > 0) I.colsq /= (double)p->ncoord;
    I.b = svm ? nullptr : db + om;
    I.r64 = dr64 + om;
    // residual rows beyond the LDS capacity go to the instance's global part
    I.drg = nullptr;
    I.lds_rows = (int32_t)std::min<int64_t>(p->nres, lds_rows);
    if (p->nres > lds_rows) {
      I.drg = ddr + og * tsz;
      og += p->nres - lds_rows;
    }
    I.m = (int32_t)p->nres;
    I.n = (int32_t)p->ncoord;
    I.tau = (int32_t)std::min<int64_t>(slots, p->ncoord);
    I.seed = set->seed + (uint64_t)i;
    double beta = p->ncoord > 1 ? 1.0 + (double)(p->omega - 1) * (double)(I.tau - 1) /
                                            ((double)p->ncoord - 1.0)
                                : 1.0;
    if (set->beta_mode == ACDC_BETA_SPECTRAL) {
      // the single solver's power method on this instance (FP64, same tau)
      acdc_settings ss{};
      ss.lambda = svm ? lambdas[(int64_t)i * path_len] : 0.0;
      ss.tau = I.tau;
      ss.seed = I.seed;
      ss.beta_mode = ACDC_BETA_SPECTRAL;
      ss.precision = ACDC_PREC_FP64;
      ss.step_mode = ACDC_STEP_SYNC;
      ss.launch_mode = ACDC_LAUNCH_PER_STEP;
      acdc_solver *sv = nullptr;
      acdc_status bs = acdc_solver_create(p, set->loss, &ss, &sv);
      if (bs == ACDC_OK) bs = acdc_get_beta(sv, &beta);
      if (sv) acdc_solver_destroy(sv);
      if (bs != ACDC_OK) return bail(bs);
    }
    I.beta = beta;
    b->n.push_back(p->ncoord);
    b->m.push_back(p->nres);
    b->xoff.push_back(on);
    b->tau.push_back(I.tau);
    b->beta.push_back(beta);
    // coordinate records: x = 0, L_j = 1/(beta ||a_j||^2) (loss_square.h:183-194;
    // MulticoreSVMDualLoss.h:50-58 for samples), times the label y_j = +-1
    // for SVM instances (the kernels read the label off L's sign)
    recs.resize(p->ncoord);
    for (int64_t j = 0; j < p->ncoord; ++j) {
      double s2 = 0;
      for (int32_t q = p->cptr[j]; q < p->cptr[j + 1]; ++q) s2 += p->cval[q] * p->cval[q];
      recs[j].x = 0.0;
      recs[j].L = s2 > 0 ? 1.0 / (beta * s2) : 0.0;
      if (svm) recs[j].L *= p->rhs[j];
      recs[j].gref = 0.0;
      recs[j].p0 = p->cptr[j];
      recs[j].len = p->cptr[j + 1] - p->cptr[j];
    }
    e = hipMemcpy(I.cs, recs.data(), recs.size() * sizeof(CoordRec), hipMemcpyHostToDevice);
    if (e == hipSuccess && p->nnz) {
      if (b->fp32) {
        pack_entries(p, ef);
        e = hipMemcpy(const_cast<void *>(I.ent), ef.data(), p->nnz * esz, hipMemcpyHostToDevice);
      } else {
        pack_entries(p, ed);
        e = hipMemcpy(const_cast<void *>(I.ent), ed.data(), p->nnz * esz, hipMemcpyHostToDevice);
      }
    }
    if (e == hipSuccess && p->nnz) {
      re.resize(p->nnz);
      for (int64_t j = 0; j < p->ncoord; ++j)
        for (int32_t q = p->cptr[j]; q < p->cptr[j + 1]; ++q) {
          re[q].col = (int32_t)j;
          re[q].row = p->cidx[q];
          re[q].val = p->cval[q];
        }
      e = hipMemcpy(const_cast<RefEntry *>(I.rent), re.data(), p->nnz * sizeof(RefEntry),
                    hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && !svm)
      e = hipMemcpy(const_cast<double *>(I.b), p->rhs.data(), p->nres * 8, hipMemcpyHostToDevice);
    onnz += p->nnz;
    on += p->ncoord;
    om += p->nres;
  }
  if (e != hipSuccess)
    return bail(bfail(ACDC_ERR_HIP, std::string("acdc_batch_create: upload: ") + hipGetErrorString(e)));

  // one work list, largest estimated cost first (longest-processing-time
  // order): an instance whose residual spills past the LDS costs more per
  // nnz than one held entirely in LDS
  std::vector<int32_t> ord(count);
  std::iota(ord.begin(), ord.end(), 0);
  auto cost = [&](int32_t i) {
    return (double)problems[i]->nnz * (problems[i]->nres <= lds_rows ? 1.0 : 1.3);
  };
  std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t c) { return cost(a) > cost(c); });
  CREATE_HIP(hipMemcpy(b->dorder, ord.data(), count * 4, hipMemcpyHostToDevice));
  CREATE_HIP(hipMemcpy(b->dinst, hinst.data(), count * sizeof(BatchInstance), hipMemcpyHostToDevice));
  CREATE_HIP(hipMemcpy(b->dlambda, lambdas, (size_t)count * path_len * 8, hipMemcpyHostToDevice));

  // grid sizes: resident blocks of each variant
  int dev = 0;
  hipDeviceProp_t prop;
  CREATE_HIP(hipGetDevice(&dev));
  CREATE_HIP(hipGetDeviceProperties(&prop, dev));
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
    b->clock_khz = khz;
  {
    const size_t shmem = (kLdsBytes / tsz) * tsz;
    CREATE_HIP(hipFuncSetAttribute(batch_kernel(b), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)shmem));
    int per_cu = 0;
    CREATE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel(b), kBatchBlock,
                                                      shmem));
    b->blocks = std::max(1, per_cu) * prop.multiProcessorCount;
  }
  CREATE_HIP(hipStreamSynchronize(b->stream));

  if (set->use_graph) {
    CREATE_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    st = enqueue_run(b);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(b->stream, &g);
    if (st != ACDC_OK) return bail(st);
    if (ce != hipSuccess)
      return bail(bfail(ACDC_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce)));
    b->graph = g;
    CREATE_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  *out = b;
  return ACDC_OK;
}

extern "C" acdc_status acdc_batch_run(acdc_batch *b, double *device_ms) {
  if (!b) return bfail(ACDC_ERR_INVALID, "acdc_batch_run: null batch");
  BHIP(hipEventRecord(b->ev0, b->stream));
  if (b->exec) {
    BHIP(hipGraphLaunch(b->exec, b->stream));
  } else {
    acdc_status st = enqueue_run(b);
    if (st != ACDC_OK) return st;
  }
  BHIP(hipEventRecord(b->ev1, b->stream));
  BHIP(hipEventSynchronize(b->ev1));
  float ms = 0;
  BHIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  if (device_ms) *device_ms = ms;
  return ACDC_OK;
}

extern "C" acdc_status acdc_batch_results(acdc_batch *b, acdc_batch_result *out, int64_t len) {
  if (!b || !out) return bfail(ACDC_ERR_INVALID, "acdc_batch_results: null argument");
  if (len != (int64_t)b->count * b->path_len)
    return bfail(ACDC_ERR_INVALID, "acdc_batch_results: len must be count * path_len");
  std::vector<BatchResult> r((size_t)len);
  BHIP(hipStreamSynchronize(b->stream));
  BHIP(hipMemcpy(r.data(), b->dres, r.size() * sizeof(BatchResult), hipMemcpyDeviceToHost));
  for (int64_t k = 0; k < len; ++k) {
    out[k].lambda = r[k].lambda;
    out[k].primal = r[k].primal;
    out[k].dual = r[k].dual;
    out[k].rel_gap = r[k].rel_gap;
    out[k].nnz_x = r[k].nnz_x;
    out[k].epochs = r[k].epochs;
    out[k].converged = r[k].converged;
    out[k].updates = r[k].updates;
    out[k].device_us = (double)r[k].ticks * 1e3 / b->clock_khz;
    out[k].ws_size = r[k].ws_size;
    out[k].status = r[k].status;
    out[k].newton_steps = r[k].newton;
  }
  return ACDC_OK;
}

extern "C" acdc_status acdc_batch_get_x(acdc_batch *b, int32_t instance, double *x, int64_t n) {
  if (!b || !x) return bfail(ACDC_ERR_INVALID, "acdc_batch_get_x: null argument");
  if (instance < 0 || instance >= b->count || n != b->n[instance])
    return bfail(ACDC_ERR_INVALID, "acdc_batch_get_x: bad instance or length");
  BHIP(hipStreamSynchronize(b->stream));
  std::vector<CoordRec> recs((size_t)n);
  BHIP(hipMemcpy(recs.data(), b->dcs + b->xoff[instance], n * sizeof(CoordRec),
                 hipMemcpyDeviceToHost));
  for (int64_t j = 0; j < n; ++j) x[j] = recs[j].x;
  return ACDC_OK;
}

extern "C" acdc_status acdc_batch_instance_info(const acdc_batch *b, int32_t instance,
                                                int64_t *tau, double *beta, int32_t *lds_resident) {
  if (!b || !tau || !beta || !lds_resident)
    return bfail(ACDC_ERR_INVALID, "acdc_batch_instance_info: null argument");
  if (instance < 0 || instance >= b->count)
    return bfail(ACDC_ERR_INVALID, "acdc_batch_instance_info: bad instance");
  *tau = b->tau[instance];
  *beta = b->beta[instance];
  *lds_resident = b->m[instance] <= (int64_t)(kLdsBytes / (b->fp32 ? 4 : 8));
  return ACDC_OK;
}

extern "C" acdc_status acdc_batch_destroy(acdc_batch *b) {
  destroy_batch(b);
  return ACDC_OK;
}
