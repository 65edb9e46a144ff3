// tune.hip -- the autotuner: picks lanes per coordinate, tau and queue
// chunking for the batched path, and lanes per coordinate for a single
// solver, by timing short runs on the device through the library's own
// entry points.
//
// Score of a batch candidate: measured progress per second, log(gap(x = 0) /
// gap after the trial) over the trial's device time, at one fixed lambda and
// seed.  Larger tau buys updates/s but shortens every step through the ESO
// beta (loss_square.h:183-194), so work rate alone would overrate it.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/acdc_hip.h"
#include "device_common.h"
#include "problem.h"

namespace {

// ||A^T b||_inf of a column-major problem (for sample rows, the SVM's
// labels: 10 / n, so that frac = 0.1 gives lambda = 1 / n)
double lambda_max(const acdc_problem *p) {
  if (p->kind == 1) return 10.0 / (double)std::max<int64_t>(p->ncoord, 1);
  double m = 0;
  for (int64_t j = 0; j < p->ncoord; ++j) {
    double g = 0;
    for (int32_t q = p->cptr[j]; q < p->cptr[j + 1]; ++q) g += p->cval[q] * p->rhs[p->cidx[q]];
    m = std::max(m, std::fabs(g));
  }
  return m;
}

acdc_status tfail(acdc_status st, const std::string &msg) {
  acdc::set_error(msg.c_str());
  return st;
}

}  // namespace

extern "C" acdc_status acdc_batch_tune(const acdc_problem *const *problems, int32_t count,
                                       const acdc_batch_settings *base,
                                       acdc_batch_settings *tuned, double *score) {
  if (!problems || !base || !tuned || count < 1)
    return tfail(ACDC_ERR_INVALID, "acdc_batch_tune: null argument or empty batch");
  // the instance that bounds the batch: most entries
  int32_t big = 0;
  const int kind = base->loss == ACDC_LOSS_SVM_DUAL_HINGE ? 1 : 0;
  for (int32_t i = 0; i < count; ++i) {
    if (!problems[i] || problems[i]->kind != kind)
      return tfail(ACDC_ERR_INVALID, kind ? "acdc_batch_tune: SVM problems must be sample rows (CSR)"
                                          : "acdc_batch_tune: Lasso and NNLS problems must be column-major");
    if (problems[i]->nnz > problems[big]->nnz) big = i;
  }
  const acdc_problem *p = problems[big];
  const double lam = 0.1 * lambda_max(p);
  const int epochs = 4;
  auto trial = [&](acdc_batch_settings s, int ep, double *ms, double *gap) -> acdc_status {
    s.refresh_period = std::max(1, ep);
    s.max_epochs = ep;
    s.gap_tol = -1.0;
    s.use_graph = 0;
    acdc_batch *b = nullptr;
    acdc_status st = acdc_batch_create(&p, 1, &lam, 1, &s, &b);
    if (st != ACDC_OK) return st;
    st = acdc_batch_run(b, ms);  // first run: code and data warm
    if (st == ACDC_OK) st = acdc_batch_run(b, ms);
    acdc_batch_result r{};
    if (st == ACDC_OK) st = acdc_batch_results(b, &r, 1);
    acdc_batch_destroy(b);
    *gap = r.rel_gap;
    return st;
  };
  double ms0 = 0, gap0 = 0;
  acdc_status st0 = trial(*base, 0, &ms0, &gap0);  // the gap at x = 0
  if (st0 != ACDC_OK) return st0;
  const bool async = base->step_mode == ACDC_STEP_ASYNC;
  const int lane_set[3] = {8, 16, 32};
  double best = -1;
  acdc_batch_settings pick = *base;
  for (int lanes : lane_set) {
    const int cap = (async ? 1 : 2) * (1024 / lanes);
    for (int div : {4, 2, 1}) {
      const int tau = std::max(1, cap / div);
      if (tau > p->ncoord) continue;
      acdc_batch_settings s = *base;
      s.lanes_per_coord = lanes;
      s.tau = tau;
      double ms = 0, gap = 0;
      const acdc_status st = trial(s, epochs, &ms, &gap);
      if (st != ACDC_OK) return st;
      const double sc = std::log(gap0 / std::max(gap, 1e-300)) / (ms * 1e-3);
      if (sc > best) {
        best = sc;
        pick.lanes_per_coord = lanes;
        pick.tau = tau;
      }
    }
  }
  // queue chunking: one instance per queue access while the batch has no
  // more instances than resident blocks (a larger chunk could only leave
  // blocks idle); otherwise the chunk with the least device time over a
  // short run of the whole batch (two lambdas, two epochs each)
  pick.queue_chunk = 1;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      count > cus) {
    std::vector<double> lams((size_t)count * 2);
    for (int32_t i = 0; i < count; ++i) {
      const double lm = lambda_max(problems[i]);
      lams[2 * (size_t)i] = 0.5 * lm;
      lams[2 * (size_t)i + 1] = 0.1 * lm;
    }
    double best_ms = -1;
    for (int chunk : {1, 2, 4, 8}) {
      acdc_batch_settings s = pick;
      s.queue_chunk = chunk;
      s.refresh_period = 2;
      s.max_epochs = 2;
      s.gap_tol = -1.0;
      s.use_graph = 0;
      acdc_batch *b = nullptr;
      acdc_status st = acdc_batch_create(problems, count, lams.data(), 2, &s, &b);
      double ms = 0;
      if (st == ACDC_OK) st = acdc_batch_run(b, &ms);
      if (b) acdc_batch_destroy(b);
      if (st != ACDC_OK) return st;
      if (best_ms < 0 || ms < best_ms) {
        best_ms = ms;
        pick.queue_chunk = chunk;
      }
    }
  }
  *tuned = pick;
  if (score) *score = best;
  return ACDC_OK;
}

extern "C" acdc_status acdc_solver_tune_lanes(const acdc_problem *p, int32_t loss,
                                              const acdc_settings *base, int32_t *lanes,
                                              double *updates_per_s) {
  if (!p || !base || !lanes) return tfail(ACDC_ERR_INVALID, "acdc_solver_tune_lanes: null argument");
  double best = -1;
  int32_t pick = 0;
  for (int32_t g : {16, 32, 64}) {
    acdc_settings s = *base;
    s.lanes_per_coord = g;
    s.refresh_period = 0;
    acdc_solver *sv = nullptr;
    acdc_status st = acdc_solver_create(p, loss, &s, &sv);
    if (st != ACDC_OK) return st;
    acdc_report r0{}, r1{};
    st = acdc_run_updates(sv, (uint64_t)p->ncoord);  // first epoch from x = 0: warm-up
    if (st == ACDC_OK) st = acdc_report_get(sv, &r0);
    if (st == ACDC_OK) st = acdc_run_updates(sv, 2 * (uint64_t)p->ncoord);
    if (st == ACDC_OK) st = acdc_report_get(sv, &r1);
    acdc_solver_destroy(sv);
    if (st != ACDC_OK) return st;
    const double ms = r1.step_ms - r0.step_ms;
    const double rate = ms > 0 ? (double)(r1.updates - r0.updates) / (ms * 1e-3) : 0.0;
    if (rate > best) {
      best = rate;
      pick = g;
    }
  }
  *lanes = pick;
  if (updates_per_s) *updates_per_s = best;
  return ACDC_OK;
}
