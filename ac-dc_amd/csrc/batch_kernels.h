// batch_kernels.h -- one persistent launch that solves many Lasso, NNLS or
// hinge-SVM dual instances, each along a warm-started lambda path.
//
// Work distribution: a device work queue (one atomic counter) hands whole
// instances to 1024-thread blocks, largest first.  Inside a block an instance
// is solved by synchronous tau-steps exactly as the single-instance kernels do
// (same sampler, same proximal rule, same anchored residual), with the block
// barrier instead of a grid barrier between the gather and the scatter half
// of a step.  The correction residual dr of the instance lives in LDS when its
// rows fit (ds_add scatters, ds_read gathers), in a block-private global
// buffer otherwise.
//
// Layout for the random-coordinate access pattern: a sampled coordinate costs
// one 32-byte record {x, L, gref, offset, length} and its column's entries as
// interleaved (row, value) pairs, so a 16-lane group reads a 128-byte line
// per chunk and a coordinate's scalars come in one sector instead of four
// scattered 8-byte reads.  The first R*G entries of a column stay in
// registers from the gather to the scatter, and the next step's records and
// columns are loaded behind the block barrier.
//
// Every refresh_period epochs the block recomputes, in FP64 and from the FP64
// values, r64 = A x - b (scattering the nonzero columns), gref = A^T r64, the
// primal, the dual point theta = -r min(1, lambda/||A^T r||_inf) and the
// relative duality gap; a (instance, lambda) pair ends at rel gap <= gap_tol
// or max_epochs, and the next lambda starts from the current x (the warm start
// of lasso_console_solver.cpp:149-172).  All instance buffers are private to
// the block that owns the instance, so block barriers order every access.
#pragma once
#include "kernels.h"

namespace acdc {

// per-coordinate state, one 32-byte record
struct alignas(32) CoordRec {
  double x;
  double L;
  double gref;
  int32_t p0;   // first entry of the column
  int32_t len;  // entries in the column
};

// One entry of the refresh's FP64 copy of a column-major matrix, in entry
// order: the column index lets a wave stream entries across column bounds.
struct alignas(16) RefEntry {
  int32_t col;
  int32_t row;
  double val;
};

struct BatchInstance {
  CoordRec *cs;          // n records
  CoordRec *ws;          // working set: copies of the records of the coordinates
                         // gap-safe screening keeps (Lasso, screening on), n slots
  int32_t *wsi;          // coordinate of each working-set slot
  const void *ent;       // Entry<T>[nnz]
  const RefEntry *rent;  // FP64 entries of the refresh, entry order
  const double *b;       // null for SVM instances (the residual is w = A^T (alpha y) / (lambda n))
  double *r64;           // FP64 residual of the refresh when it does not fit in LDS
  void *drg;             // rows [lds_rows, m) of the correction residual (T), or null
  int32_t m, n;
  int32_t tau;           // coordinates per step of this instance
  int32_t lds_rows;      // rows [0, lds_rows) of the correction residual live in LDS
  int32_t nnz;
  int32_t loss;          // kLasso, kNnls (x >= 0, no lambda) or kSvm (SVM: the
                         // coordinates are samples, x is alpha, CoordRec.L is
                         // y_j / (beta ||x_j||^2): its sign is the label)
  int32_t wperm;         // nonzero on a working set, whose records are stored in
                         // a keyed random order: the workers sweep them
                         // cyclically (see batch_async_updates)
  RefEntry *sent;        // working sets: nnz entries of scratch (the support's
                         // entries of a Newton step), or null
  double *anx;           // working sets: kWsScratchSlots x n doubles: the
                         // kAaSnapshots snapshots of x of the Anderson
                         // extrapolation, then the vectors of the support
                         // Newton steps; or null
  double scale;          // NNLS: ||A^T b||_inf, the KKT residual's scale
  double colsq;          // Lasso and NNLS: the mean ||a_j||^2 of the columns
  double beta;           // ESO beta of the records' L_j = 1/(beta ||a_j||^2)
  uint64_t seed;
};

__device__ __forceinline__ double sign_of(double v) { return (double)((v > 0.0) - (v < 0.0)); }
// a block-uniform double moved to scalar registers (a division leaves its
// result in vector registers, where it would hold two of them per lane)
__device__ __forceinline__ double uniform(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)b);
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// The coordinate step of an instance: the Lasso's soft threshold
// (MulticoreL1RegSquareLoss.h:41-51), for an NNLS instance the projected step
// onto x >= 0 (prox_delta<kNnls>), for an SVM instance the box-clipped dual
// step (MulticoreSVMDualLoss.h:119-141) with the label taken from the sign of
// the stored L.  Returns the change of x_j.  SVM instances get kernels of
// their own (the SVM template flag), the Lasso and NNLS share theirs (loss
// tested per instance), so that no kernel carries the registers of a step it
// never runs.
template <bool SVM>
__device__ __forceinline__ double batch_delta(int loss, double g, double xj, double Lj,
                                             const StepScalars &sc) {
  double c;
  if constexpr (SVM) {
    (void)loss;
    return prox_delta<kSvm>(g, xj, fabs(Lj), sign_of(Lj), sc, &c);
  } else {
    return loss == kNnls ? prox_delta<kNnls>(g, xj, Lj, 0.0, sc, &c)
                         : prox_delta<kLasso>(g, xj, Lj, 0.0, sc, &c);
  }
}
// the factor of the column in the residual update for a change d of x_j:
// d itself, or d y_j / (lambda n) for the SVM's w
template <bool SVM>
__device__ __forceinline__ double batch_coef(double d, double Lj, const StepScalars &sc) {
  if constexpr (SVM) return d * sign_of(Lj) * sc.inv_lambda_n;
  return d;
}
// step scalars of one lambda of an instance with n coordinates
template <bool SVM>
__device__ __forceinline__ StepScalars batch_scalars(double lambda, int32_t n) {
  StepScalars sc;
  sc.lambda = lambda;
  sc.lambda_n = SVM ? uniform(lambda * (double)n) : 0.0;
  sc.inv_lambda_n = SVM ? uniform(1.0 / (lambda * (double)n)) : 0.0;
  return sc;
}

struct BatchResult {
  double lambda, primal, dual, rel_gap, nnz_x;
  int32_t epochs, converged;
  uint64_t updates;
  uint64_t ticks;  // wall-clock ticks of this lambda
  double grad_inf;   // max_j |<a_j, r>| of the evaluation
  double ws_size;    // coordinates in the working set at the end (n without screening)
  int32_t status;    // 0, or kBatchNonFinite
  int32_t newton;    // support Newton steps kept for this lambda
};

// An evaluation with a NaN or an infinity (in the data or the iterate)
// abandons the instance: this and its remaining lambdas are reported with
// this status, and its block takes the next instance from the queue.
constexpr int32_t kBatchNonFinite = 1;
__device__ __forceinline__ bool finite_eval(const BatchResult &r) {
  return isfinite(r.primal) && isfinite(r.dual) && isfinite(r.rel_gap);
}
struct BatchArgs {
  const BatchInstance *inst;
  const int32_t *order;  // instances of this launch, in processing order
  int32_t count;         // entries of order
  int32_t path_len;
  const double *lambdas;  // instances x path_len
  int32_t refresh_period;
  int32_t max_epochs;
  double gap_tol;
  unsigned int *queue;
  int32_t chunk;         // queue items a block takes per queue access
  BatchResult *results;  // instances x path_len
  int32_t lds_doubles;   // FP64 rows the dynamic LDS holds (refresh residual)
  int32_t screening;     // Lasso, asynchronous workers: gap-safe working sets
  int32_t spectral;      // working sets take the spectral ESO constant of W
  int32_t ws_newton;     // working sets: CG iterations of each support Newton step (0: none)
};

// the results of lambdas l .. path_len - 1 of an abandoned instance
__device__ __forceinline__ void batch_fail_rest(const BatchArgs &a, int i, int l,
                                                const BatchResult &last, uint64_t updates) {
  if (threadIdx.x != 0) return;
  for (int k = l; k < a.path_len; ++k) {
    BatchResult *o = a.results + (int64_t)i * a.path_len + k;
    o->lambda = a.lambdas[(int64_t)i * a.path_len + k];
    o->primal = last.primal;
    o->dual = last.dual;
    o->rel_gap = last.rel_gap;
    o->nnz_x = last.nnz_x;
    o->epochs = 0;
    o->converged = 0;
    o->updates = updates;
    o->ticks = 0;
    o->grad_inf = last.grad_inf;
    o->ws_size = 0.0;
    o->status = kBatchNonFinite;
    o->newton = 0;
  }
}

// zeroes x of every coordinate record and the work-queue counters (a kernel
// rather than a memset so that the whole run can live in one captured graph)
[[maybe_unused]] static __global__ __launch_bounds__(kBlock) void k_batch_reset(CoordRec *__restrict__ cs,
                                                               int64_t ncs,
                                                               unsigned int *__restrict__ queues,
                                                               int nq) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t < ncs) cs[t].x = 0.0;
  if (t < nq) queues[t] = 0u;
}

constexpr int kBatchBlock = 1024;
constexpr int kBatchWaves = kBatchBlock / 64;

// Statistics of a refresh reduced over the block in a fixed order: an xor
// butterfly inside each wave, then thread 0 folds the 16 wave partials in
// index order.  Only thread 0 holds the totals.
constexpr int kRefreshStats = 6;  // max|g|, ||x||_1, nnz(x), ||r||^2, b.r, non-finite g
__device__ __forceinline__ void block_stats(double (&v)[kRefreshStats], double *sh) {
  v[0] = group_max<64>(v[0]);
#pragma unroll
  for (int k = 1; k < kRefreshStats; ++k) v[k] = group_sum<64>(v[k]);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < kRefreshStats; ++k) sh[k * kBatchWaves + w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kRefreshStats; ++k) {
      double t = sh[k * kBatchWaves];
      for (int u = 1; u < kBatchWaves; ++u)
        t = k == 0 ? fmax(t, sh[k * kBatchWaves + u]) : t + sh[k * kBatchWaves + u];
      v[k] = t;
    }
  }
}

// A column held by one G-lane group: lane l owns entries p0 + l + r*G.
template <typename T, int G, int R>
struct ColCache {
  int32_t p0, len;
  int32_t row[R];
  T val[R];

  __device__ __forceinline__ void load(const Entry<T> *ent, int32_t p0_, int32_t len_, int lane) {
    p0 = p0_;
    len = len_;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int32_t q = lane + r * G;
      if (q < len) {
        const Entry<T> e = gld(ent + p0 + q);
        row[r] = e.row;
        val[r] = e.val;
      } else {
        row[r] = 0;
        val[r] = (T)0;
      }
    }
  }
  // <column, dr>, partial per lane (entries past R*G read straight from memory)
  template <typename View>
  __device__ __forceinline__ T dot(const Entry<T> *ent, const View &dr, int lane) const {
    T acc = 0;
#pragma unroll
    for (int r = 0; r < R; ++r)  // padding lanes read nothing
      if (lane + r * G < len) acc += val[r] * dr.get(row[r]);
    for (int32_t q = lane + R * G; q < len; q += G) {
      const Entry<T> e = gld(ent + p0 + q);
      acc += e.val * dr.get(e.row);
    }
    return acc;
  }
};

// The correction residual of an instance: rows [0, lds_rows) in LDS, rows
// [lds_rows, m) in the instance's global buffer (the rows of the widest
// instances exceed the 159 KB of LDS by a few thousand).  FULL: every row in
// LDS.  The global part is only touched by the owning block: workgroup-scope
// atomics, agent-scope loads (not through the L1, which does not see the
// atomics performed in L2).
template <typename T, bool FULL>
struct DrView {
  T *lds;
  T *glob;
  int32_t split;
  // both paths into their own registers, then a select: a single
  // destination written by an LDS read in some lanes and a global load in
  // others would make every LDS read wait for all global loads in flight
  __device__ __forceinline__ T get(int32_t row) const {
    if (FULL) return lds[row];
    const bool here = row < split;
    const T a = lds[here ? row : 0];
    T b = (T)0;
    if (!here) b = gld_agent(&glob[row - split]);
    return here ? a : b;
  }
  __device__ __forceinline__ void add(int32_t row, T v) const {
    if (FULL || row < split)
      atomicAdd(&lds[row], v);  // ds_add on LDS
    else
      gadd_block(&glob[row - split], v);
  }
  __device__ __forceinline__ void zero(int32_t m) const {
    for (int32_t k = threadIdx.x; k < m; k += blockDim.x) {
      if (FULL || k < split)
        lds[k] = (T)0;
      else
        *gp(&glob[k - split]) = (T)0;
    }
  }
};
template <typename T, bool FULL>
__device__ __forceinline__ DrView<T, FULL> dr_view(const BatchInstance &I, T *lds) {
  return DrView<T, FULL>{lds, static_cast<T *>(I.drg), FULL ? I.m : I.lds_rows};
}

// FP64 refresh + evaluation of one instance by one block: fills out (in
// shared memory) and zeroes dr.  For an SVM instance the residual is
// w = sum_j alpha_j y_j x_j / (lambda n) (b = 0) and gref_j = <x_j, w>.  r64 lives in LDS (reusing the space of dr,
// which is rebuilt at the end) when m doubles fit, in the global r64 buffer
// otherwise.  The scatter of the nonzero x_j gives each thread its own
// columns; the gradient pass streams the (column, row, value) entries in
// wave-contiguous ranges and sums columns by segmented scans.  Statistics
// meet in thread 0 only (block_stats).
template <typename T, bool LDSR, bool RL, bool SVM>
__device__ __forceinline__ void batch_refresh_impl(const BatchInstance *Ip, double lambda, T *lds,
                                                double *sh, BatchResult *out, bool dense) {
  // a copy: through a reference every store below could alias the
  // instance's fields, which would then be re-read after each store
  const BatchInstance I = *Ip;
  const int tid = threadIdx.x;
  constexpr bool rl = RL;  // r64 in LDS
  double *rs = reinterpret_cast<double *>(lds);  // r64 in LDS when rl
  double *rg = I.r64;                            // r64 in global memory otherwise
  // a global r64 takes atomics in L2, so it is read past the L1
  auto rget = [&](int32_t row) -> double { return rl ? rs[row] : gld_agent(rg + row); };
  auto radd = [&](int32_t row, double v) {
    if (rl)
      atomicAdd(rs + row, v);
    else
      gadd_block(rg + row, v);
  };
  constexpr bool svm = SVM;
  // SVM: the factor of alpha_j y_j in w
  const double inv_ln = svm ? uniform(1.0 / (lambda * (double)I.n)) : 0.0;
  for (int k = tid; k < I.m; k += kBatchBlock) {
    const double v = svm ? 0.0 : -gld(I.b + k);
    if (rl)
      rs[k] = v;
    else
      *gp(rg + k) = v;
  }
  __syncthreads();
  // r += x_j a_j over the nonzero x_j.  ||x||_1 and nnz(x) are counted here.
  double ginf = 0, x1 = 0, nz = 0;
  double bad = 0;  // gradients that are NaN or infinite (fmax drops NaN)
  const int wv = tid >> 6, ln = tid & 63;
  // wave w's contiguous share of the entries (both streaming passes)
  const int32_t e_lo = (int32_t)((int64_t)I.nnz * wv / kBatchWaves);
  const int32_t e_hi = (int32_t)((int64_t)I.nnz * (wv + 1) / kBatchWaves);
  constexpr int U = 4;  // steps of 64 entries per wave and round
  if (!dense) {
    // few nonzeros: each thread takes its own columns, the x of four columns
    // loaded before any is tested
    constexpr int XU = 4;
    for (int j0 = tid; j0 < I.n; j0 += XU * kBatchBlock) {
      double xs[XU];
#pragma unroll
      for (int u = 0; u < XU; ++u) {
        const int j = j0 + u * kBatchBlock;
        xs[u] = j < I.n ? gld_agent(&I.cs[j].x) : 0.0;  // async workers add to x atomically
      }
#pragma unroll
      for (int u = 0; u < XU; ++u) {
        if (xs[u] == 0.0) continue;
        x1 += fabs(xs[u]);
        nz += 1.0;
        const int j = j0 + u * kBatchBlock;
        const int32_t p0 = gld(&I.cs[j].p0), len = gld(&I.cs[j].len);
        const double cf = svm ? xs[u] * sign_of(gld(&I.cs[j].L)) * inv_ln : xs[u];
        // four entries loaded before their adds (with r64 in global memory
        // the adds are atomics, and a load issued after one waits for it)
        for (int32_t q0 = 0; q0 < len; q0 += 4) {
          RefEntry e[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) e[q] = gld(I.rent + p0 + min(q0 + q, len - 1));
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (q0 + q < len) radd(e[q].row, e[q].val * cf);
        }
      }
    }
  } else {
    // many nonzeros: the statistics over the records, then every entry
    // streamed (wave-contiguous ranges, the next round's loads in flight)
    // with its column's x gathered beside it.  SVM: the column factor
    // alpha_j y_j / (lambda n) is parked in gref (which the gradient pass
    // below rewrites) and gathered instead of x.
    constexpr int XU = 8;
    for (int j0 = tid; j0 < I.n; j0 += XU * kBatchBlock) {
      double xs[XU];
#pragma unroll
      for (int u = 0; u < XU; ++u) {
        const int j = j0 + u * kBatchBlock;
        xs[u] = j < I.n ? gld_agent(&I.cs[j].x) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < XU; ++u) {
        x1 += fabs(xs[u]);
        nz += xs[u] != 0.0;
      }
      if constexpr (svm)
#pragma unroll
        for (int u = 0; u < XU; ++u) {
          const int j = j0 + u * kBatchBlock;
          if (j < I.n)
            *gp(&I.cs[j].gref) = xs[u] != 0.0 ? xs[u] * sign_of(gld(&I.cs[j].L)) * inv_ln : 0.0;
        }
    }
    if constexpr (svm) __syncthreads();  // the factors are in before the stream reads them
    if (e_lo < e_hi) {
      RefEntry nxt[U];
      auto load = [&](int32_t base) {
#pragma unroll
        for (int u = 0; u < U; ++u) nxt[u] = gld(I.rent + min(base + u * 64 + ln, e_hi - 1));
      };
      load(e_lo);
      for (int32_t base = e_lo; base < e_hi; base += 64 * U) {
        RefEntry cur[U];
#pragma unroll
        for (int u = 0; u < U; ++u) cur[u] = nxt[u];
        load(base + 64 * U);
        double xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
          xv[u] = gld_agent(svm ? &I.cs[cur[u].col].gref : &I.cs[cur[u].col].x);
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (base + u * 64 + ln < e_hi && xv[u] != 0.0) radd(cur[u].row, cur[u].val * xv[u]);
      }
    }
  }
  __syncthreads();
  // g_j = <a_j, r64> for every column, streaming the entries: wave w takes
  // the contiguous entries [E w/16, E (w+1)/16), 64 consecutive entries per
  // step and U steps per round (the next round's loads in flight while this
  // round is summed), sums the runs of equal column by a segmented scan
  // across the wave and stores each column it completes.  Two runs per wave
  // are joined by thread 0 afterwards: the wave's first column (it may have
  // begun in the previous range) and the run still open at its end.  Columns
  // without entries keep gref = 0.
  {
    double *slot = sh + 4 * wv;  // head column, head sum, tail column, tail sum
    if (ln == 0) slot[0] = -1.0;
    int first_col = -1, open_col = -1;
    double open_sum = 0.0;
    auto emit = [&](int c, double v) {
      if (c == first_col) {
        slot[0] = (double)c;
        slot[1] = v;
      } else {
        *gp(&I.cs[c].gref) = v;
        ginf = fmax(ginf, fabs(v));
        bad += isfinite(v) ? 0.0 : 1.0;
      }
    };
    RefEntry nxt[U];
    // every lane loads; a lane past the range re-reads the range's last entry
    // and is masked where the entry is used
    auto load = [&](int32_t base) {
#pragma unroll
      for (int u = 0; u < U; ++u) nxt[u] = gld(I.rent + min(base + u * 64 + ln, e_hi - 1));
    };
    if (e_lo < e_hi) {
      load(e_lo);
      first_col = __builtin_amdgcn_readfirstlane(nxt[0].col);
      for (int32_t base = e_lo; base < e_hi; base += 64 * U) {
        RefEntry cur[U];
#pragma unroll
        for (int u = 0; u < U; ++u) cur[u] = nxt[u];
        load(base + 64 * U);
        double rv[U];  // the round's gathers, all issued before the first is used
#pragma unroll
        for (int u = 0; u < U; ++u) rv[u] = rget(cur[u].row);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int32_t cb = base + u * 64;
          if (cb >= e_hi) break;  // wave-uniform
          const bool in = cb + ln < e_hi;
          const int key = in ? cur[u].col : 0x7fffffff;
          const double p = in ? cur[u].val * rv[u] : 0.0;
          const SegSum r = seg_scan_sum(p, key);
          double v = r.sum;
          const int k0 = __builtin_amdgcn_readfirstlane(key);
          if (open_col >= 0) {
            if (k0 == open_col) {
              if (key == open_col) v += open_sum;
            } else if (ln == 0) {
              emit(open_col, open_sum);  // the open run ended with the last step
            }
          }
          const int lv = min(63, e_hi - 1 - cb);  // last lane in range
          const uint64_t heads = __ballot(r.head);
          if (ln < lv && ((heads >> (ln + 1)) & 1)) emit(key, v);
          open_col = __builtin_amdgcn_readlane(key, lv);
          const long long vb = __double_as_longlong(v);
          const int lo = __builtin_amdgcn_readlane((int)vb, lv);
          const int hi = __builtin_amdgcn_readlane((int)(vb >> 32), lv);
          open_sum = __longlong_as_double(
              (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
        }
      }
    }
    if (ln == 0) {
      slot[2] = (double)open_col;
      slot[3] = open_sum;
    }
    __syncthreads();
    if (tid == 0) {  // join the partial runs, in entry order
      int c = -1;
      double v = 0.0;
      for (int k = 0; k < 2 * kBatchWaves; ++k) {
        const int kc = (int)sh[2 * k];
        if (kc < 0) continue;
        if (kc == c) {
          v += sh[2 * k + 1];
          continue;
        }
        if (c >= 0) {
          *gp(&I.cs[c].gref) = v;
          ginf = fmax(ginf, fabs(v));
          bad += isfinite(v) ? 0.0 : 1.0;
        }
        c = kc;
        v = sh[2 * k + 1];
      }
      if (c >= 0) {
        *gp(&I.cs[c].gref) = v;
        ginf = fmax(ginf, fabs(v));
        bad += isfinite(v) ? 0.0 : 1.0;
      }
    }
    __syncthreads();  // sh is free for block_stats
  }
  double r2 = 0, br = 0;
  for (int k = tid; k < I.m; k += kBatchBlock) {
    const double rk = rget(k);
    r2 += rk * rk;
    if (!svm) br += gld(I.b + k) * rk;
  }
  if (!svm && I.loss == kNnls) {  // NNLS: the KKT residual max_j |min(x_j, g_j)| takes max|g|'s slot
    ginf = 0;
    for (int j = tid; j < I.n; j += kBatchBlock)
      ginf = fmax(ginf, fabs(fmin(gld_agent(&I.cs[j].x), gld(&I.cs[j].gref))));
  }
  if constexpr (svm)  // SVM: the hinge losses sum_j max(0, 1 - y_j <x_j, w>) take b.r's slot
    for (int j = tid; j < I.n; j += kBatchBlock) {
      const double h = 1.0 - sign_of(gld(&I.cs[j].L)) * gld(&I.cs[j].gref);
      br += h > 0.0 ? h : 0.0;
    }
  double st[kRefreshStats] = {ginf, x1, nz, r2, br, bad};
  block_stats(st, sh);  // its barrier also ends every read of r
  dr_view<T, LDSR>(I, lds).zero(I.m);
  if (tid == 0) {
    out->grad_inf = st[0];
    if (!svm && I.loss == kNnls) {
      // objective 1/2 ||Ax - b||^2 (no dual is formed: 0); rel_gap is the
      // KKT residual over ||A^T b||_inf
      out->primal = 0.5 * st[3];
      out->dual = 0.0;
      out->nnz_x = st[2];
      out->rel_gap = I.scale > 0 ? st[0] / I.scale : st[0];
    } else if (svm) {
      // primal lambda/2 ||w||^2 + mean hinge, dual -lambda/2 ||w||^2 +
      // mean alpha (MulticoreSVMDualLoss.h:63-93)
      const double P = 0.5 * lambda * st[3] + st[4] / (double)I.n;
      const double D = -0.5 * lambda * st[3] + st[1] / (double)I.n;
      out->primal = P;
      out->dual = D;
      out->nnz_x = st[2];
      out->rel_gap = P != 0 ? (P - D) / fabs(P) : 0.0;
    } else {
      const double P = 0.5 * st[3] + lambda * st[1];
      const double s = st[0] > lambda ? lambda / st[0] : 1.0;
      const double D = -s * st[4] - 0.5 * s * s * st[3];
      out->primal = P;
      out->dual = D;
      out->nnz_x = st[2];
      out->rel_gap = P != 0 ? (P - D) / fabs(P) : 0.0;
    }
    if (st[5] > 0.0) out->rel_gap = __builtin_nan("");  // a gradient was not finite
  }
  __syncthreads();  // gref, dr and *out are settled for the steps
}

// r64 in LDS when its m doubles fit (then dr fits as well); one body per
// case, so that no access is issued for the case not taken
template <typename T, bool LDSR, bool SVM>
__device__ __forceinline__ void batch_refresh(const BatchInstance *Ip, double lambda, T *lds,
                                           double *sh, int lds_doubles, BatchResult *out,
                                           bool dense) {
  if (LDSR && Ip->m <= lds_doubles)
    batch_refresh_impl<T, LDSR, LDSR, SVM>(Ip, lambda, lds, sh, out, dense);
  else
    batch_refresh_impl<T, LDSR, false, SVM>(Ip, lambda, lds, sh, out, dense);
}

// ------------------------------------------------------------- working sets
// The asynchronous Lasso instances of a screened batch solve each lambda on
// a working set W of coordinates, re-formed at every full FP64 refresh from
// g = A^T r of the current x:
//   W = {j : x_j != 0} u {j : |g_j| >= thr} minus the coordinates gap-safe
//   screening proves zero,
// with thr = min(lambda, 2 lambda - lambda_prev), the sequential strong rule
// of Tibshirani et al. (JRSS-B 2012) for the step from the previous lambda
// of the path.  Since thr <= lambda, every coordinate that violates the
// optimality conditions at x (|g_j| > lambda with x_j = 0) is in W, so when
// the workers have solved the problem restricted to W, the full refresh
// either confirms the whole problem's gap or brings the violators in.
// Gap-safe screening (Fercoq, Gramfort and Salmon, "Mind the duality gap:
// safer rules for the Lasso", ICML 2015): the dual D(u) = b^T u - ||u||^2/2
// (u = -s r, s = min(1, lambda/||A^T r||_inf), the dual point of every
// evaluation here) is 1-strongly concave, so the dual optimum u* lies within
// sqrt(2 G) of u for the duality gap G = P(x) - D(u), and |<a_j, u*>| <=
// s |g_j| + ||a_j|| sqrt(2 G); below lambda, x*_j = 0 (a nonzero x*_j needs
// |<a_j, u*>| = lambda).  The workers draw uniformly from W, with the ESO
// beta of W: 1 + (omega_W - 1)(tau_W - 1)/(|W| - 1), omega_W the most
// entries of W in one row, tau_W = min(tau, |W|) workers.
//
// W is capped: when the support and the surviving candidates exceed the cap
// (twice the support, at least kWsMin), only the candidates with the largest
// |g_j| enter, the prioritised working sets of Celer (Massias, Gramfort and
// Salmon, ICML 2018).  At small lambda the strong rule keeps a large share
// of the coordinates whose |g_j| sits just under lambda, while the solution
// needs few of them; the cap keeps each epoch of W near the support's size.
// The ranking is a histogram of |g_j| / lambda over [thr / lambda, 1) in
// kWsBins bins, with every violator (|g_j| >= lambda) in a bin of its own at
// the top: the candidates of the highest bins that fit the cap enter, and
// the violators always do.
constexpr int kWsBins = 128;
constexpr int32_t kWsWorkerSpan = 8;  // coordinates of W per asynchronous worker, at least
constexpr int32_t kWsMin = 256;
static_assert(kWsBins + kBatchWaves + 1 <= 2 * kRefreshStats * kBatchWaves,
              "the histogram and the waves' counts fit the refresh scratch");

// forms W from the full refresh's g_j (in I.cs) and gap: copies the records
// of W into I.ws (coordinate order) and their coordinates into I.wsi;
// returns |W|.  cap: the largest |W| the candidates may fill (the support
// always enters whole).  shi: at least kWsBins + kBatchWaves + 1 ints of
// scratch.  The records are then put in the order of the keyed permutation
// of [0, |W|) for counter key + 2^63 (the Feistel permutation of the
// synchronous steps, apart from their counters), through I.anx as scratch.
__device__ __forceinline__ int32_t batch_working_set(const BatchInstance &I, double lambda,
                                                     double thr, double gap_abs, double ginf,
                                                     int64_t cap, uint64_t key, int *shi) {
  const int tid = threadIdx.x, wv = tid >> 6, ln = tid & 63;
  int *hist = shi, *wc = shi + kWsBins;
  const double s = ginf > lambda ? lambda / ginf : 1.0;
  // a relative margin of 1e-9 on the radius keeps rounding from screening a
  // coordinate on the boundary
  const double rad = sqrt(fmax(2.0 * gap_abs, 0.0)) * (1.0 + 1e-9) + 1e-15 * lambda;
  const double lo = thr / lambda, span = 1.0 - lo;
  // -1: not kept; kWsBins: support; else the candidate's bin
  auto classify = [&](const CoordRec &r) -> int {
    if (r.x != 0.0) return kWsBins;
    const double nrm = r.L > 0.0 ? sqrt(1.0 / (I.beta * r.L)) : 0.0;
    const double ag = fabs(r.gref);
    if (!(r.L > 0.0 && ag >= thr) || s * ag + nrm * rad < lambda) return -1;
    const double q = ag / lambda;
    if (q >= 1.0 || !(span > 0.0)) return kWsBins - 1;
    return min(kWsBins - 2, max(0, (int)((q - lo) / span * (double)(kWsBins - 1))));
  };
  auto load = [&](int32_t j) {
    CoordRec r = gld(I.cs + j);
    r.x = gld_agent(&I.cs[j].x);  // workers add to x atomically
    return r;
  };
  // pass 1: the support's size and the candidates' histogram
  for (int k = tid; k < kWsBins + kBatchWaves + 1; k += kBatchBlock) shi[k] = 0;
  __syncthreads();
  int nsup = 0;
  for (int32_t j = tid; j < I.n; j += kBatchBlock) {
    const int c = classify(load(j));
    if (c == kWsBins)
      ++nsup;
    else if (c >= 0)
      atomicAdd(hist + c, 1);
  }
  atomicAdd(wc, nsup);
  __syncthreads();
  // the lowest bin that still fits: candidates of bins >= cut enter
  if (tid == 0) {
    const int64_t room = cap - (int64_t)wc[0];
    int cut = kWsBins - 1;  // violators always enter
    int64_t acc = hist[kWsBins - 1];
    while (cut > 0 && acc + hist[cut - 1] <= room) acc += hist[--cut];
    wc[1] = cut;
  }
  __syncthreads();
  const int cut = wc[1];
  __syncthreads();  // wc is reused as the waves' counts below
  // pass 2: compaction in coordinate order
  int32_t base = 0;
  for (int32_t j0 = 0; j0 < I.n; j0 += kBatchBlock) {
    const int32_t j = j0 + tid;
    bool keep = false;
    CoordRec r{};
    if (j < I.n) {
      r = load(j);
      keep = classify(r) >= cut;
    }
    const uint64_t bal = __ballot(keep);
    const int within = __popcll(bal & ((1ull << ln) - 1ull));
    if (ln == 0) wc[wv] = __popcll(bal);
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kBatchWaves; ++k) {
      const int c = wc[k];
      off += k < wv ? c : 0;
      tot += c;
    }
    if (keep) {
      I.ws[base + off + within] = r;
      *gp(I.wsi + base + off + within) = j;
    }
    base += tot;
    __syncthreads();  // wc is rewritten by the next chunk
  }
  // pass 3: slot u takes the record of slot perm(u) (through scratch: 32
  // bytes of record and 4 of index per slot, within the kAaSnapshots x n
  // doubles of I.anx)
  if (base > 1) {
    const Perm wp = Perm::make((uint32_t)base);
    const StepKeys wk = step_keys(I.seed, key | (1ull << 63));
    CoordRec *trec = reinterpret_cast<CoordRec *>(I.anx);
    int32_t *tidx = reinterpret_cast<int32_t *>(I.anx + 4 * (int64_t)I.n);
    for (int32_t u = tid; u < base; u += kBatchBlock) {
      const uint32_t v = wp((uint32_t)u, wk);
      trec[u] = I.ws[v];
      *gp(tidx + u) = gld(I.wsi + v);
    }
    __syncthreads();
    for (int32_t u = tid; u < base; u += kBatchBlock) {
      I.ws[u] = trec[u];
      *gp(I.wsi + u) = gld(tidx + u);
    }
    __syncthreads();
  }
  return base;
}

// omega_W: the most entries one row holds among the working set's columns,
// counted into the correction residual's storage (LDS rows below the split,
// the instance's global rows above it) and zeroed again for the updates.
template <typename T, bool LDSR>
__device__ __forceinline__ int32_t batch_ws_omega(const BatchInstance &I, int32_t nw, T *lds,
                                                  int *shi) {
  const int tid = threadIdx.x;
  const int32_t split = LDSR ? I.m : I.lds_rows;
  int32_t *cl = reinterpret_cast<int32_t *>(lds);
  int32_t *cg = reinterpret_cast<int32_t *>(I.drg);
  for (int32_t k = tid; k < I.m; k += kBatchBlock) {
    if (k < split)
      cl[k] = 0;
    else
      *gp(cg + (k - split)) = 0;
  }
  __syncthreads();
  constexpr int G = 16;  // lanes per column
  const int grp = tid / G, lane = tid & (G - 1);
  for (int32_t u = grp; u < nw; u += kBatchBlock / G) {
    const int32_t p0 = gld(&I.ws[u].p0), len = gld(&I.ws[u].len);
    for (int32_t q = lane; q < len; q += G) {
      const int32_t row = gld(I.rent + p0 + q).row;
      if (row < split)
        atomicAdd(cl + row, 1);
      else
        __hip_atomic_fetch_add(gp(cg + (row - split)), 1, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  __syncthreads();
  int mx = 1;  // counted from 1, as the reference's omega (MultiCoreSolver.cpp:24)
  for (int32_t k = tid; k < I.m; k += kBatchBlock)
    mx = max(mx, k < split ? cl[k] : gld_agent(cg + (k - split)));
  mx = group_max<64>(mx);
  if ((tid & 63) == 0) shi[tid >> 6] = mx;
  __syncthreads();
  int om = shi[0];
  for (int k = 1; k < kBatchWaves; ++k) om = max(om, shi[k]);
  __syncthreads();
  dr_view<T, LDSR>(I, lds).zero(I.m);
  __syncthreads();
  return om;
}

// The spectral ESO constant of a working set, lambda_max of
// D^-1/2 A_W^T A_W D^-1/2 (D the squared column norms), by kWsPowerSteps
// steps of the power method (the rule of mSDCA.cpp:188-265, on the columns
// of W): y = sum_u a_u v_u / ||a_u|| into the correction residual's storage,
// then v'_u = <a_u, y> / ||a_u||, with v and v' in I.anx.  Returns the last
// Rayleigh estimate; zeroes the storage again.  Column norms come from the
// records (L_u = 1 / (beta ||a_u||^2) at formation).
constexpr int kWsPowerSteps = 12;

template <typename T, bool LDSR>
__device__ __forceinline__ double batch_ws_lmax(const BatchInstance &I, int32_t nw, T *lds,
                                                double *sh) {
  const int tid = threadIdx.x;
  const int32_t split = LDSR ? I.m : I.lds_rows;
  T *yl = lds;
  T *yg = static_cast<T *>(I.drg);
  double *v = I.anx, *vn = I.anx + I.n;
  auto inv_norm = [&](int32_t u) {
    const double L = gld(&I.ws[u].L);
    return L > 0.0 ? sqrt(I.beta * L) : 0.0;
  };
  // a start with no symmetry to miss the top eigenvector by: 1 + a hash of u
  for (int32_t u = tid; u < nw; u += kBatchBlock)
    *gp(v + u) = 1.0 + (double)(mix32((uint32_t)u, 0x9E3779B9u) >> 8) * 0x1p-24;
  __syncthreads();
  double est = 1.0;
  constexpr int G = 16;  // lanes per column
  const int grp = tid / G, lane = tid & (G - 1);
  for (int it = 0; it < kWsPowerSteps; ++it) {
    dr_view<T, LDSR>(I, lds).zero(I.m);
    __syncthreads();
    for (int32_t u = grp; u < nw; u += kBatchBlock / G) {
      const int32_t p0 = gld(&I.ws[u].p0), len = gld(&I.ws[u].len);
      const double c = gld_agent(v + u) * inv_norm(u);
      for (int32_t q = lane; q < len; q += G) {
        const RefEntry e = gld(I.rent + p0 + q);
        const T t = (T)(c * e.val);
        if (e.row < split)
          atomicAdd(yl + e.row, t);
        else
          __hip_atomic_fetch_add(gp(yg + (e.row - split)), t, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    __syncthreads();
    double sq = 0.0, vv = 0.0;
    for (int32_t u = grp; u < nw; u += kBatchBlock / G) {
      const int32_t p0 = gld(&I.ws[u].p0), len = gld(&I.ws[u].len);
      double acc = 0.0;
      for (int32_t q = lane; q < len; q += G) {
        const RefEntry e = gld(I.rent + p0 + q);
        acc += e.val * (double)(e.row < split ? yl[e.row] : gld_agent(yg + (e.row - split)));
      }
      acc = group_sum<G>(acc) * inv_norm(u);
      if (lane == 0) {
        const double x = gld_agent(v + u);
        *gp(vn + u) = acc;
        sq += acc * acc;
        vv += x * x;
      }
    }
    sq = group_sum<64>(sq);
    vv = group_sum<64>(vv);
    if ((tid & 63) == 0) {
      sh[tid >> 6] = sq;
      sh[kBatchWaves + (tid >> 6)] = vv;
    }
    __syncthreads();
    double tsq = 0.0, tvv = 0.0;
    for (int k = 0; k < kBatchWaves; ++k) {
      tsq += sh[k];
      tvv += sh[kBatchWaves + k];
    }
    __syncthreads();
    if (!(tsq > 0.0) || !(tvv > 0.0)) break;
    est = sqrt(tsq / tvv);  // ||A v|| / ||v|| for the Gram operator, v' = G v
    const double s = 1.0 / sqrt(tsq);
    for (int32_t u = tid; u < nw; u += kBatchBlock) *gp(v + u) = gld_agent(vn + u) * s;
    __syncthreads();
  }
  dr_view<T, LDSR>(I, lds).zero(I.m);
  __syncthreads();
  return est;
}

// FP64 refresh restricted to the working set: r64 = A x - b from W's nonzero
// x (every other x_j is zero), g_u = <a_j, r64> into I.ws[u].gref, and the
// Lasso primal, dual (dual point scaled by max_u |g_u|, the dual of the
// screened problem) and relative gap into *out.  Zeroes dr.
template <typename T, bool LDSR, bool RL>
__device__ __forceinline__ void batch_refresh_ws(const BatchInstance &I, int32_t nw,
                                                 double lambda, T *lds, double *sh,
                                                 BatchResult *out) {
  const int tid = threadIdx.x;
  double *rs = reinterpret_cast<double *>(lds);
  double *rg = I.r64;
  auto rget = [&](int32_t row) -> double { return RL ? rs[row] : gld_agent(rg + row); };
  auto radd = [&](int32_t row, double v) {
    if (RL)
      atomicAdd(rs + row, v);
    else
      gadd_block(rg + row, v);
  };
  for (int32_t k = tid; k < I.m; k += kBatchBlock) {
    const double v = -gld(I.b + k);
    if (RL)
      rs[k] = v;
    else
      *gp(rg + k) = v;
  }
  __syncthreads();
  constexpr int G = 16;  // lanes per column
  const int grp = tid / G, lane = tid & (G - 1);
  double x1 = 0, nz = 0, ginf = 0, bad = 0;
  for (int32_t u = grp; u < nw; u += kBatchBlock / G) {
    const double xv = gld_agent(&I.ws[u].x);
    if (xv == 0.0) continue;  // uniform in the group
    if (lane == 0) {
      x1 += fabs(xv);
      nz += 1.0;
    }
    const int32_t p0 = gld(&I.ws[u].p0), len = gld(&I.ws[u].len);
    for (int32_t q0 = lane; q0 < len; q0 += 4 * G) {
      RefEntry e[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) e[k] = gld(I.rent + p0 + min(q0 + k * G, len - 1));
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (q0 + k * G < len) radd(e[k].row, e[k].val * xv);
    }
  }
  __syncthreads();
  for (int32_t u = grp; u < nw; u += kBatchBlock / G) {
    const int32_t p0 = gld(&I.ws[u].p0), len = gld(&I.ws[u].len);
    double g = 0.0;
    for (int32_t q0 = lane; q0 < len; q0 += 4 * G) {
      RefEntry e[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) e[k] = gld(I.rent + p0 + min(q0 + k * G, len - 1));
      double rv[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) rv[k] = rget(e[k].row);
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (q0 + k * G < len) g += e[k].val * rv[k];
    }
    g = group_sum<G>(g);
    if (lane == 0) {
      *gp(&I.ws[u].gref) = g;
      ginf = fmax(ginf, fabs(g));
      bad += isfinite(g) ? 0.0 : 1.0;
    }
  }
  double r2 = 0, br = 0;
  for (int32_t k = tid; k < I.m; k += kBatchBlock) {
    const double rk = rget(k);
    r2 += rk * rk;
    br += gld(I.b + k) * rk;
  }
  double st[kRefreshStats] = {ginf, x1, nz, r2, br, bad};
  block_stats(st, sh);
  dr_view<T, LDSR>(I, lds).zero(I.m);
  if (tid == 0) {
    const double P = 0.5 * st[3] + lambda * st[1];
    const double sc = st[0] > lambda ? lambda / st[0] : 1.0;
    const double D = -sc * st[4] - 0.5 * sc * sc * st[3];
    out->grad_inf = st[0];
    out->primal = P;
    out->dual = D;
    out->nnz_x = st[2];
    out->rel_gap = st[5] > 0.0 ? __builtin_nan("") : (P != 0 ? (P - D) / fabs(P) : 0.0);
  }
  __syncthreads();
}

// x of the working set back into the instance's records
__device__ __forceinline__ void batch_ws_writeback(const BatchInstance &I, int32_t nw) {
  for (int32_t u = threadIdx.x; u < nw; u += kBatchBlock)
    *gp(&I.cs[gld(I.wsi + u)].x) = gld_agent(&I.ws[u].x);
  __syncthreads();
}

// ------------------------------------------------- Anderson extrapolation
// The workers on a working set repeat one cyclic order (batch_async_updates),
// so the map F from x at one snapshot to x at the next (refresh_period epochs
// of W) is the same throughout (up to the interleaving of concurrent
// updates).  Near the
// solution, where the signs of the support no longer change, F is affine,
// and Anderson acceleration (Anderson, J. ACM 1965; for coordinate descent,
// Bertrand and Massias, AISTATS 2021) extrapolates from the last
// kAaSnapshots iterates X_0 .. X_K: c minimises || sum_i c_i (X_{i+1} - X_i) ||
// subject to sum_i c_i = 1 (the regularised normal equations of the K x K
// Gram matrix of the differences), and the candidate sum_i c_i X_{i+1} is
// kept only when its FP64 objective is below that of X_K.  On the
// near-interpolating pairs of a lambda path (support close to m, where a
// coordinate epoch contracts the error very little) this cuts the epochs a
// lambda needs several times over.
constexpr int kAaDepth = 5;                 // differences per extrapolation
constexpr int kAaSnapshots = kAaDepth + 1;  // iterates it uses
constexpr int kWsRefreshPeriods = 4;        // snapshot periods per restricted refresh

// x of the working set's records into snapshot slot h (stride: the
// instance's n, the slots' size)
__device__ __forceinline__ void batch_aa_snapshot(const BatchInstance &J, int32_t nw,
                                                  int64_t stride, int h) {
  double *X = J.anx + (int64_t)h * stride;
  for (int32_t u = threadIdx.x; u < nw; u += kBatchBlock) *gp(X + u) = gld_agent(&J.cs[u].x);
  __syncthreads();
}

// snapshot slot h back into the working set's records
__device__ __forceinline__ void batch_aa_restore(const BatchInstance &J, int32_t nw,
                                                 int64_t stride, int h) {
  const double *X = J.anx + (int64_t)h * stride;
  // (snapshots are read at agent scope: a slot's older contents may sit in
  // this CU's vector cache)
  for (int32_t u = threadIdx.x; u < nw; u += kBatchBlock) *gp(&J.cs[u].x) = gld_agent(X + u);
  __syncthreads();
}

// The extrapolation from the snapshots in slots h0, h0 + 1, ... (mod
// kAaSnapshots, oldest first) into the working set's records.  Returns
// false, records untouched, when the system is singular or not finite.
// sh: at least kAaDepth (kAaDepth + 3) / 2 + 1 doubles of scratch.
__device__ __forceinline__ bool batch_aa_extrapolate(const BatchInstance &J, int32_t nw,
                                                     int64_t stride, int h0, double *sh) {
  constexpr int K = kAaDepth, NP = K * (K + 1) / 2;
  const int tid = threadIdx.x;
  __syncthreads();
  if (tid < NP + K + 1) sh[tid] = 0.0;
  __syncthreads();
  double part[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) part[p] = 0.0;
  for (int32_t u = tid; u < nw; u += kBatchBlock) {
    double xs[K + 1];
#pragma unroll
    for (int i = 0; i <= K; ++i)
      xs[i] = gld_agent(J.anx + (int64_t)((h0 + i) % kAaSnapshots) * stride + u);
    double dd[K];
#pragma unroll
    for (int i = 0; i < K; ++i) dd[i] = xs[i + 1] - xs[i];
    int p = 0;
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
      for (int b = a; b < K; ++b) part[p++] += dd[a] * dd[b];
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const double v = group_sum<64>(part[p]);
    if ((tid & 63) == 0) atomicAdd(sh + p, v);  // ds_add_f64
  }
  __syncthreads();
  if (tid == 0) {
    double g[K][K + 1];  // the Gram matrix and the right-hand side of ones
    int p = 0;
    double tr = 0.0;
    for (int a = 0; a < K; ++a)
      for (int b = a; b < K; ++b) {
        g[a][b] = g[b][a] = sh[p++];
        if (a == b) tr += g[a][a];
      }
    for (int a = 0; a < K; ++a) {
      g[a][a] += 1e-10 * tr / K;
      g[a][K] = 1.0;
    }
    bool ok = tr > 0.0;
    for (int c = 0; c < K && ok; ++c) {  // Gaussian elimination, partial pivoting
      int piv = c;
      for (int r = c + 1; r < K; ++r)
        if (fabs(g[r][c]) > fabs(g[piv][c])) piv = r;
      if (!(fabs(g[piv][c]) > 0.0)) {
        ok = false;
        break;
      }
      if (piv != c)
        for (int k = c; k <= K; ++k) {
          const double t = g[c][k];
          g[c][k] = g[piv][k];
          g[piv][k] = t;
        }
      for (int r = c + 1; r < K; ++r) {
        const double f = g[r][c] / g[c][c];
        for (int k = c; k <= K; ++k) g[r][k] -= f * g[c][k];
      }
    }
    double z[K], zs = 0.0;
    for (int r = K - 1; r >= 0 && ok; --r) {
      double v = g[r][K];
      for (int k = r + 1; k < K; ++k) v -= g[r][k] * z[k];
      z[r] = v / g[r][r];
      zs += z[r];
    }
    ok = ok && fabs(zs) > 0.0 && isfinite(zs);
    for (int i = 0; i < K && ok; ++i) {
      sh[NP + i] = z[i] / zs;
      ok = isfinite(sh[NP + i]);
    }
    sh[NP + K] = ok ? 1.0 : 0.0;
  }
  __syncthreads();
  const bool ok = sh[NP + K] != 0.0;
  if (ok) {
    double c[K];
#pragma unroll
    for (int i = 0; i < K; ++i) c[i] = sh[NP + i];
    for (int32_t u = tid; u < nw; u += kBatchBlock) {
      double v = 0.0;
#pragma unroll
      for (int i = 0; i < K; ++i)
        v += c[i] * gld_agent(J.anx + (int64_t)((h0 + i + 1) % kAaSnapshots) * stride + u);
      *gp(&J.cs[u].x) = v;
    }
  }
  __syncthreads();
  return ok;
}

// ------------------------------------------------------ support Newton steps
// Near interpolation (support close to m, or above it on the way) the
// least-squares problem on the support S is ill-conditioned: a coordinate epoch
// contracts the error by about 1 - 1/kappa of H = A_S^T A_S, so a pair needs
// tens of thousands of epochs, where conjugate gradients need about
// sqrt(kappa) products with H.  At a restricted refresh (r64 = A x - b exact,
// g_u = <a_u, r64> in the records), with s the signs of x on S, the Lasso
// objective on the orthant of s is the quadratic
//   1/2 ||A_S x_S - b||^2 + lambda s^T x_S,
// whose Newton direction d solves H d = -(g_S + lambda s).  When |S| > m (the
// iterate's support passes m on the way to a solution with at most m nonzeros)
// H is singular and the right-hand side has a part in its null space, along
// which CG steps grow without bound; the damped system (H + mu I) d = -(g_S +
// lambda s), mu = kNewtonDamping times the mean ||a_j||^2 of the instance's
// columns (the Levenberg-Marquardt damping), bounds them.  At most `iters` CG
// iterations from d = 0 approximate d.  The projected search then takes the
// first t in 1, 1/2, ..., 1/128 for which x' = P_s(x_S + t d) -- every
// coordinate that would change sign set to 0, the orthant projection of OWL-QN
// (Andrew and Gao, ICML 2007) -- has an FP64 objective below the refresh's
// primal p0. The workers then continue from x' (coordinates enter and leave
// S).  A step is only taken where the workers stall -- W's gap fell by less
// than half (kNewtonStall) since the previous restricted refresh -- so that
// well-conditioned pairs, which coordinate descent finishes in a few epochs,
// are left to it.  FP64 throughout: q = A_S p lives in the storage of r64 (LDS
// when its m doubles fit, the instance's global r64 otherwise), and p beside
// it in LDS when both fit; S is compacted into the seven slots of I.anx after
// the snapshots (x_S, d, the CG residual, p, 16 bytes of column per
// coordinate, and h = A_S^T q), and the entries of its columns are copied,
// contiguous and tagged with the compacted column, into I.sent, so that the
// products with A_S stream them: four entries in flight per thread for A_S p,
// segmented sums over wave-contiguous ranges for A_S^T q.  Returns whether x'
// was written to the records; the correction residual dr is zero on return
// either way.
constexpr int kWsScratchSlots = kAaSnapshots + 7;  // doubles of I.anx per coordinate
constexpr int kNewtonSearch = 8;                   // step lengths the search tries
constexpr double kNewtonDamping = 1e-4;            // mu over the mean ||a_j||^2
constexpr double kNewtonStall = 0.5;  // W's gap ratio between refreshes that calls a step

// the block's sum of one double per thread, in a fixed order, returned to
// every thread (so that all of them take the same branches on it)
__device__ __forceinline__ double block_sum_all(double v, double *sh) {
  v = group_sum<64>(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < kBatchWaves; ++k) t += sh[k];
  __syncthreads();
  return t;
}

// a support column of a Newton step: its slot in W, its entries in the
// instance's entries, and where its copy starts in the support's entries
struct alignas(16) NewtonCol {
  int32_t u, p0, len, at;
};

// exclusive prefix sum over the block of one int per thread (thread order);
// *total receives the block's sum.  wsum: kBatchWaves ints of scratch.
__device__ __forceinline__ int block_excl_scan(int v, int *wsum, int *total) {
  const int ln = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int s = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(s, o, 64);
    if (ln >= o) s += t;
  }
  if (ln == 63) wsum[wv] = s;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kBatchWaves; ++k) {
    const int c = wsum[k];
    off += k < wv ? c : 0;
    tot += c;
  }
  __syncthreads();
  *total = tot;
  return off + s - v;
}

// lds_doubles: the doubles the dynamic LDS holds
template <typename T, bool LDSR, bool RL>
__device__ __forceinline__ bool batch_ws_newton(const BatchInstance &J, int32_t nw, int64_t stride,
                                                double lambda, double p0, int iters, T *lds,
                                                int lds_doubles, double *sh) {
  const int tid = threadIdx.x, wv = tid >> 6, ln = tid & 63;
  double *const q = RL ? reinterpret_cast<double *>(lds) : J.r64;
  double *const xs = J.anx + (int64_t)kAaSnapshots * stride;  // x on S, compacted
  double *const dv = xs + stride;                             // d
  double *const rv = dv + stride;                             // CG residual
  double *const pv = rv + stride;                             // CG direction
  NewtonCol *const cols = reinterpret_cast<NewtonCol *>(pv + stride);
  double *const hv = pv + 3 * stride;  // h = A_S^T q
  RefEntry *const se = J.sent;         // the support's entries, column k's with col = k
  auto qget = [&](int32_t k) -> double { return RL ? q[k] : gld_agent(q + k); };
  auto qset = [&](int32_t k, double v) {
    if (RL)
      q[k] = v;
    else
      *gp(q + k) = v;
  };
  auto qadd = [&](int32_t k, double v) {
    if (RL)
      atomicAdd(q + k, v);  // ds_add_f64
    else
      gadd_block(q + k, v);
  };
  int *wc = reinterpret_cast<int *>(sh);
  // S compacted in W order (with the offsets of the columns' entry copies),
  // res = p = -(g_S + lambda s), d = 0
  int32_t ns = 0, ne = 0;
  double rr = 0.0;
  for (int32_t u0 = 0; u0 < nw; u0 += kBatchBlock) {
    const int32_t u = u0 + tid;
    const double x = u < nw ? gld_agent(&J.cs[u].x) : 0.0;  // workers add to x atomically
    const bool keep = x != 0.0;
    CoordRec r{};
    if (keep) r = gld(J.cs + u);
    int cnt = 0, tot = 0;
    const int k = ns + block_excl_scan(keep ? 1 : 0, wc, &cnt);
    const int at = ne + block_excl_scan(keep ? r.len : 0, wc, &tot);
    if (keep) {
      const double g = -(r.gref + lambda * sign_of(x));
      *gp(xs + k) = x;
      *gp(dv + k) = 0.0;
      *gp(rv + k) = g;
      *gp(pv + k) = g;
      cols[k] = NewtonCol{u, r.p0, r.len, at};
      rr += g * g;
    }
    ns += cnt;
    ne += tot;
  }
  __syncthreads();
  // p in LDS after q when both fit
  const bool pl = RL && (int64_t)J.m + ns <= lds_doubles;
  double *const pls = reinterpret_cast<double *>(lds) + J.m;
  auto pget = [&](int32_t k) -> double { return pl ? pls[k] : gld_agent(pv + k); };
  if (pl)
    for (int32_t k = tid; k < ns; k += kBatchBlock) pls[k] = gld_agent(pv + k);
  // the copy of the support's entries (16 lanes per column)
  {
    constexpr int G = 16;
    const int grp = tid / G, lane = tid & (G - 1);
    for (int32_t k = grp; k < ns; k += kBatchBlock / G) {
      const NewtonCol c = gld(cols + k);
      for (int32_t i = lane; i < c.len; i += G) {
        RefEntry e = gld(J.rent + c.p0 + i);
        e.col = k;
        se[c.at + i] = e;
      }
    }
  }
  rr = block_sum_all(rr, sh);  // its barriers also publish the stores above
  const double mu = kNewtonDamping * J.colsq;
  // q += c_k a_k over the support's entries, c_k = coef(k) (0: skipped),
  // four entries per thread in flight; returns the thread's share of
  // sum_k c_k^2 ||a_k||^2
  auto scatter = [&](auto coef) {
    double sq = 0.0;
    RefEntry nx[4];  // the next round's entries, in flight while this round is applied
    auto load = [&](int32_t i0) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int32_t i = i0 + j * kBatchBlock;
        nx[j] = i < ne ? gld(se + i) : RefEntry{0, 0, 0.0};
      }
    };
    load(tid);
    for (int32_t i0 = tid; i0 < ne; i0 += 4 * kBatchBlock) {
      RefEntry e[4];
      double c[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) e[j] = nx[j];
      load(i0 + 4 * kBatchBlock);
#pragma unroll
      for (int j = 0; j < 4; ++j) c[j] = i0 + j * kBatchBlock < ne ? coef(e[j].col) : 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (c[j] != 0.0) {
          const double v = e[j].val * c[j];
          qadd(e[j].row, v);
          sq += v * v;
        }
    }
    return sq;
  };
  // h_k = <a_k, q> for every support column, emit(k, h_k) once per column:
  // wave w streams the support's entries [ne w/16, ne (w+1)/16), 64 per step,
  // sums runs of equal column by segmented scans and emits each run it
  // completes; the wave's first run (it may have begun in the previous
  // range) and the run open at its end are joined by thread 0 afterwards
  auto seg_gather = [&](auto emit) {
    double *slot = sh + 4 * wv;  // head column, head sum, tail column, tail sum
    __syncthreads();             // sh may still be read by a block sum
    if (ln == 0) slot[0] = -1.0;
    const int32_t e_lo = (int32_t)((int64_t)ne * wv / kBatchWaves);
    const int32_t e_hi = (int32_t)((int64_t)ne * (wv + 1) / kBatchWaves);
    int first_col = -1, open_col = -1;
    double open_sum = 0.0;
    auto put = [&](int c, double v) {
      if (c == first_col) {
        slot[0] = (double)c;
        slot[1] = v;
      } else {
        emit(c, v);
      }
    };
    constexpr int UG = 4;  // steps of 64 entries per round
    RefEntry nxt[UG];
    auto load = [&](int32_t base) {
#pragma unroll
      for (int j = 0; j < UG; ++j) nxt[j] = gld(se + min(base + j * 64 + ln, e_hi - 1));
    };
    if (e_lo < e_hi) {
      load(e_lo);
      first_col = __builtin_amdgcn_readfirstlane(nxt[0].col);
      for (int32_t base = e_lo; base < e_hi; base += 64 * UG) {
        RefEntry cur[UG];
#pragma unroll
        for (int j = 0; j < UG; ++j) cur[j] = nxt[j];
        load(base + 64 * UG);
        double qv[UG];
#pragma unroll
        for (int j = 0; j < UG; ++j) qv[j] = qget(cur[j].row);
#pragma unroll
        for (int j = 0; j < UG; ++j) {
          const int32_t cb = base + j * 64;
          if (cb >= e_hi) break;  // wave-uniform
          const bool in = cb + ln < e_hi;
          const int key = in ? cur[j].col : 0x7fffffff;
          const SegSum r = seg_scan_sum(in ? cur[j].val * qv[j] : 0.0, key);
          double v = r.sum;
          const int k0 = __builtin_amdgcn_readfirstlane(key);
          if (open_col >= 0) {
            if (k0 == open_col) {
              if (key == open_col) v += open_sum;
            } else if (ln == 0) {
              put(open_col, open_sum);  // the open run ended with the last step
            }
          }
          const int lv = min(63, e_hi - 1 - cb);  // last lane in range
          const uint64_t heads = __ballot(r.head);
          if (ln < lv && ((heads >> (ln + 1)) & 1)) put(key, v);
          open_col = __builtin_amdgcn_readlane(key, lv);
          const long long vb = __double_as_longlong(v);
          const int lo = __builtin_amdgcn_readlane((int)vb, lv);
          const int hi = __builtin_amdgcn_readlane((int)(vb >> 32), lv);
          open_sum = __longlong_as_double(
              (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
        }
      }
    }
    if (ln == 0) {
      slot[2] = (double)open_col;
      slot[3] = open_sum;
    }
    __syncthreads();
    if (tid == 0) {  // join the partial runs, in entry order
      int c = -1;
      double v = 0.0;
      for (int k = 0; k < 2 * kBatchWaves; ++k) {
        const int kc = (int)sh[2 * k];
        if (kc < 0) continue;
        if (kc == c) {
          v += sh[2 * k + 1];
          continue;
        }
        if (c >= 0) emit(c, v);
        c = kc;
        v = sh[2 * k + 1];
      }
      if (c >= 0) emit(c, v);
    }
    __syncthreads();
  };
  double p2 = rr;  // ||p||^2
  const double rr0 = rr;
  int it = 0;
  while (it < iters && rr > 1e-24 * rr0) {
    for (int32_t k = tid; k < J.m; k += kBatchBlock) qset(k, 0.0);
    __syncthreads();
    // pp = sum_k p_k^2 ||a_k||^2, the curvature's scale
    double pp = scatter(pget);
    __syncthreads();
    double qq = 0.0;
    for (int32_t k = tid; k < J.m; k += kBatchBlock) {
      const double v = qget(k);
      qq += v * v;
    }
    qq = block_sum_all(qq, sh);
    pp = block_sum_all(pp, sh);
    const double curv = qq + mu * p2;  // p^T (H + mu I) p
    if (!(curv > 1e-10 * pp)) break;
    const double al = rr / curv;
    // h = A_S^T q; then d += al p, res -= al (h + mu p)
    seg_gather([&](int32_t k, double h) { *gp(hv + k) = h; });
    double rn = 0.0;
    for (int32_t k = tid; k < ns; k += kBatchBlock) {
      const double pk = pget(k);
      *gp(dv + k) = gld_agent(dv + k) + al * pk;
      const double r = gld_agent(rv + k) - al * (gld_agent(hv + k) + mu * pk);
      *gp(rv + k) = r;
      rn += r * r;
    }
    rn = block_sum_all(rn, sh);
    const double be = rn / rr;
    double pn2 = 0.0;
    for (int32_t k = tid; k < ns; k += kBatchBlock) {
      const double pk = gld_agent(rv + k) + be * pget(k);
      if (pl)
        pls[k] = pk;
      else
        *gp(pv + k) = pk;
      pn2 += pk * pk;
    }
    p2 = block_sum_all(pn2, sh);  // (its barriers publish p)
    rr = rn;
    ++it;
  }
  // the projected point of column k for step t
  auto moved = [&](int32_t k, double t) {
    const double x = gld_agent(xs + k);
    const double xn = x + t * gld_agent(dv + k);
    return sign_of(xn) == sign_of(x) ? xn : 0.0;
  };
  bool acc = false;
  for (int st = 0; it > 0 && st < kNewtonSearch && !acc; ++st) {
    const double t = ldexp(1.0, -st);
    // q = A x' - b
    for (int32_t r = tid; r < J.m; r += kBatchBlock) qset(r, -gld(J.b + r));
    __syncthreads();
    (void)scatter([&](int32_t k) { return moved(k, t); });
    double x1 = 0.0;
    for (int32_t k = tid; k < ns; k += kBatchBlock) x1 += fabs(moved(k, t));
    __syncthreads();
    double r2 = 0.0;
    for (int32_t r = tid; r < J.m; r += kBatchBlock) {
      const double v = qget(r);
      r2 += v * v;
    }
    r2 = block_sum_all(r2, sh);
    x1 = block_sum_all(x1, sh);
    // below p0 by more than the rounding of two differently ordered sums
    if (0.5 * r2 + lambda * x1 < p0 * (1.0 - 1e-14)) {
      for (int32_t k = tid; k < ns; k += kBatchBlock)
        *gp(&J.cs[gld(&cols[k].u)].x) = moved(k, t);
      acc = true;
    }
  }
  __syncthreads();
  dr_view<T, LDSR>(J, lds).zero(J.m);
  __syncthreads();
  return acc;
}

// nsteps synchronous steps of one instance; S slots per G-lane group.  The
// sampled coordinates of a wave's S * (64/G) slots are computed by its first
// lanes in one pass and broadcast; each slot's record and column are loaded
// after the first barrier of the previous step (its x writes are done then).
template <typename T, int G, int S, int R, bool LDSR, bool SVM>
__device__ __forceinline__ void batch_steps(const BatchInstance &I, double lambda, uint64_t step0,
                            uint64_t nsteps, T *lds) {
  constexpr int NG = kBatchBlock / G;  // groups in the block
  constexpr int GW = 64 / G;           // groups in a wave
  static_assert(S * GW <= 64, "one lane per slot of the wave");
  const int lane = threadIdx.x & (G - 1);
  const int grp = threadIdx.x / G;
  const int wlane = threadIdx.x & 63;
  const int wgrp0 = (threadIdx.x >> 6) * GW;  // first group of this wave
  const Perm perm = Perm::make((uint32_t)I.n);
  const Entry<T> *ent = static_cast<const Entry<T> *>(I.ent);
  const DrView<T, LDSR> dr = dr_view<T, LDSR>(I, lds);
  const StepScalars sc = batch_scalars<SVM>(lambda, I.n);
  uint32_t j[S];
  CoordRec rec[S];
  ColCache<T, G, R> col[S];
  auto prefetch = [&](uint64_t step) {
    // lane t < S*GW of the wave samples slot (wgrp0 + t % GW) + (t / GW) * NG
    const StepKeys key = step_keys(I.seed, step);
    uint32_t mine = 0;
    if (wlane < S * GW) {
      const int slot = wgrp0 + (wlane % GW) + (wlane / GW) * NG;
      if (slot < I.tau) mine = perm((uint32_t)slot, key);
    }
#pragma unroll
    for (int k = 0; k < S; ++k) {
      j[k] = (uint32_t)__shfl((int)mine, k * GW + (grp - wgrp0), 64);
      if (grp + k * NG < I.tau) {
        rec[k] = gld(I.cs + j[k]);
      } else {
        rec[k].x = rec[k].L = rec[k].gref = 0.0;
        rec[k].p0 = rec[k].len = 0;
      }
    }
#pragma unroll
    for (int k = 0; k < S; ++k) col[k].load(ent, rec[k].p0, rec[k].len, lane);
  };
  prefetch(step0);
  for (uint64_t s = 0; s < nsteps; ++s) {
    double c[S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
      c[k] = 0.0;
      if (grp + k * NG >= I.tau) continue;  // group-uniform
      T acc = col[k].dot(ent, dr, lane);
      acc = group_sum<G>(acc);
      const double g = rec[k].gref + (double)acc;
      const double d = batch_delta<SVM>(I.loss, g, rec[k].x, rec[k].L, sc);
      c[k] = batch_coef<SVM>(d, rec[k].L, sc);
      if (lane == 0) *gp(&I.cs[j[k]].x) = rec[k].x + d;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < S; ++k) {
      if (c[k] == 0.0) continue;
      const T ct = (T)c[k];
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (lane + r * G < col[k].len) dr.add(col[k].row[r], ct * col[k].val[r]);
      for (int32_t q = lane + R * G; q < col[k].len; q += G) {
        const Entry<T> e = gld(ent + col[k].p0 + q);
        dr.add(e.row, ct * e.val);
      }
    }
    if (s + 1 < nsteps) prefetch(step0 + s + 1);
    __syncthreads();
  }
}

// An instance's correction residual lives in LDS when the host gave it no
// global buffer (drg == nullptr); each block branches on that per instance,
// so one launch and one queue serve both placements.
template <typename T, int G, int S, int R, bool SVM>
__global__ __launch_bounds__(kBatchBlock) void k_batch_lasso(BatchArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  __shared__ double sh[kRefreshStats * kBatchWaves];
  __shared__ BatchResult s_res;
  __shared__ int s_item;
  int next = 0, end = 0;  // the block's current chunk of queue items
  for (;;) {
    if (next >= end) {
      if (threadIdx.x == 0) s_item = (int)atomicAdd(a.queue, (unsigned)a.chunk);
      __syncthreads();
      // readfirstlane marks the block-uniform item as such, so the
      // instance's fields and pointers live in scalar registers
      next = __builtin_amdgcn_readfirstlane(s_item);
      end = min(next + a.chunk, a.count);
      __syncthreads();
      if (next >= a.count) break;
    }
    const int item = next++;
    const int i = a.order[item];
    const BatchInstance *Ip = a.inst + i;
    const BatchInstance I = *Ip;
    const bool in_lds = I.drg == nullptr;
    // the scatter streams every entry once the last refresh of this
    // instance found more than n/64 nonzero x_j (where streaming all entries
    // costs less than the columns' dependent loads, measured)
    bool dense = false;
    auto refresh = [&](double lambda) {
      if (in_lds)
        batch_refresh<T, true, SVM>(Ip, lambda, lds, sh, a.lds_doubles, &s_res, dense);
      else
        batch_refresh<T, false, SVM>(Ip, lambda, lds, sh, a.lds_doubles, &s_res, dense);
      dense = 64.0 * s_res.nnz_x > (double)I.n;
    };
    const uint64_t steps_per_epoch = ((uint64_t)I.n + I.tau - 1) / I.tau;
    uint64_t step = 0;
    for (int l = 0; l < a.path_len; ++l) {
      const uint64_t t0 = wall_clock64();
      const double lambda = a.lambdas[(int64_t)i * a.path_len + l];
      refresh(lambda);
      double gap = s_res.rel_gap;
      bool ok = finite_eval(s_res);
      int epochs = 0;
      while (ok && gap > a.gap_tol && epochs < a.max_epochs) {
        const int ep = min(a.refresh_period, a.max_epochs - epochs);
        const uint64_t ns = steps_per_epoch * (uint64_t)ep;
        if (in_lds)
          batch_steps<T, G, S, R, true, SVM>(I, lambda, step, ns, lds);
        else
          batch_steps<T, G, S, R, false, SVM>(I, lambda, step, ns, lds);
        step += ns;
        epochs += ep;
        refresh(lambda);
        gap = s_res.rel_gap;
        ok = finite_eval(s_res);
      }
      if (!ok) {  // NaN or infinity: abandon the instance
        batch_fail_rest(a, i, l, s_res, step * (uint64_t)I.tau);
        __syncthreads();
        break;
      }
      if (threadIdx.x == 0) {
        const uint64_t ticks = wall_clock64() - t0;
        BatchResult *o = a.results + (int64_t)i * a.path_len + l;
        o->lambda = lambda;
        o->primal = s_res.primal;
        o->dual = s_res.dual;
        o->rel_gap = s_res.rel_gap;
        o->nnz_x = s_res.nnz_x;
        o->epochs = epochs;
        o->converged = gap <= a.gap_tol;
        o->updates = step * (uint64_t)I.tau;
        o->ticks = ticks;
        o->grad_inf = s_res.grad_inf;
        o->ws_size = (double)I.n;
        o->status = 0;
      }
      __syncthreads();  // s_res is rewritten by the next refresh
    }
  }
}

// ------------------------------------------------- asynchronous instances
// The multicore executor's semantics inside one block: the block's NG =
// kBatchBlock/G lane groups are I.tau workers that each draw coordinates of
// the instance with replacement from their own stream (async_coord keyed by
// the instance seed and the group) and apply updates with no barrier between
// them -- LDS atomics on the residual correction, atomic adds on x_j.  Only
// the FP64 refresh, every refresh_period epochs, stops all workers (one
// block barrier).  An update reads its 32-byte record (x_j included) two
// updates ahead and its entries one update ahead; its gathers go out before
// those loads, so that waiting for them does not wait for the prefetches
// (memory counters drain in issue order).  A worker is then bound by about
// one memory round trip per update, and a CU by its workers' count.
template <typename T, int G, int R, bool LDSR, bool TAIL, bool SVM>
__device__ __forceinline__ void batch_async_updates(const BatchInstance &I, double lambda,
                                                    uint64_t k0, uint64_t count, T *lds) {
  const int lane = threadIdx.x & (G - 1);
  const uint32_t w = threadIdx.x / G;
  if ((int32_t)w >= I.tau || count == 0) return;
  const Entry<T> *ent = static_cast<const Entry<T> *>(I.ent);
  const DrView<T, LDSR> dr = dr_view<T, LDSR>(I, lds);
  const StepScalars sc = batch_scalars<SVM>(lambda, I.n);
  const bool serial = I.tau == 1;
  // this worker's last two updates (coordinate, delta)
  uint32_t j1 = 0xffffffffu, j2 = 0xffffffffu;
  double d1 = 0.0, d2 = 0.0;
  struct Slot {
    uint32_t j;
    CoordRec r;
    int32_t ci[R];
    T cv[R];
  };
  AsyncCoordCache<G> coords;
  // On a working set (I.wperm) the workers share one cyclic order instead:
  // update q = k tau + w of the instance (k the worker's update count)
  // takes slot q mod n of W, whose records batch_working_set stored in a
  // keyed random order, the same in every epoch.  Each epoch visits every
  // coordinate of W once, the tau updates in flight are distinct
  // coordinates (the tau-nice sampling of the ESO), and the epochs repeat
  // one map of x, which the Anderson extrapolation of the working sets
  // needs.  Updates are requested in increasing k, so the slot advances by
  // tau.
  uint32_t wq_pos = 0;
  if (I.wperm) wq_pos = (uint32_t)((k0 * (uint64_t)I.tau + w) % (uint64_t)I.n);
  auto record = [&](uint64_t k, Slot &c) {
    if (I.wperm) {
      (void)k;
      c.j = wq_pos;
      wq_pos += (uint32_t)I.tau;
      if (wq_pos >= (uint32_t)I.n) wq_pos -= (uint32_t)I.n;  // tau <= n: one wrap at most
    } else {
      c.j = coords.get(I.seed, w, k0 + k, (uint32_t)I.n);
    }
    c.r = gld(I.cs + c.j);
  };
  // every lane loads (padding lanes re-read the segment's first entry, or
  // the pad entry after the last segment), so that the number of memory
  // instructions in flight is the same on every path and the wait before
  // an update's gathers only covers loads issued one update earlier
  // (a padding lane keeps the entry it loaded and is masked where the entry
  // is used -- its row is clamped into [0, m) for the gather, since an empty
  // segment's start is the next instance's first entry or the pad entry --
  // a select on the loaded value here would wait for the load at once)
  auto entries = [&](Slot &c) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int32_t q = lane + r * G;
      const Entry<T> e = gld(ent + c.r.p0 + (q < c.r.len ? q : 0));
      c.ci[r] = e.row;
      c.cv[r] = e.val;
    }
  };
  auto gather = [&](int32_t row) -> T { return dr.get((uint32_t)row < (uint32_t)I.m ? row : 0); };
  auto update = [&](uint64_t k, Slot &cur, Slot &nxt, Slot &ahead) {
    // gathers of every lane, without branches (padding lanes read a valid
    // row and are masked below), issued first: residual rows in global
    // memory then return ahead of the next updates' loads, which stay in
    // flight
    T v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = gather(cur.ci[r]);
    entries(nxt);        // update k+1 (past the last update: harmless loads)
    record(k + 2, ahead);
    // x_j from the record loaded two updates ago, plus this worker's own
    // changes to it since (the two previous updates, applied in their order,
    // which makes one worker the serial method exactly); other workers'
    // changes in between are the staleness asynchronous updates allow, and
    // the atomic add below keeps them
    double xj = cur.r.x;
    if (cur.j == j2) xj += d2;
    if (cur.j == j1) xj += d1;
    T acc = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) acc += lane + r * G < cur.r.len ? cur.cv[r] * v[r] : (T)0;
    if (TAIL)
      for (int32_t q = lane + R * G; q < cur.r.len; q += G) {
        const Entry<T> e = gld(ent + cur.r.p0 + q);
        acc += e.val * gather(e.row);
      }
    acc = group_sum<G>(acc);
    const double g = cur.r.gref + (double)acc;
    double d;
    if (!SVM && I.loss == kLasso) {
      d = batch_delta<false>(kLasso, g, xj, cur.r.L, sc);
      if (lane == 0 && d != 0.0) gadd_agent(&I.cs[cur.j].x, d);
    } else {
      // x >= 0 (NNLS) or alpha in [0, 1] (SVM) under races: the new value is
      // computed from the x_j it replaces, by compare-and-swap, and the
      // scatter takes the delta that was swapped in (one worker: the first
      // swap succeeds)
      d = 0.0;
      if (lane == 0) {
        unsigned long long *px = reinterpret_cast<unsigned long long *>(&I.cs[cur.j].x);
        double old = xj;
        for (;;) {
          const double dd = batch_delta<SVM>(kNnls, g, old, cur.r.L, sc);  // NNLS or SVM
          const double nw = old + dd;
          if (nw == old) {
            d = dd;
            break;
          }
          const unsigned long long prev =
              atomicCAS(px, __double_as_longlong(old), __double_as_longlong(nw));
          if (prev == (unsigned long long)__double_as_longlong(old)) {
            d = dd;
            break;
          }
          old = __longlong_as_double((long long)prev);
        }
      }
      d = __shfl(d, 0, G);
    }
    const double c = batch_coef<SVM>(d, cur.r.L, sc);
    j2 = j1;
    d2 = d1;
    j1 = cur.j;
    d1 = d;
    if (c != 0.0) {
      const T ct = (T)c;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (lane + r * G < cur.r.len) dr.add(cur.ci[r], ct * cur.cv[r]);
      if (TAIL)
        for (int32_t q = lane + R * G; q < cur.r.len; q += G) {
          const Entry<T> e = gld(ent + cur.r.p0 + q);
          dr.add(e.row, ct * e.val);
        }
    }
    if (serial)  // one worker: the serial method, each gather sees the last scatter
      asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
    cur.j = ahead.j;
    cur.r = ahead.r;
  };
  Slot s0, s1, ahead;
  record(0, s0);
  entries(s0);
  record(1, s1);
  uint64_t k = 0;
  for (; k + 1 < count; k += 2) {  // no conditional memory work inside
    update(k, s0, s1, ahead);
    update(k + 1, s1, s0, ahead);
  }
  if (k < count) update(k, s0, s1, ahead);
}

// WS: the working-set machinery (screening, Anderson extrapolation, support
// Newton steps) is compiled in; the kernels without it (NNLS, SVM and
// unscreened Lasso batches) keep their registers for the workers' loop
template <typename T, int G, int R, bool TAIL, bool SVM, bool WS>
__global__ __launch_bounds__(kBatchBlock) void k_batch_async(BatchArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  __shared__ double sh[kRefreshStats * kBatchWaves];
  __shared__ BatchResult s_res;
  __shared__ int s_item;
  int next = 0, end = 0;  // the block's current chunk of queue items
  for (;;) {
    if (next >= end) {
      if (threadIdx.x == 0) s_item = (int)atomicAdd(a.queue, (unsigned)a.chunk);
      __syncthreads();
      next = __builtin_amdgcn_readfirstlane(s_item);
      end = min(next + a.chunk, a.count);
      __syncthreads();
      if (next >= a.count) break;
    }
    const int item = next++;
    const int i = a.order[item];
    const BatchInstance *Ip = a.inst + i;
    const BatchInstance I = *Ip;
    const bool in_lds = I.drg == nullptr;
    double lambda_of_item = 0.0;
    // the scatter streams every entry once the last refresh of this
    // instance found more than n/64 nonzero x_j (where streaming all entries
    // costs less than the columns' dependent loads, measured)
    bool dense = false;
    auto refresh = [&](double lambda) {
      if (in_lds)
        batch_refresh<T, true, SVM>(Ip, lambda, lds, sh, a.lds_doubles, &s_res, dense);
      else
        batch_refresh<T, false, SVM>(Ip, lambda, lds, sh, a.lds_doubles, &s_res, dense);
      dense = 64.0 * s_res.nnz_x > (double)I.n;
    };
    auto updates = [&](const BatchInstance &J, uint64_t k0, uint64_t per) {
      if (in_lds)
        batch_async_updates<T, G, R, true, TAIL, SVM>(J, lambda_of_item, k0, per, lds);
      else
        batch_async_updates<T, G, R, false, TAIL, SVM>(J, lambda_of_item, k0, per, lds);
      __syncthreads();  // every worker's updates are in before the refresh
    };
    // Lasso instances with screening: working sets (strong rule, gap-safe screening)
    const bool screen = WS && !SVM && a.screening && I.loss == kLasso;
    uint64_t k = 0;      // updates per worker so far (the workers' stream offset)
    uint64_t total = 0;  // updates of the instance so far
    for (int l = 0; l < a.path_len; ++l) {
      const uint64_t t0 = wall_clock64();
      const double lambda = a.lambdas[(int64_t)i * a.path_len + l];
      lambda_of_item = lambda;
      refresh(lambda);
      double gap = s_res.rel_gap;
      bool ok = finite_eval(s_res);
      // updates of this lambda, against the cap of max_epochs epochs of n
      const double cap = (double)a.max_epochs * (double)I.n;
      double done = 0.0;
      // the strong rule's threshold for the step from the previous lambda
      const double lambda_prev = l > 0 ? a.lambdas[(int64_t)i * a.path_len + l - 1] : lambda;
      const double thr = fmin(lambda, 2.0 * lambda - lambda_prev);
      int32_t nw = I.n;          // working-set size (n: all coordinates)
      int32_t nw_last = I.n;     // the size of the last set the workers drew from
      double tol_w = a.gap_tol;  // the working set's gap target
      uint32_t wkey = 0;         // working sets formed for this lambda
      if (threadIdx.x == 0) s_res.newton = 0;  // support Newton steps kept (counted in LDS)
      while (ok && gap > a.gap_tol && done < cap) {
        int32_t w_count = 0;
        if (screen) {
          // the full refresh just run holds g = A^T r, the gap and the support
          const int64_t wcap = max((int64_t)kWsMin, 2 * (int64_t)s_res.nnz_x);
          w_count = batch_working_set(I, lambda, thr, s_res.primal - s_res.dual, s_res.grad_inf,
                                      wcap, ((uint64_t)l << 32) | wkey++,
                                      reinterpret_cast<int *>(sh));
        }
        if (screen && w_count >= 1) {
          // the working-set problem (W may hold every coordinate: the cyclic
          // sweep and the extrapolation still apply), with W's ESO beta
          nw = w_count;
          nw_last = nw;
          const int32_t om = in_lds ? batch_ws_omega<T, true>(I, nw, lds, reinterpret_cast<int *>(sh))
                                    : batch_ws_omega<T, false>(I, nw, lds, reinterpret_cast<int *>(sh));
          BatchInstance J = I;
          J.cs = I.ws;
          J.n = nw;
          // at most one worker per kWsWorkerSpan coordinates of W: on a small
          // W, many more workers update the same coordinates from stale
          // residuals (measured: pairs with |W| ~ 250 and 128 workers stalled
          // at the cap in FP64)
          J.tau = max(1, min(I.tau, nw / kWsWorkerSpan));
          J.wperm = 1;
          // the ESO constant of W: omega_W, or with the spectral rule the
          // power method's estimate on W (10% margin: it approaches
          // lambda_max from below), whichever is smaller
          double sigma = (double)om;
          if (a.spectral && I.anx && nw > 1) {
            const double lm = in_lds ? batch_ws_lmax<T, true>(I, nw, lds, sh)
                                     : batch_ws_lmax<T, false>(I, nw, lds, sh);
            sigma = fmin(sigma, fmax(1.0, 1.1 * lm));
          }
          const double bw = nw > 1 ? 1.0 + (sigma - 1.0) * (double)(J.tau - 1) / ((double)nw - 1.0)
                                   : 1.0;
          for (int32_t u = threadIdx.x; u < nw; u += kBatchBlock)
            *gp(&I.ws[u].L) = gld(&I.ws[u].L) * (I.beta / bw);
          __syncthreads();
          const double screen_gap = gap;
          double gap_w = gap;
          // a support Newton step from the restricted refresh just run
          auto newton = [&](const BatchInstance &Jw, int32_t n_w) {
            const double p0 = s_res.primal;
            const int ld = a.lds_doubles;
            if (in_lds ? I.m <= ld : false)
              return batch_ws_newton<T, true, true>(Jw, n_w, I.n, lambda, p0, a.ws_newton, lds, ld, sh);
            if (in_lds)
              return batch_ws_newton<T, true, false>(Jw, n_w, I.n, lambda, p0, a.ws_newton, lds, ld, sh);
            return batch_ws_newton<T, false, false>(Jw, n_w, I.n, lambda, p0, a.ws_newton, lds, ld, sh);
          };
          auto refresh_ws = [&]() {
            if (in_lds ? I.m <= a.lds_doubles : false)
              batch_refresh_ws<T, true, true>(J, nw, lambda, lds, sh, &s_res);
            else if (in_lds)
              batch_refresh_ws<T, true, false>(J, nw, lambda, lds, sh, &s_res);
            else
              batch_refresh_ws<T, false, false>(J, nw, lambda, lds, sh, &s_res);
            gap_w = s_res.rel_gap;
          };
          // back to the whole problem once W's gap meets its target, or has
          // fallen a hundredfold since the screening (a new screening then
          // shrinks W), or is not finite (the full refresh below reports it)
          auto w_done = [&]() {
            return !finite_eval(s_res) || gap_w <= tol_w || gap_w <= 1e-2 * screen_gap;
          };
          // snapshots of x every refresh_period epochs of W; restricted
          // refreshes (gap check, FP64 anchor) every kWsRefreshPeriods of
          // them and around each extrapolation
          int aa_n = 0, aa_h = 0;  // snapshots held, next slot
          double gap_prev = screen_gap;  // W's gap at the last restricted refresh
          int since = 0;           // periods since the last restricted refresh
          while (done < cap) {
            const uint64_t per =
                ((uint64_t)nw * (uint64_t)a.refresh_period + J.tau - 1) / J.tau;
            updates(J, k, per);
            k += per;
            done += (double)per * J.tau;
            total += per * (uint64_t)J.tau;
            const bool aa_due = I.anx && aa_n + 1 == kAaSnapshots;
            if (!I.anx || aa_due || ++since >= kWsRefreshPeriods) {
              refresh_ws();
              since = 0;
              if (w_done()) break;
              // a Newton step only where the workers stall: W's gap fell by
              // less than half since the last restricted refresh
              const bool stalled = gap_w > kNewtonStall * gap_prev;
              gap_prev = gap_w;
              if (stalled && I.anx && a.ws_newton > 0 && newton(J, nw)) {
                if (threadIdx.x == 0) ++s_res.newton;
                refresh_ws();
                gap_prev = gap_w;
                if (w_done()) break;
                aa_n = 0;  // the snapshots no longer lead to x: a new history
              }
            }
            if (I.anx) {
              const int newest = aa_h;
              batch_aa_snapshot(J, nw, I.n, newest);
              aa_h = (aa_h + 1) % kAaSnapshots;
              if (++aa_n == kAaSnapshots) {
                const double p_last = s_res.primal;
                if (batch_aa_extrapolate(J, nw, I.n, aa_h, sh)) {
                  refresh_ws();
                  since = 0;
                  if (!(finite_eval(s_res) && s_res.primal < p_last)) {
                    batch_aa_restore(J, nw, I.n, newest);  // rejected
                    refresh_ws();
                  }
                  if (w_done()) break;
                }
                // the next history starts from the current x
                batch_aa_snapshot(J, nw, I.n, aa_h);
                aa_h = (aa_h + 1) % kAaSnapshots;
                aa_n = 1;
              }
            }
          }
          batch_ws_writeback(I, nw);
          refresh(lambda);
          gap = s_res.rel_gap;
          ok = finite_eval(s_res);
          // W met its target but the whole problem did not (the dual point
          // of W is scaled by W's gradients only): aim W lower; the
          // violators outside W, if any, rank first at the next formation
          if (gap > a.gap_tol && gap_w <= tol_w) tol_w *= 0.25;
          nw = I.n;
        } else {
          nw_last = I.n;
          const uint64_t per = ((uint64_t)I.n * (uint64_t)a.refresh_period + I.tau - 1) / I.tau;
          updates(I, k, per);
          k += per;
          done += (double)per * I.tau;
          total += per * (uint64_t)I.tau;
          refresh(lambda);
          gap = s_res.rel_gap;
          ok = finite_eval(s_res);
        }
      }
      if (!ok) {  // NaN or infinity: abandon the instance
        batch_fail_rest(a, i, l, s_res, total);
        __syncthreads();
        break;
      }
      if (threadIdx.x == 0) {
        const uint64_t ticks = wall_clock64() - t0;
        BatchResult *o = a.results + (int64_t)i * a.path_len + l;
        o->lambda = lambda;
        o->primal = s_res.primal;
        o->dual = s_res.dual;
        o->rel_gap = s_res.rel_gap;
        o->nnz_x = s_res.nnz_x;
        o->epochs = (int32_t)ceil(done / (double)I.n);
        o->converged = gap <= a.gap_tol;
        o->updates = total;
        o->ticks = ticks;
        o->grad_inf = s_res.grad_inf;
        o->ws_size = (double)nw_last;
        o->status = 0;
        o->newton = s_res.newton;
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------ kernel selection
// coordinates per lane group and step of the synchronous kernel; with 4
// register chunks per column this keeps every instantiation within 128
// VGPRs (1024-thread blocks) without scratch spills
constexpr int kSlotsPerGroup = 2;

// column entries cached in registers per lane: 4 chunks, 2 for 64-lane groups
constexpr int regs_for(int lanes) { return lanes == 64 ? 2 : 4; }

// entries of a column an asynchronous worker caches in registers per lane
// (FP32: 8 chunks for 8 and 16 lanes, 4 for 32, 2 for 64; FP64: at most 4),
// the most that stays spill-free in 128 VGPRs; longer columns take the
// variant with a tail loop
template <typename T>
constexpr int async_regs_for(int lanes) {
  return lanes == 64 ? 2 : (lanes == 32 || sizeof(T) == 8 ? 4 : 8);
}

template <typename T, int G, bool SVM, bool WS>
const void *batch_fn(bool async, bool tail) {
  constexpr int RA = async_regs_for<T>(G);
  if (WS || async)  // working sets are for asynchronous workers only
    return tail ? reinterpret_cast<const void *>(&k_batch_async<T, G, RA, true, SVM, WS>)
                : reinterpret_cast<const void *>(&k_batch_async<T, G, RA, false, SVM, WS>);
  return reinterpret_cast<const void *>(&k_batch_lasso<T, G, kSlotsPerGroup, regs_for(G), SVM>);
}

// the batch kernel for a precision, lanes per coordinate (8, 16, 32 or 64),
// step mode and longest column; SVM = true and WS = true (screened Lasso,
// asynchronous) are each instantiated in a translation unit of their own and
// reached through batch_svm_kernel and batch_ws_kernel
template <bool SVM, bool WS = false>
const void *batch_kernel_for(bool fp32, int lanes, bool async, int64_t maxlen) {
  const int ra = fp32 ? async_regs_for<float>(lanes) : async_regs_for<double>(lanes);
  const bool tail = maxlen > (int64_t)ra * lanes;
  switch (lanes) {
    case 8: return fp32 ? batch_fn<float, 8, SVM, WS>(async, tail) : batch_fn<double, 8, SVM, WS>(async, tail);
    case 16: return fp32 ? batch_fn<float, 16, SVM, WS>(async, tail) : batch_fn<double, 16, SVM, WS>(async, tail);
    case 32: return fp32 ? batch_fn<float, 32, SVM, WS>(async, tail) : batch_fn<double, 32, SVM, WS>(async, tail);
    default: return fp32 ? batch_fn<float, 64, SVM, WS>(async, tail) : batch_fn<double, 64, SVM, WS>(async, tail);
  }
}
const void *batch_svm_kernel(bool fp32, int lanes, bool async, int64_t maxlen);
const void *batch_ws_kernel(bool fp32, int lanes, int64_t maxlen);

}  // namespace acdc
