// batch_kernels.h -- one persistent launch that solves many Lasso instances,
// each along a warm-started lambda path.
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

struct BatchInstance {
  CoordRec *cs;          // n records
  const void *ent;       // Entry<T>[nnz]
  const double *val64;   // FP64 values of the refresh, entry order
  const double *b;
  double *r64;           // FP64 residual of the refresh when it does not fit in LDS
  void *drg;             // global correction residual (T) when the rows do not fit in LDS
  int32_t m, n;
  int32_t tau;           // coordinates per step of this instance
  int32_t pad;
  uint64_t seed;
};

struct BatchResult {
  double lambda, primal, dual, rel_gap, nnz_x;
  int32_t epochs, converged;
  uint64_t updates;
};

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
  BatchResult *results;  // instances x path_len
  int32_t lds_doubles;   // FP64 rows the dynamic LDS holds (refresh residual)
};

// zeroes x of every coordinate record and the work-queue counters (a kernel
// rather than a memset so that the whole run can live in one captured graph)
static __global__ __launch_bounds__(kBlock) void k_batch_reset(CoordRec *__restrict__ cs,
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
constexpr int kRefreshStats = 5;  // max|g|, ||x||_1, nnz(x), ||r||^2, b.r
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
        const Entry<T> e = ent[p0 + q];
        row[r] = e.row;
        val[r] = e.val;
      } else {
        row[r] = 0;
        val[r] = (T)0;
      }
    }
  }
  // <column, v>, partial per lane (entries past R*G read straight from memory)
  template <typename V>
  __device__ __forceinline__ V dot(const Entry<T> *ent, const V *v, int lane) const {
    V acc = 0;
#pragma unroll
    for (int r = 0; r < R; ++r)  // padding lanes read nothing
      if (lane + r * G < len) acc += (V)val[r] * v[row[r]];
    for (int32_t q = lane + R * G; q < len; q += G) {
      const Entry<T> e = ent[p0 + q];
      acc += (V)e.val * v[e.row];
    }
    return acc;
  }
};

template <typename T, bool LDSR>
__device__ __forceinline__ T *dr_ptr(const BatchInstance &I, T *lds) {
  return LDSR ? lds : static_cast<T *>(I.drg);
}

template <typename V, bool LOCAL>
__device__ __forceinline__ void add_to(V *p, V v) {
  if (LOCAL)
    atomicAdd(p, v);  // ds_add on LDS
  else
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// FP64 refresh + evaluation of one instance by one block: fills out (in
// shared memory) and zeroes dr.  r64 lives in LDS (reusing the space of dr,
// which is rebuilt at the end) when m doubles fit, in the global r64 buffer
// otherwise.  The gradient pass walks the columns in rounds of
// kBatchBlock/16 x 2 columns, one 16-lane group per column.  Its statistics
// meet in thread 0 only (block_stats), which keeps the pass near 36 VGPRs so
// that inlining it costs the steps nothing.
template <typename T, bool LDSR>
__device__ __forceinline__ void batch_refresh(const BatchInstance *Ip, double lambda, T *lds,
                                           double *sh, int lds_doubles, BatchResult *out) {
  constexpr int G = 16, NG = kBatchBlock / G;
  const BatchInstance &I = *Ip;
  const int tid = threadIdx.x, lane = tid & (G - 1), grp = tid / G;
  const bool rl = LDSR && I.m <= lds_doubles;  // block-uniform
  double *r = rl ? reinterpret_cast<double *>(lds) : I.r64;
  const Entry<T> *ent = static_cast<const Entry<T> *>(I.ent);
  for (int k = tid; k < I.m; k += kBatchBlock) r[k] = -I.b[k];
  __syncthreads();
  // r += x_j a_j over the nonzero x_j: a 16-lane group per column
  for (int j = grp; j < I.n; j += NG) {
    const double xj = ld_coherent(&I.cs[j].x);  // async workers add to x atomically
    if (xj == 0.0) continue;
    const int32_t p0 = I.cs[j].p0, len = I.cs[j].len;
    for (int32_t q = lane; q < len; q += G) {
      const int32_t p = p0 + q;
      const double v = I.val64[p] * xj;
      if (rl)
        add_to<double, true>(&r[ent[p].row], v);
      else
        add_to<double, false>(&r[ent[p].row], v);
    }
  }
  __syncthreads();
  double ginf = 0, x1 = 0, nz = 0;
  // two columns per group and round, both loads issued before the reductions
  for (int j0 = grp; j0 < I.n; j0 += 2 * NG) {
    const int j1 = j0 + NG;
    const bool on1 = j1 < I.n;
    const int32_t pa = I.cs[j0].p0, la = I.cs[j0].len;
    const int32_t pb = on1 ? I.cs[j1].p0 : 0, lb = on1 ? I.cs[j1].len : 0;
    double ga = 0, gb = 0;
    for (int32_t q = lane; q < la; q += G) ga += I.val64[pa + q] * r[ent[pa + q].row];
    for (int32_t q = lane; q < lb; q += G) gb += I.val64[pb + q] * r[ent[pb + q].row];
    ga = group_sum<G>(ga);
    gb = group_sum<G>(gb);
    if (lane == 0) {
      const double xa = ld_coherent(&I.cs[j0].x);
      I.cs[j0].gref = ga;
      ginf = fmax(ginf, fabs(ga));
      x1 += fabs(xa);
      nz += xa != 0.0;
      if (on1) {
        const double xb = ld_coherent(&I.cs[j1].x);
        I.cs[j1].gref = gb;
        ginf = fmax(ginf, fabs(gb));
        x1 += fabs(xb);
        nz += xb != 0.0;
      }
    }
  }
  double r2 = 0, br = 0;
  for (int k = tid; k < I.m; k += kBatchBlock) {
    const double rk = r[k];
    r2 += rk * rk;
    br += I.b[k] * rk;
  }
  double st[kRefreshStats] = {ginf, x1, nz, r2, br};
  block_stats(st, sh);  // its barrier also ends every read of r
  T *dr = dr_ptr<T, LDSR>(I, lds);
  for (int k = tid; k < I.m; k += kBatchBlock) dr[k] = (T)0;
  if (tid == 0) {
    const double P = 0.5 * st[3] + lambda * st[1];
    const double s = st[0] > lambda ? lambda / st[0] : 1.0;
    const double D = -s * st[4] - 0.5 * s * s * st[3];
    out->primal = P;
    out->dual = D;
    out->nnz_x = st[2];
    out->rel_gap = P != 0 ? (P - D) / fabs(P) : 0.0;
  }
  __syncthreads();  // gref, dr and *out are settled for the steps
}

// nsteps synchronous steps of one instance; S slots per G-lane group.  The
// sampled coordinates of a wave's S * (64/G) slots are computed by its first
// lanes in one pass and broadcast; each slot's record and column are loaded
// after the first barrier of the previous step (its x writes are done then).
template <typename T, int G, int S, int R, bool LDSR>
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
  T *dr = dr_ptr<T, LDSR>(I, lds);
  StepScalars sc;
  sc.lambda = lambda;
  sc.lambda_n = 0;
  sc.inv_lambda_n = 0;
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
        rec[k] = I.cs[j[k]];
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
      const double d = prox_delta<kLasso>(g, rec[k].x, rec[k].L, 0.0, sc, &c[k]);
      if (lane == 0) I.cs[j[k]].x = rec[k].x + d;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < S; ++k) {
      if (c[k] == 0.0) continue;
      const T ct = (T)c[k];
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (lane + r * G < col[k].len) add_to<T, LDSR>(&dr[col[k].row[r]], ct * col[k].val[r]);
      for (int32_t q = lane + R * G; q < col[k].len; q += G) {
        const Entry<T> e = ent[col[k].p0 + q];
        add_to<T, LDSR>(&dr[e.row], ct * e.val);
      }
    }
    if (s + 1 < nsteps) prefetch(step0 + s + 1);
    __syncthreads();
  }
}

template <typename T, int G, int S, int R, bool LDSR>
__global__ __launch_bounds__(kBatchBlock) void k_batch_lasso(BatchArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  __shared__ double sh[kRefreshStats * kBatchWaves];
  __shared__ BatchResult s_res;
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(a.queue, 1u);
    __syncthreads();
    // readfirstlane marks the block-uniform item as such, so the instance's
    // fields and pointers live in scalar registers
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (item >= a.count) break;
    const int i = a.order[item];
    const BatchInstance *Ip = a.inst + i;
    const BatchInstance I = *Ip;
    const uint64_t steps_per_epoch = ((uint64_t)I.n + I.tau - 1) / I.tau;
    uint64_t step = 0;
    for (int l = 0; l < a.path_len; ++l) {
      const double lambda = a.lambdas[(int64_t)i * a.path_len + l];
      batch_refresh<T, LDSR>(Ip, lambda, lds, sh, a.lds_doubles, &s_res);
      double gap = s_res.rel_gap;
      int epochs = 0;
      while (gap > a.gap_tol && epochs < a.max_epochs) {
        const int ep = min(a.refresh_period, a.max_epochs - epochs);
        const uint64_t ns = steps_per_epoch * (uint64_t)ep;
        batch_steps<T, G, S, R, LDSR>(I, lambda, step, ns, lds);
        step += ns;
        epochs += ep;
        batch_refresh<T, LDSR>(Ip, lambda, lds, sh, a.lds_doubles, &s_res);
        gap = s_res.rel_gap;
      }
      if (threadIdx.x == 0) {
        BatchResult res = s_res;
        res.lambda = lambda;
        res.epochs = epochs;
        res.converged = gap <= a.gap_tol;
        res.updates = step * (uint64_t)I.tau;
        a.results[(int64_t)i * a.path_len + l] = res;
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
// block barrier).  An update reads its 32-byte record two updates ahead and
// its entries one update ahead, so its critical path is the LDS gather plus
// the agent-scope load of x_j (agent scope: the L1 does not hold x_j across
// the other workers' atomic updates).
template <typename T, int G, int R, bool LDSR>
__device__ __forceinline__ void batch_async_updates(const BatchInstance &I, double lambda,
                                                    uint64_t k0, uint64_t count, T *lds) {
  const int lane = threadIdx.x & (G - 1);
  const uint32_t w = threadIdx.x / G;
  if ((int32_t)w >= I.tau || count == 0) return;
  const Entry<T> *ent = static_cast<const Entry<T> *>(I.ent);
  T *dr = dr_ptr<T, LDSR>(I, lds);
  StepScalars sc;
  sc.lambda = lambda;
  sc.lambda_n = 0;
  sc.inv_lambda_n = 0;
  const bool serial = I.tau == 1;
  struct Slot {
    uint32_t j;
    CoordRec r;
    int32_t ci[R];
    T cv[R];
  };
  auto record = [&](uint64_t k, Slot &c) {
    c.j = async_coord(I.seed, w, k0 + k, (uint32_t)I.n);
    c.r = I.cs[c.j];
  };
  auto entries = [&](Slot &c) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int32_t q = lane + r * G;
      if (q < c.r.len) {
        const Entry<T> e = ent[c.r.p0 + q];
        c.ci[r] = e.row;
        c.cv[r] = e.val;
      } else {
        c.ci[r] = 0;
        c.cv[r] = (T)0;
      }
    }
  };
  auto gather = [&](int32_t row) -> T {
    return LDSR ? dr[row] : ld_coherent(&dr[row]);
  };
  auto update = [&](uint64_t k, Slot &cur, Slot &nxt, Slot &ahead) {
    if (k + 1 < count) entries(nxt);
    if (k + 2 < count) record(k + 2, ahead);
    const double xj = ld_coherent(&I.cs[cur.j].x);
    T acc = 0;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (lane + r * G < cur.r.len) acc += cur.cv[r] * gather(cur.ci[r]);
    for (int32_t q = lane + R * G; q < cur.r.len; q += G) {
      const Entry<T> e = ent[cur.r.p0 + q];
      acc += e.val * gather(e.row);
    }
    acc = group_sum<G>(acc);
    double c;
    const double d = prox_delta<kLasso>(cur.r.gref + (double)acc, xj, cur.r.L, 0.0, sc, &c);
    if (lane == 0 && d != 0.0) unsafeAtomicAdd(&I.cs[cur.j].x, d);
    if (c != 0.0) {
      const T ct = (T)c;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (lane + r * G < cur.r.len) add_to<T, LDSR>(&dr[cur.ci[r]], ct * cur.cv[r]);
      for (int32_t q = lane + R * G; q < cur.r.len; q += G) {
        const Entry<T> e = ent[cur.r.p0 + q];
        add_to<T, LDSR>(&dr[e.row], ct * e.val);
      }
    }
    if (serial)  // one worker: the serial method, each update sees the last
      asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
    cur.j = ahead.j;
    cur.r = ahead.r;
  };
  Slot s0, s1, ahead;
  record(0, s0);
  entries(s0);
  if (count > 1) record(1, s1);
  for (uint64_t k = 0; k < count; k += 2) {
    update(k, s0, s1, ahead);
    if (k + 1 < count) update(k + 1, s1, s0, ahead);
  }
}

template <typename T, int G, int R, bool LDSR>
__global__ __launch_bounds__(kBatchBlock) void k_batch_async(BatchArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *lds = reinterpret_cast<T *>(lds_raw);
  __shared__ double sh[kRefreshStats * kBatchWaves];
  __shared__ BatchResult s_res;
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(a.queue, 1u);
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (item >= a.count) break;
    const int i = a.order[item];
    const BatchInstance *Ip = a.inst + i;
    const BatchInstance I = *Ip;
    uint64_t k = 0;  // updates per worker so far
    for (int l = 0; l < a.path_len; ++l) {
      const double lambda = a.lambdas[(int64_t)i * a.path_len + l];
      batch_refresh<T, LDSR>(Ip, lambda, lds, sh, a.lds_doubles, &s_res);
      double gap = s_res.rel_gap;
      int epochs = 0;
      while (gap > a.gap_tol && epochs < a.max_epochs) {
        const int ep = min(a.refresh_period, a.max_epochs - epochs);
        const uint64_t per = ((uint64_t)I.n * (uint64_t)ep + I.tau - 1) / I.tau;
        batch_async_updates<T, G, R, LDSR>(I, lambda, k, per, lds);
        __syncthreads();  // every worker's updates are in before the refresh
        k += per;
        epochs += ep;
        batch_refresh<T, LDSR>(Ip, lambda, lds, sh, a.lds_doubles, &s_res);
        gap = s_res.rel_gap;
      }
      if (threadIdx.x == 0) {
        BatchResult res = s_res;
        res.lambda = lambda;
        res.epochs = epochs;
        res.converged = gap <= a.gap_tol;
        res.updates = k * (uint64_t)I.tau;
        a.results[(int64_t)i * a.path_len + l] = res;
      }
      __syncthreads();
    }
  }
}

}  // namespace acdc
