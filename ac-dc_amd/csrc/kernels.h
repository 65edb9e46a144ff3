// kernels.h -- gfx950 kernels of the PCDM hot path (templates; instantiated
// by the launchers in solver.hip).
//
// Data are "segments": the coordinate-major matrix holds one segment per
// coordinate (a CSC column for Lasso/NNLS, a CSR sample row for the SVM
// dual).  A group of G lanes (G = 8, 16, 32 or 64, aligned inside the
// 64-wide wavefront) owns one segment: its lanes stride over the segment with
// unit stride, so the index and value streams are read coalesced, and the
// partial sums meet in a fixed xor butterfly.  MFMA is not used: every kernel
// here is a sparse gather/scatter bound by HBM and L2 traffic.
//
// Residual representation.  FP64 path: res is the running residual r (the
// SVM's w) in FP64.  FP32 path: the residual is anchored,
//     r = r64 + dr,   g_j = gref_j + <a_j, dr>,
// where r64 = Ax - b and gref = A^T r64 are recomputed in FP64 -- from the
// FP64 values of A -- at every refresh, and dr (FP32) accumulates only the
// updates since then.  The FP32 terms are corrections whose rounding error
// shrinks with the step sizes, so the iteration converges to the solution of
// the FP64 problem (iterative refinement), not of the FP32-rounded one.
#pragma once
#include <type_traits>

#include "device_common.h"

namespace acdc {

template <typename T>
struct Seg {
  const int32_t *__restrict__ ptr;  // nseg + 1 offsets
  const int32_t *__restrict__ idx;  // nnz indices
  const T *__restrict__ val;        // nnz values
};

constexpr int kBlock = 256;  // threads per block

// Global-memory (address space 1) views of generic pointers.  Memory
// operations through a generic pointer compile to flat_* instructions, which
// count against the LDS counter as well, so every LDS access after one waits
// for the flat load's memory round trip; through these views they are
// global_* instructions with their own counter.
template <typename T>
struct GlobalOf {
  typedef __attribute__((address_space(1))) T type;
};
template <typename T>
__device__ __forceinline__ typename GlobalOf<T>::type *gp(T *p) {
  return (typename GlobalOf<T>::type *)p;
}
template <typename S>
__device__ __forceinline__ S gld(const S *p) {
  S v;
  __builtin_memcpy(&v, gp(p), sizeof(S));
  return v;
}
template <typename T>
__device__ __forceinline__ T gld_agent(const T *p) {
  return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void gadd_agent(T *p, T v) {
  __hip_atomic_fetch_add(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void gadd_block(T *p, T v) {
  __hip_atomic_fetch_add(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One (index, value) pair of a segment, interleaved so that a lane's entry
// is one 8-byte (FP32) or 16-byte (FP64) load and a segment's entries share
// cache lines.
template <typename T>
struct Entry;
template <>
struct alignas(8) Entry<float> {
  int32_t row;
  float val;
};
template <>
struct alignas(16) Entry<double> {
  int32_t row;
  int32_t pad;
  double val;
};

template <typename T>
struct StepArgs {
  Seg<T> A;
  T *res;              // FP64: running residual; FP32: correction dr
  const double *gref;  // FP32: anchored gradient A^T r64; FP64: null
  double *x;
  const double *L;
  const double *y;     // SVM labels (null otherwise)
  StepScalars sc;
  Perm perm;
  uint64_t seed;
  int64_t tau;
  int32_t *coord;
  double *coef;
  // streamed scatter (k_step_scatter_t): two halves of ncoord coefficients,
  // step s writes half s & 1 at its coordinates and clears those of step
  // s - 1 in the other half; null when the scatter is atomic
  T *cvec;
  int64_t ncoord;
};

// One coordinate's proximal step from the group's reduced partial sum; every
// lane of the group evaluates it identically.
template <typename T, int LOSS>
__device__ __forceinline__ double coord_delta(const StepArgs<T> &a, uint32_t j,
                                              double acc, double xj, double *c) {
  const double g = a.gref ? a.gref[j] + acc : acc;
  return prox_delta<LOSS>(g, xj, a.L[j], LOSS == kSvm ? a.y[j] : 0.0, a.sc, c);
}

// ------------------------------------------------------------------ phase A
// Synchronous step s, gather half: group -> slot i < tau, coordinate
// j = perm_s(i); g = <segment j, residual>; delta from the loss's proximal
// rule; x_j += delta (coordinates of one step are distinct, so no other group
// touches x_j); the scatter coefficient is staged for phase B.  A group is 8
// to 64 lanes of one wave, or the whole 256-thread block for power-law-long
// segments (SVM sample rows), where the longest segment of a step would
// otherwise keep one wave walking it while the rest of the step waits.
// Restates compute_update + the x half of performOneCoordinateUpdate
// (MulticoreL1RegSquareLoss.h:41-56, MulticoreSVMDualLoss.h:119-134) with the
// synchronous semantics of cdn_common.h:255-270.
template <typename T, int G, int LOSS>
__global__ __launch_bounds__(kBlock) void k_step_gather(StepArgs<T> a, uint64_t step) {
  static_assert(G <= 64 || G == kBlock, "a group is at most a wave, or the whole block");
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t slot = t / G;
  const int lane = threadIdx.x & (G - 1);
  if (slot >= a.tau) return;  // whole groups leave together
  const uint32_t j = a.perm((uint32_t)slot, step_keys(a.seed, step));
  const int32_t p0 = a.A.ptr[j], p1 = a.A.ptr[j + 1];
  // the coordinate's x, L, anchor and label, loaded with the segment bounds
  // rather than after the dot (one dependent round trip less per step)
  const double xj = a.x[j], Lj = a.L[j];
  const double gj = a.gref ? a.gref[j] : 0.0;
  const double yj = LOSS == kSvm ? a.y[j] : 0.0;
  // U entries per lane in flight: their indices, then their gathers (eight
  // for a block per coordinate: the step waits for its longest row, whose
  // rounds of dependent loads this halves)
  constexpr int U = G == kBlock ? 8 : 4;
  T acc = 0;
  int32_t p = p0 + lane;
  for (; p + (U - 1) * G < p1; p += U * G) {
    int32_t r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) r[u] = a.A.idx[p + u * G];
#pragma unroll
    for (int u = 0; u < U; ++u) acc += a.A.val[p + u * G] * a.res[r[u]];
  }
  for (; p < p1; p += G) acc += a.A.val[p] * a.res[a.A.idx[p]];
  if constexpr (G == kBlock) {  // a block per coordinate: the waves meet in LDS
    __shared__ T part[kBlock / 64];
    acc = group_sum<64>(acc);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    acc = part[0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) acc += part[w];
  } else {
    acc = group_sum<G>(acc);
  }
  double c;
  const double d = prox_delta<LOSS>(gj + (double)acc, xj, Lj, yj, a.sc, &c);
  if (lane == 0) {
    if (a.cvec) {
      // this slot's coordinate of the previous step leaves the other half
      // (read before this step's coordinate replaces it); no other slot of
      // this step writes there, since the other half belongs to step s - 1
      const int64_t half = (int64_t)(step & 1) * a.ncoord;
      a.cvec[(a.ncoord - half) + a.coord[slot]] = (T)0;
      a.cvec[half + j] = (T)c;
    }
    a.x[j] = xj + d;
    a.coord[slot] = (int32_t)j;
    a.coef[slot] = c;
  }
}

// ------------------------------------------------------------------ phase B
// Scatter half: residual += coef * segment, with hardware float/double
// atomics (parallel::atomic_add CAS loops in the reference,
// parallel_essentials_posix.h:62-78).  A zero coefficient -- the common case
// for Lasso coordinates that stay at zero -- skips the segment entirely.
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_step_scatter(StepArgs<T> a) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t slot = t / G;
  const int lane = threadIdx.x & (G - 1);
  if (slot >= a.tau) return;
  const double c = a.coef[slot];
  if (c == 0.0) return;
  const int32_t j = a.coord[slot];
  const T ct = (T)c;
  const int32_t p0 = a.A.ptr[j], p1 = a.A.ptr[j + 1];
  constexpr int U = G == kBlock ? 8 : 4;  // entries per lane per round trip
  int32_t p = p0 + lane;
  for (; p + (U - 1) * G < p1; p += U * G) {
    int32_t r[U];
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      r[u] = a.A.idx[p + u * G];
      v[u] = a.A.val[p + u * G];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) unsafeAtomicAdd(&a.res[r[u]], ct * v[u]);
  }
  for (; p < p1; p += G) unsafeAtomicAdd(&a.res[a.A.idx[p]], ct * a.A.val[p]);
}

// ------------------------------------------------ phase B, streamed form
// The scatter of a whole step as one pass over the transposed matrix: for
// every residual index i, res_i += sum over its entries (coordinate j, value
// a) of c_j a, with c_j the step's coefficient (zero for coordinates not
// sampled).  Taken when a step covers a large share of the entries (SVM
// minibatches of thousands of sample rows): every entry is streamed once,
// coalesced, instead of one random atomic per sampled entry.
//
// The host cuts the transposed entries into block ranges at residual-index
// boundaries, so block b owns the indices [f0, f1) outright.  The block
// copies the step's ncoord coefficients into LDS, adds every entry's
// c_j a into an LDS accumulator of its f1 - f0 indices (LDS atomics: no
// global memory operation inside the stream besides the entry loads, which
// stay in flight a batch ahead), and ends with one coalesced
// read-modify-write of res[f0, f1).
template <typename T>
struct TEntry {
  uint16_t lseg;   // residual index - f0 of the entry's block
  uint16_t coord;  // coordinate (column, or sample row)
  T val;
};

// threads per block of the streamed scatter: its LDS (the coefficients and
// the accumulator) allows one block per CU, so the block is the CU's
// sixteen waves
constexpr int kTBlock = 1024;

template <typename T>
struct TScatterArgs {
  const TEntry<T> *ent;
  const int32_t *eb;  // entries of block b: [eb[b], eb[b+1])
  const int32_t *fb;  // residual indices of block b: [fb[b], fb[b+1])
  const T *cvec;      // the step's half: ncoord coefficients
  int32_t ncoord;
  T *res;
};

template <typename T>
__global__ __launch_bounds__(kTBlock) void k_step_scatter_t(TScatterArgs<T> a) {
  extern __shared__ __align__(16) unsigned char tl_raw[];
  T *cl = reinterpret_cast<T *>(tl_raw);  // coefficients, then the accumulator
  const int32_t e0 = a.eb[blockIdx.x], e1 = a.eb[blockIdx.x + 1];
  const int32_t f0 = a.fb[blockIdx.x], f1 = a.fb[blockIdx.x + 1];
  T *acc = cl + a.ncoord;
  constexpr int SU = 16;  // loads in flight per thread before their uses
  // a batch of SU entries per thread (lanes past e1 re-read the last entry
  // and are masked where it is used)
  auto load = [&](TEntry<T>(&e)[SU], int64_t q0) {
#pragma unroll
    for (int u = 0; u < SU; ++u) e[u] = a.ent[min(q0 + u * kTBlock, (int64_t)e1 - 1)];
  };
  // the first batch is in flight while the coefficients are staged, and each
  // next batch while the current one is summed (memory counters drain in
  // issue order: waiting for a batch does not wait for the one after it)
  TEntry<T> cur[SU];
  const int64_t q_first = (int64_t)e0 + threadIdx.x;
  if (e0 < e1) load(cur, q_first);
  for (int32_t i0 = threadIdx.x; i0 < a.ncoord; i0 += SU * kTBlock) {
    T v[SU];
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      const int32_t i = i0 + u * kTBlock;
      v[u] = i < a.ncoord ? a.cvec[i] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      const int32_t i = i0 + u * kTBlock;
      if (i < a.ncoord) cl[i] = v[u];
    }
  }
  for (int32_t k = threadIdx.x; k < f1 - f0; k += kTBlock) acc[k] = (T)0;
  __syncthreads();
  for (int64_t q0 = q_first; q0 < e1; q0 += SU * kTBlock) {
    TEntry<T> nxt[SU];
    const bool more = q0 + SU * kTBlock < e1;
    if (more) load(nxt, q0 + SU * kTBlock);
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      const T c = cl[cur[u].coord];
      if (q0 + u * kTBlock < e1 && c != (T)0) atomicAdd(&acc[cur[u].lseg], c * cur[u].val);
    }
    if (more)
#pragma unroll
      for (int u = 0; u < SU; ++u) cur[u] = nxt[u];
  }
  __syncthreads();
  for (int32_t k0 = threadIdx.x; k0 < f1 - f0; k0 += SU * kTBlock) {
    T r[SU];
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      const int32_t k = k0 + u * kTBlock;
      r[u] = k < f1 - f0 ? a.res[f0 + k] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      const int32_t k = k0 + u * kTBlock;
      if (k < f1 - f0 && acc[k] != (T)0) a.res[f0 + k] = r[u] + acc[k];
    }
  }
}

// ------------------------------------------------------ grid-wide barrier
// Sense-counting barrier over all blocks of a cooperative launch.  Arrival is
// one agent-scope atomic per block; the last block bumps the generation.  The
// wait is bounded: after spin_limit polls the block raises the shared error
// flag, every waiting block sees it and returns false, and the kernel ends.
//
// No cache-wide fence is issued (an agent-scope fence writes back and
// invalidates the whole L2 of the XCD, once per block).  Instead every access
// that crosses blocks inside the persistent kernel is itself an agent-scope
// atomic (residual gathers and scatters, x_j), and each wave waits for its own
// memory operations to complete before the block arrives, so an arrival is
// only counted once the block's atomics have been performed.
// Two levels: block b arrives on sub-counter b % kSubBarriers; the last block
// of each subgroup arrives on the top counter; the last subgroup publishes the
// new generation to one release word per subgroup, which only that
// subgroup's blocks poll.  Every counter and word sits on its own 256-byte
// line, so arrivals are spread over kSubBarriers addresses and polling never
// contends with arrivals.
constexpr int kSubBarriers = 16;

struct alignas(256) SyncWord {
  unsigned int v;
  unsigned int pad[63];
};

struct GridSync {
  SyncWord top;
  SyncWord error;
  SyncWord sub[kSubBarriers];
  SyncWord release[kSubBarriers];
};

__device__ __forceinline__ void wait_memory() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

__device__ __forceinline__ unsigned int atomic_inc_agent(unsigned int *p) {
  return __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned int load_agent(const unsigned int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_agent(unsigned int *p, unsigned int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool grid_sync(GridSync *gs, unsigned int &gen,
                                          unsigned int spin_limit) {
  __shared__ unsigned int s_ok;
  wait_memory();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int want = gen + 1;
    const unsigned int nb = gridDim.x;
    const unsigned int nsub = nb < (unsigned)kSubBarriers ? nb : (unsigned)kSubBarriers;
    const unsigned int sub = blockIdx.x % nsub;
    const unsigned int in_sub = (nb - sub + nsub - 1) / nsub;  // blocks b with b % nsub == sub
    bool released = false;
    if (atomic_inc_agent(&gs->sub[sub].v) == in_sub - 1) {
      store_agent(&gs->sub[sub].v, 0u);
      if (atomic_inc_agent(&gs->top.v) == nsub - 1) {
        store_agent(&gs->top.v, 0u);
        wait_memory();
        for (unsigned int k = 0; k < nsub; ++k) store_agent(&gs->release[k].v, want);
        released = true;
      }
    }
    if (!released) {
      unsigned int spins = 0;
      while (load_agent(&gs->release[sub].v) != want) {
        ++spins;
        if ((spins & 63u) == 0 && load_agent(&gs->error.v)) break;
        if (spins > spin_limit) {
          store_agent(&gs->error.v, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    s_ok = load_agent(&gs->error.v) == 0;
  }
  __syncthreads();
  gen += 1;
  return s_ok != 0;
}

// agent-scope (device-coherent) accesses used inside the persistent kernel
template <typename T>
__device__ __forceinline__ T ld_coherent(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_coherent(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------- persistent sync steps
// nsteps synchronous steps in one cooperative launch: two grid barriers per
// step (after the gathers, after the scatters).  Each group owns at most one
// slot per step (the host sizes the grid so that tau <= groups).  The first
// R*G entries of the slot's segment stay in registers from the gather to the
// scatter, and the next step's coordinate, segment bounds and cached entries
// -- none of which depend on the residual -- are loaded before the second
// barrier, so only the residual gather and x_j remain on the critical path.
constexpr int kPersistentBlock = 1024;  // one block per CU: 256 barrier arrivals

// Sum over aligned groups of G = 128 or 256 lanes of a block (several waves
// per group): wave sums, then each group adds its waves' sums from LDS in
// wave order.  Every thread of the block must call it (block barriers).
template <int G, typename T>
__device__ __forceinline__ T block_group_sum(T v, T *red) {
  static_assert(G % 64 == 0 && G > 64, "groups of whole waves");
  v = group_sum<64>(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const int w0 = (int)(threadIdx.x / G) * (G / 64);
  T t = red[w0];
#pragma unroll
  for (int k = 1; k < G / 64; ++k) t += red[w0 + k];
  __syncthreads();  // red is rewritten by the next call
  return t;
}

// G = 8 .. 64 lanes of one wave per slot, or G = kBlock (256): four waves per
// slot (the SVM's long sample rows), whose partial sums meet in LDS.
template <typename T, int G, int LOSS, int R>
__global__ __launch_bounds__(kPersistentBlock) void k_sync_persistent(
    StepArgs<T> a, uint64_t step0, uint64_t nsteps, GridSync *gs, unsigned int spin_limit) {
  __shared__ T red[kPersistentBlock / 64];
  const int lane = threadIdx.x & (G - 1);
  const int64_t slot = ((int64_t)blockIdx.x * kPersistentBlock + threadIdx.x) / G;
  const bool active = slot < a.tau;  // group-uniform
  unsigned int gen = 0;
  uint32_t j = 0;
  int32_t p0 = 0, p1 = 0;
  int32_t ci[R];
  T cv[R];
  auto prefetch = [&](uint64_t step) {
    j = a.perm((uint32_t)slot, step_keys(a.seed, step));
    p0 = a.A.ptr[j];
    p1 = a.A.ptr[j + 1];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int32_t p = p0 + lane + r * G;
      ci[r] = p < p1 ? a.A.idx[p] : 0;
      cv[r] = p < p1 ? a.A.val[p] : (T)0;
    }
  };
  if (active) prefetch(step0);
  for (uint64_t s = 0; s < nsteps; ++s) {
    double c = 0.0;
    T acc = 0;
    if (active) {
#pragma unroll
      for (int r = 0; r < R; ++r)  // padding lanes issue no load
        if (p0 + lane + r * G < p1) acc += cv[r] * ld_coherent(&a.res[ci[r]]);
      // entries past the cached ones: four per lane in flight (their
      // indices, then their gathers), as long rows bound the step
      int32_t p = p0 + lane + R * G;
      for (; p + 3 * G < p1; p += 4 * G) {
        int32_t ix[4];
        T vv[4], g4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          ix[u] = a.A.idx[p + u * G];
          vv[u] = a.A.val[p + u * G];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) g4[u] = ld_coherent(&a.res[ix[u]]);
#pragma unroll
        for (int u = 0; u < 4; ++u) acc += vv[u] * g4[u];
      }
      for (; p < p1; p += G) acc += a.A.val[p] * ld_coherent(&a.res[a.A.idx[p]]);
    }
    if constexpr (G <= 64)
      acc = group_sum<G>(acc);
    else
      acc = block_group_sum<G>(acc, red);  // every thread, active or not
    if (active) {
      const double xj = ld_coherent(&a.x[j]);
      const double d = coord_delta<T, LOSS>(a, j, (double)acc, xj, &c);
      if (lane == 0) st_coherent(&a.x[j], xj + d);
    }
    if (!grid_sync(gs, gen, spin_limit)) return;
    if (active) {
      if (c != 0.0) {
        const T ct = (T)c;
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (p0 + lane + r * G < p1) unsafeAtomicAdd(&a.res[ci[r]], ct * cv[r]);
        int32_t p = p0 + lane + R * G;
        for (; p + 3 * G < p1; p += 4 * G) {  // four entries per lane per round trip
          int32_t ix[4];
          T vv[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            ix[u] = a.A.idx[p + u * G];
            vv[u] = a.A.val[p + u * G];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) unsafeAtomicAdd(&a.res[ix[u]], ct * vv[u]);
        }
        for (; p < p1; p += G) unsafeAtomicAdd(&a.res[a.A.idx[p]], ct * a.A.val[p]);
      }
      if (s + 1 < nsteps) prefetch(step0 + s + 1);
    }
    if (!grid_sync(gs, gen, spin_limit)) return;
  }
}

// ------------------------------------------------------------- serial path
// tau = 1: one wavefront runs nsteps consecutive coordinate updates
// (do_single_iteration_serial, loss_square.h:72-83; the SVM twin
// MulticoreSVMDualLoss.h:119-141).  Residual and x traffic uses agent-scope
// atomics and a fence per update so that every update sees all earlier ones.
template <typename T, int LOSS>
__global__ __launch_bounds__(64) void k_serial(StepArgs<T> a, uint64_t step0, uint64_t nsteps) {
  const int lane = threadIdx.x;
  for (uint64_t s = 0; s < nsteps; ++s) {
    const uint32_t j = a.perm(0u, step_keys(a.seed, step0 + s));
    const int32_t p0 = a.A.ptr[j], p1 = a.A.ptr[j + 1];
    T acc = 0;
    for (int32_t p = p0 + lane; p < p1; p += 64)
      acc += a.A.val[p] * __hip_atomic_load(&a.res[a.A.idx[p]], __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
    acc = group_sum<64>(acc);
    const double xj = __hip_atomic_load(&a.x[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double c;
    const double d = coord_delta<T, LOSS>(a, j, (double)acc, xj, &c);
    if (lane == 0)
      __hip_atomic_store(&a.x[j], xj + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (c != 0.0) {
      const T ct = (T)c;
      for (int32_t p = p0 + lane; p < p1; p += 64)
        __hip_atomic_fetch_add(&a.res[a.A.idx[p]], ct * a.A.val[p], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
  }
}

// ----------------------------------------------------- asynchronous workers
// tau workers, one G-lane group each, run `count` updates apiece with no
// synchronisation between them: the reference's multicore executor
// (MulticoreEngineExecutor::executeBulkOfIterations, :83-89; the update
// MulticoreL1RegSquareLoss.h:53-63 with parallel::atomic_add,
// parallel_essentials_posix.h:62-68) and its CUDA twin
// (randomL2L1Solver.h:15-62).  Residual and x are read with agent-scope
// loads, so a worker sees every update that has reached the device-coherent
// level, and written with hardware atomic adds (x_j += d, r += c a_j), so
// no update is lost.  At most tau updates are in flight at once.
//
// Layout for random coordinates: all per-coordinate state is one 32-byte
// record (x_j, the anchored gradient or L_j, the label, the segment's offset
// and length) and the segment is a run of interleaved (index, value) pairs,
// so an update touches one record sector and the few lines of its segment
// instead of five scattered arrays.  FP32 records keep L_j in FP32 (the
// fixed point of the proximal step does not depend on the step length);
// FP64 records keep it in FP64, so that one worker reproduces the serial
// FP64 method.
//
// Pipelining: the record and first R*G entries of update k+1 are loaded
// during update k (the record one update earlier still), just before update
// k's own gather, so a single wait per update covers every load in flight
// (vector memory counters drain in order) and the critical path of an
// update is one round trip.  With TAIL = false every segment fits the R*G
// cached entries (the host checks the longest), which keeps variable-length
// loops out of the update.  With tau = 1 the worker also waits for its
// scatter before the next gather, which makes the run the serial method.
template <typename T>
struct AsyncRec;
template <>
struct alignas(32) AsyncRec<float> {
  double x;
  double gref;  // A^T r64 at the last refresh
  float L;
  float y;      // SVM label
  int32_t p0, len;
};
template <>
struct alignas(32) AsyncRec<double> {
  double x;
  double L;
  float y;
  float pad;
  int32_t p0, len;
};
__device__ __forceinline__ double rec_L(const AsyncRec<float> &r) { return (double)r.L; }
__device__ __forceinline__ double rec_L(const AsyncRec<double> &r) { return r.L; }
__device__ __forceinline__ double rec_gref(const AsyncRec<float> &r) { return r.gref; }
__device__ __forceinline__ double rec_gref(const AsyncRec<double> &) { return 0.0; }

template <typename T>
struct AsyncArgs {
  AsyncRec<T> *rec;
  const Entry<T> *ent;  // nnz entries and two of zero padding
  int64_t nnz;
  T *res;  // FP64: residual; FP32: correction dr
  StepScalars sc;
  uint64_t seed;
  int64_t tau;
  uint32_t n;
};

constexpr int kAsyncBlock = 256;

template <typename T, int G, int LOSS, int R, bool TAIL>
__global__ __launch_bounds__(kAsyncBlock) void k_async(AsyncArgs<T> a, uint64_t k0,
                                                       uint64_t count) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t w = ((int64_t)blockIdx.x * kAsyncBlock + threadIdx.x) / G;
  if (w >= a.tau || count == 0) return;  // whole groups leave together
  const bool serial = a.tau == 1;
  struct Slot {
    uint32_t j;
    AsyncRec<T> r;
    int32_t ci[R];
    T cv[R];
  };
  AsyncCoordCache<G> coords;  // one generator call per lane serves 2G updates
  auto record = [&](uint64_t k, Slot &c) {
    c.j = coords.get(a.seed, (uint32_t)w, k0 + k, a.n);
    c.r = a.rec[c.j];
  };
  auto entries = [&](Slot &c) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int32_t q = lane + r * G;
      if (q < c.r.len) {
        const Entry<T> e = a.ent[c.r.p0 + q];
        c.ci[r] = e.row;
        c.cv[r] = e.val;
      } else {
        c.ci[r] = 0;
        c.cv[r] = (T)0;
      }
    }
  };
  // update k on cur while nxt (update k+1) gets its entries and the record
  // of update k+2 lands in ahead, which then becomes cur's next content
  auto update = [&](uint64_t k, Slot &cur, Slot &nxt, Slot &ahead) {
    if (k + 1 < count) entries(nxt);
    if (k + 2 < count) record(k + 2, ahead);
    T v[R];
#pragma unroll
    for (int r = 0; r < R; ++r)  // padding lanes issue no load
      v[r] = lane + r * G < cur.r.len ? ld_coherent(&a.res[cur.ci[r]]) : (T)0;
    const double xj = ld_coherent(&a.rec[cur.j].x);
    T acc = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) acc += cur.cv[r] * v[r];
    if (TAIL)  // entries past the cached ones, four chunks per round trip
      for (int32_t q0 = lane + R * G; q0 < cur.r.len; q0 += 4 * G) {
        Entry<T> e[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) e[u] = a.ent[cur.r.p0 + min(q0 + u * G, cur.r.len - 1)];
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = ld_coherent(&a.res[e[u].row]);
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (q0 + u * G < cur.r.len) acc += e[u].val * v[u];
      }
    acc = group_sum<G>(acc);
    double c;
    const double g = rec_gref(cur.r) + (double)acc;
    if constexpr (LOSS == kLasso) {
      const double d = prox_delta<LOSS>(g, xj, rec_L(cur.r), (double)cur.r.y, a.sc, &c);
      if (lane == 0 && d != 0.0) unsafeAtomicAdd(&a.rec[cur.j].x, d);
    } else {
      // constrained coordinates (x >= 0, alpha in [0,1]): the new value is
      // computed from the x_j it replaces, by compare-and-swap, so that two
      // workers racing on one coordinate cannot leave it infeasible; the
      // scatter applies the delta of the step that was swapped in
      double d = 0.0;
      if (lane == 0) {
        unsigned long long *px = reinterpret_cast<unsigned long long *>(&a.rec[cur.j].x);
        double old = xj;
        for (;;) {
          double cc;
          const double dd = prox_delta<LOSS>(g, old, rec_L(cur.r), (double)cur.r.y, a.sc, &cc);
          const double nw = old + dd;
          if (nw == old) {  // x_j unchanged; the residual still takes dd
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
      c = LOSS == kSvm ? d * (double)cur.r.y * a.sc.inv_lambda_n : d;
    }
    if (c != 0.0) {
      const T ct = (T)c;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (lane + r * G < cur.r.len) unsafeAtomicAdd(&a.res[cur.ci[r]], ct * cur.cv[r]);
      if (TAIL)
        for (int32_t q0 = lane + R * G; q0 < cur.r.len; q0 += 4 * G) {
          Entry<T> e[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) e[u] = a.ent[cur.r.p0 + min(q0 + u * G, cur.r.len - 1)];
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (q0 + u * G < cur.r.len) unsafeAtomicAdd(&a.res[e[u].row], ct * e[u].val);
        }
    }
    if (serial) wait_memory();
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

// ------------------------------------------------ staged asynchronous workers
// The same workers as k_async, one per 64-lane wave, with the data of each
// update staged in LDS by a loader wave so that a worker waits on nothing but
// its residual gathers.
//
// Why: a wave's vector-memory counter retires in issue order, so in k_async
// the gathers of update k wait for the record and entry prefetches of the
// next updates issued before them -- an HBM round trip per update (~2 us
// under load, against ~0.12 us for an L2-resident gather).  Here a block is
// kStagedWorkers worker waves plus one loader wave that walks the workers'
// coordinate streams ahead of them and copies, with global_load_lds_dwordx4
// (straight into LDS, no registers), each coming update's 32-byte record
// kStagedLeadRec phases before its use and, kStagedLead phases before its
// use, its entries (the bounds come from the record copied three phases
// earlier).  One s_barrier per phase keeps the rings in step: at phase p
// worker w handles its update p - kStagedLeadRec.
//
// The loader issues the same kStagedLoadsPerPhase copies every phase
// (lanes, updates and workers out of range re-load valid addresses), so
// "s_waitcnt vmcnt(2 * kStagedLoadsPerPhase)" at the end of a phase retires
// exactly the copies issued two phases before.  The copies are written in
// inline assembly: the compiler then does not see LDS writes in flight and
// does not wait for all of them before the loader's own LDS reads of the
// bounds (which the explicit waits already order).  The loader issues no
// other vector-memory instruction.
//
// A worker reads its record and entries from LDS, reads x_j and gathers the
// residual with agent-scope loads, and applies the update as k_async does
// (hardware atomics; compare-and-swap for the constrained losses).
//
// LDS: kStagedDepth entry slots -- per worker kStagedSlotBytes (FP32: 256
// entries from the even entry at or below the segment start; FP64: 128) --
// and kStagedRecDepth record slots of 1 KiB (8 records in the first 256
// bytes; the other lanes of the copy write duplicates behind them).
// Segments longer than a slot read the rest from global memory (TAIL).
constexpr int kStagedWorkers = 8;
constexpr int kStagedBlock = 64 * (kStagedWorkers + 1);
constexpr int kStagedDepth = 4;                  // entry slots
constexpr int kStagedLead = kStagedDepth - 1;    // phases from an entry copy to its use
constexpr int kStagedLeadRec = 2 * kStagedLead;  // phases from a record copy to its use
constexpr int kStagedRecDepth = 8;               // record slots (> kStagedLeadRec)
constexpr int kStagedSlotBytes = 2048;
constexpr int kStagedRecBytes = 1024;
constexpr int kStagedLds =
    kStagedDepth * kStagedWorkers * kStagedSlotBytes + kStagedRecDepth * kStagedRecBytes;
// copies per phase: one record copy (16 lanes hold the 8 records), two entry
// copies per worker
constexpr int kStagedLoadsPerPhase = 1 + 2 * kStagedWorkers;
static_assert(3 * kStagedLoadsPerPhase <= 63, "the vector-memory counter holds 63 loads");
static_assert(kStagedRecDepth > kStagedLeadRec, "a record slot outlives its use");
// entries of a segment a slot holds whatever its alignment (FP32: the copy
// starts at the even entry at or below the segment's first)
template <typename T>
constexpr int kStagedCached = kStagedSlotBytes / (int)sizeof(Entry<T>) - (sizeof(T) == 4 ? 1 : 0);

// 16 bytes per lane from global memory to LDS byte address lds + 16 * lane
// (lds wave-uniform).  The instruction has no result register.
__device__ __forceinline__ void copy16_to_lds(const void *src, uint32_t lds) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(src), "s"(lds)
               : "memory", "m0");
}

// Sum over the whole wave with VALU lane moves only (DPP inside rows of 16,
// then row and half swaps), so no LDS instruction waits behind outstanding
// loads.  Every lane adds the same two values in the same order at each
// step, so all lanes receive the bitwise-identical total.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v += dpp_move<kDppXor1>(v);
  v += dpp_move<kDppXor2>(v);
  v += dpp_move<kDppHalfMirror>(v);
  v += dpp_move<kDppMirror>(v);
  auto swap_sum = [](T x, auto swap) -> T {
    if constexpr (sizeof(T) == 4) {
      const auto r = swap(__float_as_uint(x), __float_as_uint(x));
      return __uint_as_float(r[0]) + __uint_as_float(r[1]);
    } else {
      const unsigned long long b = (unsigned long long)__double_as_longlong(x);
      const auto lo = swap((unsigned)b, (unsigned)b);
      const auto hi = swap((unsigned)(b >> 32), (unsigned)(b >> 32));
      const double d0 = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
      const double d1 = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
      return d0 + d1;
    }
  };
  // rows 0,1 <-> rows 2,3 pairs: lanes of rows 0 and 1 get r0 + r1, of rows 2 and 3 r2 + r3
  v = swap_sum(v, [](unsigned a, unsigned b) { return __builtin_amdgcn_permlane16_swap(a, b, false, false); });
  // halves: every lane gets (r0 + r1) + (r2 + r3)
  v = swap_sum(v, [](unsigned a, unsigned b) { return __builtin_amdgcn_permlane32_swap(a, b, false, false); });
  return v;
}

// The update of one worker wave once its partial dot is reduced: the
// proximal step on x_j (hardware atomic add for the Lasso, compare-and-swap
// against the x_j it replaces for the constrained losses, as in k_async) and
// the coefficient of the residual scatter.
template <typename T, int LOSS>
__device__ __forceinline__ double async_apply(const AsyncArgs<T> &a, uint32_t j, double g,
                                              double xj, double Lj, double yj, int lane) {
  double c;
  if constexpr (LOSS == kLasso) {
    const double d = prox_delta<LOSS>(g, xj, Lj, yj, a.sc, &c);
    if (lane == 0 && d != 0.0) unsafeAtomicAdd(&a.rec[j].x, d);
  } else {
    double d = 0.0;
    if (lane == 0) {
      unsigned long long *px = reinterpret_cast<unsigned long long *>(&a.rec[j].x);
      double old = xj;
      for (;;) {
        double cc;
        const double dd = prox_delta<LOSS>(g, old, Lj, yj, a.sc, &cc);
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
    d = __shfl(d, 0, 64);
    c = LOSS == kSvm ? d * yj * a.sc.inv_lambda_n : d;
  }
  return c;
}

template <typename T, int LOSS, bool TAIL>
__global__ __launch_bounds__(kStagedBlock) void k_async_staged(AsyncArgs<T> a, uint64_t k0,
                                                               uint64_t count) {
  __shared__ __attribute__((aligned(16))) unsigned char ring[kStagedLds];
  constexpr int kSlotEntries = kStagedSlotBytes / (int)sizeof(Entry<T>);
  constexpr int R = kSlotEntries / 64;  // slot entries per lane
  const int64_t nphase = (int64_t)count + kStagedLeadRec;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  unsigned char *const rec_ring = ring + kStagedDepth * kStagedWorkers * kStagedSlotBytes;
  if (count == 0) return;  // uniform: every wave leaves before any barrier

  if (wave < kStagedWorkers) {
    // ---- worker wave
    const int64_t w = (int64_t)blockIdx.x * kStagedWorkers + wave;
    const bool active = w < a.tau;
    const bool serial = a.tau == 1;
    AsyncCoordCache<64> coords;
    for (int64_t ph = 0; ph < nphase; ++ph) {
      const int64_t u = ph - kStagedLeadRec;
      if (active && u >= 0) {
        const uint32_t j = coords.get(a.seed, (uint32_t)w, k0 + (uint64_t)u, a.n);
        const AsyncRec<T> r = *(const AsyncRec<T> *)(rec_ring +
                                                     (int)(u % kStagedRecDepth) * kStagedRecBytes +
                                                     wave * 32);
        const int off = sizeof(T) == 4 ? (r.p0 & 1) : 0;
        const Entry<T> *eb =
            (const Entry<T> *)(ring + ((int)(u % kStagedDepth) * kStagedWorkers + wave) *
                                          kStagedSlotBytes) + off;
        const int32_t cached = min(r.len, kSlotEntries - off);
        int32_t ci[R];
        T cv[R], v[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {  // all slot reads at once; entries past the segment zeroed
          const Entry<T> e = eb[lane + 64 * k];
          const bool in = lane + 64 * k < cached;
          ci[k] = in ? e.row : 0;
          cv[k] = in ? e.val : (T)0;
        }
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = lane + 64 * k < cached ? ld_coherent(&a.res[ci[k]]) : (T)0;
        const double xj = ld_coherent(&a.rec[j].x);
        T acc = 0;
#pragma unroll
        for (int k = 0; k < R; ++k) acc += cv[k] * v[k];
        if (TAIL)  // entries beyond the slot, four chunks per round trip
          for (int32_t q0 = cached + lane; q0 < r.len; q0 += 256) {
            Entry<T> e[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) e[k] = a.ent[r.p0 + min(q0 + 64 * k, r.len - 1)];
            T g4[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) g4[k] = ld_coherent(&a.res[e[k].row]);
#pragma unroll
            for (int k = 0; k < 4; ++k)
              if (q0 + 64 * k < r.len) acc += e[k].val * g4[k];
          }
        acc = wave_sum(acc);
        const double c = async_apply<T, LOSS>(a, j, rec_gref(r) + (double)acc, xj, rec_L(r),
                                              (double)r.y, lane);
        if (c != 0.0) {
          const T ct = (T)c;
#pragma unroll
          for (int k = 0; k < R; ++k)
            if (lane + 64 * k < cached) unsafeAtomicAdd(&a.res[ci[k]], ct * cv[k]);
          if (TAIL)
            for (int32_t q = cached + lane; q < r.len; q += 64) {
              const Entry<T> e = a.ent[r.p0 + q];
              unsafeAtomicAdd(&a.res[e.row], ct * e.val);
            }
        }
        if (serial) wait_memory();  // one worker: the serial method
      }
      __builtin_amdgcn_s_barrier();
    }
  } else {
    // ---- loader wave.  Phase ph: copy the record of update ph
    // (kStagedLeadRec phases ahead of the workers) into record slot
    // ph mod kStagedRecDepth, then the entries of update ph - kStagedLead,
    // whose record was copied kStagedLead phases ago, into entry slot
    // (ph - kStagedLead) mod kStagedDepth.
    const uint32_t ring0 = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) unsigned char *)ring;
    const uint32_t rec0 = ring0 + kStagedDepth * kStagedWorkers * kStagedSlotBytes;
    const int rw = (lane & 15) >> 1;  // record copy: worker of this lane (two lanes per record)
    const int64_t w = (int64_t)blockIdx.x * kStagedWorkers + rw;
    const uint32_t wid = w < a.tau ? (uint32_t)w : 0u;
    const int bw = lane & (kStagedWorkers - 1);  // bounds read: worker of this lane
    const int64_t last_u = (int64_t)count - 1;
    for (int64_t ph = 0; ph < nphase; ++ph) {
      {
        const uint64_t u = (uint64_t)min(ph, last_u);
        const uint32_t j = async_coord(a.seed, wid, k0 + u, a.n);
        copy16_to_lds(reinterpret_cast<const char *>(&a.rec[j]) + 16 * (lane & 1),
                      rec0 + (uint32_t)(ph % kStagedRecDepth) * kStagedRecBytes);
      }
      const int64_t u2 = ph - kStagedLead;
      int2 b = make_int2(0, 0);  // (offset, length) of worker bw's segment
      if (u2 >= 0)
        b = *(const int2 *)(rec_ring + (int)(min(u2, last_u) % kStagedRecDepth) * kStagedRecBytes +
                            bw * 32 + 24);
      const uint32_t slot0 =
          ring0 + (uint32_t)((((u2 % kStagedDepth) + kStagedDepth) % kStagedDepth) * kStagedWorkers) *
                      kStagedSlotBytes;
#pragma unroll
      for (int i = 0; i < kStagedWorkers; ++i) {
        // bounds kept inside the entry array whatever LDS holds
        const int32_t p0 = min(max(__builtin_amdgcn_readlane(b.x, i), 0), (int32_t)a.nnz);
        const int32_t len = min(max(__builtin_amdgcn_readlane(b.y, i), 0), (int32_t)a.nnz - p0);
        const uint32_t dst = slot0 + (uint32_t)i * kStagedSlotBytes;
        if constexpr (sizeof(T) == 4) {
          // entry pairs from the even entry at or below p0; lanes past the
          // segment re-load its last pair
          const int32_t base = p0 & ~1;
          const int32_t last = max(0, (p0 + len - 1) & ~1) - base;
#pragma unroll
          for (int c = 0; c < 2; ++c)
            copy16_to_lds(a.ent + base + min(2 * (64 * c + lane), last), dst + 1024 * c);
        } else {
          const int32_t last = max(0, len - 1);
#pragma unroll
          for (int c = 0; c < 2; ++c) copy16_to_lds(a.ent + p0 + min(64 * c + lane, last), dst + 1024 * c);
        }
      }
      // retire the copies of two phases ago: the entries the workers use
      // next phase and the records whose bounds are read next phase
      asm volatile("s_waitcnt vmcnt(%0)" ::"n"(2 * kStagedLoadsPerPhase) : "memory");
      __builtin_amdgcn_s_barrier();
    }
    wait_memory();  // no copy into LDS may outlive the block
  }
}

// record set-up and the x / anchor exchange with the solver's plain arrays
template <typename T>
__global__ __launch_bounds__(kBlock) void k_rec_init(AsyncRec<T> *__restrict__ rec, int64_t n,
                                                     const int32_t *__restrict__ ptr,
                                                     const double *__restrict__ L,
                                                     const double *__restrict__ y) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  AsyncRec<T> r{};
  r.L = L[j];
  r.y = y ? (float)y[j] : 0.0f;
  r.p0 = ptr[j];
  r.len = ptr[j + 1] - ptr[j];
  rec[j] = r;
}
// x (and the FP32 anchor gref) into the records
template <typename T>
__global__ __launch_bounds__(kBlock) void k_rec_put(AsyncRec<T> *__restrict__ rec, int64_t n,
                                                    const double *__restrict__ x,
                                                    const double *__restrict__ gref) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  rec[j].x = x[j];
  if constexpr (sizeof(T) == 4) rec[j].gref = gref[j];
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_rec_get_x(const AsyncRec<T> *__restrict__ rec,
                                                      int64_t n, double *__restrict__ x) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j < n) x[j] = rec[j].x;
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pack_entries(const int32_t *__restrict__ idx,
                                                         const T *__restrict__ val, int64_t nnz,
                                                         Entry<T> *__restrict__ ent) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= nnz) return;
  Entry<T> e{};
  e.row = idx[p];
  e.val = val[p];
  ent[p] = e;
}

// ------------------------------------------------------------------- setup
// L_c = 1/(beta ||segment c||^2) in FP64, 0 for an empty segment
// (loss_square.h:183-194, MulticoreSVMDualLoss.h:48-61).
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_lipschitz(Seg<T> A, int64_t nseg, double beta,
                                                      double *__restrict__ L) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t c = t / G;
  const int lane = threadIdx.x & (G - 1);
  if (c >= nseg) return;
  double s = 0;
  for (int32_t p = A.ptr[c] + lane; p < A.ptr[c + 1]; p += G) {
    const double v = (double)A.val[p];
    s += v * v;
  }
  s = group_sum<G>(s);
  if (lane == 0) L[c] = s > 0 ? 1.0 / (beta * s) : 0.0;
}

// ------------------------------------------------------ spectral ESO beta
// Power steps on D^-1/2 A^T A D^-1/2 (segments as columns, D = diag
// ||a_c||^2), as mSDCA.cpp:188-265: y = sum_c a_c (w_c v_c scale) with
// w_c = 1/||a_c|| (0 for an empty segment), then v_c = w_c <a_c, y>.
template <int G>
__global__ __launch_bounds__(kBlock) void k_pm_scatter(Seg<double> A, int64_t nseg,
                                                       const double *__restrict__ v,
                                                       const double *__restrict__ w, double scale,
                                                       double *__restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t c = t / G;
  const int lane = threadIdx.x & (G - 1);
  if (c >= nseg) return;
  const double a = v[c] * w[c] * scale;
  if (a == 0.0) return;
  for (int32_t p = A.ptr[c] + lane; p < A.ptr[c + 1]; p += G)
    unsafeAtomicAdd(&y[A.idx[p]], a * A.val[p]);
}

// v_c = w_c <a_c, y> over a grid-stride of segments; partial[block] = the
// block's sum of v_c^2
template <int G>
__global__ __launch_bounds__(kBlock) void k_pm_gather(Seg<double> A, int64_t nseg,
                                                      const double *__restrict__ y,
                                                      const double *__restrict__ w,
                                                      double *__restrict__ v,
                                                      double *__restrict__ partial) {
  __shared__ double sh[kBlock / 64];
  const int lane = threadIdx.x & (G - 1);
  const int64_t groups = (int64_t)gridDim.x * (kBlock / G);
  double sq = 0;
  for (int64_t c = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / G; c < nseg; c += groups) {
    double g = 0;
    for (int32_t p = A.ptr[c] + lane; p < A.ptr[c + 1]; p += G) g += A.val[p] * y[A.idx[p]];
    g = group_sum<G>(g) * w[c];
    if (lane == 0) {
      v[c] = g;
      sq += g * g;
    }
  }
  sq = group_sum<64>(sq);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = sq;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int k = 0; k < kBlock / 64; ++k) t += sh[k];
    partial[blockIdx.x] = t;
  }
}

// ----------------------------------------------------------------- refresh
// r64 = -b (square losses) or 0 (SVM w), then the nonzero coordinates are
// scattered in with FP64 atomics (column-scatter form of
// MulticoreSquareLoss::recomputeResiduals, MulticoreSquareLoss.h:42-53, and of
// MulticoreSVMDualLoss.h:31-46 with alpha clipped to [0,1]).  Only segments
// of nonzero coordinates are read, so a sparse x costs O(n + nnz(active)).
[[maybe_unused]] static __global__ __launch_bounds__(kBlock) void k_res_init(int64_t nres, const double *__restrict__ b,
                                                     double *__restrict__ r64) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k < nres) r64[k] = b ? -b[k] : 0.0;
}

template <typename T, int G, int LOSS>
__global__ __launch_bounds__(kBlock) void k_refresh_scatter(Seg<T> A, int64_t nseg,
                                                            const double *__restrict__ x,
                                                            const double *__restrict__ y,
                                                            double inv_lambda_n,
                                                            double *__restrict__ r64) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t c = t / G;
  const int lane = threadIdx.x & (G - 1);
  if (c >= nseg) return;
  double coefx = x[c];
  if constexpr (LOSS == kSvm) {
    coefx = coefx > 1.0 ? 1.0 : (coefx < 0.0 ? 0.0 : coefx);
    coefx *= y[c] * inv_lambda_n;
  }
  if (coefx == 0.0) return;
  for (int32_t p = A.ptr[c] + lane; p < A.ptr[c + 1]; p += G)
    unsafeAtomicAdd(&r64[A.idx[p]], (double)A.val[p] * coefx);
}

// The same r64 as k_refresh_scatter (SVM: w from every alpha), gathered by
// residual index from an FP64 transposed copy At (segment i: the entries of
// residual index i, coordinates as indices): r64_i = sum c_j a_ij with the
// coefficients c of k_svm_coef.  One G-lane group per index, no atomics;
// taken where nearly every coordinate is nonzero (SVM duals), so that the
// refresh reads every entry once instead of adding each with an atomic.
template <int G>
__global__ __launch_bounds__(kBlock) void k_seg_dot(Seg<double> At, int64_t nseg,
                                                    const double *__restrict__ c,
                                                    double *__restrict__ out) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t ngroups = (int64_t)gridDim.x * (kBlock / G);
  for (int64_t i = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / G; i < nseg; i += ngroups) {
    const int32_t p0 = At.ptr[i], p1 = At.ptr[i + 1];
    double acc = 0.0;
    int32_t p = p0 + lane;
    for (; p + 3 * G < p1; p += 4 * G) {  // four entries in flight, then their gathers
      int32_t j[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) j[u] = At.idx[p + u * G];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc += At.val[p + u * G] * c[j[u]];
    }
    for (; p < p1; p += G) acc += At.val[p] * c[At.idx[p]];
    acc = group_sum<G>(acc);
    if (lane == 0) out[i] = acc;
  }
}

// c_j = clip(alpha_j, 0, 1) y_j / (lambda n): the factor of sample j in w
// (the SVM branch of k_refresh_scatter)
[[maybe_unused]] static __global__ __launch_bounds__(kBlock) void k_svm_coef(
    int64_t n, const double *__restrict__ x, const double *__restrict__ y, double inv_lambda_n,
    double *__restrict__ c) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const double a = x[j] > 1.0 ? 1.0 : (x[j] < 0.0 ? 0.0 : x[j]);
  c[j] = a * y[j] * inv_lambda_n;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(T *__restrict__ v, int64_t n, T value) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k < n) v[k] = value;
}

// -------------------------------------------------------------- evaluation
// Per-coordinate FP64 statistics with the FP64 residual r64, one partial row
// of kStats doubles per block (fixed order), then k_reduce_rows folds the
// rows in a fixed tree:
//   Lasso: {max|g_j|, sum|x_j|, nnz(x)}
//   NNLS:  {max|min(x_j,g_j)|, sum x_j g_j, min g_j, max|g_j|, sum|x_j|, nnz}
//   SVM:   {sum hinge, #(y_i m_i > 0), sum alpha_i, nnz(alpha)}
// with g = A^T r (the Lasso dual point needs ||A^T r||_inf) and
// m_i = <x_i, w> (MulticoreSVMDualLoss.h:72-88).  If gref is not null the
// per-coordinate g (m_i for the SVM) is stored there: the anchor of the FP32
// path.
constexpr int kStats = 8;

template <int LOSS>
struct EvalMasks;
template <>
struct EvalMasks<kLasso> {
  static constexpr uint32_t max = 1u << 0, min = 0;
};
template <>
struct EvalMasks<kNnls> {
  static constexpr uint32_t max = (1u << 0) | (1u << 3), min = 1u << 2;
};
template <>
struct EvalMasks<kSvm> {
  static constexpr uint32_t max = 0, min = 0;
};

__device__ __forceinline__ double fold(double a, double o, int s, uint32_t maxmask,
                                       uint32_t minmask) {
  if (maxmask & (1u << s)) return o > a ? o : a;
  if (minmask & (1u << s)) return o < a ? o : a;
  return a + o;
}

template <typename T, int G, int LOSS>
__global__ __launch_bounds__(kBlock) void k_eval_coord(Seg<T> A, int64_t nseg,
                                                       const double *__restrict__ r64,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ y,
                                                       double *__restrict__ gref,
                                                       double *__restrict__ partial) {
  // Grid-stride over segments, software-pipelined: the bounds, first
  // R*G entries and x_c of the next segment load while the
  // current one gathers r64, so a segment costs about one round trip.
  constexpr int R = 4;
  double v[kStats];
  for (int s = 0; s < kStats; ++s) v[s] = 0;
  if (LOSS == kNnls) v[2] = INFINITY;
  const int lane = threadIdx.x & (G - 1);
  const int64_t ngroups = (int64_t)gridDim.x * (kBlock / G);
  int64_t c = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / G;
  int32_t p0 = 0, p1 = 0;
  int32_t ci[R];
  T cv[R];
  double xc = 0, yc = 0;
  auto load = [&](int64_t col) {
    p0 = A.ptr[col];
    p1 = A.ptr[col + 1];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int32_t p = p0 + lane + r * G;
      ci[r] = p < p1 ? A.idx[p] : 0;
      cv[r] = p < p1 ? A.val[p] : (T)0;
    }
    xc = x[col];
    yc = LOSS == kSvm ? y[col] : 0.0;
  };
  if (c < nseg) load(c);
  for (; c < nseg; c += ngroups) {
    double g = 0;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (p0 + lane + r * G < p1) g += (double)cv[r] * r64[ci[r]];
    for (int32_t p = p0 + lane + R * G; p < p1; p += G) g += (double)A.val[p] * r64[A.idx[p]];
    const double xcur = xc, ycur = yc;
    if (c + ngroups < nseg) load(c + ngroups);
    g = group_sum<G>(g);
    if (lane == 0) {
      if (gref) gref[c] = g;
      if constexpr (LOSS == kLasso) {
        v[0] = fmax(v[0], fabs(g));
        v[1] += fabs(xcur);
        v[2] += xcur != 0.0;
      } else if constexpr (LOSS == kNnls) {
        v[0] = fmax(v[0], fabs(fmin(xcur, g)));
        v[1] += xcur * g;
        v[2] = fmin(v[2], g);
        v[3] = fmax(v[3], fabs(g));
        v[4] += fabs(xcur);
        v[5] += xcur != 0.0;
      } else {
        const double al = xcur > 1.0 ? 1.0 : (xcur < 0.0 ? 0.0 : xcur);
        const double ym = ycur * g;
        const double h = 1.0 - ym;
        v[0] += h > 0 ? h : 0.0;
        v[1] += ym > 0;
        v[2] += al;
        v[3] += al != 0.0;
      }
    }
  }
  __shared__ double sh[kBlock / G][kStats];
  const int grp = threadIdx.x / G;
  if (lane == 0)
    for (int s = 0; s < kStats; ++s) sh[grp][s] = v[s];
  __syncthreads();
  if (threadIdx.x < kStats) {
    const int s = threadIdx.x;
    double acc = sh[0][s];
    for (int g2 = 1; g2 < kBlock / G; ++g2)
      acc = fold(acc, sh[g2][s], s, EvalMasks<LOSS>::max, EvalMasks<LOSS>::min);
    partial[(int64_t)blockIdx.x * kStats + s] = acc;
  }
}

// Per-residual-entry statistics: {sum r_k^2, sum b_k r_k}.
[[maybe_unused]] static __global__ __launch_bounds__(kBlock) void k_eval_res(int64_t nres, const double *__restrict__ r64,
                                                     const double *__restrict__ b,
                                                     double *__restrict__ partial) {
  double v0 = 0, v1 = 0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < nres; k += stride) {
    const double r = r64[k];
    v0 += r * r;
    if (b) v1 += b[k] * r;
  }
  v0 = group_sum<64>(v0);
  v1 = group_sum<64>(v1);
  __shared__ double sh[kBlock / 64][2];
  if ((threadIdx.x & 63) == 0) {
    sh[threadIdx.x / 64][0] = v0;
    sh[threadIdx.x / 64][1] = v1;
  }
  __syncthreads();
  if (threadIdx.x < kStats) {
    double acc = 0;
    if (threadIdx.x < 2)
      for (int w = 0; w < kBlock / 64; ++w) acc += sh[w][threadIdx.x];
    partial[(int64_t)blockIdx.x * kStats + threadIdx.x] = acc;
  }
}

// Folds nrows partial rows into out[kStats] with one block: thread t folds
// rows t, t+256, ... in order, then a fixed pairwise tree.
// Folds nrows partial rows of kStats into out[0..kStats) in a fixed order:
// thread t reads the flattened rows coalesced, from element t with a stride
// of kReduceBlock (a multiple of kStats, so every element it reads is stat
// t % kStats), then a tree over the threads of each stat.
constexpr int kReduceBlock = 1024;
static_assert(kReduceBlock % kStats == 0, "a thread folds one stat");
[[maybe_unused]] static __global__ __launch_bounds__(kReduceBlock) void k_reduce_rows(
    const double *__restrict__ partial, int64_t nrows, uint32_t maxmask, uint32_t minmask,
    double *__restrict__ out) {
  __shared__ double sh[kReduceBlock];
  const int t = threadIdx.x, s = t % kStats;
  double v = (minmask & (1u << s)) ? INFINITY : ((maxmask & (1u << s)) ? -INFINITY : 0.0);
  const int64_t total = nrows * kStats;
#pragma unroll 8
  for (int64_t k = t; k < total; k += kReduceBlock) v = fold(v, partial[k], s, maxmask, minmask);
  sh[t] = v;
  __syncthreads();
  for (int w = kReduceBlock / 2; w >= kStats; w >>= 1) {
    if (t < w) sh[t] = fold(sh[t], sh[t + w], s, maxmask, minmask);
    __syncthreads();
  }
  if (t < kStats) out[t] = sh[t];
}

// --------------------------------------------------------------- sampling
// out[s*tau + i] = perm_{seed, step0+s}(i): the device side of the host replay.
[[maybe_unused]] static __global__ __launch_bounds__(kBlock) void k_sample(Perm perm, uint64_t seed, uint64_t step0,
                                                   uint64_t nsteps, int64_t tau,
                                                   int32_t *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)nsteps * tau) return;
  const uint64_t s = (uint64_t)(t / tau);
  const uint32_t i = (uint32_t)(t % tau);
  out[t] = (int32_t)perm(i, step_keys(seed, step0 + s));
}

// out[s*tau + w] = coordinate of update step0+s of asynchronous worker w.
[[maybe_unused]] static __global__ __launch_bounds__(kBlock) void k_sample_async(uint32_t n, uint64_t seed,
                                                         uint64_t step0, uint64_t nsteps,
                                                         int64_t tau, int32_t *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)nsteps * tau) return;
  const uint64_t s = (uint64_t)(t / tau);
  const uint32_t w = (uint32_t)(t % tau);
  out[t] = (int32_t)async_coord(seed, w, step0 + s, n);
}

// v *= f (the SVM's w, its anchor and margins scale with 1/lambda).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale(T *__restrict__ v, int64_t n, double f) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t < n) v[t] = (T)((double)v[t] * f);
}

}  // namespace acdc
