// kernels.h -- gfx950 kernels of the PCDM hot path (templates; instantiated
// by the launchers in solver.hip).
//
// Data are "segments": the coordinate-major matrix holds one segment per
// coordinate (a CSC column for Lasso/NNLS, a CSR sample row for the SVM
// dual) and the residual-major matrix one segment per residual entry (a CSR
// row of A; a feature column of the SVM data).  A group of G lanes (G = 8,
// 16, 32 or 64, aligned inside the 64-wide wavefront) owns one segment: its
// lanes stride over the segment with unit stride, so the index and value
// streams are read coalesced, and the partial sums meet in a fixed xor
// butterfly.  MFMA is not used: every kernel here is a sparse gather/scatter
// bound by HBM and L2 traffic.
#pragma once
#include "device_common.h"

namespace acdc {

template <typename T>
struct Seg {
  const int32_t *__restrict__ ptr;  // nseg + 1 offsets
  const int32_t *__restrict__ idx;  // nnz indices
  const T *__restrict__ val;        // nnz values
};

constexpr int kBlock = 256;  // threads per block of the grid-wide kernels

// ------------------------------------------------------------------ phase A
// Synchronous step s, gather half: group -> slot i < tau, coordinate
// j = perm_s(i); g = <segment j, residual>; delta from the loss's proximal
// rule; x_j += delta (coordinates of one step are distinct, so no other group
// touches x_j); the scatter coefficient is staged for phase B.
// Restates compute_update + the x half of performOneCoordinateUpdate
// (MulticoreL1RegSquareLoss.h:41-56, MulticoreSVMDualLoss.h:119-134) with the
// synchronous semantics of cdn_common.h:255-270.
template <typename T, int G, int LOSS>
__global__ __launch_bounds__(kBlock) void k_step_gather(
    Seg<T> A, const T *__restrict__ res, double *__restrict__ x,
    const double *__restrict__ L, const double *__restrict__ y, StepScalars sc,
    Perm perm, uint64_t seed, uint64_t step, int64_t tau,
    int32_t *__restrict__ coord, double *__restrict__ coef) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t slot = t / G;
  const int lane = threadIdx.x & (G - 1);
  if (slot >= tau) return;  // whole groups leave together
  const StepKeys key = step_keys(seed, step);
  const uint32_t j = perm((uint32_t)slot, key);
  const int32_t p0 = A.ptr[j], p1 = A.ptr[j + 1];
  T acc = 0;
  for (int32_t p = p0 + lane; p < p1; p += G) acc += A.val[p] * res[A.idx[p]];
  acc = group_sum<G>(acc);
  if (lane == 0) {
    const double xj = x[j];
    double c;
    const double d =
        prox_delta<LOSS>((double)acc, xj, L[j], LOSS == kSvm ? y[j] : 0.0, sc, &c);
    x[j] = xj + d;
    coord[slot] = (int32_t)j;
    coef[slot] = c;
  }
}

// ------------------------------------------------------------------ phase B
// Scatter half: residual += coef * segment, with hardware float/double
// atomics (parallel::atomic_add CAS loops in the reference,
// parallel_essentials_posix.h:62-78).  A zero coefficient -- the common case
// for Lasso coordinates that stay at zero -- skips the segment entirely.
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_step_scatter(
    Seg<T> A, T *__restrict__ res, const int32_t *__restrict__ coord,
    const double *__restrict__ coef, int64_t tau) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t slot = t / G;
  const int lane = threadIdx.x & (G - 1);
  if (slot >= tau) return;
  const double c = coef[slot];
  if (c == 0.0) return;
  const int32_t j = coord[slot];
  const T ct = (T)c;
  const int32_t p0 = A.ptr[j], p1 = A.ptr[j + 1];
  for (int32_t p = p0 + lane; p < p1; p += G)
    unsafeAtomicAdd(&res[A.idx[p]], ct * A.val[p]);
}

// ------------------------------------------------------------- serial path
// tau = 1: one wavefront runs nsteps consecutive coordinate updates
// (do_single_iteration_serial, loss_square.h:72-83; the SVM twin
// MulticoreSVMDualLoss.h:119-141).  Residual and x traffic uses agent-scope
// atomics and a fence per update so that every update sees all earlier ones.
template <typename T, int LOSS>
__global__ __launch_bounds__(64) void k_serial(
    Seg<T> A, T *__restrict__ res, double *__restrict__ x,
    const double *__restrict__ L, const double *__restrict__ y, StepScalars sc,
    Perm perm, uint64_t seed, uint64_t step0, uint64_t nsteps) {
  const int lane = threadIdx.x;
  for (uint64_t s = 0; s < nsteps; ++s) {
    const StepKeys key = step_keys(seed, step0 + s);
    const uint32_t j = perm(0u, key);
    const int32_t p0 = A.ptr[j], p1 = A.ptr[j + 1];
    T acc = 0;
    for (int32_t p = p0 + lane; p < p1; p += 64)
      acc += A.val[p] * __hip_atomic_load(&res[A.idx[p]], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
    acc = group_sum<64>(acc);
    const double xj =
        __hip_atomic_load(&x[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double c;
    const double d =
        prox_delta<LOSS>((double)acc, xj, L[j], LOSS == kSvm ? y[j] : 0.0, sc, &c);
    if (lane == 0)
      __hip_atomic_store(&x[j], xj + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (c != 0.0) {
      const T ct = (T)c;
      for (int32_t p = p0 + lane; p < p1; p += 64)
        __hip_atomic_fetch_add(&res[A.idx[p]], ct * A.val[p], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
  }
}

// ------------------------------------------------------------------- setup
// L_c = 1/(beta ||segment c||^2) in FP64, 0 for an empty segment
// (loss_square.h:183-194, MulticoreSVMDualLoss.h:48-61).
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_lipschitz(Seg<T> A, int64_t nseg,
                                                      double beta,
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

// ----------------------------------------------------------------- refresh
// FP64 residual from the coordinates, one group per residual entry over the
// residual-major copy (no atomics, fixed summation order):
//   Lasso/NNLS  r_k = sum_j a_kj x_j - b_k          (MulticoreSquareLoss.h:42-53)
//   SVM         w_f = sum_i x_if y_i alpha_i/(lambda n), alpha clipped to [0,1]
//               (MulticoreSVMDualLoss.h:31-46)
// res64 receives the FP64 value, resT (if not null) its rounding to T.
template <typename T, int G, int LOSS>
__global__ __launch_bounds__(kBlock) void k_refresh(
    Seg<T> R, int64_t nres, const double *__restrict__ x,
    const double *__restrict__ b, const double *__restrict__ y,
    double inv_lambda_n, double *__restrict__ res64, T *__restrict__ resT) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t k = t / G;
  const int lane = threadIdx.x & (G - 1);
  if (k >= nres) return;
  double s = 0;
  for (int32_t p = R.ptr[k] + lane; p < R.ptr[k + 1]; p += G) {
    const int32_t c = R.idx[p];
    double coefx;
    if constexpr (LOSS == kSvm) {
      double a = x[c];
      a = a > 1.0 ? 1.0 : (a < 0.0 ? 0.0 : a);
      coefx = a * y[c] * inv_lambda_n;
    } else {
      coefx = x[c];
    }
    s += (double)R.val[p] * coefx;
  }
  s = group_sum<G>(s);
  if (lane == 0) {
    const double r = (LOSS == kSvm) ? s : s - b[k];
    res64[k] = r;
    if (resT) resT[k] = (T)r;
  }
}

// -------------------------------------------------------------- evaluation
// Per-coordinate FP64 statistics with the FP64 residual, reduced to one
// partial row of kStats doubles per block (fixed order; the host adds the
// rows in order).
//   Lasso: {max|g_j|, sum|x_j|, nnz(x)}
//   NNLS:  {max|min(x_j,g_j)|, sum x_j g_j, min g_j, max|g_j|, sum|x_j|, nnz}
//   SVM:   {sum hinge, #(y_i m_i > 0), sum alpha_i, nnz(alpha)}
// with g = A^T r (MulticoreL1RegSquareLoss.h:30-39 needs only ||x||_1; the
// gradient feeds the dual point) and m_i = <x_i, w> (MulticoreSVMDualLoss.h:72-88).
constexpr int kStats = 8;

template <int G>
__device__ __forceinline__ void block_reduce_stats(double (&v)[kStats],
                                                   const uint32_t maxmask,
                                                   const uint32_t minmask,
                                                   double *__restrict__ out) {
  __shared__ double sh[kBlock / G][kStats];
  const int grp = threadIdx.x / G;
  const int lane = threadIdx.x & (G - 1);
  if (lane == 0)
    for (int s = 0; s < kStats; ++s) sh[grp][s] = v[s];
  __syncthreads();
  if (threadIdx.x < kStats) {
    const int s = threadIdx.x;
    double a = sh[0][s];
    for (int g = 1; g < kBlock / G; ++g) {
      const double o = sh[g][s];
      if (maxmask & (1u << s))
        a = o > a ? o : a;
      else if (minmask & (1u << s))
        a = o < a ? o : a;
      else
        a += o;
    }
    out[(int64_t)blockIdx.x * kStats + s] = a;
  }
}

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

template <typename T, int G, int LOSS>
__global__ __launch_bounds__(kBlock) void k_eval_coord(
    Seg<T> A, int64_t nseg, const double *__restrict__ res64,
    const double *__restrict__ x, const double *__restrict__ y,
    double *__restrict__ partial) {
  double v[kStats];
  for (int s = 0; s < kStats; ++s) v[s] = 0;
  if (LOSS == kNnls) v[2] = INFINITY;
  const int lane = threadIdx.x & (G - 1);
  const int64_t ngroups = (int64_t)gridDim.x * (kBlock / G);
  for (int64_t c = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / G; c < nseg;
       c += ngroups) {
    double g = 0;
    for (int32_t p = A.ptr[c] + lane; p < A.ptr[c + 1]; p += G)
      g += (double)A.val[p] * res64[A.idx[p]];
    g = group_sum<G>(g);
    if (lane == 0) {
      const double xc = x[c];
      if constexpr (LOSS == kLasso) {
        v[0] = fmax(v[0], fabs(g));
        v[1] += fabs(xc);
        v[2] += xc != 0.0;
      } else if constexpr (LOSS == kNnls) {
        v[0] = fmax(v[0], fabs(fmin(xc, g)));
        v[1] += xc * g;
        v[2] = fmin(v[2], g);
        v[3] = fmax(v[3], fabs(g));
        v[4] += fabs(xc);
        v[5] += xc != 0.0;
      } else {
        const double a = xc > 1.0 ? 1.0 : (xc < 0.0 ? 0.0 : xc);
        const double ym = y[c] * g;
        const double h = 1.0 - ym;
        v[0] += h > 0 ? h : 0.0;
        v[1] += ym > 0;
        v[2] += a;
        v[3] += a != 0.0;
      }
    }
  }
  block_reduce_stats<G>(v, EvalMasks<LOSS>::max, EvalMasks<LOSS>::min, partial);
}

// Per-residual-entry statistics: {sum r_k^2, sum b_k r_k}.
__global__ __launch_bounds__(kBlock) void k_eval_res(
    int64_t nres, const double *__restrict__ res64,
    const double *__restrict__ b, double *__restrict__ partial) {
  double v[kStats];
  for (int s = 0; s < kStats; ++s) v[s] = 0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < nres; k += stride) {
    const double r = res64[k];
    v[0] += r * r;
    if (b) v[1] += b[k] * r;
  }
  // every lane holds data here: reduce the wave first, then the block
  for (int s = 0; s < 2; ++s) v[s] = group_sum<64>(v[s]);
  __shared__ double sh[kBlock / 64][2];
  if ((threadIdx.x & 63) == 0) {
    sh[threadIdx.x / 64][0] = v[0];
    sh[threadIdx.x / 64][1] = v[1];
  }
  __syncthreads();
  if (threadIdx.x < kStats) {
    double a = 0;
    if (threadIdx.x < 2)
      for (int w = 0; w < kBlock / 64; ++w) a += sh[w][threadIdx.x];
    partial[(int64_t)blockIdx.x * kStats + threadIdx.x] = a;
  }
}

// --------------------------------------------------------------- sampling
// out[s*tau + i] = perm_{seed, step0+s}(i): the device side of the host replay.
__global__ __launch_bounds__(kBlock) void k_sample(Perm perm, uint64_t seed,
                                                   uint64_t step0, uint64_t nsteps,
                                                   int64_t tau,
                                                   int32_t *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)nsteps * tau) return;
  const uint64_t s = (uint64_t)(t / tau);
  const uint32_t i = (uint32_t)(t % tau);
  out[t] = (int32_t)perm(i, step_keys(seed, step0 + s));
}

// Rescale the residual (SVM w) by a factor: w scales with 1/lambda.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale(T *__restrict__ v, int64_t n,
                                                  double f) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t < n) v[t] = (T)((double)v[t] * f);
}

}  // namespace acdc
