// device_common.h -- shared device/host helpers of the gfx950 PCDM kernels:
// Philox4x32-10 counter-based generator, the keyed per-step permutation that
// turns step s into tau distinct coordinates, and wave-level reductions.
//
// The reference samples with per-thread GSL mt19937 streams seeded by thread
// id (cpp/src/utils/randomNumbersUtil.h:83-125, used at
// cpp/src/solver/MulticoreEngineExecutor.h:86) and tau-distinct sets by
// rejection (cpp/src/experiments/cdn_common.h:238-248).  Here coordinate i of
// step s is perm_{key(seed,s)}(i): a pure function of (seed, s, i) that any
// wave can evaluate independently and the host can replay bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace acdc {

// records the text returned by acdc_last_error (defined in solver.hip)
void set_error(const char *msg);

#define ACDC_HD __host__ __device__ __forceinline__

// Philox4x32-10, J. Salmon et al., "Parallel random numbers: as easy as
// 1, 2, 3" (SC'11): multipliers and Weyl increments from the paper.
struct Philox {
  static constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  static constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;

  ACDC_HD static void round_fn(uint32_t &c0, uint32_t &c1, uint32_t &c2,
                               uint32_t &c3, uint32_t k0, uint32_t k1) {
    const uint64_t a = (uint64_t)M0 * c0;
    const uint64_t bb = (uint64_t)M1 * c2;
    const uint32_t n0 = (uint32_t)(bb >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(a >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)bb;
    c3 = (uint32_t)a;
    c0 = n0;
    c2 = n2;
  }

  ACDC_HD static void gen(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int i = 0; i < 10; ++i) {
      round_fn(c[0], c[1], c[2], c[3], k0, k1);
      k0 += W0;
      k1 += W1;
    }
  }
};

// Tag that separates the coordinate stream from the data-generation streams.
constexpr uint32_t kSamplerTag = 0x50434453u;  // "PCDS"

struct StepKeys {
  uint32_t k[4];
};

ACDC_HD StepKeys step_keys(uint64_t seed, uint64_t step) {
  uint32_t c[4] = {(uint32_t)step, (uint32_t)(step >> 32), kSamplerTag, 0u};
  Philox::gen(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  StepKeys s;
  s.k[0] = c[0];
  s.k[1] = c[1];
  s.k[2] = c[2];
  s.k[3] = c[3];
  return s;
}

// Asynchronous workers sample with replacement, each from its own stream,
// like the reference's per-thread generators (randomNumbersUtil.h:83-125
// drawn by gsl_rng_uniform_int, MulticoreEngineExecutor.h:86).  Update k of
// worker w takes 64 bits of Philox4x32-10 at counter (k/2, w, "PCDA") --
// two updates per block of output -- and maps them to [0, n) by the high
// word of the 64 x 32-bit product (multiply-shift; bias below n / 2^64).
constexpr uint32_t kAsyncTag = 0x50434441u;  // "PCDA"

// floor((hi 2^32 + lo) n / 2^64), exact in 64-bit pieces
ACDC_HD uint32_t mulshift64(uint32_t lo, uint32_t hi, uint32_t n) {
  return (uint32_t)(((uint64_t)hi * n + (((uint64_t)lo * n) >> 32)) >> 32);
}

// coordinates of updates 2 blk and 2 blk + 1 of a worker (one Philox block)
ACDC_HD void async_pair(uint64_t seed, uint32_t worker, uint64_t blk, uint32_t n, uint32_t &even,
                        uint32_t &odd) {
  uint32_t c[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), worker, kAsyncTag};
  Philox::gen(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  even = mulshift64(c[0], c[1], n);
  odd = mulshift64(c[2], c[3], n);
}

ACDC_HD uint32_t async_coord(uint64_t seed, uint32_t worker, uint64_t k, uint32_t n) {
  uint32_t e, o;
  async_pair(seed, worker, k >> 1, n, e, o);
  return (k & 1) ? o : e;
}

// A G-lane group's cache of its worker's coordinates: lane l holds the
// Philox block B + l (updates 2(B+l), 2(B+l)+1), so one generator call per
// lane serves 2G updates, handed out by a lane shuffle.  Updates must be
// requested in increasing order by every group of the wave together (the
// refill branch is then uniform).
template <int G>
struct AsyncCoordCache {
  uint64_t B = ~0ull >> 1;  // first request refills
  uint32_t even = 0, odd = 0;
  __device__ __forceinline__ uint32_t get(uint64_t seed, uint32_t worker, uint64_t u, uint32_t n) {
    const uint64_t blk = u >> 1;
    if (blk - B >= (uint64_t)G) {
      B = blk & ~(uint64_t)(G - 1);
      async_pair(seed, worker, B + (threadIdx.x & (G - 1)), n, even, odd);
    }
    return (uint32_t)__shfl((int)((u & 1) ? odd : even), (int)(blk - B), G);
  }
};

// Feistel round function: a 32-bit multiply/xor-shift integer mixer.
ACDC_HD uint32_t mix32(uint32_t x, uint32_t k) {
  uint32_t h = x ^ k;
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}

// Keyed bijection of [0, n): a four-round alternating-xor Feistel network on
// [0, 2^b), 2^b >= n, with cycle walking back into [0, n).  Each round xors a
// function of one half into the other, so every round -- and the network --
// is invertible.  The walk ends because the orbit of i < n returns to i.
struct Perm {
  uint32_t n, hb, ma, mb;

  ACDC_HD static Perm make(uint32_t n) {
    uint32_t b = 1;
    while (b < 32 && ((uint64_t)1 << b) < (uint64_t)n) ++b;
    Perm p;
    p.n = n;
    p.hb = b >> 1;
    const uint32_t ha = b - p.hb;
    p.mb = p.hb == 0 ? 0u : (uint32_t)(((uint64_t)1 << p.hb) - 1);
    p.ma = (uint32_t)(((uint64_t)1 << ha) - 1);
    return p;
  }

  ACDC_HD uint32_t operator()(uint32_t i, const StepKeys &key) const {
    uint32_t v = i;
    do {
      uint32_t lo = v & mb;
      uint32_t hi = hb >= 32 ? 0u : (v >> hb);
      hi ^= mix32(lo, key.k[0]) & ma;
      lo ^= mix32(hi, key.k[1]) & mb;
      hi ^= mix32(lo, key.k[2]) & ma;
      lo ^= mix32(hi, key.k[3]) & mb;
      v = (hi << hb) | lo;
    } while (v >= n);
    return v;
  }
};

// ---------------------------------------------------------------- reductions
// DPP lane moves within rows of 16 lanes (VALU latency, no LDS crossbar):
// quad_perm [1,0,3,2] (lane ^ 1), quad_perm [2,3,0,1] (lane ^ 2),
// row_half_mirror (lane i <- 7 - i within 8), row_mirror (i <- 15 - i).
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;

template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over aligned groups of G lanes (G = 8..64) of one 64-wide wavefront;
// every lane of the group receives the same total.  Inside a row of 16 the
// partial sums combine by DPP moves -- pairs, quads, the mirrored quad, the
// mirrored half -- so each lane adds the same two values in one of two
// orders, and addition commutes exactly; wider groups then add across rows
// with xor shuffles.  The order is fixed, so results are reproducible.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
  static_assert(G == 8 || G == 16 || G == 32 || G == 64, "groups of 8..64 lanes");
  v += dpp_move<kDppXor1>(v);
  v += dpp_move<kDppXor2>(v);
  v += dpp_move<kDppHalfMirror>(v);
  if (G >= 16) v += dpp_move<kDppMirror>(v);
  if (G >= 32) v += __shfl_xor(v, 16, 64);
  if (G >= 64) v += __shfl_xor(v, 32, 64);
  return v;
}

// DPP move in which lanes of rows outside ROWS, and lanes whose source lies
// outside their row (row shifts) or the wave (wave shifts), keep `old`.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_keep(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_keep(double old, double v) {
  const long long o = __double_as_longlong(old), b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)b, CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), CTRL, ROWS, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Segmented inclusive sum across one 64-wide wavefront: lane l receives the
// sum of v over the lanes l' <= l whose key equals its own, for keys that
// never decrease with the lane index (equal keys are then contiguous).
// Hillis-Steele steps of 1, 2, 4, 8 lanes inside each row of 16 (DPP row
// shifts), then the last lane of row 0 and of row 2 into the row after it
// (row_bcast:15), then lane 31 into rows 2 and 3 (row_bcast:31): after the
// steps inside rows, lane l holds the run from max(start of its key, start of
// its row), and each broadcast extends runs whose key continues across the
// row boundary.  `head` is true where the key differs from the previous
// lane's (lane 0 always), read off the same moves.
struct SegSum {
  double sum;
  bool head;
};
__device__ __forceinline__ SegSum seg_scan_sum(double v, int key) {
  constexpr int kNone = -2;  // keys are >= -1
  constexpr int kShr1 = 0x111, kShr2 = 0x112, kShr4 = 0x114, kShr8 = 0x118;
  constexpr int kBcast15 = 0x142, kBcast31 = 0x143;
  const int lane = threadIdx.x & 63;
  int prev;
  {
    const int k2 = dpp_keep<kShr1, 0xF>(kNone, key);
    const double v2 = dpp_keep<kShr1, 0xF>(0.0, v);
    prev = k2;
    if (k2 == key) v += v2;
  }
#define ACDC_SEG_STEP(CTRL, ROWS)                        \
  {                                                      \
    const int k2 = dpp_keep<CTRL, ROWS>(kNone, key);     \
    const double v2 = dpp_keep<CTRL, ROWS>(0.0, v);      \
    if (k2 == key) v += v2;                              \
  }
  ACDC_SEG_STEP(kShr2, 0xF)
  ACDC_SEG_STEP(kShr4, 0xF)
  ACDC_SEG_STEP(kShr8, 0xF)
#undef ACDC_SEG_STEP
  {
    const int k2 = dpp_keep<kBcast15, 0xA>(kNone, key);  // lane 15 -> row 1, lane 47 -> row 3
    const double v2 = dpp_keep<kBcast15, 0xA>(0.0, v);
    if (lane == 16 || lane == 48) prev = k2;
    if (k2 == key) v += v2;
  }
  {
    const int k2 = dpp_keep<kBcast31, 0xC>(kNone, key);  // lane 31 -> rows 2 and 3
    const double v2 = dpp_keep<kBcast31, 0xC>(0.0, v);
    if (lane == 32) prev = k2;
    if (k2 == key) v += v2;
  }
  return SegSum{v, lane == 0 || prev != key};
}

template <int G, typename T>
__device__ __forceinline__ T group_max(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, G);
    v = o > v ? o : v;
  }
  return v;
}

template <int G, typename T>
__device__ __forceinline__ T group_min(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, G);
    v = o < v ? o : v;
  }
  return v;
}

// ---------------------------------------------------------- coordinate steps
// The three proximal rules.  g is the coordinate's partial derivative of the
// smooth part evaluated at the step's state (the column.residual dot for the
// square losses, the sample.w dot for the SVM), L the reciprocal ESO-scaled
// Lipschitz constant.  Returns delta; *coef receives the factor that
// multiplies the coordinate's stored values in the residual scatter.
enum LossId : int { kLasso = 0, kSvm = 2, kNnls = 10 };

struct StepScalars {
  double lambda;     // Lasso lambda / SVM lambda
  double lambda_n;   // SVM: lambda * n
  double inv_lambda_n;
};

template <int LOSS>
__device__ __forceinline__ double prox_delta(double g, double xj, double Lj,
                                             double yj, const StepScalars &s,
                                             double *coef) {
  double d;
  if constexpr (LOSS == kLasso) {
    // sgn(v)(|v| - L lambda)_+ - x: MulticoreL1RegSquareLoss.h:41-51 with
    // compute_soft_treshold, treshhold_functions.h:27-33
    const double t = Lj * s.lambda;
    const double v = xj - Lj * g;
    double a = fabs(v) - t;
    a = a < 0.0 ? 0.0 : a;
    const double sg = (double)((v > 0.0) - (v < 0.0));
    d = sg * a - xj;
    *coef = d;
  } else if constexpr (LOSS == kNnls) {
    // projected coordinate step onto x >= 0 (no reference counterpart)
    const double v = xj - Lj * g;
    d = (v > 0.0 ? v : 0.0) - xj;
    *coef = d;
  } else {
    // box-clipped dual step: MulticoreSVMDualLoss.h:119-141
    d = (1.0 - g * yj) * s.lambda_n * Lj;
    if (d < -xj)
      d = -xj;
    else if (d > 1.0 - xj)
      d = 1.0 - xj;
    *coef = d * yj * s.inv_lambda_n;
  }
  return d;
}

}  // namespace acdc
