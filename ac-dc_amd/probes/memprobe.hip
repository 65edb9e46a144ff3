// memprobe.hip -- device probes behind the roofline figures of the
// asynchronous PCDM kernel: how fast a wave can gather random 4-byte residual
// entries, and where those gathers are served from, for each load scope the
// kernels may use.
//
//   latency   each lane of one wave follows a random cycle through an int32
//             array from its own start (64 dependent gathers per step, as a
//             wave gathers a column's residual rows), ns per step, for
//             working sets that stay in
//             L2 (the 200,000-row residual of the 200,000 x 1,000,000 Lasso,
//             800 KB), that fit the memory-side cache (64 MiB) and that
//             stream from HBM (1 GiB)
//   gather    every lane of 16 waves per CU keeps 8 independent random
//             4-byte gathers in flight over the same working sets:
//             gathers per second for the whole device
//   atomic    random FP32 atomic adds into the 800 KB array at workgroup and
//             agent scope: atomics per second
//
// Load scopes: "plain" (non-atomic global load), "wg" (relaxed atomic load at
// workgroup scope), "agent" (relaxed atomic load at agent scope, the residual
// gathers of the solver).  Usage: acdc_memprobe [latency|gather|atomic|all]
// Prints one line per measurement: "<test> <scope> <bytes> <value> <unit>".
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#define CHECK(x)                                                                   \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                  \
      std::exit(1);                                                                \
    }                                                                              \
  } while (0)

enum Scope { kPlain = 0, kWg = 1, kAgent = 2 };

template <int S>
__device__ __forceinline__ int ld(const int *p) {
  typedef __attribute__((address_space(1))) const int gint;
  if constexpr (S == kPlain) return *(gint *)p;
  if constexpr (S == kWg) return __hip_atomic_load((gint *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  return __hip_atomic_load((gint *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// each lane walks `steps` links of the cycle in `next` (a permutation of
// [0, n), so every index stays in bounds) from starts[lane]
template <int S>
__global__ void k_chase(const int *next, int steps, const int *starts, int *out) {
  int i = starts[threadIdx.x];
  for (int s = 0; s < steps; ++s) i = ld<S>(next + i);
  out[threadIdx.x] = i;
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// each lane: `rounds` rounds of 8 independent gathers at hashed indices
// below n (index = high word of hash * n), summed so no load is dead
template <int S>
__global__ __launch_bounds__(256) void k_gather(const int *a, uint32_t n, int rounds, int *out) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  int acc = 0;
  uint32_t h = hash32(t * 0x9E3779B9u + 1u);
  for (int r = 0; r < rounds; ++r) {
    int v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      h = hash32(h + (uint32_t)u);
      v[u] = ld<S>(a + (uint32_t)(((uint64_t)h * n) >> 32));
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += v[u];
  }
  if (acc == 0x7fffffff) out[0] = acc;  // keeps the loads live
}

template <int S>
__global__ __launch_bounds__(256) void k_atomic(float *a, uint32_t n, int rounds) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  uint32_t h = hash32(t * 0x9E3779B9u + 7u);
  for (int r = 0; r < rounds; ++r) {
    h = hash32(h + 1u);
    float *p = a + (uint32_t)(((uint64_t)h * n) >> 32);
    if constexpr (S == kWg)
      __hip_atomic_fetch_add(p, 1.0f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else
      __hip_atomic_fetch_add(p, 1.0f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

static const char *scope_name(int s) { return s == kPlain ? "plain" : s == kWg ? "wg" : "agent"; }

struct Timer {
  hipEvent_t a, b;
  Timer() {
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
  }
  ~Timer() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  void start() { CHECK(hipEventRecord(a, 0)); }
  float stop_ms() {
    CHECK(hipEventRecord(b, 0));
    CHECK(hipEventSynchronize(b));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, a, b));
    return ms;
  }
};

static void latency(int *dout) {
  const size_t sets[] = {200000, 16u << 20, 256u << 20};  // ints: 800 KB, 64 MiB, 1 GiB
  for (size_t n : sets) {
    std::vector<int> perm(n), next(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(12345);
    std::shuffle(perm.begin(), perm.end(), rng);
    for (size_t k = 0; k < n; ++k) next[perm[k]] = perm[(k + 1) % n];  // one cycle through all
    int *d = nullptr, *st = nullptr;
    CHECK(hipMalloc(&d, n * sizeof(int)));
    CHECK(hipMemcpy(d, next.data(), n * sizeof(int), hipMemcpyHostToDevice));
    std::vector<int> starts(128);  // two sets of 64 lane starts along the cycle
    for (int l = 0; l < 128; ++l) starts[l] = perm[(size_t)l * (n / 128)];
    CHECK(hipMalloc(&st, starts.size() * sizeof(int)));
    CHECK(hipMemcpy(st, starts.data(), starts.size() * sizeof(int), hipMemcpyHostToDevice));
    const int steps = 100000;
    Timer tm;
    for (int s = 0; s < 3; ++s) {
      auto run = [&](auto kern) {
        kern<<<1, 64>>>(d, steps / 10, st + 64, dout);  // warm up
        CHECK(hipDeviceSynchronize());
        tm.start();
        kern<<<1, 64>>>(d, steps, st, dout);
        return tm.stop_ms();
      };
      float ms = s == kPlain ? run(k_chase<kPlain>) : s == kWg ? run(k_chase<kWg>) : run(k_chase<kAgent>);
      std::printf("latency %s %zu %.1f ns\n", scope_name(s), n * sizeof(int), ms * 1e6 / steps);
      std::fflush(stdout);
    }
    CHECK(hipFree(st));
    CHECK(hipFree(d));
  }
}

static void gather(int *dout, int cus) {
  const size_t sets[] = {200000, 16u << 20, 256u << 20};
  for (size_t n : sets) {
    int *d = nullptr;
    CHECK(hipMalloc(&d, n * sizeof(int)));
    CHECK(hipMemset(d, 1, n * sizeof(int)));
    const int blocks = cus * 4, rounds = 256;  // 16 waves per CU
    const double loads = (double)blocks * 256 * rounds * 8;
    Timer tm;
    for (int s = 0; s < 3; ++s) {
      auto run = [&](auto kern) {
        kern<<<blocks, 256>>>(d, (uint32_t)n, rounds / 8, dout);
        CHECK(hipDeviceSynchronize());
        tm.start();
        kern<<<blocks, 256>>>(d, (uint32_t)n, rounds, dout);
        return tm.stop_ms();
      };
      float ms = s == kPlain ? run(k_gather<kPlain>) : s == kWg ? run(k_gather<kWg>) : run(k_gather<kAgent>);
      std::printf("gather %s %zu %.4g gathers/s\n", scope_name(s), n * sizeof(int), loads / (ms * 1e-3));
      std::fflush(stdout);
    }
    CHECK(hipFree(d));
  }
}

static void atomic(int cus) {
  const uint32_t n = 200000;
  float *d = nullptr;
  CHECK(hipMalloc(&d, n * sizeof(float)));
  CHECK(hipMemset(d, 0, n * sizeof(float)));
  const int blocks = cus * 4, rounds = 256;
  const double ops = (double)blocks * 256 * rounds;
  Timer tm;
  for (int s = kWg; s <= kAgent; ++s) {
    auto run = [&](auto kern) {
      kern<<<blocks, 256>>>(d, n, rounds / 8);
      CHECK(hipDeviceSynchronize());
      tm.start();
      kern<<<blocks, 256>>>(d, n, rounds);
      return tm.stop_ms();
    };
    float ms = s == kWg ? run(k_atomic<kWg>) : run(k_atomic<kAgent>);
    std::printf("atomic %s %zu %.4g atomics/s\n", scope_name(s), (size_t)n * 4, ops / (ms * 1e-3));
    std::fflush(stdout);
  }
  CHECK(hipFree(d));
}

int main(int argc, char **argv) {
  const char *what = argc > 1 ? argv[1] : "all";
  int dev = 0, cus = 0;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int *dout = nullptr;
  CHECK(hipMalloc(&dout, 64 * sizeof(int)));
  const bool all = !std::strcmp(what, "all");
  if (all || !std::strcmp(what, "latency")) latency(dout);
  if (all || !std::strcmp(what, "gather")) gather(dout, cus);
  if (all || !std::strcmp(what, "atomic")) atomic(cus);
  CHECK(hipFree(dout));
  return 0;
}
