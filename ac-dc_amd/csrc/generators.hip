// generators.hip -- host-side seeded generators of the benchmark instances
// (include/acdc_gen.h).  Host code only; it shares the Philox4x32-10
// definition with the device sampler.  Columns (or rows) are generated in
// parallel with std::thread: every column draws from its own counter range,
// so the result does not depend on the thread count.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/acdc_gen.h"
#include "device_common.h"

namespace {

// purpose tags of the generator streams
enum : uint32_t {
  kTagCount = 0x434f554eu,  // column/row lengths
  kTagIndex = 0x494e4458u,  // row/feature indices
  kTagValue = 0x56414c55u,  // values
  kTagXstar = 0x58535452u,  // planted solution positions/values
  kTagNoise = 0x4e4f4953u,  // observation noise / label flips
};

// A counter-based stream: draw i of (seed, tag, sub) is word i%4 of
// Philox4x32-10 at counter (i/4, tag, sub).
class Stream {
 public:
  Stream(uint64_t seed, uint32_t tag, uint64_t sub)
      : seed_(seed), tag_(tag), sub_(sub) {}

  uint32_t u32() {
    if (pos_ == 4) {
      buf_[0] = (uint32_t)blk_;
      buf_[1] = (uint32_t)(blk_ >> 32);
      buf_[2] = tag_ ^ (uint32_t)(sub_ >> 32);
      buf_[3] = (uint32_t)sub_;
      acdc::Philox::gen(buf_, (uint32_t)seed_, (uint32_t)(seed_ >> 32));
      ++blk_;
      pos_ = 0;
    }
    return buf_[pos_++];
  }
  // uniform double in [0,1) with 53 random bits
  double uniform() {
    const uint64_t a = u32() >> 5, b = u32() >> 6;
    return (double)(a * 67108864ull + b) * (1.0 / 9007199254740992.0);
  }
  // standard normal by the Box-Muller transform
  double normal() {
    if (have_spare_) {
      have_spare_ = false;
      return spare_;
    }
    const double u1 = 1.0 - uniform();  // (0,1]
    const double u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double t = 6.283185307179586 * u2;
    spare_ = r * std::sin(t);
    have_spare_ = true;
    return r * std::cos(t);
  }
  // unbiased integer in [0, n) (Lemire's multiply-shift with rejection)
  uint32_t below(uint32_t n) {
    uint64_t m = (uint64_t)u32() * n;
    uint32_t l = (uint32_t)m;
    if (l < n) {
      const uint32_t t = (uint32_t)(-n) % n;
      while (l < t) {
        m = (uint64_t)u32() * n;
        l = (uint32_t)m;
      }
    }
    return (uint32_t)(m >> 32);
  }

 private:
  uint64_t seed_;
  uint32_t tag_;
  uint64_t sub_;
  uint64_t blk_ = 0;
  uint32_t buf_[4] = {0, 0, 0, 0};
  int pos_ = 4;
  bool have_spare_ = false;
  double spare_ = 0;
};

// k distinct values of [0, n) in increasing order (Floyd's algorithm)
void distinct_sorted(Stream &st, uint32_t n, uint32_t k, std::vector<int32_t> &out) {
  out.clear();
  if (k >= n) {
    for (uint32_t i = 0; i < n; ++i) out.push_back((int32_t)i);
    return;
  }
  for (uint32_t i = n - k; i < n; ++i) {
    const uint32_t t = st.below(i + 1);
    const bool seen = std::find(out.begin(), out.end(), (int32_t)t) != out.end();
    out.push_back(seen ? (int32_t)i : (int32_t)t);
  }
  std::sort(out.begin(), out.end());
}

template <class F>
void parallel_for(int64_t n, F f) {
  unsigned nt = std::thread::hardware_concurrency();
  nt = std::max(1u, std::min(nt, 32u));
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const int64_t lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([=]() {
      for (int64_t i = lo; i < hi; ++i) f(i);
    });
  }
  for (auto &x : th) x.join();
}

// discrete sampler on [lo, hi] with P(k) ~ k^-s (inverse CDF by bisection)
struct ZipfTable {
  int32_t lo = 1;
  std::vector<double> cdf;
  ZipfTable(int32_t lo_, int32_t hi_, double s) : lo(lo_) {
    double acc = 0;
    for (int32_t k = lo_; k <= hi_; ++k) {
      acc += std::pow((double)k, -s);
      cdf.push_back(acc);
    }
    for (double &c : cdf) c /= acc;
  }
  int32_t draw(double u) const {
    const auto it = std::upper_bound(cdf.begin(), cdf.end(), u);
    const size_t i = std::min((size_t)(it - cdf.begin()), cdf.size() - 1);
    return lo + (int32_t)i;
  }
};

// Zipf(s) on [1, inf) (Devroye's rejection method for the zeta law), then
// clipped to [lo, hi]
int32_t zipf_clipped(Stream &st, double s, int32_t lo, int32_t hi) {
  const double b = std::pow(2.0, s - 1.0);
  for (;;) {
    const double u = 1.0 - st.uniform(), v = st.uniform();
    const double xx = std::floor(std::pow(u, -1.0 / (s - 1.0)));  // u >= 2^-53: finite
    // t - 1 through expm1/log1p so that x (t - 1) stays exact for huge x
    const double tm1 = std::expm1((s - 1.0) * std::log1p(1.0 / xx));
    const double t = 1.0 + tm1;
    if (v * xx * tm1 / (b - 1.0) <= t / b) {
      const double k = std::min(std::max(xx, (double)lo), (double)hi);
      return (int32_t)k;
    }
  }
}

}  // namespace

struct acdc_dataset {
  acdc_dataset_info info{};
  std::vector<int32_t> ptr, idx;
  std::vector<double> vals, rhs, xstar;
};

namespace {
acdc_status gfail(acdc_status st, const char *msg) {
  acdc::set_error(msg);
  return st;
}
}  // namespace

extern "C" acdc_status acdc_gen_lasso(const acdc_gen_lasso_cfg *cfg, acdc_dataset **out) {
  if (!cfg || !out) return gfail(ACDC_ERR_INVALID, "acdc_gen_lasso: null argument");
  const int64_t m = cfg->m, n = cfg->n;
  if (m < 1 || n < 1 || m >= INT32_MAX || n >= INT32_MAX || cfg->xstar_nnz < 0 ||
      cfg->xstar_nnz > n || !(cfg->noise >= 0) || !(cfg->lambda_frac >= 0))
    return gfail(ACDC_ERR_INVALID, "acdc_gen_lasso: bad sizes");
  if (cfg->nnz_law == ACDC_NNZ_FIXED) {
    if (cfg->nnz_fixed < 0) return gfail(ACDC_ERR_INVALID, "acdc_gen_lasso: nnz_fixed < 0");
  } else if (cfg->nnz_law == ACDC_NNZ_ZIPF_TRUNC || cfg->nnz_law == ACDC_NNZ_ZIPF_CLIP) {
    if (cfg->nnz_lo < 1 || cfg->nnz_hi < cfg->nnz_lo || !(cfg->zipf_s > 1.0))
      return gfail(ACDC_ERR_INVALID, "acdc_gen_lasso: Zipf law needs 1 <= lo <= hi and s > 1");
  } else {
    return gfail(ACDC_ERR_INVALID, "acdc_gen_lasso: unknown nnz law");
  }
  acdc_dataset *d = new (std::nothrow) acdc_dataset();
  if (!d) return gfail(ACDC_ERR_NOMEM, "acdc_gen_lasso: out of memory");
  try {
    const uint64_t seed = cfg->seed;
    std::vector<int32_t> cnt(n);
    const ZipfTable zt(std::max(1, cfg->nnz_lo), std::max(cfg->nnz_lo, cfg->nnz_hi),
                       cfg->zipf_s > 1 ? cfg->zipf_s : 2.0);
    parallel_for(n, [&](int64_t j) {
      Stream st(seed, kTagCount, (uint64_t)j);
      int32_t k;
      if (cfg->nnz_law == ACDC_NNZ_FIXED)
        k = cfg->nnz_fixed;
      else if (cfg->nnz_law == ACDC_NNZ_ZIPF_TRUNC)
        k = zt.draw(st.uniform());
      else
        k = zipf_clipped(st, cfg->zipf_s, cfg->nnz_lo, cfg->nnz_hi);
      cnt[j] = (int32_t)std::min<int64_t>(k, m);
    });
    d->ptr.resize(n + 1);
    d->ptr[0] = 0;
    int64_t tot = 0;
    for (int64_t j = 0; j < n; ++j) {
      tot += cnt[j];
      if (tot > INT32_MAX) {
        delete d;
        return gfail(ACDC_ERR_INVALID, "acdc_gen_lasso: more than 2^31-1 nonzeros");
      }
      d->ptr[j + 1] = (int32_t)tot;
    }
    d->idx.resize(tot);
    d->vals.resize(tot);
    parallel_for(n, [&](int64_t j) {
      Stream si(seed, kTagIndex, (uint64_t)j), sv(seed, kTagValue, (uint64_t)j);
      std::vector<int32_t> rows;
      distinct_sorted(si, (uint32_t)m, (uint32_t)cnt[j], rows);
      double nrm = 0;
      const int64_t p0 = d->ptr[j];
      for (size_t q = 0; q < rows.size(); ++q) {
        const double v = sv.normal();
        d->idx[p0 + q] = rows[q];
        d->vals[p0 + q] = v;
        nrm += v * v;
      }
      if (nrm > 0) {
        const double inv = 1.0 / std::sqrt(nrm);
        for (size_t q = 0; q < rows.size(); ++q) d->vals[p0 + q] *= inv;
      }
    });
    // planted solution
    d->xstar.assign(n, 0.0);
    {
      Stream sx(seed, kTagXstar, 0), svx(seed, kTagXstar, 1);
      std::vector<int32_t> pos;
      distinct_sorted(sx, (uint32_t)n, (uint32_t)cfg->xstar_nnz, pos);
      for (int32_t j : pos)
        d->xstar[j] = cfg->xstar_normal ? svx.normal() : (svx.u32() & 1u ? 1.0 : -1.0);
    }
    // b = A x* + noise
    d->rhs.assign(m, 0.0);
    for (int64_t j = 0; j < n; ++j) {
      const double xj = d->xstar[j];
      if (xj == 0) continue;
      for (int32_t p = d->ptr[j]; p < d->ptr[j + 1]; ++p) d->rhs[d->idx[p]] += d->vals[p] * xj;
    }
    {
      Stream sn(seed, kTagNoise, 0);
      for (int64_t i = 0; i < m; ++i) d->rhs[i] += cfg->noise * sn.normal();
    }
    // lambda_max = ||A^T b||_inf
    std::vector<double> colmax(n);
    parallel_for(n, [&](int64_t j) {
      double g = 0;
      for (int32_t p = d->ptr[j]; p < d->ptr[j + 1]; ++p) g += d->vals[p] * d->rhs[d->idx[p]];
      colmax[j] = std::fabs(g);
    });
    double lmax = 0;
    for (double g : colmax) lmax = std::max(lmax, g);
    d->info.kind = 0;
    d->info.m = m;
    d->info.n = n;
    d->info.nnz = tot;
    d->info.lambda_max = lmax;
    d->info.lambda = cfg->lambda_frac * lmax;
    d->info.xstar_nnz = cfg->xstar_nnz;
  } catch (const std::bad_alloc &) {
    delete d;
    return gfail(ACDC_ERR_NOMEM, "acdc_gen_lasso: out of memory");
  }
  *out = d;
  return ACDC_OK;
}

extern "C" acdc_status acdc_gen_svm(const acdc_gen_svm_cfg *cfg, acdc_dataset **out) {
  if (!cfg || !out) return gfail(ACDC_ERR_INVALID, "acdc_gen_svm: null argument");
  const int64_t ns = cfg->n_samples, nf = cfg->n_features;
  if (ns < 1 || nf < 1 || ns >= INT32_MAX || nf >= INT32_MAX || !(cfg->row_median > 0) ||
      !(cfg->row_sigma >= 0) || cfg->wstar_nnz < 0 || cfg->wstar_nnz > nf ||
      !(cfg->flip_frac >= 0 && cfg->flip_frac <= 1))
    return gfail(ACDC_ERR_INVALID, "acdc_gen_svm: bad configuration");
  acdc_dataset *d = new (std::nothrow) acdc_dataset();
  if (!d) return gfail(ACDC_ERR_NOMEM, "acdc_gen_svm: out of memory");
  try {
    const uint64_t seed = cfg->seed;
    std::vector<int32_t> cnt(ns);
    const double mu = std::log(cfg->row_median);
    parallel_for(ns, [&](int64_t i) {
      Stream st(seed, kTagCount, (uint64_t)i);
      double k = std::round(std::exp(mu + cfg->row_sigma * st.normal()));
      k = std::min(std::max(k, 1.0), (double)nf);
      cnt[i] = (int32_t)k;
    });
    d->ptr.resize(ns + 1);
    d->ptr[0] = 0;
    int64_t tot = 0;
    for (int64_t i = 0; i < ns; ++i) {
      tot += cnt[i];
      if (tot > INT32_MAX) {
        delete d;
        return gfail(ACDC_ERR_INVALID, "acdc_gen_svm: more than 2^31-1 nonzeros");
      }
      d->ptr[i + 1] = (int32_t)tot;
    }
    d->idx.resize(tot);
    d->vals.resize(tot);
    parallel_for(ns, [&](int64_t i) {
      Stream si(seed, kTagIndex, (uint64_t)i), sv(seed, kTagValue, (uint64_t)i);
      std::vector<int32_t> cols;
      distinct_sorted(si, (uint32_t)nf, (uint32_t)cnt[i], cols);
      double nrm = 0;
      const int64_t p0 = d->ptr[i];
      for (size_t q = 0; q < cols.size(); ++q) {
        const double v = std::exp(sv.normal());
        d->idx[p0 + q] = cols[q];
        d->vals[p0 + q] = v;
        nrm += v * v;
      }
      const double inv = 1.0 / std::sqrt(nrm);
      for (size_t q = 0; q < cols.size(); ++q) d->vals[p0 + q] *= inv;
    });
    d->xstar.assign(nf, 0.0);
    {
      Stream sx(seed, kTagXstar, 0), svx(seed, kTagXstar, 1);
      std::vector<int32_t> pos;
      distinct_sorted(sx, (uint32_t)nf, (uint32_t)cfg->wstar_nnz, pos);
      for (int32_t f : pos) d->xstar[f] = svx.normal();
    }
    d->rhs.resize(ns);
    parallel_for(ns, [&](int64_t i) {
      double s = 0;
      for (int32_t p = d->ptr[i]; p < d->ptr[i + 1]; ++p) s += d->vals[p] * d->xstar[d->idx[p]];
      double y = s < 0 ? -1.0 : 1.0;
      Stream sn(seed, kTagNoise, (uint64_t)i);
      if (sn.uniform() < cfg->flip_frac) y = -y;
      d->rhs[i] = y;
    });
    d->info.kind = 1;
    d->info.m = nf;
    d->info.n = ns;
    d->info.nnz = tot;
    d->info.lambda = 1.0 / (double)ns;
    d->info.lambda_max = 0;
    d->info.xstar_nnz = cfg->wstar_nnz;
  } catch (const std::bad_alloc &) {
    delete d;
    return gfail(ACDC_ERR_NOMEM, "acdc_gen_svm: out of memory");
  }
  *out = d;
  return ACDC_OK;
}

extern "C" acdc_status acdc_gen_ttd(int32_t cells, acdc_dataset **out) {
  if (!out) return gfail(ACDC_ERR_INVALID, "acdc_gen_ttd: null argument");
  if (cells < 1 || cells > 400) return gfail(ACDC_ERR_INVALID, "acdc_gen_ttd: cells must lie in [1, 400]");
  acdc_dataset *d = new (std::nothrow) acdc_dataset();
  if (!d) return gfail(ACDC_ERR_NOMEM, "acdc_gen_ttd: out of memory");
  try {
    const int64_t K = cells, P = K + 1;
    const int64_t nodes = P * P * P;
    auto node = [&](int64_t i, int64_t j, int64_t k) { return i + P * (j + P * k); };
    // free-node numbering: the z = 0 face is fixed
    std::vector<int64_t> freeid(nodes, -1);
    int64_t nfree = 0;
    for (int64_t v = 0; v < nodes; ++v)
      if (v / (P * P) != 0) freeid[v] = nfree++;
    // the 13 lexicographically positive Chebyshev-1 directions
    std::vector<std::array<int, 3>> dirs;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if (dx == 0 && dy == 0 && dz == 0) continue;
          const bool pos = dz > 0 || (dz == 0 && (dy > 0 || (dy == 0 && dx > 0)));
          if (pos) dirs.push_back({dx, dy, dz});
        }
    d->ptr.push_back(0);
    for (int64_t k = 0; k < P; ++k)
      for (int64_t j = 0; j < P; ++j)
        for (int64_t i = 0; i < P; ++i)
          for (const auto &dv : dirs) {
            const int64_t i2 = i + dv[0], j2 = j + dv[1], k2 = k + dv[2];
            if (i2 < 0 || i2 >= P || j2 < 0 || j2 >= P || k2 < 0 || k2 >= P) continue;
            const int64_t a = node(i, j, k), b = node(i2, j2, k2);
            if (freeid[a] < 0 && freeid[b] < 0) continue;  // both ends fixed
            const double len2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
            std::vector<std::pair<int32_t, double>> ent;
            for (int c = 0; c < 3; ++c) {
              if (dv[c] == 0) continue;
              const double beta = dv[c] / len2;
              // compression positive: +d on the lower end a, -d on the upper end b
              if (freeid[a] >= 0) ent.push_back({(int32_t)(c * nfree + freeid[a]), beta});
              if (freeid[b] >= 0) ent.push_back({(int32_t)(c * nfree + freeid[b]), -beta});
            }
            std::sort(ent.begin(), ent.end());
            for (auto &e : ent) {
              d->idx.push_back(e.first);
              d->vals.push_back(e.second);
            }
            d->ptr.push_back((int32_t)d->idx.size());
          }
    const int64_t m = 3 * nfree;
    d->rhs.assign(m, 0.0);
    for (int64_t j = 0; j < P; ++j)
      for (int64_t i = 0; i < P; ++i) d->rhs[2 * nfree + freeid[node(i, j, K)]] = -1.0;
    d->info.kind = 0;
    d->info.m = m;
    d->info.n = (int64_t)d->ptr.size() - 1;
    d->info.nnz = (int64_t)d->idx.size();
    d->info.lambda = 0;
    d->info.lambda_max = 0;
    d->info.xstar_nnz = 0;
    d->xstar.assign(d->info.n, 0.0);
  } catch (const std::bad_alloc &) {
    delete d;
    return gfail(ACDC_ERR_NOMEM, "acdc_gen_ttd: out of memory");
  }
  *out = d;
  return ACDC_OK;
}

extern "C" acdc_status acdc_dataset_info_get(const acdc_dataset *d, acdc_dataset_info *out) {
  if (!d || !out) return gfail(ACDC_ERR_INVALID, "acdc_dataset_info_get: null argument");
  *out = d->info;
  return ACDC_OK;
}

extern "C" acdc_status acdc_dataset_copy(const acdc_dataset *d, int32_t *ptr, int32_t *idx,
                                         double *vals, double *rhs, double *xstar) {
  if (!d) return gfail(ACDC_ERR_INVALID, "acdc_dataset_copy: null dataset");
  if (ptr) std::memcpy(ptr, d->ptr.data(), d->ptr.size() * 4);
  if (idx) std::memcpy(idx, d->idx.data(), d->idx.size() * 4);
  if (vals) std::memcpy(vals, d->vals.data(), d->vals.size() * 8);
  if (rhs) std::memcpy(rhs, d->rhs.data(), d->rhs.size() * 8);
  if (xstar) std::memcpy(xstar, d->xstar.data(), d->xstar.size() * 8);
  return ACDC_OK;
}

extern "C" acdc_status acdc_dataset_problem(const acdc_dataset *d, acdc_problem **out) {
  if (!d || !out) return gfail(ACDC_ERR_INVALID, "acdc_dataset_problem: null argument");
  if (d->info.kind == 0)
    return acdc_problem_create_csc(d->info.m, d->info.n, d->info.nnz, d->ptr.data(),
                                   d->idx.data(), d->vals.data(), d->rhs.data(), out);
  return acdc_problem_create_csr(d->info.n, d->info.m, d->info.nnz, d->ptr.data(),
                                 d->idx.data(), d->vals.data(), d->rhs.data(), out);
}

extern "C" acdc_status acdc_dataset_destroy(acdc_dataset *d) {
  delete d;
  return ACDC_OK;
}
