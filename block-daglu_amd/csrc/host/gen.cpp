// Seeded synthetic workloads with a hidden block-triangular structure.
//
// Each diagonal block of size s >= 2 holds a directed cycle (i -> i+1 mod s)
// plus random in-block entries per row; blocks of size 1 hold only the
// diagonal; coupling entries go only to columns of earlier blocks, so the
// coupling is acyclic and DM/BTF recovers the blocks. Off-diagonal values are
// N(0,1); the diagonal is f * sum_{j != i} |a_ij| + U(0.1, 1). Independent row
// and column permutations hide the structure. All randomness comes from
// SplitMix64 substreams, so any host produces the same bits.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../../include/bdlu_host.h"

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed, uint64_t purpose) : s(seed * 0x9e3779b97f4a7c15ull ^ (purpose + 0x632be59bd9b4e019ull)) {
    next();
    next();
  }
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  int64_t below(int64_t m) { return static_cast<int64_t>(uniform() * static_cast<double>(m)); }
  double normal() {
    // Box-Muller, one value per call.
    double u1 = uniform();
    while (u1 <= 0.0) u1 = uniform();
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  int poisson(double lam) {
    const double L = std::exp(-lam);
    int k = 0;
    double p = 1.0;
    do {
      ++k;
      p *= uniform();
    } while (p > L);
    return k - 1;
  }
};

enum Purpose : uint64_t { kSizes = 1, kPattern, kValues, kPerm, kRhs, kCoupling, kPlace, kPerturb, kScale };

struct Workload {
  int32_t n = 0;
  int32_t nrhs = 1;
  std::vector<int32_t> row, col;
  std::vector<double> val;
  std::vector<double> rhs;
};

// Adjust sizes by +-1 within [lo, hi] until they sum to `target`.
void fit_sizes(std::vector<int32_t>& sz, int64_t target, int32_t lo, int32_t hi, Rng& rng) {
  int64_t sum = 0;
  for (int32_t v : sz) sum += v;
  const int64_t nb = static_cast<int64_t>(sz.size());
  int guard = 0;
  while (sum != target && nb > 0 && guard < 100000000) {
    const int64_t b = rng.below(nb);
    if (sum < target && sz[b] < hi) {
      ++sz[b];
      ++sum;
    } else if (sum > target && sz[b] > lo) {
      --sz[b];
      --sum;
    }
    ++guard;
  }
}

Workload build(const bdlu_gen_spec& sp) {
  Workload w;
  Rng rs(sp.seed, kSizes), rp(sp.seed, kPattern), rv(sp.seed, kValues), rperm(sp.seed, kPerm),
      rr(sp.seed, kRhs), rc(sp.seed, kCoupling), rplace(sp.seed, kPlace);
  // Small block sizes.
  std::vector<int32_t> small(static_cast<std::size_t>(sp.n_small_blocks));
  for (auto& s : small) {
    double v = 1.0;
    switch (sp.size_dist) {
      case BDLU_GEN_GEOMETRIC: {  // mean 1/p, support 1,2,...
        const double u = std::max(rs.uniform(), 1e-300);
        v = std::floor(std::log(u) / std::log(1.0 - sp.size_param)) + 1.0;
        break;
      }
      case BDLU_GEN_UNIFORM:
        v = static_cast<double>(sp.size_lo + rs.below(static_cast<int64_t>(sp.size_param) - sp.size_lo + 1));
        break;
      case BDLU_GEN_LOGUNIFORM:
        v = std::floor(std::exp(rs.uniform(0.0, std::log(sp.size_param))));
        break;
    }
    s = static_cast<int32_t>(std::min<double>(std::max<double>(v, sp.size_lo), sp.size_hi));
  }
  if (sp.n_small_total > 0) fit_sizes(small, sp.n_small_total, sp.size_lo, sp.size_hi, rs);
  // Large blocks, inserted at random block positions.
  std::vector<int32_t> sizes(small.begin(), small.end());
  std::vector<int> kind(sizes.size(), 0);  // 0 small, 1 large
  for (int32_t g = 0; g < sp.n_large_blocks; ++g) {
    int32_t s = sp.large_size_lo;
    if (sp.large_size_hi > sp.large_size_lo) s += static_cast<int32_t>(rs.below(sp.large_size_hi - sp.large_size_lo + 1));
    const int64_t pos = rplace.below(static_cast<int64_t>(sizes.size()) + 1);
    sizes.insert(sizes.begin() + pos, s);
    kind.insert(kind.begin() + pos, 1);
  }
  const int32_t nb = static_cast<int32_t>(sizes.size());
  std::vector<int32_t> off(static_cast<std::size_t>(nb) + 1, 0);
  for (int32_t b = 0; b < nb; ++b) off[b + 1] = off[b] + sizes[b];
  const int32_t n = off[nb];
  w.n = n;

  // Zipf(1) CDF for the large blocks' extra columns (ranked per block).
  std::vector<double> zcdf;
  std::vector<int32_t> cols_row;
  std::vector<double> vals_row;
  std::vector<int32_t> rank;
  for (int32_t b = 0; b < nb; ++b) {
    const int32_t s = sizes[b];
    const bool large = kind[b] == 1;
    if (large && sp.large_fill == BDLU_GEN_ZIPF) {
      zcdf.assign(s, 0.0);
      double acc = 0.0;
      for (int32_t r = 0; r < s; ++r) {
        acc += 1.0 / (r + 1.0);
        zcdf[r] = acc;
      }
      for (auto& z : zcdf) z /= acc;
      rank.resize(s);
      for (int32_t r = 0; r < s; ++r) rank[r] = r;
      for (int32_t r = s - 1; r > 0; --r) std::swap(rank[r], rank[rp.below(r + 1)]);
    }
    for (int32_t i = 0; i < s; ++i) {
      const int32_t gi = off[b] + i;
      cols_row.clear();
      vals_row.clear();
      auto has = [&](int32_t c) { return std::find(cols_row.begin(), cols_row.end(), c) != cols_row.end(); };
      if (s >= 2) {
        cols_row.push_back(off[b] + (i + 1) % s);
        if (large && sp.large_fill == BDLU_GEN_BERNOULLI) {
          for (int32_t j = 0; j < s; ++j) {
            if (j == i || j == (i + 1) % s) continue;
            if (rp.uniform() < sp.large_fill_param) cols_row.push_back(off[b] + j);
          }
        } else {
          const int cap = s - 2;
          const int k = std::min(rp.poisson(sp.extra_per_row), cap);
          int placed = 0, tries = 0;
          while (placed < k && tries < 64 * (k + 1)) {
            ++tries;
            int32_t j;
            if (large && sp.large_fill == BDLU_GEN_ZIPF) {
              const double u = rp.uniform();
              const int32_t r = static_cast<int32_t>(std::lower_bound(zcdf.begin(), zcdf.end(), u) - zcdf.begin());
              j = rank[std::min(r, s - 1)];
            } else {
              j = static_cast<int32_t>(rp.below(s));
            }
            const int32_t gj = off[b] + j;
            if (j == i || has(gj)) continue;
            cols_row.push_back(gj);
            ++placed;
          }
        }
      }
      // Coupling to earlier blocks.
      if (off[b] > 0) {
        int k = 0;
        if (sp.coupling_kind == BDLU_GEN_POISSON) k = rc.poisson(sp.coupling_rate);
        else k = rc.uniform() < sp.coupling_rate ? 1 : 0;
        for (int t = 0; t < k; ++t) {
          const int32_t gj = static_cast<int32_t>(rc.below(off[b]));
          if (!has(gj)) cols_row.push_back(gj);
        }
      }
      double rowsum = 0.0;
      for (std::size_t t = 0; t < cols_row.size(); ++t) {
        const double v = rv.normal();
        vals_row.push_back(v);
        rowsum += std::abs(v);
      }
      const double d = sp.diag_factor * rowsum + rv.uniform(0.1, 1.0);
      w.row.push_back(gi);
      w.col.push_back(gi);
      w.val.push_back(d);
      for (std::size_t t = 0; t < cols_row.size(); ++t) {
        w.row.push_back(gi);
        w.col.push_back(cols_row[t]);
        w.val.push_back(vals_row[t]);
      }
    }
  }
  // Hide the structure (unless asked to keep the generated order; the values and
  // the pattern come from their own streams, so they do not change).
  std::vector<int32_t> pr(n), pc(n);
  for (int32_t i = 0; i < n; ++i) pr[i] = pc[i] = i;
  if (!sp.keep_order) {
    for (int32_t i = n - 1; i > 0; --i) std::swap(pr[i], pr[rperm.below(i + 1)]);
    for (int32_t i = n - 1; i > 0; --i) std::swap(pc[i], pc[rperm.below(i + 1)]);
  }
  for (std::size_t k = 0; k < w.row.size(); ++k) {
    w.row[k] = pr[w.row[k]];
    w.col[k] = pc[w.col[k]];
  }
  if (sp.scale_exp > 0.0) {
    Rng rsc(sp.scale_seed ? sp.scale_seed : sp.seed, kScale);
    std::vector<double> dr(n), dc(n);
    for (auto& x : dr) x = std::pow(10.0, rsc.uniform(-sp.scale_exp, sp.scale_exp));
    for (auto& x : dc) x = std::pow(10.0, rsc.uniform(-sp.scale_exp, sp.scale_exp));
    for (std::size_t k = 0; k < w.row.size(); ++k) w.val[k] *= dr[w.row[k]] * dc[w.col[k]];
  }
  w.nrhs = std::max(1, sp.nrhs);
  w.rhs.resize(static_cast<std::size_t>(n) * w.nrhs);
  for (auto& x : w.rhs) x = rr.uniform(-1.0, 1.0);
  return w;
}

}  // namespace

extern "C" {

void bdlu_gen_config_spec(int cfg, int pattern, bdlu_gen_spec* sp) {
  std::memset(sp, 0, sizeof(*sp));
  sp->extra_per_row = 4.0;
  sp->nrhs = 1;
  switch (cfg) {
    case 0:
      sp->seed = 0;
      sp->n_small_blocks = 200;
      sp->size_dist = BDLU_GEN_GEOMETRIC;
      sp->size_param = 1.0 / 25.0;
      sp->size_lo = 1;
      sp->size_hi = 200;
      sp->n_small_total = 5000;
      sp->coupling_kind = BDLU_GEN_POISSON;
      sp->coupling_rate = 1.0;
      sp->diag_factor = 0.5;
      break;
    case 1:
    case 3:
      sp->seed = 1;
      sp->n_small_blocks = 9000;
      sp->size_dist = BDLU_GEN_UNIFORM;
      sp->size_param = 39;
      sp->size_lo = 1;
      sp->size_hi = 40;
      sp->n_small_total = 180000;
      sp->n_large_blocks = 1;
      sp->large_size_lo = sp->large_size_hi = 20000;
      sp->large_fill = BDLU_GEN_ZIPF;
      sp->coupling_kind = BDLU_GEN_BERNOULLI;
      sp->coupling_rate = 0.5;
      sp->diag_factor = 0.6;
      if (cfg == 3) {
        sp->scale_exp = 4.0;
        sp->scale_seed = 3;
      }
      break;
    case 2:
      sp->seed = 2;
      sp->n_small_blocks = 50000;
      sp->size_dist = BDLU_GEN_LOGUNIFORM;
      sp->size_param = 65.0;
      sp->size_lo = 1;
      sp->size_hi = 64;
      sp->n_large_blocks = 8;
      sp->large_size_lo = 512;
      sp->large_size_hi = 2048;
      sp->large_fill = BDLU_GEN_BERNOULLI;
      sp->large_fill_param = 0.3;
      sp->coupling_kind = BDLU_GEN_POISSON;
      sp->coupling_rate = 1.0;
      sp->diag_factor = 0.5;
      sp->nrhs = 4;
      break;
    case 4: {
      sp->seed = 4000 + static_cast<uint64_t>(pattern);
      Rng r(sp->seed, kSizes + 100);
      const int32_t np = static_cast<int32_t>(std::llround(std::exp(r.uniform(std::log(1e3), std::log(1e5)))));
      const int32_t giant = std::max(2, np / 10);
      sp->n_small_blocks = std::max(1, (np - giant) / 20);
      sp->size_dist = BDLU_GEN_UNIFORM;
      sp->size_param = 39;
      sp->size_lo = 1;
      sp->size_hi = 40;
      sp->n_small_total = np - giant;
      sp->n_large_blocks = 1;
      sp->large_size_lo = sp->large_size_hi = giant;
      sp->large_fill = BDLU_GEN_ZIPF;
      sp->coupling_kind = BDLU_GEN_BERNOULLI;
      sp->coupling_rate = 0.5;
      sp->diag_factor = 0.6;
      break;
    }
    default:
      break;
  }
}

void* bdlu_gen_create(const bdlu_gen_spec* sp) {
  try {
    return new Workload(build(*sp));
  } catch (...) {
    return nullptr;
  }
}

void bdlu_gen_info(const void* h, int32_t* n, int64_t* nnz, int32_t* nrhs) {
  const Workload* w = static_cast<const Workload*>(h);
  *n = w->n;
  *nnz = static_cast<int64_t>(w->row.size());
  *nrhs = w->nrhs;
}

void bdlu_gen_copy(const void* h, int32_t* rows, int32_t* cols, double* vals, double* rhs) {
  const Workload* w = static_cast<const Workload*>(h);
  if (rows) std::copy(w->row.begin(), w->row.end(), rows);
  if (cols) std::copy(w->col.begin(), w->col.end(), cols);
  if (vals) std::copy(w->val.begin(), w->val.end(), vals);
  if (rhs) std::copy(w->rhs.begin(), w->rhs.end(), rhs);
}

void bdlu_gen_free(void* h) { delete static_cast<Workload*>(h); }

void bdlu_gen_perturb(double* vals, int64_t nnz, uint64_t seed, double rel) {
  Rng r(seed, kPerturb);
  for (int64_t k = 0; k < nnz; ++k) vals[k] *= 1.0 + r.uniform(-rel, rel);
}

}  // extern "C"
