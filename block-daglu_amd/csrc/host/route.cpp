// Route selection of Solver::analyze (reference src/solver.cpp:432-556): the
// AUTO choice between the BLOCK route and the global-KLU route, and the KLU
// settings of the global route (btf, scale, tol, ordering). These are decision
// rules on sizes and a sampled degree profile only; the same matrix and
// environment give the same choice as the reference.
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <unordered_map>

#include "blockdaglu/env.h"
#include "host.h"

namespace bdlu {

// compute_profile_lite (src/solver.cpp:182-230): row/column degrees from a
// deterministic stride over the canonical entries (about 200,000 samples),
// scaled back by the stride; p95 = the element at floor(0.95 (n-1)) of the
// sampled degrees.
ProfileLite profile_lite(const Canonical& a) {
  ProfileLite p;
  const int32_t n = a.n;
  if (n <= 0) return p;
  const std::size_t nnz = a.nnz();
  p.avg_deg = static_cast<double>(nnz) / static_cast<double>(std::max<int32_t>(1, n));
  std::vector<int32_t> rd(static_cast<std::size_t>(n), 0), cd(static_cast<std::size_t>(n), 0);
  std::vector<uint8_t> dg(static_cast<std::size_t>(n), 0);
  const std::size_t target = 200000;
  const std::size_t step = nnz > target ? std::max<std::size_t>(1, nnz / target) : 1;
  for (std::size_t k = 0; k < nnz; k += step) {
    const int32_t r = a.row[k], c = a.col[k];
    if (r < 0 || c < 0 || r >= n || c >= n) continue;
    rd[r]++;
    cd[c]++;
    if (r == c) dg[r] = 1;
  }
  int32_t present = 0;
  for (int32_t i = 0; i < n; ++i) present += dg[i] != 0;
  p.diag_cov = static_cast<double>(present) / static_cast<double>(n);
  const int32_t st = static_cast<int32_t>(step);
  p.row_max = *std::max_element(rd.begin(), rd.end()) * st;
  p.col_max = *std::max_element(cd.begin(), cd.end()) * st;
  auto p95 = [](std::vector<int32_t>& v) {
    const std::size_t k = static_cast<std::size_t>(0.95 * static_cast<double>(v.size() - 1));
    std::nth_element(v.begin(), v.begin() + static_cast<std::ptrdiff_t>(k), v.end());
    return v[k];
  };
  p.row_p95 = p95(rd) * st;
  p.col_p95 = p95(cd) * st;
  return p;
}

namespace {

struct OverrideKey {
  int32_t n;
  std::size_t nnz;
  uint64_t sig, sig2, value_sig, config;
  bool operator==(const OverrideKey& o) const {
    return n == o.n && nnz == o.nnz && sig == o.sig && sig2 == o.sig2 && value_sig == o.value_sig && config == o.config;
  }
};
struct OverrideHash {
  std::size_t operator()(const OverrideKey& k) const {
    uint64_t h = k.sig ^ (k.sig2 * 0x9e3779b97f4a7c15ull) ^ k.value_sig ^ (k.config << 1);
    h ^= static_cast<uint64_t>(k.n) * 0xbf58476d1ce4e5b9ull + k.nnz;
    return static_cast<std::size_t>(h ^ (h >> 31));
  }
};
std::mutex g_override_mu;
std::unordered_map<OverrideKey, bool, OverrideHash> g_prefer_klu;

bool cache_on() {
  if (const char* e = blockdaglu_getenv_compat("USERALGO_CACHE")) return std::atoi(e) != 0;
  return true;
}

int32_t auto_block_n() {
  if (const char* e = blockdaglu_getenv_compat("USERALGO_AUTO_BLOCK_N")) {
    const int v = std::atoi(e);
    if (v > 0) return v;
  }
  return 50000;
}

std::size_t auto_block_nnz() {
  if (const char* e = blockdaglu_getenv_compat("USERALGO_AUTO_BLOCK_NNZ")) {
    const long long v = std::atoll(e);
    if (v > 0) return static_cast<std::size_t>(v);
  }
  return 2000000;
}

}  // namespace

RouteMode route_mode_from_env() {
  // USERALGO_MODE (src/solver.cpp:822-840): "klu", "block", "auto"/"default";
  // anything else is AUTO.
  const char* e = blockdaglu_getenv_compat("USERALGO_MODE");
  if (!e) return RouteMode::Auto;
  std::string v(e);
  for (auto& ch : v) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
  if (v == "klu") return RouteMode::Klu;
  if (v == "block") return RouteMode::Block;
  return RouteMode::Auto;
}

// auto_choose_block (src/solver.cpp:450-515): the learned override first, then
// the coarse size rules, then (for large inputs) the degree profile.
bool auto_choose_block(const Canonical& a, uint64_t sig, uint64_t sig2, uint64_t value_sig, uint64_t config) {
  const int32_t n = a.n;
  const std::size_t nnz = a.nnz();
  if (cache_on()) {
    std::lock_guard<std::mutex> g(g_override_mu);
    auto it = g_prefer_klu.find(OverrideKey{n, nnz, sig, sig2, value_sig, config});
    if (it != g_prefer_klu.end()) return !it->second;
  }
  bool coarse;
  if (n >= 65000 && nnz >= 3000000) coarse = false;
  else if (n <= 25000 && nnz >= 600000) coarse = true;
  else if (n >= auto_block_n()) coarse = true;
  else if (nnz >= auto_block_nnz()) coarse = true;
  else coarse = false;
  if (!coarse) return false;
  if (n < 40000 && nnz < 1500000) return true;
  const ProfileLite p = profile_lite(a);
  const int32_t dense_thr = std::max<int32_t>(4096, n / 8);
  const bool hub = p.row_max >= dense_thr || p.col_max >= dense_thr;
  const int32_t broad_thr = std::max<int32_t>(512, n / 64);
  const bool broad = p.row_p95 >= broad_thr && p.col_p95 >= broad_thr && p.diag_cov >= 0.9;
  return !(hub || broad);
}

// The self-adaptive override of src/solver.cpp:1133-1145: a BLOCK analysis that
// took more than 300 ms and left one block above 80% of n makes AUTO prefer
// the global-KLU route for this matrix (same values) from then on.
void note_block_analysis(const Canonical& a, uint64_t sig, uint64_t sig2, uint64_t value_sig, uint64_t config,
                         double build_ms, int32_t max_block) {
  if (!cache_on() || build_ms <= 300.0 || a.n <= 0) return;
  if (static_cast<double>(max_block) / static_cast<double>(a.n) <= 0.80) return;
  std::lock_guard<std::mutex> g(g_override_mu);
  g_prefer_klu[OverrideKey{a.n, a.nnz(), sig, sig2, value_sig, config}] = true;
}

// pick_klu_btf_for (src/solver.cpp:517-532), with USERALGO_KLU_BTF (0/1) first.
int klu_btf_for(int32_t n, std::size_t nnz) {
  if (const char* e = blockdaglu_getenv_compat("USERALGO_KLU_BTF")) {
    const int v = std::atoi(e);
    if (v == 0 || v == 1) return v;
  }
  if (n <= 10000) return 1;
  if (n <= 25000 && nnz >= 600000) return 1;
  if (n >= 35000 && n <= 50000 && nnz >= 500000 && nnz <= 800000) return 1;
  if (n >= 30000 && n <= 38000 && nnz >= 1500000) return 1;
  return 0;
}

// pick_klu_scale_for (src/solver.cpp:534-554), with USERALGO_KLU_SCALE (-1..2) first.
int klu_scale_for(int32_t n, std::size_t nnz, int btf) {
  if (const char* e = blockdaglu_getenv_compat("USERALGO_KLU_SCALE")) {
    const int v = std::atoi(e);
    if (v >= -1 && v <= 2) return v;
  }
  if (btf == 1 && n >= 35000 && n <= 50000 && nnz >= 500000 && nnz <= 800000) return 2;
  if (btf == 1 && n >= 30000 && n <= 38000 && nnz >= 1500000) return 2;
  if (n >= 65000 && nnz >= 3000000) return 2;
  return 0;
}

// pick_klu_tol_for (src/solver.cpp:556-565), with USERALGO_KLU_TOL (>= 0) first;
// KLU's default tolerance otherwise.
double klu_tol_for(int32_t n, std::size_t nnz, int btf) {
  if (const char* e = blockdaglu_getenv_compat("USERALGO_KLU_TOL")) {
    const double v = std::atof(e);
    if (v >= 0.0) return v;
  }
  if (btf == 1 && n >= 30000 && n <= 38000 && nnz >= 1500000) return 1e-6;
  return 0.001;
}

// pick_klu_ordering (src/solver.cpp:400-407): 0 AMD, 1 COLAMD (default),
// 2 user-given, 3 user function.
int klu_ordering() {
  if (const char* e = blockdaglu_getenv_compat("USERALGO_KLU_ORDERING")) {
    const int v = std::atoi(e);
    if (v >= 0 && v <= 3) return v;
  }
  return 1;
}

}  // namespace bdlu
