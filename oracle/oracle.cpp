// ORACLE — test infrastructure, not product code.
//
// A CPU restatement of the reference BLOCK route (CaptSteven/block-daglu),
// used only by tests/, by __graft_entry__.smoke() as the checker, and by
// bench.py's cpu_baseline leg. Nothing in the product links or calls it.
//
// Each function names the reference code it follows (paths relative to
// blockdaglu/; "ubtf:" = src/ubtf_block_solver.cpp). Third-party numerics the
// reference takes from SuiteSparse (vcpkg tag 2024.05.24, not vendored, not
// present here) are restated from their published algorithms:
//   btf_order   (maxtrans + Tarjan SCC)            -> oracle_btf_order
//   klu_analyze ordering=1 (COLAMD)                -> oracle::colamd
//               (oracle/colamd_oracle.cpp, the oracle's own restatement; pinned
//               against the COLAMD that SuperLU bundles, through scipy)
//   klu_factor  (left-looking Gilbert-Peierls, threshold partial pivoting with
//               diagonal preference and diagonal tracking, tol 0.001, scale=1
//               row sums)                         -> klu_like_factor
//   klu_solve                                      -> klu_like_solve
// Pinning (what checks each function) is listed in the docstrings of the test files.
//
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off, no fast-math).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

namespace oracle {
std::vector<int32_t> colamd(int32_t n, const int32_t* Ap, const int32_t* Ai);
}

namespace oracle {

using Coo = std::vector<std::tuple<int, int, double>>;

enum Err { kNone = 0, kNotPrepared = 1, kInvalidDimension = 2, kInvalidIndex = 3, kNonFinite = 4,
           kInvalidRhs = 5, kBackend = 6, kAnalysis = 7, kFactor = 8, kSolve = 9 };

// ---- src/solver.cpp:642-722 ------------------------------------------------------
int validate_and_canonicalize(Coo& coo, int32_t hint, int32_t& n) {
  if (hint < -1) return kInvalidDimension;
  int32_t max_index = -1;
  for (const auto& t : coo) {
    const int r = std::get<0>(t), c = std::get<1>(t);
    const double v = std::get<2>(t);
    if (r < 0 || c < 0) return kInvalidIndex;
    if (!std::isfinite(v)) return kNonFinite;
    if (hint >= 0 && (r >= hint || c >= hint)) return kInvalidIndex;
    max_index = std::max<int32_t>(max_index, std::max<int32_t>(r, c));
  }
  if (hint >= 0) n = hint;
  else {
    if (max_index == std::numeric_limits<int32_t>::max()) return kInvalidDimension;
    n = max_index + 1;
  }
  if (n <= 0) return kInvalidDimension;
  std::stable_sort(coo.begin(), coo.end(), [](const auto& a, const auto& b) {
    if (std::get<0>(a) != std::get<0>(b)) return std::get<0>(a) < std::get<0>(b);
    return std::get<1>(a) < std::get<1>(b);
  });
  Coo out;
  out.reserve(coo.size());
  for (std::size_t i = 0; i < coo.size();) {
    const int r = std::get<0>(coo[i]), c = std::get<1>(coo[i]);
    long double sum = 0.0L;
    std::size_t j = i;
    while (j < coo.size() && std::get<0>(coo[j]) == r && std::get<1>(coo[j]) == c) {
      sum += static_cast<long double>(std::get<2>(coo[j]));
      ++j;
    }
    const double v = static_cast<double>(sum);
    if (!std::isfinite(v)) return kNonFinite;
    if (v != 0.0) out.emplace_back(r, c, v);
    i = j;
  }
  coo.swap(out);
  return kNone;
}

// ---- src/solver.cpp:90-136 (FNV-1a / splitmix mixing) -----------------------------
uint64_t fnv(uint64_t h, uint64_t x) { return (h ^ x) * 1099511628211ull; }
void signatures(int32_t n, const Coo& coo, uint64_t* s1, uint64_t* s2, uint64_t* s3) {
  uint64_t h = fnv(fnv(1469598103934665603ull, static_cast<uint64_t>(n)), coo.size());
  for (const auto& t : coo)
    h = fnv(h, (static_cast<uint64_t>(static_cast<uint32_t>(std::get<0>(t))) << 32) ^
                   static_cast<uint64_t>(static_cast<uint32_t>(std::get<1>(t))));
  *s1 = h;
  h = 0x6a09e667f3bcc909ull ^ static_cast<uint64_t>(n);
  for (const auto& t : coo) {
    uint64_t x = ((static_cast<uint64_t>(static_cast<uint32_t>(std::get<0>(t))) << 32) ^
                  static_cast<uint64_t>(static_cast<uint32_t>(std::get<1>(t)))) ^ 0x9e3779b97f4a7c15ull;
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    h ^= x + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
  }
  *s2 = h ^ static_cast<uint64_t>(coo.size());
  h = 0xbb67ae8584caa73bull;
  for (const auto& t : coo) {
    uint64_t bits;
    const double v = std::get<2>(t);
    std::memcpy(&bits, &v, 8);
    h = fnv(h, bits);
  }
  *s3 = h ^ static_cast<uint64_t>(coo.size());
}

// ---- SuiteSparse BTF restated ----------------------------------------------------
// maxtrans: for each column k in order, a depth-first search for an augmenting
// path; each column first scans its rows once for an unmatched one (cheap
// assignment, pointer kept across searches), then recurses through matched rows.
int32_t maxtrans(int32_t n, const std::vector<int32_t>& Ap, const std::vector<int32_t>& Ai,
                 std::vector<int32_t>& match) {
  match.assign(n, -1);
  std::vector<int32_t> cheap(Ap.begin(), Ap.end() - 1), visited(n, -1);
  struct Frame { int32_t col, pos, row; };
  std::vector<Frame> st;
  int32_t nmatch = 0;
  for (int32_t k = 0; k < n; ++k) {
    st.clear();
    st.push_back({k, -1, -1});
    bool found = false;
    while (!st.empty()) {
      Frame& f = st.back();
      const int32_t j = f.col;
      if (visited[j] != k) {
        visited[j] = k;
        int32_t p = cheap[j];
        int32_t hit = -1;
        for (; p < Ap[j + 1]; ++p) {
          if (match[Ai[p]] < 0) {
            hit = Ai[p];
            ++p;
            break;
          }
        }
        cheap[j] = p;
        if (hit >= 0) {
          f.row = hit;
          found = true;
          break;
        }
        f.pos = Ap[j];
      }
      bool pushed = false;
      for (int32_t p = f.pos; p < Ap[j + 1]; ++p) {
        const int32_t i = Ai[p];
        const int32_t j2 = match[i];
        if (visited[j2] != k) {
          f.pos = p + 1;
          f.row = i;
          st.push_back({j2, -1, -1});
          pushed = true;
          break;
        }
      }
      if (!pushed) st.pop_back();
    }
    if (found) {
      for (const Frame& f : st) match[f.row] = f.col;
      ++nmatch;
    }
  }
  return nmatch;
}

// Tarjan SCC on the graph of A(:,Q) (node j -> rows of column Q[j]); blocks in
// the order found, natural order inside blocks; Q <- Q*P.
int32_t strongcomp(int32_t n, const std::vector<int32_t>& Ap, const std::vector<int32_t>& Ai,
                   std::vector<int32_t>& Q, std::vector<int32_t>& P, std::vector<int32_t>& R) {
  std::vector<int32_t> index(n, -1), low(n, 0), comp(n, -1), stack;
  struct Frame { int32_t node, pos; };
  std::vector<Frame> dfs;
  int32_t counter = 0, nb = 0;
  auto unflip = [](int32_t j) { return j < -1 ? -j - 2 : j; };
  for (int32_t root = 0; root < n; ++root) {
    if (index[root] >= 0) continue;
    dfs.push_back({root, -1});
    while (!dfs.empty()) {
      Frame& f = dfs.back();
      const int32_t j = f.node;
      const int32_t col = unflip(Q[j]);
      if (f.pos < 0) {
        index[j] = low[j] = counter++;
        stack.push_back(j);
        f.pos = Ap[col];
      }
      bool pushed = false;
      while (f.pos < Ap[col + 1]) {
        const int32_t i = Ai[f.pos++];
        if (index[i] < 0) {
          dfs.push_back({i, -1});
          pushed = true;
          break;
        }
        if (comp[i] < 0) low[j] = std::min(low[j], index[i]);
      }
      if (pushed) continue;
      if (low[j] == index[j]) {
        while (true) {
          const int32_t v = stack.back();
          stack.pop_back();
          comp[v] = nb;
          if (v == j) break;
        }
        ++nb;
      }
      dfs.pop_back();
      if (!dfs.empty()) low[dfs.back().node] = std::min(low[dfs.back().node], low[j]);
    }
  }
  R.assign(nb + 1, 0);
  for (int32_t i = 0; i < n; ++i) R[comp[i] + 1]++;
  for (int32_t b = 0; b < nb; ++b) R[b + 1] += R[b];
  P.assign(n, 0);
  std::vector<int32_t> next(R.begin(), R.end() - 1);
  for (int32_t i = 0; i < n; ++i) P[next[comp[i]]++] = i;
  std::vector<int32_t> q2(n);
  for (int32_t k = 0; k < n; ++k) q2[k] = Q[P[k]];
  Q.swap(q2);
  return nb;
}

int32_t btf_order(int32_t n, const std::vector<int32_t>& Ap, const std::vector<int32_t>& Ai, std::vector<int32_t>& P,
                  std::vector<int32_t>& Q, std::vector<int32_t>& R, int32_t& nmatch) {
  nmatch = maxtrans(n, Ap, Ai, Q);
  if (nmatch < n) {
    std::vector<char> used(n, 0);
    for (int32_t i = 0; i < n; ++i)
      if (Q[i] >= 0) used[Q[i]] = 1;
    std::vector<int32_t> free_cols;  // ascending
    for (int32_t j = 0; j < n; ++j)
      if (!used[j]) free_cols.push_back(j);
    std::size_t t = 0;
    for (int32_t i = 0; i < n; ++i)
      if (Q[i] < 0 && t < free_cols.size()) Q[i] = -free_cols[t++] - 2;
  }
  return strongcomp(n, Ap, Ai, Q, P, R);
}

// ---- ubtf:393-476 ---------------------------------------------------------------
std::vector<int32_t> merge_blocks(const std::vector<int32_t>& R, int32_t n, int32_t merge_max) {
  if (R.size() < 3) return {0, n};
  const int32_t small_max = std::max<int32_t>(1, std::min<int32_t>(merge_max, 32));
  std::vector<int32_t> out{0};
  int32_t acc = 0, rs = 0, re = 0;
  auto flush = [&]() {
    if (acc <= 0) return;
    if (out.back() != rs) out.push_back(rs);
    if (out.back() != re) out.push_back(re);
    acc = 0;
  };
  for (std::size_t b = 0; b + 1 < R.size(); ++b) {
    const int32_t lo = R[b], hi = R[b + 1], sz = hi - lo;
    if (sz <= small_max) {
      if (acc == 0) { rs = lo; re = hi; acc = sz; continue; }
      if (acc + sz <= merge_max) { acc += sz; re = hi; continue; }
      flush();
      rs = lo; re = hi; acc = sz;
      continue;
    }
    flush();
    if (out.back() != lo) out.push_back(lo);
    if (out.back() != hi) out.push_back(hi);
  }
  flush();
  if (out.back() != n) out.push_back(n);
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

// ---- factor storage --------------------------------------------------------------
struct BlockFactor {
  int32_t start = 0, size = 0;
  bool ok = false;
  bool dense = false;
  std::vector<double> lu;   // row-major (ubtf:1455)
  std::vector<int32_t> piv;
  // KLU-like
  std::vector<int32_t> q, prow;
  std::vector<int32_t> Lp, Li;  std::vector<double> Lx;
  std::vector<int32_t> Up, Ui;  std::vector<double> Ux, Ud;
  std::vector<double> rs;
};

// ---- ubtf:1441-1498: dense LU, partial pivoting ----------------------------------
bool dense_factor(BlockFactor& bf, const std::vector<std::tuple<int, int, double>>& coo) {
  const int m = bf.size;
  bf.dense = true;
  bf.lu.assign(static_cast<std::size_t>(m) * m, 0.0);
  bf.piv.assign(m, 0);
  for (const auto& t : coo) bf.lu[static_cast<std::size_t>(std::get<0>(t)) * m + std::get<1>(t)] += std::get<2>(t);
  double* a = bf.lu.data();
  for (int k = 0; k < m; ++k) {
    int p = k;
    double best = std::abs(a[static_cast<std::size_t>(k) * m + k]);
    for (int i = k + 1; i < m; ++i) {
      const double v = std::abs(a[static_cast<std::size_t>(i) * m + k]);
      if (v > best) { best = v; p = i; }
    }
    if (best == 0.0) return false;
    bf.piv[k] = p;
    if (p != k)
      for (int j = 0; j < m; ++j) std::swap(a[static_cast<std::size_t>(k) * m + j], a[static_cast<std::size_t>(p) * m + j]);
    const double akk = a[static_cast<std::size_t>(k) * m + k];
    for (int i = k + 1; i < m; ++i) {
      double& aik = a[static_cast<std::size_t>(i) * m + k];
      aik /= akk;
      const double l = aik;
      for (int j = k + 1; j < m; ++j) a[static_cast<std::size_t>(i) * m + j] -= l * a[static_cast<std::size_t>(k) * m + j];
    }
  }
  return true;
}

// ---- ubtf:2064-2100 --------------------------------------------------------------
bool dense_solve(const BlockFactor& bf, double* rhs) {
  const int m = bf.size;
  const double* a = bf.lu.data();
  for (int k = 0; k < m; ++k)
    if (bf.piv[k] != k) std::swap(rhs[k], rhs[bf.piv[k]]);
  for (int i = 0; i < m; ++i) {
    double s = rhs[i];
    for (int j = 0; j < i; ++j) s -= a[static_cast<std::size_t>(i) * m + j] * rhs[j];
    rhs[i] = s;
  }
  for (int i = m - 1; i >= 0; --i) {
    double s = rhs[i];
    for (int j = i + 1; j < m; ++j) s -= a[static_cast<std::size_t>(i) * m + j] * rhs[j];
    const double d = a[static_cast<std::size_t>(i) * m + i];
    if (d == 0.0) return false;
    rhs[i] = s / d;
  }
  return true;
}

// ---- ubtf:1749-1825 via klu_analyze/klu_factor semantics --------------------------
// Local CSC with rows in entry order (coo_to_csc_i32 on the block COO, ubtf:1749);
// COLAMD column order; symmetric permutation (the "diagonal" of position j is
// local row q[j]); scale=1 row sums when size >= 4096 (ubtf:556-568); tol 0.001.
// Updates of column j are applied in ascending pivot position; candidates are
// scanned in ascending row order and the first largest |x| is the partial-pivot
// choice; the diagonal is taken when |x_diag| >= tol * max.
bool klu_like_factor(BlockFactor& bf, const std::vector<std::tuple<int, int, double>>& coo, double tol,
                     int32_t scale_min = 4096) {
  const int32_t s = bf.size;
  std::vector<int32_t> Ap(s + 1, 0), Ai(coo.size());
  std::vector<double> Ax(coo.size());
  for (const auto& t : coo) Ap[std::get<1>(t) + 1]++;
  for (int32_t j = 0; j < s; ++j) Ap[j + 1] += Ap[j];
  {
    std::vector<int32_t> next(Ap.begin(), Ap.end() - 1);
    for (const auto& t : coo) {
      const int32_t p = next[std::get<1>(t)]++;
      Ai[p] = std::get<0>(t);
      Ax[p] = std::get<2>(t);
    }
  }
  bf.q = oracle::colamd(s, Ap.data(), Ai.data());
  bf.rs.assign(s, 1.0);
  if (s >= scale_min) {
    std::fill(bf.rs.begin(), bf.rs.end(), 0.0);
    for (std::size_t p = 0; p < Ai.size(); ++p) bf.rs[Ai[p]] += std::abs(Ax[p]);
    for (auto& v : bf.rs)
      if (v == 0.0) v = 1.0;
  }
  std::vector<int32_t> pinv(s, -1), diag_of(s), col_of_diag(s);
  for (int32_t j = 0; j < s; ++j) {
    diag_of[j] = bf.q[j];
    col_of_diag[bf.q[j]] = j;
  }
  std::vector<double> x(s, 0.0);
  std::vector<int32_t> mark(s, -1), vis(s, -1), reach, cand, stk, pstk;
  bf.Lp.assign(1, 0);
  bf.Up.assign(1, 0);
  bf.prow.clear();
  bf.Ud.clear();
  for (int32_t j = 0; j < s; ++j) {
    const int32_t col = bf.q[j];
    reach.clear();
    cand.clear();
    for (int32_t p = Ap[col]; p < Ap[col + 1]; ++p) {
      const int32_t i = Ai[p];
      x[i] = Ax[p] / bf.rs[i];
      if (pinv[i] < 0) {
        if (mark[i] != j) { mark[i] = j; cand.push_back(i); }
        continue;
      }
      if (vis[pinv[i]] == j) continue;
      // DFS over L(:,k) rows that are pivotal.
      stk.assign(1, pinv[i]);
      pstk.assign(1, bf.Lp[pinv[i]]);
      vis[pinv[i]] = j;
      while (!stk.empty()) {
        const int32_t k = stk.back();
        int32_t& pp = pstk.back();
        bool down = false;
        while (pp < bf.Lp[k + 1]) {
          const int32_t r = bf.Li[pp++];
          const int32_t kk = pinv[r];
          if (kk < 0) {
            if (mark[r] != j) { mark[r] = j; cand.push_back(r); }
            continue;
          }
          if (vis[kk] == j) continue;
          vis[kk] = j;
          stk.push_back(kk);
          pstk.push_back(bf.Lp[kk]);
          down = true;
          break;
        }
        if (!down) {
          reach.push_back(k);
          stk.pop_back();
          pstk.pop_back();
        }
      }
    }
    std::sort(reach.begin(), reach.end());
    for (int32_t k : reach) {
      const double xk = x[bf.prow[k]];
      for (int32_t p = bf.Lp[k]; p < bf.Lp[k + 1]; ++p) x[bf.Li[p]] -= bf.Lx[p] * xk;
    }
    std::sort(cand.begin(), cand.end());
    double amax = -1.0;
    int32_t imax = -1;
    for (int32_t i : cand)
      if (std::abs(x[i]) > amax) { amax = std::abs(x[i]); imax = i; }
    if (imax < 0 || amax == 0.0) return false;
    const int32_t d = diag_of[j];
    int32_t piv = imax;
    if (pinv[d] < 0 && std::abs(x[d]) >= tol * amax) piv = d;
    if (piv != d && pinv[d] < 0) {
      const int32_t kbar = col_of_diag[piv];
      diag_of[kbar] = d;
      col_of_diag[d] = kbar;
    }
    const double pv = x[piv];
    bf.prow.push_back(piv);
    pinv[piv] = j;
    bf.Ud.push_back(pv);
    for (int32_t i : cand) {
      if (i != piv) {
        bf.Li.push_back(i);
        bf.Lx.push_back(x[i] / pv);
      }
      x[i] = 0.0;
    }
    bf.Lp.push_back(static_cast<int32_t>(bf.Li.size()));
    for (int32_t k : reach) {
      bf.Ui.push_back(k);
      bf.Ux.push_back(x[bf.prow[k]]);
      x[bf.prow[k]] = 0.0;
    }
    bf.Up.push_back(static_cast<int32_t>(bf.Ui.size()));
  }
  return true;
}

// klu_solve semantics: x = Q U^{-1} L^{-1} P R^{-1} b, in place on the block rhs.
bool klu_like_solve(const BlockFactor& bf, double* rhs) {
  const int32_t s = bf.size;
  std::vector<double> y(rhs, rhs + s);
  for (int32_t i = 0; i < s; ++i) y[i] /= bf.rs[i];
  std::vector<double> z(s);
  for (int32_t k = 0; k < s; ++k) {
    const double t = y[bf.prow[k]];
    z[k] = t;
    for (int32_t p = bf.Lp[k]; p < bf.Lp[k + 1]; ++p) y[bf.Li[p]] -= bf.Lx[p] * t;
  }
  for (int32_t j = s - 1; j >= 0; --j) {
    if (bf.Ud[j] == 0.0) return false;
    z[j] /= bf.Ud[j];
    const double t = z[j];
    for (int32_t p = bf.Up[j]; p < bf.Up[j + 1]; ++p) z[bf.Ui[p]] -= bf.Ux[p] * t;
  }
  for (int32_t j = 0; j < s; ++j) rhs[bf.q[j]] = z[j];
  return true;
}

struct Solver {
  Coo coo;
  int32_t hint = -1, n = 0;
  bool validated = false, decomposed = false, factored = false;
  int err = 0;
  // decomposition (ubtf.h:13-44)
  int32_t nblocks_raw = 0, nmatch = 0, nblocks = 0;
  std::vector<int32_t> P, Q, invP, invQ, R;
  std::vector<int32_t> off_ptr, off_col;
  std::vector<double> off_val;
  std::vector<std::vector<std::tuple<int, int, double>>> blocks;
  std::vector<BlockFactor> fac;
  double t_analyze = 0, t_factor = 0, t_solve = 0;
  int threads_used = 1;
  // Global-KLU route: every block by klu_like_factor with klu_tol, whole rows
  // scaled by rs (empty: unscaled).
  int route = 0;
  double klu_tol = 0.001;
  std::vector<double> rs;
};

int analyze(Solver& s) {
  using Clock = std::chrono::steady_clock;
  const auto t0 = Clock::now();
  if (!s.validated) {
    const int e = validate_and_canonicalize(s.coo, s.hint, s.n);
    if (e) return e;
    s.validated = true;
  }
  const int32_t n = s.n;
  // ubtf:582-589: drop |v| < 1e-20 iff n >= 50000.
  const double drop = n >= 50000 ? 1e-20 : 0.0;
  std::vector<int32_t> Ap(n + 1, 0);
  std::vector<int32_t> Ai;
  for (const auto& t : s.coo)
    if (!(drop > 0 && std::abs(std::get<2>(t)) < drop)) Ap[std::get<1>(t) + 1]++;
  for (int32_t j = 0; j < n; ++j) Ap[j + 1] += Ap[j];
  Ai.resize(Ap[n]);
  {
    std::vector<int32_t> next(Ap.begin(), Ap.end() - 1);
    for (const auto& t : s.coo)
      if (!(drop > 0 && std::abs(std::get<2>(t)) < drop)) Ai[next[std::get<1>(t)]++] = std::get<0>(t);
  }
  std::vector<int32_t> Rraw;
  s.nblocks_raw = btf_order(n, Ap, Ai, s.P, s.Q, Rraw, s.nmatch);
  for (auto& q : s.Q) q = q < -1 ? -q - 2 : q;  // ubtf:703-708
  s.R = merge_blocks(Rraw, n, 64);
  s.nblocks = static_cast<int32_t>(s.R.size()) - 1;
  s.invP.assign(n, -1);
  s.invQ.assign(n, -1);
  for (int32_t k = 0; k < n; ++k) {
    s.invP[s.P[k]] = k;
    s.invQ[s.Q[k]] = k;
  }
  std::vector<int32_t> blk(n);
  for (int32_t b = 0; b < s.nblocks; ++b)
    for (int32_t k = s.R[b]; k < s.R[b + 1]; ++k) blk[k] = b;
  // ubtf:739-853
  std::vector<int32_t> cnt(n + 1, 0);
  s.blocks.assign(s.nblocks, {});
  for (const auto& t : s.coo) {
    if (drop > 0 && std::abs(std::get<2>(t)) < drop) continue;
    const int32_t rp = s.invP[std::get<0>(t)], cp = s.invQ[std::get<1>(t)];
    if (blk[cp] < blk[rp]) return kAnalysis;
    if (blk[cp] > blk[rp]) cnt[rp + 1]++;
  }
  for (int32_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  s.off_ptr = cnt;
  s.off_col.resize(cnt[n]);
  s.off_val.resize(cnt[n]);
  std::vector<int32_t> next(cnt.begin(), cnt.end() - 1);
  for (const auto& t : s.coo) {
    if (drop > 0 && std::abs(std::get<2>(t)) < drop) continue;
    const int32_t rp = s.invP[std::get<0>(t)], cp = s.invQ[std::get<1>(t)];
    const int32_t br = blk[rp], bc = blk[cp];
    if (bc == br) {
      s.blocks[br].emplace_back(rp - s.R[br], cp - s.R[br], std::get<2>(t));
    } else {
      const int32_t p = next[rp]++;
      s.off_col[p] = cp;
      s.off_val[p] = std::get<2>(t);
    }
  }
  s.decomposed = true;
  s.factored = false;
  s.t_analyze = std::chrono::duration<double>(Clock::now() - t0).count();
  return kNone;
}

// ---- global-KLU route: src/solver.cpp:878-968 (analyze), :1190-1242 (factor),
// :1472-1479 (solve), through klu_analyze / klu_factor / klu_solve ------------------
// klu_analyze with ordering = 1: the CSC of the full canonical pattern
// (coo_to_csc_i32, no drop rule), BTF when btf = 1 (btf_order, Q unflipped, no
// merging) or the whole matrix as one block, then COLAMD on every block (in
// klu_like_factor). klu_factor with scale = 1 / 2 divides every row by the sum /
// maximum of |a_ij| over the whole row (klu_scale; an all-zero row gets 1),
// including the entries outside the diagonal blocks, and factors each block with
// the threshold rule at tol; a block of size 1 pivots on its entry.
int analyze_klu(Solver& s, int btf, int scale, double tol) {
  using Clock = std::chrono::steady_clock;
  const auto t0 = Clock::now();
  if (!s.validated) {
    const int e = validate_and_canonicalize(s.coo, s.hint, s.n);
    if (e) return e;
    s.validated = true;
  }
  const int32_t n = s.n;
  s.route = 1;
  s.klu_tol = tol;
  s.rs.clear();
  if (scale > 0) {
    s.rs.assign(n, 0.0);
    for (const auto& t : s.coo) {
      const double a = std::abs(std::get<2>(t));
      double& r = s.rs[std::get<0>(t)];
      r = scale == 1 ? r + a : std::max(r, a);
    }
    for (double& r : s.rs)
      if (r == 0.0) r = 1.0;
  }
  auto val = [&](const std::tuple<int, int, double>& t) {
    return s.rs.empty() ? std::get<2>(t) : std::get<2>(t) / s.rs[std::get<0>(t)];
  };
  if (btf) {
    std::vector<int32_t> Ap(n + 1, 0), Ai(s.coo.size());
    for (const auto& t : s.coo) Ap[std::get<1>(t) + 1]++;
    for (int32_t j = 0; j < n; ++j) Ap[j + 1] += Ap[j];
    std::vector<int32_t> next(Ap.begin(), Ap.end() - 1);
    for (const auto& t : s.coo) Ai[next[std::get<1>(t)]++] = std::get<0>(t);
    s.nblocks_raw = btf_order(n, Ap, Ai, s.P, s.Q, s.R, s.nmatch);
    for (auto& q : s.Q) q = q < -1 ? -q - 2 : q;
  } else {
    s.P.resize(n);
    s.Q.resize(n);
    for (int32_t k = 0; k < n; ++k) s.P[k] = s.Q[k] = k;
    s.R = {0, n};
    s.nblocks_raw = 1;
    s.nmatch = n;
  }
  s.nblocks = static_cast<int32_t>(s.R.size()) - 1;
  s.invP.assign(n, -1);
  s.invQ.assign(n, -1);
  for (int32_t k = 0; k < n; ++k) {
    s.invP[s.P[k]] = k;
    s.invQ[s.Q[k]] = k;
  }
  std::vector<int32_t> blk(n);
  for (int32_t b = 0; b < s.nblocks; ++b)
    for (int32_t k = s.R[b]; k < s.R[b + 1]; ++k) blk[k] = b;
  std::vector<int32_t> cnt(n + 1, 0);
  s.blocks.assign(s.nblocks, {});
  for (const auto& t : s.coo) {
    const int32_t rp = s.invP[std::get<0>(t)], cp = s.invQ[std::get<1>(t)];
    if (blk[cp] < blk[rp]) return kAnalysis;
    if (blk[cp] > blk[rp]) cnt[rp + 1]++;
  }
  for (int32_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  s.off_ptr = cnt;
  s.off_col.resize(cnt[n]);
  s.off_val.resize(cnt[n]);
  std::vector<int32_t> next(cnt.begin(), cnt.end() - 1);
  for (const auto& t : s.coo) {
    const int32_t rp = s.invP[std::get<0>(t)], cp = s.invQ[std::get<1>(t)];
    const int32_t br = blk[rp], bc = blk[cp];
    if (bc == br) {
      s.blocks[br].emplace_back(rp - s.R[br], cp - s.R[br], val(t));
    } else {
      const int32_t p = next[rp]++;
      s.off_col[p] = cp;
      s.off_val[p] = val(t);
    }
  }
  s.decomposed = true;
  s.factored = false;
  s.t_analyze = std::chrono::duration<double>(Clock::now() - t0).count();
  return kNone;
}

// src/solver.cpp:366-390 thread policy (work threads for factor).
int work_threads(const Solver& s, int hw) {
  const int cap = std::min(hw, 16);
  int t = std::min(cap, 4);
  int32_t mx = 0;
  for (int32_t b = 0; b < s.nblocks; ++b) mx = std::max(mx, s.R[b + 1] - s.R[b]);
  const double ratio = s.n > 0 ? static_cast<double>(mx) / s.n : 1.0;
  if (ratio > 0.80) t = std::min(t, 2);
  if (s.n >= 50000 && s.nblocks >= 250 && ratio < 0.20) t = std::min(cap, 6);
  if (s.nblocks > 0) t = std::min(t, s.nblocks);
  return std::max(t, 1);
}

int factor(Solver& s, int threads) {
  using Clock = std::chrono::steady_clock;
  const auto t0 = Clock::now();
  if (!s.decomposed) return kNotPrepared;
  s.fac.assign(s.nblocks, BlockFactor{});
  if (threads <= 0) {
    const unsigned hw = std::thread::hardware_concurrency();
    threads = work_threads(s, hw ? static_cast<int>(hw) : 8);
  }
  s.threads_used = threads;
  std::mutex mu;
  std::size_t next = 0;
  const std::size_t chunk = s.nblocks > 2048 ? 16 : 1;
  auto worker = [&]() {
    while (true) {
      std::size_t lo;
      {
        std::lock_guard<std::mutex> g(mu);
        if (next >= static_cast<std::size_t>(s.nblocks)) return;
        lo = next;
        next += chunk;
      }
      const std::size_t hi = std::min<std::size_t>(lo + chunk, s.nblocks);
      for (std::size_t b = lo; b < hi; ++b) {
        BlockFactor& bf = s.fac[b];
        bf.start = s.R[b];
        bf.size = s.R[b + 1] - s.R[b];
        if (s.route == 1)
          bf.ok = klu_like_factor(bf, s.blocks[b], s.klu_tol, std::numeric_limits<int32_t>::max());
        else
          bf.ok = bf.size <= 64 ? dense_factor(bf, s.blocks[b]) : klu_like_factor(bf, s.blocks[b], 0.001);
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < threads; ++t) th.emplace_back(worker);
  worker();
  for (auto& x : th) x.join();
  s.t_factor = std::chrono::duration<double>(Clock::now() - t0).count();
  for (const auto& bf : s.fac)
    if (!bf.ok) return kFactor;
  s.factored = true;
  return kNone;
}

// ubtf:2220-2414 (reverse block order) + src/solver.cpp:1486-1536 refinement.
bool solve_ubtf(const Solver& s, std::vector<double>& b) {
  const int32_t n = s.n;
  std::vector<double> bp(n), xp(n, 0.0), rhs;
  for (int32_t i = 0; i < n; ++i) bp[i] = s.rs.empty() ? b[s.P[i]] : b[s.P[i]] / s.rs[s.P[i]];
  for (int32_t bid = s.nblocks - 1; bid >= 0; --bid) {
    const BlockFactor& bf = s.fac[bid];
    rhs.resize(bf.size);
    for (int32_t r = bf.start; r < bf.start + bf.size; ++r) {
      double v = bp[r];
      for (int32_t p = s.off_ptr[r]; p < s.off_ptr[r + 1]; ++p) v -= s.off_val[p] * xp[s.off_col[p]];
      rhs[r - bf.start] = v;
    }
    if (!(bf.dense ? dense_solve(bf, rhs.data()) : klu_like_solve(bf, rhs.data()))) return false;
    for (int32_t i = 0; i < bf.size; ++i) xp[bf.start + i] = rhs[i];
  }
  for (int32_t i = 0; i < n; ++i) b[s.Q[i]] = xp[i];
  return true;
}

int solve(Solver& s, double* bptr, int refine_iters) {
  using Clock = std::chrono::steady_clock;
  const auto t0 = Clock::now();
  if (!s.factored) return kNotPrepared;
  std::vector<double> b(bptr, bptr + s.n);
  for (double v : b)
    if (!std::isfinite(v)) return kNonFinite;
  const std::vector<double> b0 = b;
  if (!solve_ubtf(s, b)) return kSolve;
  double bn = 0.0;
  for (double v : b0) bn += v * v;
  bn = std::sqrt(bn);
  for (int it = 0; it < refine_iters && bn != 0.0; ++it) {
    std::vector<double> r = b0;
    for (const auto& t : s.coo) r[std::get<0>(t)] -= std::get<2>(t) * b[std::get<1>(t)];
    double rn = 0.0;
    for (double v : r) rn += v * v;
    if (std::sqrt(rn) / bn < 1e-12) break;
    if (!solve_ubtf(s, r)) return kSolve;
    for (int32_t i = 0; i < s.n; ++i) b[i] += r[i];
  }
  std::copy(b.begin(), b.end(), bptr);
  s.t_solve = std::chrono::duration<double>(Clock::now() - t0).count();
  return kNone;
}

}  // namespace oracle

extern "C" {

void* oracle_create(int64_t nnz, const int32_t* rows, const int32_t* cols, const double* vals, int32_t dim) {
  auto* s = new oracle::Solver;
  s->coo.resize(static_cast<std::size_t>(nnz));
  for (int64_t k = 0; k < nnz; ++k) s->coo[k] = std::make_tuple(rows[k], cols[k], vals[k]);
  s->hint = dim;
  return s;
}
void oracle_free(void* h) { delete static_cast<oracle::Solver*>(h); }
int oracle_analyze(void* h) { return oracle::analyze(*static_cast<oracle::Solver*>(h)); }
int oracle_analyze_klu(void* h, int btf, int scale, double tol) {
  return oracle::analyze_klu(*static_cast<oracle::Solver*>(h), btf, scale, tol);
}
int oracle_factor(void* h, int threads) { return oracle::factor(*static_cast<oracle::Solver*>(h), threads); }
int oracle_solve(void* h, double* b, int refine_iters) {
  return oracle::solve(*static_cast<oracle::Solver*>(h), b, refine_iters);
}
// Replaces the values of the stored (canonical) matrix, pattern unchanged.
void oracle_set_values(void* h, const double* canonical_vals) {
  auto* s = static_cast<oracle::Solver*>(h);
  for (std::size_t k = 0; k < s->coo.size(); ++k) std::get<2>(s->coo[k]) = canonical_vals[k];
  s->decomposed = s->factored = false;
}
int32_t oracle_n(void* h) { return static_cast<oracle::Solver*>(h)->n; }
int64_t oracle_nnz(void* h) { return static_cast<int64_t>(static_cast<oracle::Solver*>(h)->coo.size()); }
void oracle_canonical(void* h, int32_t* rows, int32_t* cols, double* vals) {
  const auto* s = static_cast<oracle::Solver*>(h);
  for (std::size_t k = 0; k < s->coo.size(); ++k) {
    rows[k] = std::get<0>(s->coo[k]);
    cols[k] = std::get<1>(s->coo[k]);
    vals[k] = std::get<2>(s->coo[k]);
  }
}
void oracle_info(void* h, int32_t* nblocks_raw, int32_t* nblocks, int32_t* nmatch, int32_t* threads) {
  const auto* s = static_cast<oracle::Solver*>(h);
  *nblocks_raw = s->nblocks_raw;
  *nblocks = s->nblocks;
  *nmatch = s->nmatch;
  *threads = s->threads_used;
}
void oracle_perm(void* h, int32_t* P, int32_t* Q, int32_t* R) {
  const auto* s = static_cast<oracle::Solver*>(h);
  std::copy(s->P.begin(), s->P.end(), P);
  std::copy(s->Q.begin(), s->Q.end(), Q);
  std::copy(s->R.begin(), s->R.end(), R);
}
// Dense block b: row-major LU (size^2) and pivots; returns size or -1.
int32_t oracle_dense_factor(void* h, int32_t b, double* lu, int32_t* piv) {
  const auto* s = static_cast<oracle::Solver*>(h);
  if (b < 0 || b >= static_cast<int32_t>(s->fac.size()) || !s->fac[b].dense) return -1;  // not factored yet
  const auto& f = s->fac[b];
  if (lu) std::copy(f.lu.begin(), f.lu.end(), lu);
  if (piv) std::copy(f.piv.begin(), f.piv.end(), piv);
  return f.size;
}
// Sparse (KLU-like) block b: column order q (size) and pivot rows (size).
int32_t oracle_sparse_order(void* h, int32_t b, int32_t* q, int32_t* prow) {
  const auto* s = static_cast<oracle::Solver*>(h);
  if (b < 0 || b >= static_cast<int32_t>(s->fac.size()) || s->fac[b].dense || s->fac[b].q.empty()) return -1;
  const auto& f = s->fac[b];
  if (q) std::copy(f.q.begin(), f.q.end(), q);
  if (prow) std::copy(f.prow.begin(), f.prow.end(), prow);
  return f.size;
}
void oracle_timings(void* h, double* t3) {
  const auto* s = static_cast<oracle::Solver*>(h);
  t3[0] = s->t_analyze;
  t3[1] = s->t_factor;
  t3[2] = s->t_solve;
}
void oracle_signatures(void* h, uint64_t* s1, uint64_t* s2, uint64_t* s3) {
  const auto* s = static_cast<oracle::Solver*>(h);
  oracle::signatures(s->n, s->coo, s1, s2, s3);
}
// Standalone btf_order on a CSC pattern (unflipped Q); returns nblocks.
int32_t oracle_btf_order(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t* P, int32_t* Q, int32_t* R,
                         int32_t* nmatch) {
  std::vector<int32_t> ap(Ap, Ap + n + 1), ai(Ai, Ai + Ap[n]), p, q, r;
  const int32_t nb = oracle::btf_order(n, ap, ai, p, q, r, *nmatch);
  for (int32_t k = 0; k < n; ++k) {
    P[k] = p[k];
    Q[k] = q[k] < -1 ? -q[k] - 2 : q[k];
  }
  std::copy(r.begin(), r.end(), R);
  return nb;
}
void oracle_colamd(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t* order) {
  const std::vector<int32_t> q = oracle::colamd(n, Ap, Ai);
  std::copy(q.begin(), q.end(), order);
}

int32_t oracle_merge(const int32_t* R, int32_t nb, int32_t n, int32_t merge_max, int32_t* out) {
  std::vector<int32_t> r(R, R + nb + 1);
  const auto m = oracle::merge_blocks(r, n, merge_max);
  std::copy(m.begin(), m.end(), out);
  return static_cast<int32_t>(m.size()) - 1;
}

}  // extern "C"
