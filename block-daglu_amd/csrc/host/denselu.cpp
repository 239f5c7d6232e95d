// Dense threshold-pivoting LU used once per pattern to fix the pivot order of a
// sparse block's dense Schur tail (the device then refactors on that order).
// Pivot rule = KLU's (tol, diagonal preference, diagonal tracking), the rule
// the reference gets from klu_factor for blocks above the dense threshold.
// Blocked right-looking: unblocked panel of width 64 with the pivot search,
// row swaps across the matrix, triangular solve for the U row block, and a
// threaded rank-64 update of the trailing matrix.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "symbolic.h"

namespace bdlu {

namespace {
constexpr int32_t kPanel = 64;

// C(rows r0.., cols c0..c1) -= A(rows r0.., cols k0..k0+kb) * B(rows k0..k0+kb, cols c0..c1)
void trailing_update(int32_t m, double* a, int32_t k0, int32_t kb, int32_t c0, int32_t c1) {
  const int32_t r0 = k0 + kb;
  const std::size_t ld = static_cast<std::size_t>(m);
  int32_t j = c0;
  for (; j + 4 <= c1; j += 4) {
    double* C0 = a + ld * j;
    double* C1 = C0 + ld;
    double* C2 = C1 + ld;
    double* C3 = C2 + ld;
    for (int32_t k = k0; k < r0; ++k) {
      const double* A = a + ld * k;
      const double b0 = C0[k], b1 = C1[k], b2 = C2[k], b3 = C3[k];
      for (int32_t i = r0; i < m; ++i) {
        const double x = A[i];
        C0[i] -= x * b0;
        C1[i] -= x * b1;
        C2[i] -= x * b2;
        C3[i] -= x * b3;
      }
    }
  }
  for (; j < c1; ++j) {
    double* C = a + ld * j;
    for (int32_t k = k0; k < r0; ++k) {
      const double* A = a + ld * k;
      const double b = C[k];
      for (int32_t i = r0; i < m; ++i) C[i] -= A[i] * b;
    }
  }
}
}  // namespace

bool dense_threshold_lu(int32_t m, double* a, const std::vector<int32_t>& diag_in, double tol,
                        std::vector<int32_t>& perm) {
  const std::size_t ld = static_cast<std::size_t>(m);
  perm.resize(m);
  // pos[row] = current position of original row `row` (rows are 0..m-1 here).
  std::vector<int32_t> pos(m);
  for (int32_t i = 0; i < m; ++i) perm[i] = pos[i] = i;
  std::vector<int32_t> diag(diag_in), col_of_diag(m, -1);
  for (int32_t t = 0; t < m; ++t)
    if (diag[t] >= 0) col_of_diag[diag[t]] = t;
  std::vector<char> pivotal(m, 0);
  unsigned hw = std::thread::hardware_concurrency();
  const int32_t nthreads = static_cast<int32_t>(std::max(1u, std::min(hw ? hw : 1u, 16u)));

  for (int32_t k0 = 0; k0 < m; k0 += kPanel) {
    const int32_t kb = std::min(kPanel, m - k0);
    // Unblocked panel: columns k0..k0+kb over rows k0..m-1.
    for (int32_t k = k0; k < k0 + kb; ++k) {
      double* col = a + ld * k;
      double amax = -1.0;
      int32_t imax = -1;
      for (int32_t i = k; i < m; ++i) {
        const double v = std::abs(col[i]);
        if (v > amax || (v == amax && imax >= 0 && perm[i] < perm[imax])) {
          amax = v;
          imax = i;
        }
      }
      if (imax < 0 || amax == 0.0) return false;
      int32_t ip = imax;
      const int32_t d = diag[k];
      if (d >= 0 && !pivotal[d]) {
        const int32_t dp = pos[d];
        if (std::abs(col[dp]) >= tol * amax) ip = dp;
      }
      const int32_t prow = perm[ip];
      if (prow != d && d >= 0 && !pivotal[d]) {
        const int32_t kbar = col_of_diag[prow];
        if (kbar >= 0) {
          diag[kbar] = d;
          col_of_diag[d] = kbar;
        }
      }
      pivotal[prow] = 1;
      if (ip != k) {
        // Swap whole rows k and ip (all columns).
        for (int32_t j = 0; j < m; ++j) std::swap(a[ld * j + k], a[ld * j + ip]);
        std::swap(perm[k], perm[ip]);
        pos[perm[k]] = k;
        pos[perm[ip]] = ip;
      }
      const double piv = col[k];
      for (int32_t i = k + 1; i < m; ++i) col[i] /= piv;
      // Rank-1 update of the rest of the panel.
      for (int32_t j = k + 1; j < k0 + kb; ++j) {
        double* cj = a + ld * j;
        const double b = cj[k];
        if (b == 0.0) continue;
        for (int32_t i = k + 1; i < m; ++i) cj[i] -= col[i] * b;
      }
    }
    const int32_t c0 = k0 + kb;
    if (c0 >= m) break;
    // U row block: solve L11 * U12 = A12 (unit lower, columns c0..m-1).
    for (int32_t j = c0; j < m; ++j) {
      double* cj = a + ld * j;
      for (int32_t k = k0; k < k0 + kb; ++k) {
        const double b = cj[k];
        if (b == 0.0) continue;
        const double* ak = a + ld * k;
        for (int32_t i = k + 1; i < k0 + kb; ++i) cj[i] -= ak[i] * b;
      }
    }
    // Trailing update, threaded over column ranges.
    const int32_t ncols = m - c0;
    const int32_t nt = std::min<int32_t>(nthreads, std::max<int32_t>(1, ncols / 64));
    if (nt <= 1) {
      trailing_update(m, a, k0, kb, c0, m);
    } else {
      std::vector<std::thread> th;
      const int32_t chunk = (ncols + nt - 1) / nt;
      for (int32_t t = 0; t < nt; ++t) {
        const int32_t lo = c0 + t * chunk, hi = std::min(m, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back(trailing_update, m, a, k0, kb, lo, hi);
      }
      for (auto& x : th) x.join();
    }
  }
  return true;
}

}  // namespace bdlu
