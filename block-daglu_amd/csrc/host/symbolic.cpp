// Symbolic analysis + first pivoted factorization of the sparse leading columns
// of one diagonal block (the KLU branch of factor_one_block, ubtf:1749-1825).
#include "symbolic.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "blockdaglu/env.h"
#include "host.h"

namespace bdlu {

SymbolicOptions symbolic_options_from_env() {
  SymbolicOptions o;
  // BLOCK_KLU_TOL / BLOCK_KLU_TOL_BIG: pivot tolerance override (ubtf:536-554).
  if (const char* v = blockdaglu_getenv_compat("USERALGO_BLOCK_KLU_TOL")) {
    const double d = std::atof(v);
    if (d >= 0.0) o.tol = std::min(1.0, d);
  }
  if (const char* v = std::getenv("BLOCKDAGLU_GPU_TAIL_TOL")) {
    const double d = std::atof(v);
    if (d > 0.0) o.tail_tol = std::min(1.0, d);
  }
  if (const char* v = std::getenv("BLOCKDAGLU_GPU_TAIL_DENSITY")) {
    const double d = std::atof(v);
    if (d > 0.0) o.tail_density = d;
  }
  if (const char* v = std::getenv("BLOCKDAGLU_GPU_TAIL_MIN")) {
    const int x = std::atoi(v);
    if (x > 0) o.tail_min = x;
  }
  if (const char* v = blockdaglu_getenv_compat("USERALGO_BLOCK_KLU_BIG_N")) {
    const int x = std::atoi(v);
    if (x > 0) o.scale_min_n = x;
  }
  return o;
}

namespace {
// Pivot positions reachable from the rows of one column through the graph of
// L(:, 0:nk): node k -> rows of L(:,k) that are pivotal. Non-pivotal rows met
// on the way are appended to `cand` (once, via `row_mark`).
struct Reach {
  std::vector<int32_t> stack, pstack, visit_mark, row_mark;
  int32_t stamp = 0;
  void init(int32_t s) {
    stack.resize(s);
    pstack.resize(s);
    visit_mark.assign(s, -1);
    row_mark.assign(s, -1);
  }
  template <class RowIt>
  void run(RowIt rb, RowIt re, const std::vector<int32_t>& pinv, const std::vector<int32_t>& l_ptr,
           const std::vector<int32_t>& l_row, std::vector<int32_t>& reach, std::vector<int32_t>* cand) {
    ++stamp;
    reach.clear();
    if (cand) cand->clear();
    auto touch_row = [&](int32_t i) {
      if (pinv[i] < 0) {
        if (cand && row_mark[i] != stamp) {
          row_mark[i] = stamp;
          cand->push_back(i);
        }
      }
    };
    for (RowIt it = rb; it != re; ++it) {
      const int32_t i = *it;
      const int32_t k0 = pinv[i];
      if (k0 < 0) {
        touch_row(i);
        continue;
      }
      if (visit_mark[k0] == stamp) continue;
      int32_t head = 0;
      stack[0] = k0;
      visit_mark[k0] = stamp;
      pstack[0] = l_ptr[k0];
      while (head >= 0) {
        const int32_t k = stack[head];
        int32_t p = pstack[head];
        const int32_t pe = l_ptr[k + 1];
        bool pushed = false;
        for (; p < pe; ++p) {
          const int32_t r = l_row[p];
          const int32_t kk = pinv[r];
          if (kk < 0) {
            touch_row(r);
            continue;
          }
          if (visit_mark[kk] == stamp) continue;
          pstack[head] = p + 1;
          visit_mark[kk] = stamp;
          stack[++head] = kk;
          pstack[head] = l_ptr[kk];
          pushed = true;
          break;
        }
        if (!pushed) {
          reach.push_back(k);
          --head;
        }
      }
    }
    std::sort(reach.begin(), reach.end());
  }
};
}  // namespace

bool analyze_sparse_block(int32_t s, const std::vector<int32_t>& rows, const std::vector<int32_t>& cols,
                          const std::vector<double>& vals, const SymbolicOptions& opt, SparseBlockPlan& out,
                          SparseFactorValues* values) {
  out = SparseBlockPlan{};
  out.s = s;
  const int32_t ne = static_cast<int32_t>(rows.size());
  // Local CSC, rows in entry order inside a column (the reference builds it with
  // coo_to_csc_i32 from the block's canonical-order COO, ubtf:1749).
  std::vector<int32_t> Ap(static_cast<std::size_t>(s) + 1, 0), Ai(ne), Ae(ne);
  for (int32_t t = 0; t < ne; ++t) Ap[cols[t] + 1]++;
  for (int32_t j = 0; j < s; ++j) Ap[j + 1] += Ap[j];
  {
    std::vector<int32_t> next(Ap.begin(), Ap.end() - 1);
    for (int32_t t = 0; t < ne; ++t) {
      const int32_t p = next[cols[t]]++;
      Ai[p] = rows[t];
      Ae[p] = t;
    }
  }
  out.q = colamd_order(s, Ap.data(), Ai.data());
  out.scale = s >= opt.scale_min_n;
  std::vector<double> rs(s, 1.0);
  if (out.scale) {
    std::fill(rs.begin(), rs.end(), 0.0);
    // Each row's terms in ascending local column order, the order klu_scale
    // visits the block's CSC (a sum's last bits depend on its order).
    for (int32_t c = 0; c < s; ++c)
      for (int32_t p = Ap[c]; p < Ap[c + 1]; ++p) {
        const double a = std::abs(vals[Ae[p]]);
        double& r = rs[Ai[p]];
        r = opt.scale_kind == 2 ? std::max(r, a) : r + a;
      }
    for (int32_t i = 0; i < s; ++i)
      if (rs[i] == 0.0) rs[i] = 1.0;
  }

  // Entries grouped by pivot-column position.
  out.a_ptr.assign(static_cast<std::size_t>(s) + 1, 0);
  for (int32_t j = 0; j < s; ++j) out.a_ptr[j + 1] = out.a_ptr[j] + (Ap[out.q[j] + 1] - Ap[out.q[j]]);
  out.a_row.resize(ne);
  out.a_ent.resize(ne);
  for (int32_t j = 0; j < s; ++j) {
    const int32_t c = out.q[j];
    int32_t w = out.a_ptr[j];
    for (int32_t p = Ap[c]; p < Ap[c + 1]; ++p, ++w) {
      out.a_row[w] = Ai[p];
      out.a_ent[w] = Ae[p];
    }
  }

  std::vector<int32_t> pinv(s, -1), diag_of(s), col_of_diag(s);
  for (int32_t j = 0; j < s; ++j) {
    diag_of[j] = out.q[j];
    col_of_diag[out.q[j]] = j;
  }
  std::vector<double> x(s, 0.0);
  std::vector<double> lx;
  out.l_ptr.assign(1, 0);
  out.u_ptr.assign(1, 0);
  Reach reach_state;
  reach_state.init(s);
  std::vector<int32_t> reach, cand;
  int32_t k0 = s;
  for (int32_t j = 0; j < s; ++j) {
    const int32_t pa = out.a_ptr[j], pe = out.a_ptr[j + 1];
    reach_state.run(out.a_row.begin() + pa, out.a_row.begin() + pe, pinv, out.l_ptr, out.l_row, reach, &cand);
    for (int32_t p = pa; p < pe; ++p) x[out.a_row[p]] = vals[out.a_ent[p]] / rs[out.a_row[p]];
    for (int32_t k : reach) {
      const double xk = x[out.prow[k]];
      for (int32_t p = out.l_ptr[k]; p < out.l_ptr[k + 1]; ++p) x[out.l_row[p]] -= lx[p] * xk;
    }
    const int32_t left = s - j;
    if (left >= opt.tail_min && static_cast<double>(cand.size()) > opt.tail_density * left) {
      for (int32_t k : reach) x[out.prow[k]] = 0.0;
      for (int32_t i : cand) x[i] = 0.0;
      k0 = j;
      break;
    }
    std::sort(cand.begin(), cand.end());
    double amax = -1.0;
    int32_t imax = -1;
    for (int32_t i : cand) {
      const double a = std::abs(x[i]);
      if (a > amax) {
        amax = a;
        imax = i;
      }
    }
    if (imax < 0 || amax == 0.0) return false;
    const int32_t diag = diag_of[j];
    int32_t piv = imax;
    if (pinv[diag] < 0 && std::abs(x[diag]) >= opt.tol * amax) {
      // diag is a candidate whenever it is non-pivotal and in the pattern; a
      // zero x[diag] outside the pattern cannot pass the test since amax > 0.
      piv = diag;
    }
    if (piv != diag) {
      ++out.offdiag_pivots;
      if (pinv[diag] < 0) {
        const int32_t kbar = col_of_diag[piv];
        diag_of[kbar] = diag;
        col_of_diag[diag] = kbar;
      }
    }
    const double pivot = x[piv];
    out.prow.push_back(piv);
    pinv[piv] = j;
    for (int32_t i : cand) {
      if (i == piv) continue;
      out.l_row.push_back(i);
      lx.push_back(x[i] / pivot);
    }
    out.l_ptr.push_back(static_cast<int32_t>(out.l_row.size()));
    for (int32_t k : reach) {
      out.u_idx.push_back(k);
      out.updates += out.l_ptr[k + 1] - out.l_ptr[k];
      if (values) values->ux.push_back(x[out.prow[k]]);
    }
    out.u_ptr.push_back(static_cast<int32_t>(out.u_idx.size()));
    if (values) values->udiag.push_back(pivot);
    for (int32_t k : reach) x[out.prow[k]] = 0.0;
    for (int32_t i : cand) x[i] = 0.0;
  }
  out.k0 = k0;
  if (values) {
    values->lx = lx;
    values->rs = rs;
  }
  out.m = s - k0;

  // Tail: remaining rows, starting diagonal bookkeeping, and U(0:k0, j) patterns.
  for (int32_t i = 0; i < s; ++i)
    if (pinv[i] < 0) out.trow.push_back(i);
  std::vector<int32_t> tail_index(s, -1);
  for (int32_t t = 0; t < out.m; ++t) tail_index[out.trow[t]] = t;
  out.tdiag.resize(out.m);
  for (int32_t t = 0; t < out.m; ++t) {
    const int32_t d = diag_of[k0 + t];
    out.tdiag[t] = pinv[d] < 0 ? tail_index[d] : -1;
  }
  // Schur complement of the tail (host copy, only to fix the tail pivot order).
  const std::size_t mm = static_cast<std::size_t>(out.m);
  std::vector<double> S(mm * mm, 0.0);
  for (int32_t j = k0; j < s; ++j) {
    const int32_t pa = out.a_ptr[j], pe = out.a_ptr[j + 1];
    reach_state.run(out.a_row.begin() + pa, out.a_row.begin() + pe, pinv, out.l_ptr, out.l_row, reach, nullptr);
    for (int32_t p = pa; p < pe; ++p) x[out.a_row[p]] = vals[out.a_ent[p]] / rs[out.a_row[p]];
    for (int32_t k : reach) {
      out.u_idx.push_back(k);
      out.updates += out.l_ptr[k + 1] - out.l_ptr[k];
      const double xk = x[out.prow[k]];
      for (int32_t p = out.l_ptr[k]; p < out.l_ptr[k + 1]; ++p) x[out.l_row[p]] -= lx[p] * xk;
    }
    out.u_ptr.push_back(static_cast<int32_t>(out.u_idx.size()));
    double* col = S.data() + mm * static_cast<std::size_t>(j - k0);
    for (int32_t r = 0; r < out.m; ++r) {
      col[r] = x[out.trow[r]];
      x[out.trow[r]] = 0.0;
    }
    for (int32_t k : reach) x[out.prow[k]] = 0.0;
  }
  const double tail_tol = opt.tail_tol > 0.0 ? opt.tail_tol : opt.tol;
  if (out.m > 0 && !dense_threshold_lu(out.m, S.data(), out.tdiag, tail_tol, out.tperm)) return false;

  // Refactor levels of the sparse columns.
  std::vector<int32_t> lev(k0, 0);
  int32_t maxlev = -1;
  for (int32_t j = 0; j < k0; ++j) {
    int32_t l = 0;
    for (int32_t p = out.u_ptr[j]; p < out.u_ptr[j + 1]; ++p) l = std::max(l, lev[out.u_idx[p]] + 1);
    lev[j] = l;
    maxlev = std::max(maxlev, l);
  }
  out.levels = maxlev + 1;
  out.lev_ptr.assign(static_cast<std::size_t>(out.levels) + 1, 0);
  for (int32_t j = 0; j < k0; ++j) out.lev_ptr[lev[j] + 1]++;
  for (int32_t l = 0; l < out.levels; ++l) out.lev_ptr[l + 1] += out.lev_ptr[l];
  out.lev_col.resize(k0);
  {
    std::vector<int32_t> next(out.lev_ptr.begin(), out.lev_ptr.end() - 1);
    for (int32_t j = 0; j < k0; ++j) out.lev_col[next[lev[j]]++] = j;
  }

  const double m = out.m;
  out.nnz_l = static_cast<int64_t>(out.l_row.size()) + static_cast<int64_t>(m * (m - 1) / 2);
  out.nnz_u = static_cast<int64_t>(out.u_idx.size()) + k0 + static_cast<int64_t>(m * (m + 1) / 2);
  out.flops = 2.0 * static_cast<double>(out.updates) + static_cast<double>(out.l_row.size()) +
              (2.0 / 3.0) * m * m * m;
  return true;
}

}  // namespace bdlu
