// Column approximate minimum degree ordering of the diagonal blocks larger
// than the dense threshold: the role SuiteSparse COLAMD plays inside
// klu_analyze with ordering = 1 (ubtf_block_solver.cpp:493-499, :1804), called
// with default knobs (dense row/column cut 10, aggressive absorption on).
//
// Restated from the published method (Davis, Gilbert, Larimore, Ng, "A column
// approximate minimum degree ordering algorithm", ACM TOMS 30(3), 2004, and
// the COLAMD 2.x distribution it describes), including the details that fix
// the order on ties:
//   * columns are taken with sorted, duplicate-free row lists; rows list their
//     columns in ascending order;
//   * empty columns, then dense columns (more than max(16, 10 sqrt(min(m,n)))
//     entries), then columns left empty by dense rows (more than
//     max(16, 10 sqrt(n)) live columns) are ordered last, each group scanned
//     from the highest column index down;
//   * initial score of a column = sum over its live rows of (row degree - 1),
//     capped at n; degree lists are filled from the highest column index down,
//     inserting at the head, and the pivot is the head of the lowest non-empty
//     list;
//   * the pivot row is the union of the pivot column's live rows, columns in
//     the order met; all those rows die and the pivot row takes the index of
//     the first row stored in the pivot column;
//   * set differences |row \ pivot row| with aggressive absorption of rows
//     whose difference reaches zero, columns compacted in place, mass
//     elimination of columns left with no row;
//   * supercolumns: columns of the pivot row hashed by (sum of row indices)
//     mod (n + 1) into buckets (each insertion at the front), identical columns
//     (same length, same partial score, same row sequence) merged into the
//     first of the bucket that precedes them;
//   * final score = partial score + pivot row degree - thickness, capped at
//     n - k - thickness, columns re-inserted at the head of their lists;
//   * absorbed (non-principal) columns take the positions of their principal
//     column's group in ascending column order, the principal the last one.
// The workspace layout and garbage collection of the original only move
// storage and never reorder entries, so separate per-row and per-column lists
// give the same order.
#include <algorithm>
#include <cmath>

#include "host.h"

namespace bdlu {

namespace {

constexpr int32_t kNone = -1;

struct ColamdState {
  int32_t n = 0;
  // columns
  std::vector<std::vector<int32_t>> col_rows;
  std::vector<int32_t> thickness, score, order, parent, prev, next, hash, hash_next;
  std::vector<uint8_t> col_dead, col_principal;  // dead: out of the live set; principal: not absorbed
  // rows
  std::vector<std::vector<int32_t>> row_cols;
  std::vector<int32_t> row_degree;
  std::vector<uint8_t> row_dead;
  // degree lists
  std::vector<int32_t> head;

  void list_insert(int32_t c, int32_t sc) {
    const int32_t nx = head[sc];
    next[c] = nx;
    prev[c] = kNone;
    if (nx != kNone) prev[nx] = c;
    head[sc] = c;
  }
  void list_remove(int32_t c) {
    const int32_t pv = prev[c], nx = next[c];
    if (pv == kNone) head[score[c]] = nx;
    else next[pv] = nx;
    if (nx != kNone) prev[nx] = pv;
  }
  void kill_principal(int32_t c) {
    col_dead[c] = 1;
    col_principal[c] = 1;
  }
};

}  // namespace

std::vector<int32_t> colamd_order(int32_t n, const int32_t* Ap, const int32_t* Ai) {
  std::vector<int32_t> perm;
  if (n <= 0) return perm;
  ColamdState S;
  S.n = n;
  S.col_rows.assign(n, {});
  S.row_cols.assign(n, {});
  // Sorted, duplicate-free column lists; row lists in ascending column order.
  for (int32_t c = 0; c < n; ++c) {
    std::vector<int32_t>& rows = S.col_rows[c];
    rows.assign(Ai + Ap[c], Ai + Ap[c + 1]);
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    for (int32_t r : rows) S.row_cols[r].push_back(c);
  }
  S.thickness.assign(n, 1);
  S.score.assign(n, 0);
  S.order.assign(n, kNone);
  S.parent.assign(n, kNone);
  S.prev.assign(n, kNone);
  S.next.assign(n, kNone);
  S.hash.assign(n, 0);
  S.hash_next.assign(n, kNone);
  S.col_dead.assign(n, 0);
  S.col_principal.assign(n, 1);
  S.row_degree.assign(n, 0);
  S.row_dead.assign(n, 0);
  for (int32_t r = 0; r < n; ++r) S.row_degree[r] = static_cast<int32_t>(S.row_cols[r].size());

  // ---- initial scoring ------------------------------------------------------
  const auto dense_cut = [](double alpha, int32_t m) {
    return static_cast<int32_t>(std::max(16.0, alpha * std::sqrt(static_cast<double>(m))));
  };
  const int32_t dense_row = dense_cut(10.0, n);
  const int32_t dense_col = dense_cut(10.0, n);  // min(rows, cols) = n for a square block
  int32_t n_col2 = n;
  for (int32_t c = n - 1; c >= 0; --c)
    if (S.col_rows[c].empty()) {
      S.order[c] = --n_col2;
      S.kill_principal(c);
    }
  for (int32_t c = n - 1; c >= 0; --c) {
    if (S.col_dead[c]) continue;
    if (static_cast<int32_t>(S.col_rows[c].size()) > dense_col) {
      S.order[c] = --n_col2;
      for (int32_t r : S.col_rows[c]) S.row_degree[r]--;
      S.kill_principal(c);
    }
  }
  for (int32_t r = 0; r < n; ++r)
    if (S.row_degree[r] > dense_row || S.row_degree[r] == 0) S.row_dead[r] = 1;
  for (int32_t c = n - 1; c >= 0; --c) {
    if (S.col_dead[c]) continue;
    std::vector<int32_t>& rows = S.col_rows[c];
    int32_t sc = 0, w = 0;
    for (int32_t r : rows) {
      if (S.row_dead[r]) continue;
      rows[w++] = r;
      sc = std::min(sc + S.row_degree[r] - 1, n);
    }
    rows.resize(w);
    if (w == 0) {
      S.order[c] = --n_col2;
      S.kill_principal(c);
    } else {
      S.score[c] = sc;
    }
  }
  S.head.assign(static_cast<std::size_t>(n) + 1, kNone);
  int32_t min_score = n;
  for (int32_t c = n - 1; c >= 0; --c) {
    if (S.col_dead[c]) continue;
    S.list_insert(c, S.score[c]);
    min_score = std::min(min_score, S.score[c]);
  }

  // ---- elimination ------------------------------------------------------------
  std::vector<int32_t> in_pivot_row(n, 0);   // column already in this pivot row
  std::vector<int32_t> seen(n, -1);          // row -> step at which its set difference was started
  std::vector<int32_t> setdiff(n, 0);        // row -> |row \ pivot row| (thickness-weighted)
  std::vector<int32_t> bucket(static_cast<std::size_t>(n) + 1, kNone);
  std::vector<int32_t> prow;
  min_score = 0;
  for (int32_t k = 0, step = 0; k < n_col2; ++step) {
    while (min_score < n && S.head[min_score] == kNone) ++min_score;
    const int32_t pc = S.head[min_score];
    {
      const int32_t nx = S.next[pc];
      S.head[min_score] = nx;
      if (nx != kNone) S.prev[nx] = kNone;
    }
    S.order[pc] = k;
    k += S.thickness[pc];

    // Pivot row: union of the live rows of the pivot column.
    prow.clear();
    int32_t prow_degree = 0;
    in_pivot_row[pc] = 1;
    for (int32_t r : S.col_rows[pc]) {
      if (S.row_dead[r]) continue;
      for (int32_t c : S.row_cols[r]) {
        if (S.col_dead[c] || in_pivot_row[c]) continue;
        in_pivot_row[c] = 1;
        prow.push_back(c);
        prow_degree += S.thickness[c];
      }
    }
    in_pivot_row[pc] = 0;
    for (int32_t r : S.col_rows[pc]) S.row_dead[r] = 1;
    const int32_t pivot_row = prow.empty() ? kNone : S.col_rows[pc].front();

    // Set differences, with aggressive absorption.
    for (int32_t c : prow) {
      in_pivot_row[c] = 0;
      S.list_remove(c);
      for (int32_t r : S.col_rows[c]) {
        if (S.row_dead[r]) continue;
        if (seen[r] != step) {
          seen[r] = step;
          setdiff[r] = S.row_degree[r];
        }
        setdiff[r] -= S.thickness[c];
        if (setdiff[r] == 0) S.row_dead[r] = 1;
      }
    }
    // Partial scores, column compaction, mass elimination, hash buckets.
    for (int32_t c : prow) {
      std::vector<int32_t>& rows = S.col_rows[c];
      uint32_t h = 0;  // unsigned 32-bit sum, wrapping as the original's
      int32_t sc = 0, w = 0;
      for (int32_t r : rows) {
        if (S.row_dead[r]) continue;
        rows[w++] = r;
        h += static_cast<uint32_t>(r);
        sc = std::min(sc + setdiff[r], n);
      }
      rows.resize(w);
      if (w == 0) {
        S.kill_principal(c);
        prow_degree -= S.thickness[c];
        S.order[c] = k;
        k += S.thickness[c];
      } else {
        S.score[c] = sc;
        const int32_t hv = static_cast<int32_t>(h % (static_cast<uint32_t>(n) + 1u));
        S.hash[c] = hv;
        S.hash_next[c] = bucket[hv];
        bucket[hv] = c;
      }
    }
    // Supercolumn detection.
    for (int32_t c0 : prow) {
      if (S.col_dead[c0]) continue;
      const int32_t hv = S.hash[c0];
      for (int32_t sup = bucket[hv]; sup != kNone; sup = S.hash_next[sup]) {
        const std::vector<int32_t>& ref = S.col_rows[sup];
        int32_t before = sup;
        for (int32_t c = S.hash_next[sup]; c != kNone; c = S.hash_next[c]) {
          if (S.col_rows[c].size() != ref.size() || S.score[c] != S.score[sup] || S.col_rows[c] != ref) {
            before = c;
            continue;
          }
          S.thickness[sup] += S.thickness[c];
          S.parent[c] = sup;
          S.col_dead[c] = 1;
          S.col_principal[c] = 0;
          S.order[c] = kNone;
          S.hash_next[before] = S.hash_next[c];
        }
      }
      bucket[hv] = kNone;
    }
    S.kill_principal(pc);

    // Final scores; the pivot row joins every remaining column.
    int32_t w = 0;
    for (int32_t c : prow) {
      if (S.col_dead[c]) continue;
      prow[w++] = c;
      S.col_rows[c].push_back(pivot_row);
      const int32_t max_score = n - k - S.thickness[c];
      const int32_t sc = std::min(S.score[c] + prow_degree - S.thickness[c], max_score);
      S.score[c] = sc;
      S.list_insert(c, sc);
      min_score = std::min(min_score, sc);
    }
    prow.resize(w);
    if (prow_degree > 0) {
      S.row_cols[pivot_row] = prow;
      S.row_degree[pivot_row] = prow_degree;
      S.row_dead[pivot_row] = 0;
    }
  }

  // ---- absorbed columns: the slots of their principal's group -------------------
  for (int32_t i = 0; i < n; ++i) {
    if (S.col_principal[i] || S.order[i] != kNone) continue;
    int32_t p = i;
    do p = S.parent[p];
    while (!S.col_principal[p]);
    S.order[i] = S.order[p];
    S.parent[i] = p;
    S.order[p] += 1;
  }
  perm.assign(n, 0);
  for (int32_t c = 0; c < n; ++c) perm[S.order[c]] = c;
  return perm;
}

}  // namespace bdlu
