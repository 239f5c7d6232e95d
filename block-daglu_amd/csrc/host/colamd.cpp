// Column approximate-minimum-degree ordering for the diagonal blocks larger than
// the dense threshold. In the reference this is SuiteSparse COLAMD, reached
// through klu_analyze with ordering = 1 (ubtf:493-499, :1804). This is a
// restatement of the published COLAMD method (Davis, Gilbert, Larimore, Ng,
// ACM TOMS 30(3), 2004) in a compact form:
//   * dense columns (> max(16, 10*sqrt(s)) entries) and empty columns are
//     ordered last, ascending; dense rows (> max(16, 10*sqrt(s)) live columns)
//     and empty rows are ignored;
//   * the pivot is a live column of minimum approximate score, ties broken by
//     the smallest column index;
//   * eliminating column c merges every row of c into one new row (the pivot
//     row), rows wholly contained in it are absorbed (aggressive absorption),
//     and every column of the pivot row is re-scored with the approximate
//     external degree  (|pivot row| - 1) + sum over its other rows of
//     |row \ pivot row|, capped by the number of live columns.
// Supercolumn detection and COLAMD's degree-list tie order are not reproduced,
// so the sequence can differ from SuiteSparse's on ties.
#include <algorithm>
#include <cmath>
#include <queue>

#include "host.h"

namespace bdlu {

std::vector<int32_t> colamd_order(int32_t s, const int32_t* Ap, const int32_t* Ai) {
  std::vector<int32_t> order;
  order.reserve(s);
  if (s <= 0) return order;
  const int32_t dense_cut = std::max<int32_t>(16, static_cast<int32_t>(10.0 * std::sqrt(static_cast<double>(s))));

  // Column -> rows, row -> columns.
  std::vector<char> col_alive(s, 1);
  std::vector<int32_t> col_cnt(s);
  for (int32_t c = 0; c < s; ++c) col_cnt[c] = Ap[c + 1] - Ap[c];
  std::vector<char> col_killed(s, 0);
  for (int32_t c = 0; c < s; ++c)
    if (col_cnt[c] == 0 || col_cnt[c] > dense_cut) col_killed[c] = 1;

  std::vector<std::vector<int32_t>> rows_of;  // row/element -> live columns
  rows_of.resize(s);
  for (int32_t c = 0; c < s; ++c) {
    if (col_killed[c]) continue;
    for (int32_t p = Ap[c]; p < Ap[c + 1]; ++p) rows_of[Ai[p]].push_back(c);
  }
  std::vector<char> row_alive(s, 1);
  for (int32_t r = 0; r < s; ++r)
    if (rows_of[r].empty() || static_cast<int32_t>(rows_of[r].size()) > dense_cut) row_alive[r] = 0;

  std::vector<std::vector<int32_t>> cols_of(s);  // column -> live rows/elements
  for (int32_t c = 0; c < s; ++c) {
    if (col_killed[c]) continue;
    for (int32_t p = Ap[c]; p < Ap[c + 1]; ++p)
      if (row_alive[Ai[p]]) cols_of[c].push_back(Ai[p]);
  }

  int32_t live_cols = 0;
  for (int32_t c = 0; c < s; ++c) live_cols += col_killed[c] ? 0 : 1;

  using Item = std::pair<int64_t, int32_t>;  // (score, column)
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
  std::vector<int64_t> score(s, 0);
  for (int32_t c = 0; c < s; ++c) {
    if (col_killed[c]) continue;
    int64_t sc = 0;
    for (int32_t r : cols_of[c]) sc += static_cast<int64_t>(rows_of[r].size()) - 1;
    score[c] = std::min<int64_t>(sc, live_cols - 1);
    heap.push({score[c], c});
  }

  std::vector<int32_t> col_mark(s, -1);
  std::vector<int32_t> row_overlap;  // indexed by row/element id
  std::vector<int32_t> row_stamp;
  row_overlap.assign(rows_of.size(), 0);
  row_stamp.assign(rows_of.size(), -1);
  std::vector<int32_t> touched;
  std::vector<int32_t> pivot_cols;

  for (int32_t step = 0; !heap.empty();) {
    const Item top = heap.top();
    heap.pop();
    const int32_t c = top.second;
    if (col_killed[c] || !col_alive[c] || top.first != score[c]) continue;
    col_alive[c] = 0;
    order.push_back(c);
    --live_cols;

    // Pivot row: union of the live rows of c, minus c.
    pivot_cols.clear();
    for (int32_t r : cols_of[c]) {
      if (!row_alive[r]) continue;
      for (int32_t c2 : rows_of[r]) {
        if (c2 == c || !col_alive[c2] || col_mark[c2] == step) continue;
        col_mark[c2] = step;
        pivot_cols.push_back(c2);
      }
      row_alive[r] = 0;
    }
    cols_of[c].clear();
    cols_of[c].shrink_to_fit();
    const int32_t e = static_cast<int32_t>(rows_of.size());
    rows_of.push_back(pivot_cols);
    row_alive.push_back(pivot_cols.empty() ? 0 : 1);
    row_overlap.push_back(0);
    row_stamp.push_back(-1);

    // |r intersect pivot row| for the other live rows of the pivot columns.
    touched.clear();
    for (int32_t c2 : pivot_cols) {
      for (int32_t r : cols_of[c2]) {
        if (!row_alive[r]) continue;
        if (row_stamp[r] != step) {
          row_stamp[r] = step;
          row_overlap[r] = 0;
          touched.push_back(r);
        }
        row_overlap[r]++;
      }
    }
    // Aggressive absorption of rows contained in the pivot row.
    for (int32_t r : touched)
      if (static_cast<int32_t>(rows_of[r].size()) == row_overlap[r]) row_alive[r] = 0;

    const int64_t esize = static_cast<int64_t>(pivot_cols.size());
    for (int32_t c2 : pivot_cols) {
      auto& lst = cols_of[c2];
      std::size_t w = 0;
      int64_t sc = esize - 1;
      for (std::size_t t = 0; t < lst.size(); ++t) {
        const int32_t r = lst[t];
        if (!row_alive[r]) continue;
        lst[w++] = r;
        sc += static_cast<int64_t>(rows_of[r].size()) - row_overlap[r];
      }
      lst.resize(w);
      lst.push_back(e);
      sc = std::min<int64_t>(sc, std::max<int32_t>(live_cols - 1, 0));
      score[c2] = sc;
      heap.push({sc, c2});
    }
    // Absorbed rows' column lists are no longer needed.
    for (int32_t r : touched)
      if (!row_alive[r]) {
        rows_of[r].clear();
        rows_of[r].shrink_to_fit();
      }
    ++step;
  }
  for (int32_t c = 0; c < s; ++c)
    if (col_killed[c]) order.push_back(c);
  return order;
}

}  // namespace bdlu
