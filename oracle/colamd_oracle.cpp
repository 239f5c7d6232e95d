// ORACLE — test infrastructure, not product code.
//
// The oracle's own COLAMD (the ordering klu_analyze uses with ordering = 1,
// ubtf_block_solver.cpp:493-499, :1804), restated from the published method:
// Davis, Gilbert, Larimore, Ng, "A column approximate minimum degree ordering
// algorithm", ACM TOMS 30(3), 2004, and the COLAMD 2.x distribution it
// describes (default knobs: dense cut 10, aggressive absorption). Written
// independently of the product's ordering, on the original's memory model: one
// integer workspace holding the column lists, then the row lists, then the
// pivot rows appended as elimination proceeds (grown instead of garbage
// collected), and row marks relative to a moving tag (set difference = mark -
// tag). Pinned in tests/ against the COLAMD bundled with SuperLU (scipy).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace oracle {

namespace {

constexpr int32_t EMPTY = -1;

struct Col {
  int32_t start = 0, length = 0;
  int32_t thickness = 1;  // < 0 while tagged in the pivot row
  int32_t parent = EMPTY;
  int32_t score = 0;
  int32_t order = EMPTY;
  int32_t prev = EMPTY, next = EMPTY;  // degree list
  int32_t hash = 0, hash_next = EMPTY;
  bool alive = true, principal = true;
};

struct Row {
  int32_t start = 0, length = 0, degree = 0;
  int64_t mark = 0;  // < 0: dead
};

}  // namespace

std::vector<int32_t> colamd(int32_t n, const int32_t* Ap, const int32_t* Ai) {
  std::vector<int32_t> p(n);
  if (n <= 0) return p;
  std::vector<Col> C(n);
  std::vector<Row> R(n);
  std::vector<int32_t> A;
  // Column form, sorted and duplicate-free (a jumbled input is rebuilt so).
  for (int32_t c = 0; c < n; ++c) {
    std::vector<int32_t> rows(Ai + Ap[c], Ai + Ap[c + 1]);
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    C[c].start = static_cast<int32_t>(A.size());
    C[c].length = static_cast<int32_t>(rows.size());
    A.insert(A.end(), rows.begin(), rows.end());
    for (int32_t r : rows) R[r].length++;
  }
  // Row form after it, columns ascending.
  {
    int32_t at = static_cast<int32_t>(A.size());
    for (int32_t r = 0; r < n; ++r) {
      R[r].start = at;
      at += R[r].length;
      R[r].degree = R[r].length;
    }
    A.resize(at);
    std::vector<int32_t> fill(n);
    for (int32_t r = 0; r < n; ++r) fill[r] = R[r].start;
    for (int32_t c = 0; c < n; ++c)
      for (int32_t q = C[c].start; q < C[c].start + C[c].length; ++q) A[fill[A[q]]++] = c;
  }
  auto dense = [](int32_t m) { return static_cast<int32_t>(std::max(16.0, 10.0 * std::sqrt(double(m)))); };
  const int32_t dense_rows = dense(n), dense_cols = dense(n);
  int32_t last = n;  // columns ordered last take positions last-1, last-2, ...
  auto kill_principal = [&](int32_t c) {
    C[c].alive = false;
    C[c].principal = true;
  };
  for (int32_t c = n - 1; c >= 0; --c)
    if (C[c].length == 0) {
      C[c].order = --last;
      kill_principal(c);
    }
  for (int32_t c = n - 1; c >= 0; --c)
    if (C[c].alive && C[c].length > dense_cols) {
      C[c].order = --last;
      for (int32_t q = C[c].start; q < C[c].start + C[c].length; ++q) R[A[q]].degree--;
      kill_principal(c);
    }
  for (int32_t r = 0; r < n; ++r)
    if (R[r].degree > dense_rows || R[r].degree == 0) R[r].mark = -1;
  for (int32_t c = n - 1; c >= 0; --c) {
    if (!C[c].alive) continue;
    int32_t w = C[c].start, score = 0;
    for (int32_t q = C[c].start; q < C[c].start + C[c].length; ++q) {
      const int32_t r = A[q];
      if (R[r].mark < 0) continue;
      A[w++] = r;
      score = std::min(score + R[r].degree - 1, n);
    }
    C[c].length = w - C[c].start;
    if (C[c].length == 0) {
      C[c].order = --last;
      kill_principal(c);
    } else {
      C[c].score = score;
    }
  }
  std::vector<int32_t> head(static_cast<std::size_t>(n) + 1, EMPTY);
  auto push = [&](int32_t c) {
    const int32_t s = C[c].score;
    C[c].prev = EMPTY;
    C[c].next = head[s];
    if (head[s] != EMPTY) C[head[s]].prev = c;
    head[s] = c;
  };
  for (int32_t c = n - 1; c >= 0; --c)
    if (C[c].alive) push(c);

  std::vector<int32_t> buckets(static_cast<std::size_t>(n) + 1, EMPTY);
  int64_t tag = 1;
  int32_t max_deg = 0;
  for (int32_t r = 0; r < n; ++r)
    if (R[r].mark >= 0) max_deg = std::max(max_deg, R[r].degree);
  int32_t min_score = 0;
  for (int32_t k = 0; k < last;) {
    while (head[min_score] == EMPTY && min_score < n) ++min_score;
    const int32_t piv = head[min_score];
    head[min_score] = C[piv].next;
    if (C[piv].next != EMPTY) C[C[piv].next].prev = EMPTY;
    C[piv].order = k;
    const int32_t piv_thick = C[piv].thickness;
    k += piv_thick;
    // Pivot row, appended to the workspace.
    const int32_t prow_start = static_cast<int32_t>(A.size());
    int32_t prow_degree = 0;
    C[piv].thickness = -piv_thick;
    for (int32_t q = C[piv].start; q < C[piv].start + C[piv].length; ++q) {
      const int32_t r = A[q];
      if (R[r].mark < 0) continue;
      for (int32_t t = R[r].start; t < R[r].start + R[r].length; ++t) {
        const int32_t c = A[t];
        if (C[c].alive && C[c].thickness > 0) {
          prow_degree += C[c].thickness;
          C[c].thickness = -C[c].thickness;
          A.push_back(c);
        }
      }
    }
    C[piv].thickness = piv_thick;
    max_deg = std::max(max_deg, prow_degree);
    for (int32_t q = C[piv].start; q < C[piv].start + C[piv].length; ++q) R[A[q]].mark = -1;
    const int32_t prow_len = static_cast<int32_t>(A.size()) - prow_start;
    const int32_t prow_id = prow_len > 0 ? A[C[piv].start] : EMPTY;
    // Set differences (mark - tag), absorbing rows that fall inside the pivot row.
    for (int32_t t = prow_start; t < prow_start + prow_len; ++t) {
      const int32_t c = A[t];
      C[c].thickness = -C[c].thickness;
      if (C[c].prev == EMPTY) head[C[c].score] = C[c].next;
      else C[C[c].prev].next = C[c].next;
      if (C[c].next != EMPTY) C[C[c].next].prev = C[c].prev;
      for (int32_t q = C[c].start; q < C[c].start + C[c].length; ++q) {
        Row& row = R[A[q]];
        if (row.mark < 0) continue;
        int64_t d = row.mark - tag;
        if (d < 0) d = row.degree;
        d -= C[c].thickness;
        if (d == 0) row.mark = -1;
        else row.mark = d + tag;
      }
    }
    // Column scores so far, compaction, mass elimination, hashing.
    for (int32_t t = prow_start; t < prow_start + prow_len; ++t) {
      const int32_t c = A[t];
      uint32_t h = 0;
      int64_t score = 0;
      int32_t w = C[c].start;
      for (int32_t q = C[c].start; q < C[c].start + C[c].length; ++q) {
        const int32_t r = A[q];
        if (R[r].mark < 0) continue;
        A[w++] = r;
        h += static_cast<uint32_t>(r);
        score = std::min<int64_t>(score + (R[r].mark - tag), n);
      }
      C[c].length = w - C[c].start;
      if (C[c].length == 0) {
        kill_principal(c);
        prow_degree -= C[c].thickness;
        C[c].order = k;
        k += C[c].thickness;
      } else {
        C[c].score = static_cast<int32_t>(score);
        C[c].hash = static_cast<int32_t>(h % static_cast<uint32_t>(n + 1));
        C[c].hash_next = buckets[C[c].hash];
        buckets[C[c].hash] = c;
      }
    }
    // Supercolumns: identical columns within a bucket merge into the earlier one.
    for (int32_t t = prow_start; t < prow_start + prow_len; ++t) {
      const int32_t c = A[t];
      if (!C[c].alive) continue;
      const int32_t hv = C[c].hash;
      for (int32_t s = buckets[hv]; s != EMPTY; s = C[s].hash_next) {
        int32_t prev = s;
        for (int32_t o = C[s].hash_next; o != EMPTY; o = C[o].hash_next) {
          bool same = C[o].length == C[s].length && C[o].score == C[s].score;
          for (int32_t i = 0; same && i < C[s].length; ++i) same = A[C[s].start + i] == A[C[o].start + i];
          if (!same) {
            prev = o;
            continue;
          }
          C[s].thickness += C[o].thickness;
          C[o].parent = s;
          C[o].alive = false;
          C[o].principal = false;
          C[o].order = EMPTY;
          C[prev].hash_next = C[o].hash_next;
        }
      }
      buckets[hv] = EMPTY;
    }
    kill_principal(piv);
    tag += max_deg + 1;
    // Final scores; append the pivot row to each remaining column. A column's
    // storage had at least one dead row removed above, so the append fits in
    // place when the column is not the last list of the workspace.
    int32_t w = prow_start;
    for (int32_t t = prow_start; t < prow_start + prow_len; ++t) {
      const int32_t c = A[t];
      if (!C[c].alive) continue;
      A[w++] = c;
      A[C[c].start + C[c].length++] = prow_id;
      const int32_t cap = n - k - C[c].thickness;
      C[c].score = std::min(C[c].score + prow_degree - C[c].thickness, cap);
      push(c);
      min_score = std::min(min_score, C[c].score);
    }
    if (prow_degree > 0) {
      R[prow_id].start = prow_start;
      R[prow_id].length = w - prow_start;
      R[prow_id].degree = prow_degree;
      R[prow_id].mark = 0;
    }
    A.resize(w);
  }
  // Absorbed columns: their principal's slots, ascending, principal last.
  for (int32_t i = 0; i < n; ++i) {
    if (C[i].principal || C[i].order != EMPTY) continue;
    int32_t par = i;
    while (!C[par].principal) par = C[par].parent;
    C[i].order = C[par].order++;
    C[i].parent = par;
  }
  for (int32_t c = 0; c < n; ++c) p[C[c].order] = c;
  return p;
}

}  // namespace oracle
