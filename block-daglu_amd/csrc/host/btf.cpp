// DM/BTF ordering: maximum transversal + strongly connected components.
//
// The reference calls SuiteSparse BTF's btf_order (ubtf:656-699; SuiteSparse
// from vcpkg tag 2024.05.24, not vendored). This file restates the published
// algorithms with the traversal orders that fix the output permutation:
//   maxtrans   — Duff's depth-first augmenting-path search with a cheap
//                assignment pass; columns k = 0..n-1 are matched in order; the
//                cheap pass scans each column once, front to back; the DFS scans
//                adjacency front to back and resumes where it left off.
//   strongcomp — Tarjan's SCC on the graph of A(:,Q): node j has edges to the
//                row indices of column Q[j]; DFS roots in increasing order,
//                adjacency front to back. Blocks are numbered in the order found
//                (that order makes P*A*Q block upper triangular); nodes keep
//                their natural (increasing) order inside a block.
#include <algorithm>

#include "host.h"

namespace bdlu {

namespace {
// One augmenting-path search for column k. Returns 1 found, 0 not found,
// -1 work limit reached.
int augment(int32_t k, const int32_t* Ap, const int32_t* Ai, int32_t* Match, int32_t* Cheap, int32_t* Flag,
            int32_t* Istack, int32_t* Jstack, int32_t* Pstack, double* work, double maxwork) {
  const bool limited = maxwork > 0;
  bool found = false;
  int32_t i = kEmpty;
  int32_t head = 0;
  Jstack[0] = k;
  while (head >= 0) {
    const int32_t j = Jstack[head];
    const int32_t pend = Ap[j + 1];
    if (Flag[j] != k) {
      Flag[j] = k;
      int32_t p = Cheap[j];
      for (; p < pend && !found; ++p) {
        i = Ai[p];
        found = (Match[i] == kEmpty);
      }
      Cheap[j] = p;
      if (found) {
        Istack[head] = i;
        break;
      }
      Pstack[head] = Ap[j];
    }
    if (limited && *work > maxwork) return -1;
    const int32_t pstart = Pstack[head];
    int32_t p = pstart;
    for (; p < pend; ++p) {
      i = Ai[p];
      const int32_t j2 = Match[i];
      if (Flag[j2] != k) {
        Pstack[head] = p + 1;
        Istack[head] = i;
        Jstack[++head] = j2;
        break;
      }
    }
    *work += static_cast<double>(p - pstart + 1);
    if (p == pend) --head;
  }
  if (found) {
    for (int32_t h = head; h >= 0; --h) Match[Istack[h]] = Jstack[h];
  }
  return found ? 1 : 0;
}
}  // namespace

int32_t btf_maxtrans(int32_t nrow, int32_t ncol, const int32_t* Ap, const int32_t* Ai, double maxwork,
                     double* work, int32_t* Match) {
  *work = 0;
  if (nrow <= 0 || ncol <= 0) return 0;
  std::vector<int32_t> Cheap(ncol), Flag(ncol, kEmpty), Istack(ncol), Jstack(ncol), Pstack(ncol);
  for (int32_t j = 0; j < ncol; ++j) Cheap[j] = Ap[j];
  for (int32_t i = 0; i < nrow; ++i) Match[i] = kEmpty;
  if (maxwork > 0) maxwork *= static_cast<double>(Ap[ncol]);
  int32_t nmatch = 0;
  bool limit_hit = false;
  for (int32_t k = 0; k < ncol; ++k) {
    const int r = augment(k, Ap, Ai, Match, Cheap.data(), Flag.data(), Istack.data(), Jstack.data(),
                          Pstack.data(), work, maxwork);
    if (r == 1) ++nmatch;
    else if (r < 0) limit_hit = true;
  }
  if (limit_hit) *work = -1;
  return nmatch;
}

int32_t btf_strongcomp(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t* Q, int32_t* P, int32_t* R) {
  constexpr int32_t kUnvisited = -2;
  constexpr int32_t kUnassigned = -1;
  std::vector<int32_t> Time(n, kEmpty), Low(n, kEmpty), Flag(n, kUnvisited);
  std::vector<int32_t> Cstack(n), Jstack(n), Pstack(n);
  int32_t nblocks = 0;
  int32_t timestamp = 0;
  for (int32_t root = 0; root < n; ++root) {
    if (Flag[root] != kUnvisited) continue;
    int32_t chead = -1;
    int32_t jhead = 0;
    Jstack[0] = root;
    while (jhead >= 0) {
      const int32_t j = Jstack[jhead];
      const int32_t jj = Q ? btf_unflip(Q[j]) : j;
      const int32_t pend = Ap[jj + 1];
      if (Flag[j] == kUnvisited) {
        Cstack[++chead] = j;
        ++timestamp;
        Time[j] = timestamp;
        Low[j] = timestamp;
        Flag[j] = kUnassigned;
        Pstack[jhead] = Ap[jj];
      }
      int32_t p = Pstack[jhead];
      for (; p < pend; ++p) {
        const int32_t i = Ai[p];
        if (Flag[i] == kUnvisited) {
          Pstack[jhead] = p + 1;
          Jstack[++jhead] = i;
          break;
        } else if (Flag[i] == kUnassigned) {
          Low[j] = std::min(Low[j], Time[i]);
        }
      }
      if (p == pend) {
        --jhead;
        if (Low[j] == Time[j]) {
          while (true) {
            const int32_t i = Cstack[chead--];
            Flag[i] = nblocks;
            if (i == j) break;
          }
          ++nblocks;
        }
        if (jhead >= 0) {
          const int32_t parent = Jstack[jhead];
          Low[parent] = std::min(Low[parent], Low[j]);
        }
      }
    }
  }
  // Block boundaries, then the permutation keeping natural order inside blocks.
  std::vector<int32_t> count(static_cast<std::size_t>(nblocks) + 1, 0);
  for (int32_t i = 0; i < n; ++i) count[Flag[i]]++;
  R[0] = 0;
  for (int32_t b = 0; b < nblocks; ++b) R[b + 1] = R[b] + count[b];
  std::vector<int32_t> next(R, R + nblocks);
  for (int32_t i = 0; i < n; ++i) P[next[Flag[i]]++] = i;
  if (Q) {
    std::vector<int32_t> qp(n);
    for (int32_t k = 0; k < n; ++k) qp[k] = Q[P[k]];
    std::copy(qp.begin(), qp.end(), Q);
  }
  return nblocks;
}

int32_t btf_order(int32_t n, const int32_t* Ap, const int32_t* Ai, double maxwork, double* work, int32_t* P,
                  int32_t* Q, int32_t* R, int32_t* nmatch) {
  *nmatch = btf_maxtrans(n, n, Ap, Ai, maxwork, work, Q);
  if (*nmatch < n) {
    // Complete the matching: unmatched rows take unmatched columns, flagged.
    std::vector<char> used(n, 0);
    for (int32_t i = 0; i < n; ++i)
      if (Q[i] != kEmpty) used[Q[i]] = 1;
    std::vector<int32_t> badcol;
    for (int32_t j = n - 1; j >= 0; --j)
      if (!used[j]) badcol.push_back(j);
    for (int32_t i = 0; i < n; ++i) {
      if (Q[i] == kEmpty && !badcol.empty()) {
        const int32_t j = badcol.back();
        badcol.pop_back();
        Q[i] = btf_flip(j);
      }
    }
  }
  return btf_strongcomp(n, Ap, Ai, Q, P, R);
}

}  // namespace bdlu
