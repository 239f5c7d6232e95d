// UBTF decomposition of the canonical matrix (reference ubtf:570-938):
//   CSC of the pattern (with the |v| < drop rule, ubtf:336-342, :582-589)
//   -> btf_order -> unflip Q -> merge small blocks (ubtf:393-476)
//   -> inverse permutations, strictly-upper inter-block CSR and per-block local
//      COO, both in canonical-COO order (ubtf:739-853)
//   -> block-DAG levels (ubtf:857-909).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>

#include "blockdaglu/env.h"
#include "host.h"

namespace bdlu {

DecompOptions decomp_options_from_env() {
  DecompOptions o;
  if (const char* v = blockdaglu_getenv_compat("USERALGO_DROP_ABS_LT")) {
    const double d = std::atof(v);
    o.drop_abs_lt = d >= 0.0 ? d : -1.0;
  }
  if (const char* v = blockdaglu_getenv_compat("USERALGO_BTF_MAXWORK")) {
    const double d = std::atof(v);
    o.btf_maxwork = d > 0.0 ? d : 0.0;
  }
  if (const char* v = blockdaglu_getenv_compat("USERALGO_BLOCK_MERGE")) o.merge = std::atoi(v) != 0;
  // merge cap: BLOCK_MERGE_MAX_SIZE, else DENSE_BLOCK_MAX, else 64 (ubtf:407-418)
  int32_t mm = 0;
  if (const char* v = blockdaglu_getenv_compat("USERALGO_BLOCK_MERGE_MAX_SIZE")) {
    const int x = std::atoi(v);
    if (x > 0) mm = x;
  }
  if (mm <= 0) {
    if (const char* v = blockdaglu_getenv_compat("USERALGO_DENSE_BLOCK_MAX")) {
      const int x = std::atoi(v);
      if (x > 0) mm = x;
    }
  }
  o.merge_max = mm > 0 ? mm : 64;
  return o;
}

std::vector<int32_t> merge_block_boundaries(const std::vector<int32_t>& R, int32_t n, int32_t merge_max) {
  if (R.size() < 2) return {0, n};
  const int32_t nb = static_cast<int32_t>(R.size()) - 1;
  if (nb <= 1) return {0, n};
  // Blocks up to min(merge_max, 32) are "small" and are chained into runs whose
  // total stays <= merge_max; larger blocks stay intact and end a run.
  const int32_t small_max = std::max<int32_t>(1, std::min<int32_t>(merge_max, 32));
  std::vector<int32_t> out;
  out.reserve(R.size());
  out.push_back(0);
  auto emit = [&](int32_t x) {
    if (out.back() != x) out.push_back(x);
  };
  int32_t run_len = 0, run_begin = 0, run_end = 0;
  auto close_run = [&]() {
    if (run_len <= 0) return;
    emit(run_begin);
    emit(run_end);
    run_len = 0;
  };
  for (int32_t b = 0; b < nb; ++b) {
    const int32_t lo = R[b], hi = R[b + 1], sz = hi - lo;
    if (sz <= small_max) {
      if (run_len > 0 && run_len + sz <= merge_max) {
        run_len += sz;
        run_end = hi;
        continue;
      }
      close_run();
      run_begin = lo;
      run_end = hi;
      run_len = sz;
      continue;
    }
    close_run();
    emit(lo);
    emit(hi);
  }
  close_run();
  emit(n);
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

Decomposition build_decomposition(const Canonical& a, const DecompOptions& opt) {
  Decomposition d;
  const int32_t n = a.n;
  d.n = n;
  double drop = opt.drop_abs_lt;
  if (drop < 0.0) drop = (n >= 50000) ? 1e-20 : 0.0;
  d.drop_abs_lt = drop;
  const std::size_t nnz = a.nnz();
  auto kept = [&](std::size_t k) { return !(drop > 0.0 && std::abs(a.val[k]) < drop); };

  // CSC of the kept pattern; rows ascending inside a column because the
  // canonical order is row-major (coo_to_csc_i32, sparse_formats.h:23-53).
  std::vector<int32_t> Ap(static_cast<std::size_t>(n) + 1, 0);
  for (std::size_t k = 0; k < nnz; ++k)
    if (kept(k)) Ap[a.col[k] + 1]++;
  for (int32_t j = 0; j < n; ++j) Ap[j + 1] += Ap[j];
  std::vector<int32_t> Ai(static_cast<std::size_t>(Ap[n]));
  {
    std::vector<int32_t> next(Ap.begin(), Ap.end() - 1);
    for (std::size_t k = 0; k < nnz; ++k)
      if (kept(k)) Ai[next[a.col[k]]++] = a.row[k];
  }

  std::vector<int32_t> P(n), Q(n), R(static_cast<std::size_t>(n) + 1);
  int32_t nb = 1;
  if (opt.btf) {
    double work = 0.0;
    int32_t nmatch = 0;
    nb = btf_order(n, Ap.data(), Ai.data(), opt.btf_maxwork, &work, P.data(), Q.data(), R.data(), &nmatch);
    if (nb <= 0) throw std::runtime_error("btf_order failed");
    if (opt.btf_maxwork > 0.0 && nmatch < n) {
      nb = btf_order(n, Ap.data(), Ai.data(), 0.0, &work, P.data(), Q.data(), R.data(), &nmatch);
      if (nb <= 0) throw std::runtime_error("btf_order retry failed");
    }
    d.nmatch = nmatch;
    for (int32_t k = 0; k < n; ++k) Q[k] = btf_unflip(Q[k]);
  } else {
    // KLU with btf = 0: the matrix is one block in its own order.
    for (int32_t k = 0; k < n; ++k) P[k] = Q[k] = k;
    R[0] = 0;
    R[1] = n;
    d.nmatch = n;
  }
  d.nblocks_raw = nb;

  std::vector<int32_t> Rraw(R.begin(), R.begin() + nb + 1);
  d.R = opt.merge ? merge_block_boundaries(Rraw, n, opt.merge_max) : Rraw;
  d.nblocks = static_cast<int32_t>(d.R.size()) - 1;
  d.P = std::move(P);
  d.Q = std::move(Q);
  d.invP.assign(n, -1);
  d.invQ.assign(n, -1);
  for (int32_t k = 0; k < n; ++k) {
    d.invP[d.P[k]] = k;
    d.invQ[d.Q[k]] = k;
  }
  std::vector<int32_t> blk(n);
  for (int32_t b = 0; b < d.nblocks; ++b)
    for (int32_t k = d.R[b]; k < d.R[b + 1]; ++k) blk[k] = b;

  if (opt.order_small_max >= 2) {
    // Column order of each block of 2..order_small_max (klu_analyze's block
    // ordering): the block's local CSC takes the rows of each column in the
    // order of the matrix's CSC (ascending original row, which is canonical
    // order), then the order is applied to rows and columns alike.
    auto target = [&](int32_t b) {
      const int32_t s = d.R[b + 1] - d.R[b];
      return s >= 2 && s <= opt.order_small_max;
    };
    std::vector<int32_t> cptr(static_cast<std::size_t>(n) + 1, 0);
    for (std::size_t k = 0; k < nnz; ++k) {
      if (!kept(k)) continue;
      const int32_t rp = d.invP[a.row[k]], cp = d.invQ[a.col[k]];
      if (blk[rp] == blk[cp] && target(blk[cp])) cptr[cp + 1]++;
    }
    for (int32_t j = 0; j < n; ++j) cptr[j + 1] += cptr[j];
    std::vector<int32_t> crow(static_cast<std::size_t>(cptr[n]));
    {
      std::vector<int32_t> next(cptr.begin(), cptr.end() - 1);
      for (std::size_t k = 0; k < nnz; ++k) {
        if (!kept(k)) continue;
        const int32_t rp = d.invP[a.row[k]], cp = d.invQ[a.col[k]];
        if (blk[rp] == blk[cp] && target(blk[cp])) crow[next[cp]++] = rp - d.R[blk[cp]];
      }
    }
    d.tie_key.resize(n);
    std::vector<int32_t> P2(d.P), Q2(d.Q), lp;
    for (int32_t b = 0; b < d.nblocks; ++b) {
      const int32_t lo = d.R[b], s = d.R[b + 1] - lo;
      if (!target(b)) {
        for (int32_t k = 0; k < s; ++k) d.tie_key[lo + k] = k;
        continue;
      }
      lp.assign(static_cast<std::size_t>(s) + 1, 0);
      for (int32_t j = 0; j <= s; ++j) lp[j] = cptr[lo + j] - cptr[lo];
      const std::vector<int32_t> q = colamd_order(s, lp.data(), crow.data() + cptr[lo]);
      for (int32_t k = 0; k < s; ++k) {
        P2[lo + k] = d.P[lo + q[k]];
        Q2[lo + k] = d.Q[lo + q[k]];
        d.tie_key[lo + k] = q[k];
      }
    }
    d.P = std::move(P2);
    d.Q = std::move(Q2);
    for (int32_t k = 0; k < n; ++k) {
      d.invP[d.P[k]] = k;
      d.invQ[d.Q[k]] = k;
    }
  }

  // Pass 1: counts (and the UBTF check, ubtf:739-804).
  std::vector<int32_t> rowcnt(n, 0);
  std::vector<int64_t> dcnt(static_cast<std::size_t>(d.nblocks), 0);
  bool lower = false;
  for (std::size_t k = 0; k < nnz; ++k) {
    if (!kept(k)) continue;
    const int32_t rp = d.invP[a.row[k]], cp = d.invQ[a.col[k]];
    const int32_t br = blk[rp], bc = blk[cp];
    if (bc < br) {
      lower = true;
      continue;
    }
    if (bc == br) dcnt[br]++;
    else rowcnt[rp]++;
  }
  if (lower) throw std::runtime_error("permutation is not UBTF (found lower-block entries)");

  // Pass 2: fill, canonical-COO order (ubtf:806-853).
  d.off_ptr.assign(static_cast<std::size_t>(n) + 1, 0);
  for (int32_t i = 0; i < n; ++i) d.off_ptr[i + 1] = d.off_ptr[i] + rowcnt[i];
  d.off_col.resize(static_cast<std::size_t>(d.off_ptr[n]));
  d.off_src.resize(static_cast<std::size_t>(d.off_ptr[n]));
  d.dblk_ptr.assign(static_cast<std::size_t>(d.nblocks) + 1, 0);
  for (int32_t b = 0; b < d.nblocks; ++b) d.dblk_ptr[b + 1] = d.dblk_ptr[b] + dcnt[b];
  d.dblk_r.resize(static_cast<std::size_t>(d.dblk_ptr[d.nblocks]));
  d.dblk_c.resize(d.dblk_r.size());
  d.dblk_src.resize(d.dblk_r.size());
  {
    std::vector<int32_t> onext(d.off_ptr.begin(), d.off_ptr.end() - 1);
    std::vector<int64_t> dnext(d.dblk_ptr.begin(), d.dblk_ptr.end() - 1);
    for (std::size_t k = 0; k < nnz; ++k) {
      if (!kept(k)) continue;
      const int32_t rp = d.invP[a.row[k]], cp = d.invQ[a.col[k]];
      const int32_t br = blk[rp], bc = blk[cp];
      if (bc == br) {
        const int64_t p = dnext[br]++;
        d.dblk_r[p] = rp - d.R[br];
        d.dblk_c[p] = cp - d.R[br];
        d.dblk_src[p] = static_cast<int64_t>(k);
      } else {
        const int32_t p = onext[rp]++;
        d.off_col[p] = cp;
        d.off_src[p] = static_cast<int64_t>(k);
      }
    }
  }

  if (opt.build_levels) {
    // deps(b) = blocks c > b with an entry in block row b; level from the last
    // block down (ubtf:857-909).
    d.level.assign(static_cast<std::size_t>(d.nblocks), 0);
    std::vector<int32_t> seen(static_cast<std::size_t>(d.nblocks), -1);
    d.max_level = 0;
    for (int32_t b = d.nblocks - 1; b >= 0; --b) {
      int32_t lev = 0;
      for (int32_t r = d.R[b]; r < d.R[b + 1]; ++r) {
        for (int32_t p = d.off_ptr[r]; p < d.off_ptr[r + 1]; ++p) {
          const int32_t bc = blk[d.off_col[p]];
          if (seen[bc] == b) continue;
          seen[bc] = b;
          lev = std::max(lev, d.level[bc] + 1);
        }
      }
      d.level[b] = lev;
      d.max_level = std::max(d.max_level, lev);
    }
    d.level_ptr.assign(static_cast<std::size_t>(d.max_level) + 2, 0);
    for (int32_t b = 0; b < d.nblocks; ++b) d.level_ptr[d.level[b] + 1]++;
    for (int32_t l = 0; l <= d.max_level; ++l) d.level_ptr[l + 1] += d.level_ptr[l];
    d.level_blocks.resize(static_cast<std::size_t>(d.nblocks));
    std::vector<int32_t> next(d.level_ptr.begin(), d.level_ptr.end() - 1);
    for (int32_t b = 0; b < d.nblocks; ++b) d.level_blocks[next[d.level[b]]++] = b;
  }
  return d;
}

}  // namespace bdlu
