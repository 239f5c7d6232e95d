// Device plan assembly (see plan.h). The value arena holds, in order: every
// dense block's s*s column-major values, every sparse block's entries in
// pivot-column order, and the strictly-upper inter-block CSR values. vmap
// routes each canonical entry to its arena slot (reference placement:
// ubtf:823-853 for the CSR and block COO, ubtf:1455-1462 for dense blocks).
#include "plan.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <stdexcept>
#include <mutex>
#include <thread>

#include "blockdaglu/env.h"

namespace bdlu {

int32_t dense_block_max_from_env() {
  // DENSE_BLOCK_MAX (ubtf:1441-1450); the device dense kernel holds one block
  // per 64-wide wavefront, so larger settings are capped at 64 and the larger
  // blocks take the sparse-block path.
  int32_t v = 64;
  if (const char* e = blockdaglu_getenv_compat("USERALGO_DENSE_BLOCK_MAX")) {
    const int x = std::atoi(e);
    if (x >= 1) v = x;
  }
  return std::min<int32_t>(v, 64);
}

std::vector<double> klu_row_scale(const Canonical& a, int scale) {
  std::vector<double> rs;
  if (scale <= 0) return rs;
  rs.assign(static_cast<std::size_t>(a.n), 0.0);
  // Canonical order is row-major with ascending columns, so each row's terms
  // are combined in the column order klu_scale visits them.
  for (std::size_t k = 0; k < a.nnz(); ++k) {
    const double v = std::abs(a.val[k]);
    double& r = rs[a.row[k]];
    r = scale == 1 ? r + v : std::max(r, v);
  }
  for (double& r : rs)
    if (r == 0.0) r = 1.0;
  return rs;
}

namespace {
// Pivoting options of the sparse-block analysis: the BLOCK route's (env), or
// the global-KLU route's tolerance with the row scaling applied to the whole
// matrix beforehand (so no per-block scaling).
SymbolicOptions sparse_options(const HostPlan& p) {
  SymbolicOptions so = symbolic_options_from_env();
  if (p.route == 1) {
    so.tol = p.klu.tol;
    so.scale_min_n = std::numeric_limits<int32_t>::max();
  }
  return so;
}

// Symbolic analysis of the sparse blocks, largest first, on worker threads.
bool analyze_all_sparse(const Canonical& a, HostPlan& p, PlanStatus& status) {
  const Decomposition& d = p.dec;
  const SymbolicOptions so = sparse_options(p);
  const bool scaled = p.route == 1 && !p.rs.empty();
  p.sparse.assign(p.sparse_blocks.size(), SparseBlockPlan{});
  std::vector<int32_t> order(p.sparse_blocks.size());
  for (std::size_t i = 0; i < order.size(); ++i) order[i] = static_cast<int32_t>(i);
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
    const int32_t bx = p.sparse_blocks[x], by = p.sparse_blocks[y];
    return (d.R[bx + 1] - d.R[bx]) > (d.R[by + 1] - d.R[by]);
  });
  std::vector<char> ok(order.size(), 1);
  {
    std::size_t next = 0;
    std::mutex mu;
    auto worker = [&]() {
      while (true) {
        std::size_t i;
        {
          std::lock_guard<std::mutex> g(mu);
          if (next >= order.size()) return;
          i = static_cast<std::size_t>(order[next++]);
        }
        const int32_t b = p.sparse_blocks[i];
        const int32_t s = d.R[b + 1] - d.R[b];
        std::vector<int32_t> rr, cc;
        std::vector<double> vv;
        for (int64_t q = d.dblk_ptr[b]; q < d.dblk_ptr[b + 1]; ++q) {
          rr.push_back(d.dblk_r[q]);
          cc.push_back(d.dblk_c[q]);
          const int64_t e = d.dblk_src[q];
          vv.push_back(scaled ? a.val[e] / p.rs[a.row[e]] : a.val[e]);
        }
        ok[i] = analyze_sparse_block(s, rr, cc, vv, so, p.sparse[i]) ? 1 : 0;
      }
    };
    unsigned hw = std::thread::hardware_concurrency();
    const std::size_t nt = std::min<std::size_t>(order.size(), std::max(1u, std::min(hw ? hw : 1u, 8u)));
    std::vector<std::thread> th;
    for (std::size_t t = 1; t < nt; ++t) th.emplace_back(worker);
    worker();
    for (auto& x : th) x.join();
  }
  p.flops_sparse = 0.0;
  p.nnz_l = p.nnz_u = 0;
  for (std::size_t i = 0; i < ok.size(); ++i) {
    if (!ok[i]) {
      status.ok = false;
      status.bad_block = p.sparse_blocks[i];
      return false;
    }
    p.flops_sparse += p.sparse[i].flops;
    p.nnz_l += p.sparse[i].nnz_l;
    p.nnz_u += p.sparse[i].nnz_u;
  }
  return true;
}
}  // namespace

PlanStatus reanalyze_sparse_blocks(const Canonical& a, HostPlan& p) {
  PlanStatus status;
  std::vector<std::vector<int32_t>> old_q(p.sparse.size());
  for (std::size_t i = 0; i < p.sparse.size(); ++i) old_q[i] = p.sparse[i].q;
  if (p.route == 1) p.rs = klu_row_scale(a, p.klu.scale);  // the scale factors follow the values
  if (!analyze_all_sparse(a, p, status)) return status;
  for (std::size_t i = 0; i < p.sparse.size(); ++i)
    if (p.sparse[i].q != old_q[i]) throw std::runtime_error("column order changed on re-analysis");
  finalize_desc(p);
  return status;
}

PlanStatus build_host_plan(const Canonical& a, HostPlan& p, const KluRoute* klu) {
  const auto t0 = std::chrono::steady_clock::now();
  PlanStatus status;
  p = HostPlan{};
  p.n = a.n;
  p.nnz = static_cast<int64_t>(a.nnz());
  if (klu) {
    // Global KLU: the full pattern (no drop rule, coo_to_csc_i32 at
    // src/solver.cpp:896), KLU's BTF or none, no merging, every block of
    // 2..64 in its COLAMD order (those blocks are factored by the dense kernel
    // under KLU's pivot rule; larger ones by the sparse-block analysis).
    p.route = 1;
    p.klu = *klu;
    p.rs = klu_row_scale(a, klu->scale);
    DecompOptions o;
    o.drop_abs_lt = 0.0;
    o.btf_maxwork = 0.0;
    o.merge = false;
    o.btf = klu->btf != 0;
    o.order_small_max = 64;
    p.dec = build_decomposition(a, o);
    p.dense_max = 64;
  } else {
    p.dec = build_decomposition(a, decomp_options_from_env());
    p.dense_max = dense_block_max_from_env();
  }
  const Decomposition& d = p.dec;
  const int32_t nb = d.nblocks;
  p.kind.assign(nb, 0);
  for (int32_t b = 0; b < nb; ++b) {
    const int32_t s = d.R[b + 1] - d.R[b];
    p.max_block = std::max(p.max_block, s);
    if (s > p.dense_max) {
      p.kind[b] = 1;
      p.sparse_blocks.push_back(b);
    }
  }
  if (!analyze_all_sparse(a, p, status)) return status;

  // Value routing.
  p.vmap.assign(p.nnz, -1);
  int64_t off = 0;
  for (int32_t b = 0; b < nb; ++b) {
    if (p.kind[b] != 0) continue;
    const int32_t s = d.R[b + 1] - d.R[b];
    p.dense_block.push_back(b);
    p.dense_val0.push_back(off);
    for (int64_t q = d.dblk_ptr[b]; q < d.dblk_ptr[b + 1]; ++q)
      p.vmap[d.dblk_src[q]] = off + static_cast<int64_t>(d.dblk_c[q]) * s + d.dblk_r[q];
    off += static_cast<int64_t>(s) * s;
    p.flops_dense += (2.0 / 3.0) * s * s * s;
  }
  p.sparse_val0.resize(p.sparse_blocks.size());
  for (std::size_t i = 0; i < p.sparse_blocks.size(); ++i) {
    const int32_t b = p.sparse_blocks[i];
    const SparseBlockPlan& sp = p.sparse[i];
    p.sparse_val0[i] = off;
    const int64_t base = d.dblk_ptr[b];
    const int32_t ne = sp.a_ptr[sp.s];
    for (int32_t q = 0; q < ne; ++q) p.vmap[d.dblk_src[base + sp.a_ent[q]]] = off + q;
    off += ne;
  }
  p.off_val0 = off;
  for (std::size_t q = 0; q < d.off_src.size(); ++q) p.vmap[d.off_src[q]] = off + static_cast<int64_t>(q);
  off += static_cast<int64_t>(d.off_src.size());
  p.arena_len = off;

  // Canonical CSR (canonical order is row-major, so entry e is CSR slot e).
  p.csr_ptr.assign(static_cast<std::size_t>(a.n) + 1, 0);
  for (std::size_t k = 0; k < a.nnz(); ++k) p.csr_ptr[a.row[k] + 1]++;
  for (int32_t i = 0; i < a.n; ++i) p.csr_ptr[i + 1] += p.csr_ptr[i];
  p.csr_col = a.col;

  // Block dependencies from the inter-block CSR.
  std::vector<int32_t> blk(a.n);
  for (int32_t b = 0; b < nb; ++b)
    for (int32_t k = d.R[b]; k < d.R[b + 1]; ++k) blk[k] = b;
  p.dep_ptr.assign(1, 0);
  std::vector<int32_t> seen(nb, -1);
  for (int32_t b = 0; b < nb; ++b) {
    for (int32_t r = d.R[b]; r < d.R[b + 1]; ++r)
      for (int32_t q = d.off_ptr[r]; q < d.off_ptr[r + 1]; ++q) {
        const int32_t c = blk[d.off_col[q]];
        if (seen[c] == b) continue;
        seen[c] = b;
        p.dep_blk.push_back(c);
      }
    p.dep_ptr.push_back(static_cast<int32_t>(p.dep_blk.size()));
  }
  p.analysis_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  finalize_desc(p);
  return status;
}

void finalize_desc(HostPlan& p) {
  const Decomposition& d = p.dec;
  p.sdesc.resize(p.sparse_blocks.size());
  for (std::size_t i = 0; i < p.sparse_blocks.size(); ++i) {
    const SparseBlockPlan& sp = p.sparse[i];
    bdlu_sparse_desc& s = p.sdesc[i];
    s.block = p.sparse_blocks[i];
    s.s = sp.s;
    s.k0 = sp.k0;
    s.m = sp.m;
    s.scale = sp.scale ? 1 : 0;
    s.tol = sparse_options(p).tol;
    s.a_val0 = p.sparse_val0[i];
    s.a_ptr = sp.a_ptr.data();
    s.a_row = sp.a_row.data();
    s.q = sp.q.data();
    s.prow = sp.prow.data();
    s.l_ptr = sp.l_ptr.data();
    s.l_row = sp.l_row.data();
    s.u_ptr = sp.u_ptr.data();
    s.u_idx = sp.u_idx.data();
    s.lev_ptr = sp.lev_ptr.data();
    s.lev_col = sp.lev_col.data();
    s.nlev = sp.levels;
    s.trow = sp.trow.data();
    s.tperm = sp.tperm.data();
  }
  bdlu_plan_desc& q = p.desc;
  q.n = p.n;
  q.nnz = p.nnz;
  q.vmap = p.vmap.data();
  q.arena_len = p.arena_len;
  q.csr_ptr = p.csr_ptr.data();
  q.csr_col = p.csr_col.data();
  q.P = d.P.data();
  q.Q = d.Q.data();
  q.nblocks = d.nblocks;
  q.R = d.R.data();
  q.off_ptr = d.off_ptr.data();
  q.off_col = d.off_col.data();
  q.off_val0 = p.off_val0;
  q.dep_ptr = p.dep_ptr.data();
  q.dep_blk = p.dep_blk.data();
  q.ndense = static_cast<int32_t>(p.dense_block.size());
  q.dense_block = p.dense_block.data();
  q.dense_val0 = p.dense_val0.data();
  q.nsparse = static_cast<int32_t>(p.sdesc.size());
  q.sparse = p.sdesc.data();
  q.dense_rule = p.route == 1 ? 1 : 0;
  q.dense_tol = p.route == 1 ? p.klu.tol : 1.0;
  q.dense_key = p.route == 1 ? d.tie_key.data() : nullptr;
  q.row_scale = p.route == 1 ? std::max(0, p.klu.scale) : 0;
}

}  // namespace bdlu
