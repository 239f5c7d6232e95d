#pragma once
// Device plan assembly: turns the host analysis of one sparsity pattern into
// the plain arrays of bdlu_plan_desc (include/bdlu_gpu.h).
#include <cstdint>
#include <memory>
#include <vector>

#include "../../../include/bdlu_gpu.h"
#include "host.h"
#include "symbolic.h"

namespace bdlu {

// Settings of the global-KLU route (reference src/solver.cpp:878-968,
// :1190-1242: klu_analyze + klu_factor on the whole matrix, btf/scale/tol from
// pick_klu_btf_for / pick_klu_scale_for / pick_klu_tol_for).
struct KluRoute {
  int btf = 0;         // 1: KLU's BTF blocks (no merge); 0: the matrix is one block
  int scale = 0;       // KLU scale: <= 0 none, 1 row sums, 2 row maxima of |a| over whole rows
  double tol = 0.001;  // threshold pivoting tolerance (diagonal preference)
};

struct HostPlan {
  int32_t route = 0;   // 0: BLOCK route, 1: global-KLU route
  KluRoute klu;        // route 1 only
  std::vector<double> rs;  // route 1 with scale > 0: row scale factor of each original row
  int32_t n = 0;
  int64_t nnz = 0;
  Decomposition dec;
  int32_t dense_max = 64;
  std::vector<int32_t> kind;            // per block: 0 dense, 1 sparse
  std::vector<int32_t> sparse_blocks;   // block ids of sparse blocks
  std::vector<SparseBlockPlan> sparse;  // parallel to sparse_blocks
  // desc arrays
  std::vector<int64_t> vmap;
  int64_t arena_len = 0;
  std::vector<int64_t> csr_ptr;
  std::vector<int32_t> csr_col;
  std::vector<int32_t> dep_ptr, dep_blk;
  std::vector<int32_t> dense_block;
  std::vector<int64_t> dense_val0;
  int64_t off_val0 = 0;
  std::vector<int64_t> sparse_val0;
  std::vector<bdlu_sparse_desc> sdesc;
  bdlu_plan_desc desc{};
  // statistics
  int32_t max_block = 0;
  double flops_sparse = 0.0, flops_dense = 0.0;
  int64_t nnz_l = 0, nnz_u = 0;
  double analysis_seconds = 0.0;
};

struct PlanStatus {
  bool ok = true;
  int32_t bad_block = -1;  // sparse block whose analysis failed (singular)
};

// Builds the full host plan for a canonical matrix: the BLOCK route (klu ==
// nullptr) or the global-KLU route with the given settings. Throws
// std::runtime_error on decomposition failure; returns ok=false if a sparse
// block is singular.
PlanStatus build_host_plan(const Canonical& a, HostPlan& out, const KluRoute* klu = nullptr);
// KLU's row scale factors (klu_scale semantics): scale 1 sums, scale 2 maxima
// of |a_ij| over each row, in column order; rows without a nonzero get 1.
std::vector<double> klu_row_scale(const Canonical& a, int scale);
// Re-runs the per-block symbolic analysis (pivot choice) of every sparse block
// with the current values, keeping the decomposition, the column orders and the
// value routing (the column order depends on the pattern only).
PlanStatus reanalyze_sparse_blocks(const Canonical& a, HostPlan& p);
// Fills out.desc (pointers into out's own vectors).
void finalize_desc(HostPlan& p);

int32_t dense_block_max_from_env();

}  // namespace bdlu
