// Factor export (reference src/solver.cpp:1541-1725): a global sparse LU with
// the settings the reference passes to KLU for it. The factorization is the
// host restatement of KLU's left-looking kernel (symbolic.cpp) run on the whole
// matrix with the dense-tail switch turned off, so every column is a sparse
// column with threshold partial pivoting.
#include "export.h"

#include <climits>
#include <cstdlib>
#include <string>

#include "blockdaglu/env.h"
#include "symbolic.h"

namespace bdlu {

namespace {

// KLU scale for the export (btf = 0): USERALGO_KLU_SCALE when set to -1..2,
// else max-row scaling only for very large matrices (src/solver.cpp:417-545).
int export_scale(int32_t n, std::size_t nnz) {
  if (const char* e = blockdaglu_getenv_compat("USERALGO_KLU_SCALE")) {
    const int v = std::atoi(e);
    if (v >= -1 && v <= 2) return v;
  }
  return (n >= 65000 && nnz >= 3000000) ? 2 : 0;
}

// KLU pivot tolerance: USERALGO_KLU_TOL when >= 0, else KLU's default 0.001
// (src/solver.cpp:401-409, 547-556 with btf = 0).
double export_tol() {
  if (const char* e = blockdaglu_getenv_compat("USERALGO_KLU_TOL")) {
    const double v = std::atof(e);
    if (v >= 0.0) return v;
  }
  return 0.001;
}

}  // namespace

bool export_factors(const Canonical& a, ExportFactors& out, std::string& message) {
  out = ExportFactors{};
  const int32_t n = a.n;
  out.n = n;
  if (n == 0) return true;
  SymbolicOptions opt;
  opt.tol = export_tol();
  opt.tail_min = INT_MAX;  // never switch to a dense tail
  const int scale = export_scale(n, a.nnz());
  opt.scale_min_n = scale >= 1 ? 0 : INT_MAX;
  opt.scale_kind = scale;
  SparseBlockPlan plan;
  SparseFactorValues fv;
  if (!analyze_sparse_block(n, a.row, a.col, a.val, opt, plan, &fv) || plan.k0 != n) {
    message = "sparse LU failed for the factor export (a column has no admissible pivot)";
    return false;
  }
  out.P = plan.prow;
  out.Q = plan.q;
  out.Rs = fv.rs;
  std::vector<int32_t> pinv(n);
  for (int32_t k = 0; k < n; ++k) pinv[out.P[k]] = k;
  // L by columns: the unit diagonal, then the entries below it (row positions).
  out.L.reserve(static_cast<std::size_t>(n) + fv.lx.size());
  for (int32_t k = 0; k < n; ++k) {
    out.L.emplace_back(k, k, 1.0);
    for (int32_t p = plan.l_ptr[k]; p < plan.l_ptr[k + 1]; ++p) out.L.emplace_back(pinv[plan.l_row[p]], k, fv.lx[p]);
  }
  // U by columns: the entries above the diagonal in elimination order, then the pivot.
  out.U.reserve(static_cast<std::size_t>(n) + fv.ux.size());
  for (int32_t j = 0; j < n; ++j) {
    for (int32_t p = plan.u_ptr[j]; p < plan.u_ptr[j + 1]; ++p) out.U.emplace_back(plan.u_idx[p], j, fv.ux[p]);
    out.U.emplace_back(j, j, fv.udiag[j]);
  }
  return true;
}

}  // namespace bdlu
