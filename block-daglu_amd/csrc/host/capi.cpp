// C-ABI of the host library (include/bdlu_host.h).
#include <algorithm>
#include <cstring>
#include <string>
#include <tuple>
#include <vector>

#include "../../../include/bdlu_host.h"
#include "blockdaglu/serve.h"
#include "blockdaglu/solver.h"
#include "host.h"
#include "plan.h"

using blockdaglu::CooMatrix;
using blockdaglu::Solver;

namespace blockdaglu {
class ServeEngine;
struct ServeResult;
ServeEngine* serve_engine_create(int device, int slots, uint64_t budget, int precision);
void serve_engine_destroy(ServeEngine* e);
int32_t serve_engine_add_pattern(ServeEngine* e, const CooMatrix& m, int32_t dim);
int32_t serve_engine_pattern_n(ServeEngine* e, int32_t p);
int64_t serve_engine_submit(ServeEngine* e, int32_t p, const double* v, int64_t nv, const double* B, int32_t k);
int64_t serve_engine_submit_device(ServeEngine* e, int32_t p, const double* d_v, const double* d_B, double* d_X,
                                   int32_t k);
bool serve_engine_wait(ServeEngine* e, int64_t id, double* X, ServeResult& res);
SolverError serve_engine_last_error(ServeEngine* e);
}  // namespace blockdaglu

extern "C" {

void* bdlu_solver_create(void) {
  try {
    return new Solver();
  } catch (...) {
    return nullptr;
  }
}

void bdlu_solver_destroy(void* s) { delete static_cast<Solver*>(s); }

int bdlu_solver_set_matrix(void* s, int64_t nnz, const int32_t* rows, const int32_t* cols, const double* vals,
                           int32_t dimension) {
  if (!s || nnz < 0 || (nnz > 0 && (!rows || !cols || !vals))) return 0;
  CooMatrix m;
  m.reserve(static_cast<std::size_t>(nnz));
  for (int64_t k = 0; k < nnz; ++k) m.emplace_back(rows[k], cols[k], vals[k]);
  return static_cast<Solver*>(s)->set_matrix(std::move(m), dimension) ? 1 : 0;
}

int bdlu_solver_analyze(void* s) { return s && static_cast<Solver*>(s)->analyze() ? 1 : 0; }
int bdlu_solver_factor(void* s) { return s && static_cast<Solver*>(s)->factor() ? 1 : 0; }

int bdlu_solver_solve(void* s, double* b, int64_t len) {
  if (!s || len < 0 || (len > 0 && !b)) return 0;
  std::vector<double> v(b, b + len);
  const bool ok = static_cast<Solver*>(s)->solve_in_place(v);
  if (ok) std::memcpy(b, v.data(), sizeof(double) * static_cast<std::size_t>(len));
  return ok ? 1 : 0;
}

int bdlu_solver_solve_multi(void* s, double* B, int64_t n, int32_t nrhs) {
  if (!s || !B || n < 0 || nrhs < 1) return 0;
  std::vector<double> v(B, B + n * nrhs);
  const bool ok = static_cast<Solver*>(s)->solve_multi_in_place(v, nrhs);
  if (ok) std::memcpy(B, v.data(), sizeof(double) * v.size());
  return ok ? 1 : 0;
}

int bdlu_solver_factor_device(void* s, const double* d_values) {
  return s && static_cast<Solver*>(s)->factor_device(d_values) ? 1 : 0;
}
int bdlu_solver_solve_device(void* s, double* d_B, int32_t nrhs, int64_t ldb) {
  return s && static_cast<Solver*>(s)->solve_device(d_B, nrhs, ldb) ? 1 : 0;
}
void* bdlu_solver_device_context(void* s) { return s ? static_cast<Solver*>(s)->device_context() : nullptr; }
void* bdlu_solver_device_stream(void* s) { return s ? static_cast<Solver*>(s)->device_stream() : nullptr; }
uint64_t bdlu_solver_device_plan(void* s) { return s ? static_cast<Solver*>(s)->device_plan() : 0; }

int bdlu_solver_last_error(void* s) { return s ? static_cast<int>(static_cast<Solver*>(s)->last_error()) : -1; }

const char* bdlu_solver_last_error_message(void* s) {
  return s ? static_cast<Solver*>(s)->last_error_message().c_str() : "null solver";
}

int32_t bdlu_solver_n(void* s) { return s ? static_cast<Solver*>(s)->n() : -1; }
const char* bdlu_solver_selected_mode(void* s) { return s ? static_cast<Solver*>(s)->selected_mode() : ""; }

void bdlu_solver_set_precision(void* s, int precision) {
  if (s)
    static_cast<Solver*>(s)->set_precision(precision == 1 ? blockdaglu::Precision::Fp32Ir
                                                           : blockdaglu::Precision::Fp64);
}

void bdlu_solver_report(void* s, bdlu_report* out) {
  if (!s || !out) return;
  const blockdaglu::EngineReport& r = static_cast<Solver*>(s)->report();
  out->nblocks = r.nblocks;
  out->dense_blocks = r.dense_blocks;
  out->sparse_blocks = r.sparse_blocks;
  out->max_block = r.max_block;
  out->fp64_fallback_blocks = r.fp64_fallback_blocks;
  out->repivot_blocks = r.repivot_blocks;
  out->ir_iterations = r.ir_iterations;
  out->plan_reused = r.plan_reused ? 1 : 0;
  out->rel_residual = r.rel_residual;
  out->backward_error = r.backward_error;
  out->componentwise_error = r.componentwise_error;
  out->ir_fp64_refactor = r.ir_fp64_refactor ? 1 : 0;
}

int32_t bdlu_solver_sparse_order(void* s, int32_t b, int32_t* q, int32_t* prow) {
  if (!s) return -1;
  std::vector<int32_t> qv, pv;
  const int32_t k0 = static_cast<Solver*>(s)->sparse_block_order(b, qv, pv);
  if (k0 < 0) return -1;
  if (q) std::copy(qv.begin(), qv.end(), q);
  if (prow) std::copy(pv.begin(), pv.end(), prow);
  return k0;
}

int32_t bdlu_solver_nblocks(void* s) {
  if (!s) return -1;
  const auto& R = static_cast<Solver*>(s)->block_bounds();
  return R.empty() ? 0 : static_cast<int32_t>(R.size()) - 1;
}

void bdlu_solver_copy_perm(void* s, int32_t* P, int32_t* Q, int32_t* R) {
  if (!s) return;
  const Solver* x = static_cast<Solver*>(s);
  if (P) std::copy(x->block_row_perm().begin(), x->block_row_perm().end(), P);
  if (Q) std::copy(x->block_col_perm().begin(), x->block_col_perm().end(), Q);
  if (R) std::copy(x->block_bounds().begin(), x->block_bounds().end(), R);
}

int64_t bdlu_solver_export(void* s, int which, int32_t* rows, int32_t* cols, double* vals) {
  if (!s) return -1;
  Solver* x = static_cast<Solver*>(s);
  CooMatrix m;
  switch (which) {
    case 0: m = x->get_rearranged_matrix(); break;
    case 1: m = x->getP(); break;
    case 2: m = x->getQ(); break;
    case 3: m = x->getL(); break;
    case 4: m = x->getU(); break;
    default: return -1;
  }
  if (x->last_error() != blockdaglu::SolverError::None) return -1;
  if (rows && cols && vals)
    for (std::size_t k = 0; k < m.size(); ++k) {
      rows[k] = std::get<0>(m[k]);
      cols[k] = std::get<1>(m[k]);
      vals[k] = std::get<2>(m[k]);
    }
  return static_cast<int64_t>(m.size());
}

// ---- host analysis queries ------------------------------------------------------
struct HostAnalysis {
  bdlu::Canonical canon;
  bdlu::HostPlan plan;
  bool plan_ok = false;
};

void* bdlu_host_analyze(int64_t nnz, const int32_t* rows, const int32_t* cols, const double* vals, int32_t dimension,
                        int* error_code) {
  auto* h = new HostAnalysis;
  std::vector<std::tuple<int, int, double>> coo(static_cast<std::size_t>(nnz));
  for (int64_t k = 0; k < nnz; ++k) coo[k] = std::make_tuple(rows[k], cols[k], vals[k]);
  std::string msg;
  const bdlu::CanonStatus st = bdlu::canonicalize(coo, dimension, h->canon, msg);
  int code = 0;
  switch (st) {
    case bdlu::CanonStatus::Ok: break;
    case bdlu::CanonStatus::InvalidDimension: code = BDLU_ERR_INVALID_DIMENSION; break;
    case bdlu::CanonStatus::InvalidIndex: code = BDLU_ERR_INVALID_INDEX; break;
    default: code = BDLU_ERR_NON_FINITE; break;
  }
  if (code == 0) {
    try {
      h->plan_ok = bdlu::build_host_plan(h->canon, h->plan).ok;
      if (!h->plan_ok) code = BDLU_ERR_FACTORIZATION_FAILED;
    } catch (...) {
      code = BDLU_ERR_ANALYSIS_FAILED;
    }
  }
  if (error_code) *error_code = code;
  if (code != 0 && code != BDLU_ERR_FACTORIZATION_FAILED) {
    delete h;
    return nullptr;
  }
  return h;
}

void* bdlu_host_analyze_klu(int64_t nnz, const int32_t* rows, const int32_t* cols, const double* vals,
                            int32_t dimension, int btf, int scale, double tol, int* error_code) {
  auto* h = new HostAnalysis;
  std::vector<std::tuple<int, int, double>> coo(static_cast<std::size_t>(nnz));
  for (int64_t k = 0; k < nnz; ++k) coo[k] = std::make_tuple(rows[k], cols[k], vals[k]);
  std::string msg;
  int code = bdlu::canonicalize(coo, dimension, h->canon, msg) == bdlu::CanonStatus::Ok ? 0 : BDLU_ERR_INVALID_INDEX;
  if (code == 0) {
    bdlu::KluRoute klu;
    klu.btf = btf;
    klu.scale = scale;
    klu.tol = tol;
    try {
      h->plan_ok = bdlu::build_host_plan(h->canon, h->plan, &klu).ok;
      if (!h->plan_ok) code = BDLU_ERR_FACTORIZATION_FAILED;
    } catch (...) {
      code = BDLU_ERR_ANALYSIS_FAILED;
    }
  }
  if (error_code) *error_code = code;
  if (code != 0 && code != BDLU_ERR_FACTORIZATION_FAILED) {
    delete h;
    return nullptr;
  }
  return h;
}

int bdlu_route_for(int64_t nnz, const int32_t* rows, const int32_t* cols, const double* vals, int32_t dimension,
                   int32_t* out4, double* tol) {
  if (!out4) return -1;
  std::vector<std::tuple<int, int, double>> coo(static_cast<std::size_t>(nnz));
  for (int64_t k = 0; k < nnz; ++k) coo[k] = std::make_tuple(rows[k], cols[k], vals[k]);
  std::string msg;
  bdlu::Canonical a;
  if (bdlu::canonicalize(coo, dimension, a, msg) != bdlu::CanonStatus::Ok) return BDLU_ERR_INVALID_INDEX;
  const bdlu::RouteMode mode = bdlu::route_mode_from_env();
  bool klu = mode == bdlu::RouteMode::Klu;
  if (mode == bdlu::RouteMode::Auto)
    klu = !bdlu::auto_choose_block(a, bdlu::matrix_sig(a.n, a), bdlu::matrix_sig2(a.n, a), bdlu::matrix_value_sig(a),
                                   0);
  out4[0] = klu ? 1 : 0;
  out4[1] = bdlu::klu_btf_for(a.n, a.nnz());
  out4[2] = bdlu::klu_scale_for(a.n, a.nnz(), out4[1]);
  out4[3] = bdlu::klu_ordering();
  if (tol) *tol = bdlu::klu_tol_for(a.n, a.nnz(), out4[1]);
  return 0;
}

void bdlu_host_free(void* h) { delete static_cast<HostAnalysis*>(h); }

void bdlu_host_info(void* hp, int32_t* n, int32_t* nblocks_raw, int32_t* nblocks, int32_t* nmatch, int64_t* nnz,
                    int64_t* nnz_off, int32_t* nlevels) {
  const HostAnalysis* h = static_cast<HostAnalysis*>(hp);
  const bdlu::Decomposition& d = h->plan.dec;
  if (n) *n = h->canon.n;
  if (nblocks_raw) *nblocks_raw = d.nblocks_raw;
  if (nblocks) *nblocks = d.nblocks;
  if (nmatch) *nmatch = d.nmatch;
  if (nnz) *nnz = static_cast<int64_t>(h->canon.nnz());
  if (nnz_off) *nnz_off = static_cast<int64_t>(d.off_col.size());
  if (nlevels) *nlevels = d.max_level + 1;
}

void bdlu_host_copy_perm(void* hp, int32_t* P, int32_t* Q, int32_t* R) {
  const bdlu::Decomposition& d = static_cast<HostAnalysis*>(hp)->plan.dec;
  if (P) std::copy(d.P.begin(), d.P.end(), P);
  if (Q) std::copy(d.Q.begin(), d.Q.end(), Q);
  if (R) std::copy(d.R.begin(), d.R.end(), R);
}

void bdlu_host_block_kinds(void* hp, int32_t* kind, int32_t* size) {
  const bdlu::HostPlan& p = static_cast<HostAnalysis*>(hp)->plan;
  for (int32_t b = 0; b < p.dec.nblocks; ++b) {
    if (kind) kind[b] = p.kind.empty() ? 0 : p.kind[b];
    if (size) size[b] = p.dec.R[b + 1] - p.dec.R[b];
  }
}

static const bdlu::SparseBlockPlan* find_sparse(void* hp, int32_t b) {
  const bdlu::HostPlan& p = static_cast<HostAnalysis*>(hp)->plan;
  for (std::size_t i = 0; i < p.sparse_blocks.size(); ++i)
    if (p.sparse_blocks[i] == b && i < p.sparse.size()) return &p.sparse[i];
  return nullptr;
}

int bdlu_host_sparse_stats(void* hp, int32_t b, int32_t* st, double* fs) {
  const bdlu::SparseBlockPlan* sp = find_sparse(hp, b);
  if (!sp) return 0;
  if (st) {
    st[0] = sp->s;
    st[1] = sp->k0;
    st[2] = sp->m;
    st[3] = sp->levels;
    st[4] = sp->offdiag_pivots;
    st[5] = sp->scale ? 1 : 0;
  }
  if (fs) {
    fs[0] = static_cast<double>(sp->nnz_l);
    fs[1] = static_cast<double>(sp->nnz_u);
    fs[2] = static_cast<double>(sp->updates);
    fs[3] = sp->flops;
  }
  return 1;
}

int bdlu_host_sparse_order(void* hp, int32_t b, int32_t* q, int32_t* prow) {
  const bdlu::SparseBlockPlan* sp = find_sparse(hp, b);
  if (!sp) return 0;
  if (q) std::copy(sp->q.begin(), sp->q.end(), q);
  if (prow) {
    std::copy(sp->prow.begin(), sp->prow.end(), prow);
    for (int32_t t = 0; t < sp->m; ++t) prow[sp->k0 + t] = sp->trow[sp->tperm[t]];
  }
  return 1;
}

void bdlu_host_signatures(void* hp, uint64_t* sig, uint64_t* sig2, uint64_t* value_sig) {
  const HostAnalysis* h = static_cast<HostAnalysis*>(hp);
  if (sig) *sig = bdlu::matrix_sig(h->canon.n, h->canon);
  if (sig2) *sig2 = bdlu::matrix_sig2(h->canon.n, h->canon);
  if (value_sig) *value_sig = bdlu::matrix_value_sig(h->canon);
}

}  // extern "C"

// ---- serving loop ------------------------------------------------------------
extern "C" {

void* bdlu_serve_create(int device, int slots, uint64_t hbm_budget_bytes, int precision) {
  try {
    return blockdaglu::serve_engine_create(device, slots, hbm_budget_bytes, precision);
  } catch (...) {
    return nullptr;
  }
}

void bdlu_serve_destroy(void* s) { blockdaglu::serve_engine_destroy(static_cast<blockdaglu::ServeEngine*>(s)); }

int32_t bdlu_serve_add_pattern(void* s, int64_t nnz, const int32_t* rows, const int32_t* cols, const double* vals,
                               int32_t dimension) {
  if (!s || nnz < 0 || (nnz > 0 && (!rows || !cols || !vals))) return -static_cast<int32_t>(BDLU_ERR_INVALID_INDEX);
  auto* e = static_cast<blockdaglu::ServeEngine*>(s);
  try {
    CooMatrix m;
    m.reserve(static_cast<std::size_t>(nnz));
    for (int64_t k = 0; k < nnz; ++k) m.emplace_back(rows[k], cols[k], vals[k]);
    const int32_t id = blockdaglu::serve_engine_add_pattern(e, m, dimension);
    return id >= 0 ? id : -static_cast<int32_t>(blockdaglu::serve_engine_last_error(e));
  } catch (...) {
    return -static_cast<int32_t>(BDLU_ERR_ANALYSIS_FAILED);
  }
}

int32_t bdlu_serve_pattern_n(void* s, int32_t pattern) {
  return s ? blockdaglu::serve_engine_pattern_n(static_cast<blockdaglu::ServeEngine*>(s), pattern) : -1;
}

int64_t bdlu_serve_submit(void* s, int32_t pattern, const double* vals, int64_t nnz, const double* B, int32_t nrhs) {
  if (!s || !vals || !B) return -static_cast<int64_t>(BDLU_ERR_INVALID_RHS);
  auto* e = static_cast<blockdaglu::ServeEngine*>(s);
  try {
    const int64_t id = blockdaglu::serve_engine_submit(e, pattern, vals, nnz, B, nrhs);
    return id >= 0 ? id : -static_cast<int64_t>(blockdaglu::serve_engine_last_error(e));
  } catch (...) {
    return -static_cast<int64_t>(BDLU_ERR_BACKEND_UNAVAILABLE);
  }
}

int64_t bdlu_serve_submit_device(void* s, int32_t pattern, const double* d_vals, const double* d_B, double* d_X,
                                 int32_t nrhs) {
  if (!s || !d_vals || !d_B || !d_X) return -static_cast<int64_t>(BDLU_ERR_INVALID_RHS);
  auto* e = static_cast<blockdaglu::ServeEngine*>(s);
  try {
    const int64_t id = blockdaglu::serve_engine_submit_device(e, pattern, d_vals, d_B, d_X, nrhs);
    return id >= 0 ? id : -static_cast<int64_t>(blockdaglu::serve_engine_last_error(e));
  } catch (...) {
    return -static_cast<int64_t>(BDLU_ERR_BACKEND_UNAVAILABLE);
  }
}

int bdlu_serve_wait(void* s, int64_t request, double* X, bdlu_serve_result* out) {
  if (!s) return BDLU_ERR_INVALID_INDEX;
  blockdaglu::ServeResult r;
  blockdaglu::serve_engine_wait(static_cast<blockdaglu::ServeEngine*>(s), request, X, r);
  if (out) {
    out->error = static_cast<int32_t>(r.error);
    out->slot = r.slot;
    out->ir_iterations = r.ir_iterations;
    out->fp64_fallback_blocks = r.fp64_fallback_blocks;
    out->repivots = r.repivots;
    out->rel_residual = r.rel_residual;
    out->backward_error = r.backward_error;
    out->componentwise_error = r.componentwise_error;
    out->t_submit_ms = r.t_submit_ms;
    out->t_start_ms = r.t_start_ms;
    out->t_done_ms = r.t_done_ms;
    out->ir_fp64_refactor = r.ir_fp64_refactor;
  }
  return static_cast<int>(r.error);
}

void bdlu_host_colamd(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t* order) {
  if (n <= 0 || !Ap || !Ai || !order) return;
  const std::vector<int32_t> q = bdlu::colamd_order(n, Ap, Ai);
  std::copy(q.begin(), q.end(), order);
}

}  // extern "C"
