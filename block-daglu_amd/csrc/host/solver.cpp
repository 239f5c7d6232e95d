// blockdaglu::Solver on the MI355X engine. Stage semantics follow the
// reference (src/solver.cpp): set_matrix is storage only (:745-800); analyze
// validates and canonicalizes (:642-722), computes the signatures (:817-819)
// and builds the BLOCK decomposition (:970-1182); factor factors every
// diagonal block and fails with FactorizationFailed if any block fails
// (:1244-1317, BLOCK route); solve_in_place solves and optionally refines
// (:1352-1539).
//
// Difference by design: the device plan (decomposition + symbolic analysis) is
// cached by sparsity pattern, not by pattern+values (the reference keys its
// decomposition cache with the value signature, :974-976), so a new set of
// values on a known pattern goes straight to a device refactorization.
// Resident plans carry device factors, so a plan is leased by one Solver at a
// time; a Solver returns it to the cache on set_matrix or destruction.
#include "blockdaglu/solver.h"

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <sstream>
#include <stdexcept>

#include "../../../include/bdlu_gpu.h"
#include "blockdaglu/env.h"
#include "export.h"
#include "host.h"
#include "plan.h"

namespace blockdaglu {

namespace {

struct PlanKey {
  int32_t n = 0;
  int64_t nnz = 0;
  uint64_t sig = 0, sig2 = 0, drop = 0, config = 0;
  int32_t route = 0;  // 0 BLOCK, 1 global KLU
  bool operator==(const PlanKey& o) const {
    return n == o.n && nnz == o.nnz && sig == o.sig && sig2 == o.sig2 && drop == o.drop && config == o.config &&
           route == o.route;
  }
};

struct Resident {
  PlanKey key;
  uint64_t id = 0;
  std::unique_ptr<bdlu::HostPlan> host;
};

std::mutex g_mu;
bdlu_ctx* g_ctx = nullptr;
int g_ctx_status = -1;
std::deque<std::unique_ptr<Resident>> g_free;

bdlu_ctx* context(std::string& msg) {
  std::lock_guard<std::mutex> g(g_mu);
  if (g_ctx) return g_ctx;
  if (g_ctx_status > 0) {
    msg = "no usable GPU for the BLOCK engine";
    return nullptr;
  }
  int dev = 0;
  if (const char* e = std::getenv("BLOCKDAGLU_GPU_DEVICE")) dev = std::atoi(e);
  uint64_t budget = 0;
  if (const char* e = std::getenv("BLOCKDAGLU_GPU_HBM_BUDGET_GB")) budget = static_cast<uint64_t>(std::atof(e) * 1e9);
  const int rc = bdlu_gpu_create(dev, budget, &g_ctx);
  g_ctx_status = rc;
  if (rc != BDLU_OK) {
    g_ctx = nullptr;
    msg = "no usable GPU for the BLOCK engine (bdlu_gpu_create failed)";
    return nullptr;
  }
  return g_ctx;
}

bool cache_enabled() {
  if (const char* e = blockdaglu_getenv_compat("USERALGO_CACHE")) return std::atoi(e) != 0;
  return true;
}
std::size_t cache_max() {
  if (const char* e = blockdaglu_getenv_compat("USERALGO_CACHE_MAX")) {
    const int v = std::atoi(e);
    if (v > 0) return static_cast<std::size_t>(v);
  }
  return 8;
}

std::unique_ptr<Resident> lease(const PlanKey& k) {
  if (!cache_enabled()) return nullptr;  // USERALGO_CACHE=0: every analyze() analyses its own values
  std::lock_guard<std::mutex> g(g_mu);
  for (auto it = g_free.begin(); it != g_free.end(); ++it)
    if ((*it)->key == k) {
      std::unique_ptr<Resident> r = std::move(*it);
      g_free.erase(it);
      return r;
    }
  return nullptr;
}

void give_back(std::unique_ptr<Resident> r) {
  if (!r) return;
  std::unique_ptr<Resident> victim;
  {
    std::lock_guard<std::mutex> g(g_mu);
    if (!cache_enabled() || !g_ctx) {
      victim = std::move(r);
    } else {
      g_free.push_back(std::move(r));
      if (g_free.size() > cache_max()) {
        victim = std::move(g_free.front());
        g_free.pop_front();
      }
    }
  }
  if (victim && g_ctx) bdlu_gpu_plan_release(g_ctx, victim->id);
}

// Drops the oldest cached plan; returns false if none.
bool evict_one() {
  std::unique_ptr<Resident> victim;
  {
    std::lock_guard<std::mutex> g(g_mu);
    if (g_free.empty()) return false;
    victim = std::move(g_free.front());
    g_free.pop_front();
  }
  if (g_ctx) bdlu_gpu_plan_release(g_ctx, victim->id);
  return true;
}

uint64_t gpu_config_sig() {
  uint64_t h = bdlu::cache_config_sig(1);
  static const char* const kKeys[] = {"BLOCKDAGLU_GPU_TAIL_DENSITY", "BLOCKDAGLU_GPU_TAIL_MIN", "BLOCKDAGLU_GPU_TAIL_TOL",
                                      "USERALGO_BLOCK_KLU_TOL", "USERALGO_BLOCK_KLU_BIG_N"};
  for (const char* k : kKeys) {
    const char* v = blockdaglu_getenv_compat(k);
    for (const char* p = k; *p; ++p) h = (h ^ static_cast<unsigned char>(*p)) * 1099511628211ull;
    if (!v) {
      h = (h ^ 0xff) * 1099511628211ull;
      continue;
    }
    for (const char* p = v; *p; ++p) h = (h ^ static_cast<unsigned char>(*p)) * 1099511628211ull;
  }
  return h;
}

}  // namespace

struct SolverImpl {
  CooMatrix original;
  int32_t dim_hint = -1;
  bool prepared = false, validated = false, decomposed = false, factored = false;
  bdlu::Canonical canon;
  PlanKey key;
  std::unique_ptr<Resident> plan;
  bool singular_at_analysis = false;
  // Route of the last analyze (src/solver.cpp:820-860): the requested mode and
  // whether the global-KLU route was selected.
  bdlu::RouteMode mode = bdlu::RouteMode::Auto;
  bool use_klu = false;
  uint64_t sig = 0, sig2 = 0;
  std::unique_ptr<bdlu::ExportFactors> exported;  // factor export, built on first request
  ~SolverImpl() { give_back(std::move(plan)); }
};

Solver::Solver() : impl_(new SolverImpl) {}
Solver::~Solver() = default;

bool Solver::clear_error_ok() {
  last_error_ = SolverError::None;
  last_error_message_.clear();
  return true;
}

bool Solver::fail(SolverError error, std::string message) {
  last_error_ = error;
  last_error_message_ = std::move(message);
  return false;
}

bool Solver::set_matrix(CooMatrix m) { return set_matrix(std::move(m), -1); }

bool Solver::set_matrix(CooMatrix m, int32_t dimension) {
  clear_error_ok();
  if (dimension < -1) return fail(SolverError::InvalidDimension, "matrix dimension must be non-negative");
  SolverImpl& s = *impl_;
  give_back(std::move(s.plan));
  s.original = std::move(m);
  s.dim_hint = dimension;
  s.prepared = true;
  s.validated = s.decomposed = s.factored = false;
  s.singular_at_analysis = false;
  s.use_klu = false;
  s.exported.reset();
  s.canon = bdlu::Canonical{};
  n_ = dimension >= 0 ? dimension : -1;
  report_ = EngineReport{};
  return true;
}

// Validation and canonicalization of the stored matrix (reference
// src/solver.cpp:642-722), once per set_matrix.
bool Solver::validate() {
  SolverImpl& s = *impl_;
  if (!s.validated) {
    std::string msg;
    const bdlu::CanonStatus st = bdlu::canonicalize(s.original, s.dim_hint, s.canon, msg);
    switch (st) {
      case bdlu::CanonStatus::Ok:
        break;
      case bdlu::CanonStatus::InvalidDimension:
        return fail(SolverError::InvalidDimension, msg);
      case bdlu::CanonStatus::InvalidIndex:
        return fail(SolverError::InvalidIndex, msg);
      case bdlu::CanonStatus::NonFinite:
      case bdlu::CanonStatus::DuplicateOverflow:
        return fail(SolverError::NonFiniteValue, msg);
    }
    s.validated = true;
  }
  n_ = s.canon.n;
  return true;
}

const char* Solver::selected_mode() const { return impl_->use_klu ? "KLU" : "BLOCK"; }

namespace {
// Resident plan for key k: the Solver's own if it matches, a cached one, or a
// new one built by `build` and uploaded (evicting cached plans while the HBM
// budget cannot take it). Returns BDLU_OK, or the build/upload failure with a
// message; *singular is set when the host analysis found a sparse block with
// no admissible pivot (the plan is then not uploaded).
template <class Build>
int acquire_plan(bdlu_ctx* ctx, SolverImpl& s, const PlanKey& k, Build&& build, bool* reused, bool* singular,
                 double* build_ms, std::string& msg) {
  *singular = false;
  *build_ms = 0.0;
  if (s.plan && s.plan->key == k) {
    *reused = true;
    return BDLU_OK;
  }
  give_back(std::move(s.plan));
  s.plan = lease(k);
  *reused = static_cast<bool>(s.plan);
  if (s.plan) return BDLU_OK;
  auto r = std::make_unique<Resident>();
  r->key = k;
  r->host = std::make_unique<bdlu::HostPlan>();
  bdlu::PlanStatus ps;
  const auto t0 = std::chrono::steady_clock::now();
  try {
    ps = build(*r->host);
  } catch (const std::exception& e) {
    msg = e.what();
    return BDLU_E_ANALYSIS;
  }
  *build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (!ps.ok) {
    *singular = true;
    return BDLU_OK;
  }
  int rc = BDLU_OK;
  while ((rc = bdlu_gpu_plan_upload(ctx, &r->host->desc, &r->id)) == BDLU_E_BUDGET) {
    if (!evict_one()) break;
  }
  if (rc != BDLU_OK) {
    msg = bdlu_gpu_last_error(ctx);
    return rc;
  }
  s.plan = std::move(r);
  return BDLU_OK;
}
}  // namespace

bool Solver::analyze() {
  clear_error_ok();
  SolverImpl& s = *impl_;
  if (!s.prepared) return fail(SolverError::NotPrepared, "set_matrix must be called before analyze");
  if (!validate()) return false;
  s.sig = bdlu::matrix_sig(s.canon.n, s.canon);
  s.sig2 = bdlu::matrix_sig2(s.canon.n, s.canon);
  std::string msg;
  bdlu_ctx* ctx = context(msg);
  if (!ctx) return fail(SolverError::BackendUnavailable, msg);
  report_ = EngineReport{};
  s.singular_at_analysis = false;
  s.decomposed = s.factored = false;
  // Mode (src/solver.cpp:820-860): USERALGO_MODE forces a route; AUTO decides
  // from the matrix (src/solver.cpp:432-515).
  s.mode = bdlu::route_mode_from_env();
  s.use_klu = s.mode == bdlu::RouteMode::Klu;
  const uint64_t config = gpu_config_sig();
  uint64_t value_sig = 0;
  if (s.mode == bdlu::RouteMode::Auto) {
    value_sig = bdlu::matrix_value_sig(s.canon);
    s.use_klu = !bdlu::auto_choose_block(s.canon, s.sig, s.sig2, value_sig, config);
  }
  if (s.use_klu) return analyze_klu();

  PlanKey k;
  k.n = s.canon.n;
  k.nnz = static_cast<int64_t>(s.canon.nnz());
  k.sig = s.sig;
  k.sig2 = s.sig2;
  k.config = config;
  {
    // Values below the drop threshold change the pattern (ubtf:582-589).
    double drop = bdlu::decomp_options_from_env().drop_abs_lt;
    if (drop < 0.0) drop = (k.n >= 50000) ? 1e-20 : 0.0;
    uint64_t h = 0;
    if (drop > 0.0)
      for (std::size_t e = 0; e < s.canon.nnz(); ++e)
        if (std::abs(s.canon.val[e]) < drop) h = (h ^ (e + 1)) * 1099511628211ull;
    k.drop = h;
  }
  bool reused = false, singular = false;
  double build_ms = 0.0;
  const int rc = acquire_plan(
      ctx, s, k, [&](bdlu::HostPlan& hp) { return bdlu::build_host_plan(s.canon, hp); }, &reused, &singular,
      &build_ms, msg);
  if (rc == BDLU_E_ANALYSIS) {
    // AUTO: a failed block analysis falls back to global KLU (src/solver.cpp:1026-1070).
    if (s.mode == bdlu::RouteMode::Auto) {
      s.use_klu = true;
      return analyze_klu();
    }
    return fail(SolverError::AnalysisFailed, msg);
  }
  if (rc != BDLU_OK) return fail(SolverError::BackendUnavailable, msg);
  report_.plan_reused = reused;
  s.key = k;
  if (singular) {
    // No admissible pivot sequence for a sparse block: the reference's KLU
    // factorization of that block fails, which surfaces in factor().
    s.singular_at_analysis = true;
    s.decomposed = true;
    return true;
  }
  const bdlu::HostPlan& hp = *s.plan->host;
  if (s.mode == bdlu::RouteMode::Auto) {
    // An incomplete structural match makes AUTO take global KLU
    // (src/solver.cpp:1078-1131); an expensive analysis that left one block
    // above 80% of n makes it prefer global KLU next time (:1133-1145).
    if (hp.dec.nmatch < k.n) {
      s.use_klu = true;
      return analyze_klu();
    }
    if (!reused) bdlu::note_block_analysis(s.canon, s.sig, s.sig2, value_sig, config, build_ms, hp.max_block);
  }
  report_.nblocks = hp.dec.nblocks;
  report_.sparse_blocks = static_cast<int32_t>(hp.sparse_blocks.size());
  report_.dense_blocks = report_.nblocks - report_.sparse_blocks;
  report_.max_block = hp.max_block;
  s.decomposed = true;
  return true;
}

// The global-KLU route's analysis (src/solver.cpp:878-968): KLU settings from
// pick_klu_btf_for / pick_klu_scale_for / pick_klu_tol_for, then KLU's BTF (or
// none), per-block COLAMD, and the pivot rows of a first factorization, as a
// device plan like the BLOCK route's.
bool Solver::analyze_klu() {
  SolverImpl& s = *impl_;
  s.use_klu = true;
  std::string msg;
  bdlu_ctx* ctx = context(msg);
  if (!ctx) return fail(SolverError::BackendUnavailable, msg);
  if (bdlu::klu_ordering() != 1)
    return fail(SolverError::BackendUnavailable,
                "KLU ordering other than COLAMD (USERALGO_KLU_ORDERING=1) is not available in this build");
  bdlu::KluRoute klu;
  const std::size_t nnz = s.canon.nnz();
  klu.btf = bdlu::klu_btf_for(s.canon.n, nnz);
  klu.scale = bdlu::klu_scale_for(s.canon.n, nnz, klu.btf);
  klu.tol = bdlu::klu_tol_for(s.canon.n, nnz, klu.btf);
  PlanKey k;
  k.n = s.canon.n;
  k.nnz = static_cast<int64_t>(nnz);
  k.sig = s.sig;
  k.sig2 = s.sig2;
  k.route = 1;
  {
    uint64_t h = gpu_config_sig();
    uint64_t tb = 0;
    std::memcpy(&tb, &klu.tol, sizeof(tb));
    for (uint64_t x : {static_cast<uint64_t>(klu.btf), static_cast<uint64_t>(klu.scale + 8), tb})
      h = (h ^ x) * 1099511628211ull;
    k.config = h;
  }
  bool reused = false, singular = false;
  double build_ms = 0.0;
  const int rc = acquire_plan(
      ctx, s, k, [&](bdlu::HostPlan& hp) { return bdlu::build_host_plan(s.canon, hp, &klu); }, &reused, &singular,
      &build_ms, msg);
  if (rc == BDLU_E_ANALYSIS) return fail(SolverError::AnalysisFailed, "global KLU symbolic analysis failed: " + msg);
  if (rc != BDLU_OK) return fail(SolverError::BackendUnavailable, msg);
  report_.plan_reused = reused;
  s.key = k;
  s.singular_at_analysis = singular;
  s.decomposed = true;
  s.factored = false;
  if (!singular) {
    const bdlu::HostPlan& hp = *s.plan->host;
    report_.nblocks = hp.dec.nblocks;
    report_.sparse_blocks = static_cast<int32_t>(hp.sparse_blocks.size());
    report_.dense_blocks = report_.nblocks - report_.sparse_blocks;
    report_.max_block = hp.max_block;
  }
  return true;
}

bool Solver::factor() {
  clear_error_ok();
  return factor_values(impl_->canon.val.data(), false);
}

bool Solver::factor_device(const double* d_values) {
  clear_error_ok();
  if (!d_values) return fail(SolverError::InvalidIndex, "null device value pointer");
  return factor_values(d_values, true);
}

// One refactorization on the resident plan, with the re-analysis the fixed
// pivot order may need (BDLU_E_REPIVOT: re-analyse this pattern with the
// current values and retry once). Returns the engine status.
int Solver::refactor_plan(const double* values, bool on_device, int prec, bdlu_factor_stats& fs) {
  SolverImpl& s = *impl_;
  std::string msg;
  bdlu_ctx* ctx = context(msg);
  if (!ctx) return BDLU_E_BACKEND;
  const int where = on_device ? BDLU_MEM_DEVICE : BDLU_MEM_HOST;
  int rc = bdlu_gpu_refactor(ctx, s.plan->id, values, where, prec, &fs);
  if (rc != BDLU_E_REPIVOT) return rc;
  if (on_device && bdlu_gpu_plan_values(ctx, s.plan->id, s.canon.val.data()) != BDLU_OK) return BDLU_E_BACKEND;
  std::unique_ptr<bdlu::HostPlan> host = std::move(s.plan->host);
  bdlu::PlanStatus ps;
  try {
    ps = bdlu::reanalyze_sparse_blocks(s.canon, *host);
  } catch (const std::exception&) {
    ps.ok = false;
  }
  // The device copy of the old analysis is stale either way.
  bdlu_gpu_plan_release(ctx, s.plan->id);
  s.plan.reset();
  if (!ps.ok) {
    fs.failed_blocks = std::max(1, fs.failed_blocks);
    return BDLU_E_FACTOR;
  }
  auto r = std::make_unique<Resident>();
  r->key = s.key;
  r->host = std::move(host);
  s.plan = std::move(r);
  int up = BDLU_OK;
  while ((up = bdlu_gpu_plan_upload(ctx, &s.plan->host->desc, &s.plan->id)) == BDLU_E_BUDGET) {
    if (!evict_one()) break;
  }
  if (up != BDLU_OK) {
    s.plan.reset();
    return up;
  }
  const int32_t repivots = fs.repivot_blocks;
  rc = bdlu_gpu_refactor(ctx, s.plan->id, s.canon.val.data(), BDLU_MEM_HOST, prec, &fs);
  fs.repivot_blocks = repivots;
  return rc;
}

bool Solver::factor_values(const double* values, bool on_device) {
  SolverImpl& s = *impl_;
  if (!s.decomposed) return fail(SolverError::NotPrepared, "analyze must succeed before factor");
  std::string msg;
  bdlu_ctx* ctx = context(msg);
  if (!ctx) return fail(SolverError::BackendUnavailable, msg);
  const int prec = precision_ == Precision::Fp32Ir ? BDLU_FP32 : BDLU_FP64;
  bdlu_factor_stats fs{};
  s.factored = false;
  int rc = BDLU_E_FACTOR;
  if (s.plan && !s.singular_at_analysis) rc = refactor_plan(values, on_device, prec, fs);
  else fs.failed_blocks = 1;
  report_.fp64_fallback_blocks = fs.fp64_fallback_blocks;
  report_.repivot_blocks = fs.repivot_blocks;
  const bool failed = rc == BDLU_E_FACTOR || rc == BDLU_E_REPIVOT;
  if (failed && !s.use_klu && s.mode == bdlu::RouteMode::Auto) {
    // AUTO stays robust: the same canonical matrix is refactored on the
    // global-KLU route (src/solver.cpp:1264-1316).
    if (on_device && s.plan && !s.singular_at_analysis &&
        bdlu_gpu_plan_values(ctx, s.plan->id, s.canon.val.data()) != BDLU_OK)
      return fail(SolverError::BackendUnavailable, bdlu_gpu_last_error(ctx));
    if (!analyze_klu()) {
      if (last_error_ == SolverError::AnalysisFailed)
        last_error_message_ = "global KLU analysis failed during AUTO fallback";
      return false;
    }
    bdlu_factor_stats kfs{};
    int krc = BDLU_E_FACTOR;
    if (s.plan && !s.singular_at_analysis) krc = refactor_plan(s.canon.val.data(), false, prec, kfs);
    report_.fp64_fallback_blocks = kfs.fp64_fallback_blocks;
    report_.repivot_blocks = kfs.repivot_blocks;
    if (krc == BDLU_E_FACTOR || krc == BDLU_E_REPIVOT)
      return fail(SolverError::FactorizationFailed, "global KLU factorization failed during AUTO fallback");
    if (krc != BDLU_OK) return fail(SolverError::BackendUnavailable, bdlu_gpu_last_error(ctx));
    s.factored = true;
    return true;
  }
  if (failed) {
    if (!s.plan && !s.singular_at_analysis) s.decomposed = false;
    if (s.use_klu) return fail(SolverError::FactorizationFailed, "global KLU numeric factorization failed");
    std::ostringstream m;
    m << "block factorization failed for " << std::max(1, fs.failed_blocks) << " diagonal block(s)";
    return fail(SolverError::FactorizationFailed, m.str());
  }
  if (rc != BDLU_OK) {
    if (!s.plan) s.decomposed = false;
    return fail(SolverError::BackendUnavailable, bdlu_gpu_last_error(ctx));
  }
  s.factored = true;
  return true;
}

namespace {
bdlu_ir_opts ir_options(Precision p) {
  bdlu_ir_opts o{};
  // Default: the reference's stop rule ||r||_2/||b||_2 < 1e-12 (src/solver.cpp:1508-1521);
  // BLOCKDAGLU_GPU_IR_STOP=backward stops on normwise backward error <= 1e-12 instead,
  // =componentwise on componentwise backward error <= 1e-12.
  o.stop_rule = 0;
  if (const char* e = std::getenv("BLOCKDAGLU_GPU_IR_STOP")) {
    if (std::string(e) == "backward") o.stop_rule = 1;
    if (std::string(e) == "componentwise") o.stop_rule = 2;
  }
  o.tol = 1e-12;
  int iters = 0;
  if (const char* e = blockdaglu_getenv_compat("USERALGO_REFINE_ITERS")) {
    const int v = std::atoi(e);
    if (v >= 0) iters = v;
  }
  if (p == Precision::Fp32Ir) {
    int cap = 30;
    if (const char* e = std::getenv("BLOCKDAGLU_GPU_IR_MAX")) cap = std::max(1, std::atoi(e));
    iters = std::max(iters, cap);
    // FP32 factors that cannot reach the stop rule are redone in FP64.
    o.fp64_fallback = 1;
  }
  o.max_iters = iters;
  return o;
}
}  // namespace

bool Solver::solve_in_place(std::vector<double>& b) {
  clear_error_ok();
  if (!impl_->factored) return fail(SolverError::NotPrepared, "factor must succeed before solve");
  if (static_cast<int32_t>(b.size()) != n_)
    return fail(SolverError::InvalidRightHandSide, "right-hand side size does not match matrix dimension");
  return solve_multi_in_place(b, 1);
}

bool Solver::solve_multi_in_place(std::vector<double>& B, int32_t nrhs) {
  clear_error_ok();
  SolverImpl& s = *impl_;
  if (!s.factored) return fail(SolverError::NotPrepared, "factor must succeed before solve");
  if (nrhs < 1 || static_cast<int64_t>(B.size()) != static_cast<int64_t>(n_) * nrhs)
    return fail(SolverError::InvalidRightHandSide, "right-hand side size does not match matrix dimension");
  for (double v : B)
    if (!std::isfinite(v)) return fail(SolverError::NonFiniteValue, "right-hand side contains NaN or infinity");
  std::string msg;
  bdlu_ctx* ctx = context(msg);
  if (!ctx) return fail(SolverError::BackendUnavailable, msg);
  bdlu_ir_opts ir = ir_options(precision_);
  bdlu_solve_stats ss{};
  const int rc = bdlu_gpu_solve(ctx, s.plan->id, B.data(), nrhs, n_, BDLU_MEM_HOST, &ir, &ss);
  if (rc != BDLU_OK) return fail(SolverError::SolveFailed, std::string("block solve failed: ") + bdlu_gpu_last_error(ctx));
  report_.ir_iterations = ss.iterations;
  report_.ir_fp64_refactor = ss.fp64_refactor != 0;
  report_.rel_residual = ss.rel_residual;
  report_.backward_error = ss.backward_error;
  report_.componentwise_error = ss.componentwise_error;
  for (double v : B)
    if (!std::isfinite(v)) return fail(SolverError::SolveFailed, "block solve produced a non-finite value");
  return true;
}

bool Solver::solve_device(double* d_B, int32_t nrhs, int64_t ldb) {
  clear_error_ok();
  SolverImpl& s = *impl_;
  if (!s.factored) return fail(SolverError::NotPrepared, "factor must succeed before solve");
  if (!d_B || nrhs < 1 || ldb < n_) return fail(SolverError::InvalidRightHandSide, "bad device right-hand side");
  std::string msg;
  bdlu_ctx* ctx = context(msg);
  if (!ctx) return fail(SolverError::BackendUnavailable, msg);
  bdlu_ir_opts ir = ir_options(precision_);
  bdlu_solve_stats ss{};
  const int rc = bdlu_gpu_solve(ctx, s.plan->id, d_B, nrhs, ldb, BDLU_MEM_DEVICE, &ir, &ss);
  if (rc != BDLU_OK) return fail(SolverError::SolveFailed, std::string("block solve failed: ") + bdlu_gpu_last_error(ctx));
  report_.ir_iterations = ss.iterations;
  report_.ir_fp64_refactor = ss.fp64_refactor != 0;
  report_.rel_residual = ss.rel_residual;
  report_.backward_error = ss.backward_error;
  report_.componentwise_error = ss.componentwise_error;
  return true;
}

void* Solver::device_context() const {
  std::string msg;
  return context(msg);
}

uint64_t Solver::device_plan() const { return impl_->plan ? impl_->plan->id : 0; }

void* Solver::device_stream() const {
  std::string msg;
  bdlu_ctx* c = context(msg);
  return c ? bdlu_gpu_stream(c) : nullptr;
}

// Exports (reference src/solver.cpp:1598-1725): a global sparse LU of the
// matrix given to set_matrix, built on the first request and kept until the
// next set_matrix; they need neither analyze() nor a device.
bool Solver::ensure_export() {
  SolverImpl& s = *impl_;
  if (s.exported) return true;
  if (!s.prepared) return fail(SolverError::NotPrepared, "set_matrix must be called before exporting factors");
  if (!validate()) return false;
  auto e = std::make_unique<bdlu::ExportFactors>();
  std::string msg;
  try {
    if (!bdlu::export_factors(s.canon, *e, msg)) return fail(SolverError::ExportFailed, msg);
  } catch (const std::exception& ex) {
    return fail(SolverError::ExportFailed, ex.what());
  }
  s.exported = std::move(e);
  return true;
}

// P (R \ A) Q entry by entry, in the canonical entry order; R holds the row
// scale factors of the original rows (all 1 unless the export scales).
CooMatrix Solver::get_rearranged_matrix() {
  clear_error_ok();
  if (!ensure_export()) return {};
  const bdlu::ExportFactors& e = *impl_->exported;
  const bdlu::Canonical& a = impl_->canon;
  std::vector<int32_t> pinv(e.n), qinv(e.n);
  for (int32_t k = 0; k < e.n; ++k) {
    pinv[e.P[k]] = k;
    qinv[e.Q[k]] = k;
  }
  CooMatrix out;
  out.reserve(a.nnz());
  for (std::size_t t = 0; t < a.nnz(); ++t) {
    const double r = e.Rs.empty() ? 1.0 : e.Rs[a.row[t]];
    out.emplace_back(pinv[a.row[t]], qinv[a.col[t]], r != 0.0 ? a.val[t] / r : a.val[t]);
  }
  return out;
}

// Q as a matrix: entry (Q[k], k) = 1, so (A Q)(:, k) = A(:, Q[k]).
CooMatrix Solver::getQ() {
  clear_error_ok();
  if (!ensure_export()) return {};
  const bdlu::ExportFactors& e = *impl_->exported;
  CooMatrix out;
  out.reserve(e.n);
  for (int32_t k = 0; k < e.n; ++k) out.emplace_back(e.Q[k], k, 1.0);
  return out;
}

// P as a matrix: entry (k, P[k]) = 1, so (P A)(k, :) = A(P[k], :).
CooMatrix Solver::getP() {
  clear_error_ok();
  if (!ensure_export()) return {};
  const bdlu::ExportFactors& e = *impl_->exported;
  CooMatrix out;
  out.reserve(e.n);
  for (int32_t k = 0; k < e.n; ++k) out.emplace_back(k, e.P[k], 1.0);
  return out;
}

CooMatrix Solver::getL() {
  clear_error_ok();
  if (!ensure_export()) return {};
  return impl_->exported->L;
}

CooMatrix Solver::getU() {
  clear_error_ok();
  if (!ensure_export()) return {};
  return impl_->exported->U;
}

int32_t Solver::sparse_block_order(int32_t block, std::vector<int32_t>& q, std::vector<int32_t>& prow) const {
  if (!impl_->plan) return -1;
  const bdlu::HostPlan& hp = *impl_->plan->host;
  for (std::size_t i = 0; i < hp.sparse_blocks.size(); ++i) {
    if (hp.sparse_blocks[i] != block) continue;
    const bdlu::SparseBlockPlan& sp = hp.sparse[i];
    q = sp.q;
    prow = sp.prow;
    for (int32_t t = 0; t < sp.m; ++t) prow.push_back(sp.trow[sp.tperm[t]]);
    return sp.k0;
  }
  return -1;
}

const std::vector<int32_t>& Solver::block_row_perm() const {
  static const std::vector<int32_t> empty;
  return impl_->plan ? impl_->plan->host->dec.P : empty;
}
const std::vector<int32_t>& Solver::block_col_perm() const {
  static const std::vector<int32_t> empty;
  return impl_->plan ? impl_->plan->host->dec.Q : empty;
}
const std::vector<int32_t>& Solver::block_bounds() const {
  static const std::vector<int32_t> empty;
  return impl_->plan ? impl_->plan->host->dec.R : empty;
}

}  // namespace blockdaglu
