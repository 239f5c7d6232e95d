// Serving loop: registered patterns, plans resident per slot, FIFO admission of
// requests onto slots (one engine context and HIP stream per slot, so requests
// in different slots overlap on the device). No counterpart in the reference;
// each request runs the same refactor + solve (+ refinement) as Solver does.
#include "blockdaglu/serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <deque>
#include <list>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>

#include "../../../include/bdlu_gpu.h"
#include "blockdaglu/env.h"
#include "host.h"
#include "plan.h"

namespace blockdaglu {
namespace {

using Clock = std::chrono::steady_clock;

struct Pattern {
  int32_t n = 0;
  int64_t nnz_in = 0;
  std::vector<int64_t> slot;  // input entry -> canonical index, -1 if not part of the plan
  bdlu::Canonical canon;
  std::unique_ptr<bdlu::HostPlan> plan;
  bool singular = false;
};

struct Request {
  int32_t pattern = 0;
  std::vector<double> vals;  // canonical order
  std::vector<double> B;     // right-hand sides in, solutions out (column-major n x nrhs)
  int32_t nrhs = 1;
  ServeResult res;
  bool done = false;
};

int refine_cap(Precision p) {
  int iters = 0;
  if (const char* e = blockdaglu_getenv_compat("USERALGO_REFINE_ITERS")) iters = std::max(0, std::atoi(e));
  if (p == Precision::Fp32Ir) {
    int cap = 30;
    if (const char* e = std::getenv("BLOCKDAGLU_GPU_IR_MAX")) cap = std::max(1, std::atoi(e));
    iters = std::max(iters, cap);
  }
  return iters;
}

}  // namespace

class ServeEngine {
 public:
  ServeEngine(int device, int slots, uint64_t budget, Precision p)
      : device_(device), slots_(std::max(1, slots)), budget_(budget), precision_(p), t0_(Clock::now()) {
    for (int w = 0; w < slots_; ++w) workers_.emplace_back([this, w] { worker(w); });
  }
  ~ServeEngine() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }

  int32_t add_pattern(const CooMatrix& m, int32_t dimension) {
    auto pt = std::make_unique<Pattern>();
    std::vector<std::tuple<int, int, double>> coo(m.begin(), m.end());
    std::string msg;
    const bdlu::CanonStatus st = bdlu::canonicalize(coo, dimension, pt->canon, msg);
    if (st != bdlu::CanonStatus::Ok) {
      last_error_ = st == bdlu::CanonStatus::NonFinite ? SolverError::NonFiniteValue : SolverError::InvalidIndex;
      return -1;
    }
    pt->n = pt->canon.n;
    pt->nnz_in = static_cast<int64_t>(m.size());
    // Input entry -> canonical entry: entries sorted by (row, col) (stable), each
    // (row, col) group mapped to the canonical entry with that key if it exists
    // (groups whose registration sum was exactly zero are not in the pattern).
    std::vector<int64_t> ord(m.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
      return std::get<0>(m[a]) != std::get<0>(m[b]) ? std::get<0>(m[a]) < std::get<0>(m[b])
                                                     : std::get<1>(m[a]) < std::get<1>(m[b]);
    });
    pt->slot.assign(m.size(), -1);
    std::size_t ci = 0;
    for (std::size_t k = 0; k < ord.size(); ++k) {
      const auto& e = m[ord[k]];
      while (ci < pt->canon.nnz() && (pt->canon.row[ci] < std::get<0>(e) ||
                                      (pt->canon.row[ci] == std::get<0>(e) && pt->canon.col[ci] < std::get<1>(e))))
        ++ci;
      if (ci < pt->canon.nnz() && pt->canon.row[ci] == std::get<0>(e) && pt->canon.col[ci] == std::get<1>(e))
        pt->slot[ord[k]] = static_cast<int64_t>(ci);
    }
    pt->plan = std::make_unique<bdlu::HostPlan>();
    try {
      const bdlu::PlanStatus ps = bdlu::build_host_plan(pt->canon, *pt->plan);
      pt->singular = !ps.ok;
    } catch (const std::exception&) {
      last_error_ = SolverError::AnalysisFailed;
      return -1;
    }
    std::lock_guard<std::mutex> g(mu_);
    patterns_.push_back(std::move(pt));
    return static_cast<int32_t>(patterns_.size()) - 1;
  }

  int32_t pattern_n(int32_t p) const {
    std::lock_guard<std::mutex> g(mu_);
    return p >= 0 && p < static_cast<int32_t>(patterns_.size()) ? patterns_[p]->n : -1;
  }

  int64_t submit(int32_t pattern, const double* values, int64_t nvals, const double* B, int32_t nrhs) {
    const Pattern* pt = nullptr;
    {
      std::lock_guard<std::mutex> g(mu_);
      if (pattern < 0 || pattern >= static_cast<int32_t>(patterns_.size())) {
        last_error_ = SolverError::InvalidIndex;
        return -1;
      }
      pt = patterns_[pattern].get();
    }
    if (nvals != pt->nnz_in || nrhs < 1) {
      last_error_ = SolverError::InvalidRightHandSide;
      return -1;
    }
    auto r = std::make_unique<Request>();
    r->pattern = pattern;
    r->nrhs = nrhs;
    // Canonical values: duplicates summed in long double (as the canonical form).
    std::vector<long double> acc(pt->canon.nnz(), 0.0L);
    for (int64_t k = 0; k < nvals; ++k)
      if (pt->slot[k] >= 0) acc[pt->slot[k]] += static_cast<long double>(values[k]);
    r->vals.resize(acc.size());
    for (std::size_t k = 0; k < acc.size(); ++k) r->vals[k] = static_cast<double>(acc[k]);
    r->B.assign(B, B + static_cast<int64_t>(pt->n) * nrhs);
    std::lock_guard<std::mutex> g(mu_);
    r->res.t_submit_ms = now_ms();
    const int64_t id = next_id_++;
    queue_.push_back(r.get());
    requests_.emplace(id, std::move(r));
    cv_.notify_one();
    return id;
  }

  // Blocks until request `id` is done; its solutions go to X (raw pointer, n*nrhs)
  // or Xv (resized), whichever is given.
  bool wait(int64_t id, double* X, std::vector<double>* Xv, ServeResult& out) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = requests_.find(id);
    if (it == requests_.end()) {
      last_error_ = SolverError::InvalidIndex;
      return false;
    }
    Request* r = it->second.get();
    done_cv_.wait(lk, [&] { return r->done; });
    out = r->res;
    if (X) std::copy(r->B.begin(), r->B.end(), X);
    if (Xv) *Xv = std::move(r->B);
    requests_.erase(it);
    return out.error == SolverError::None;
  }

  SolverError last_error() const { return last_error_; }

 private:
  double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0_).count(); }

  // Uploads a plan into this slot's context, evicting its least recently used
  // resident plans while the slot's share of the budget is exceeded.
  int upload(bdlu_ctx* ctx, const bdlu::HostPlan& hp, std::unordered_map<int32_t, uint64_t>& resident,
             std::list<int32_t>& lru, uint64_t* id) {
    int rc;
    while ((rc = bdlu_gpu_plan_upload(ctx, &hp.desc, id)) == BDLU_E_BUDGET && !lru.empty()) {
      const int32_t victim = lru.back();
      lru.pop_back();
      bdlu_gpu_plan_release(ctx, resident[victim]);
      resident.erase(victim);
    }
    return rc;
  }

  void run(bdlu_ctx* ctx, std::unordered_map<int32_t, uint64_t>& resident, std::list<int32_t>& lru, Request& r) {
    const Pattern* pt;
    {
      std::lock_guard<std::mutex> g(mu_);
      pt = patterns_[r.pattern].get();
    }
    if (pt->singular) {
      r.res.error = SolverError::FactorizationFailed;
      return;
    }
    uint64_t id = 0;
    auto it = resident.find(r.pattern);
    if (it == resident.end()) {
      if (upload(ctx, *pt->plan, resident, lru, &id) != BDLU_OK) {
        r.res.error = SolverError::BackendUnavailable;
        return;
      }
      resident[r.pattern] = id;
      lru.push_front(r.pattern);
    } else {
      id = it->second;
      lru.remove(r.pattern);
      lru.push_front(r.pattern);
    }
    const int prec = precision_ == Precision::Fp32Ir ? BDLU_FP32 : BDLU_FP64;
    bdlu_factor_stats fs{};
    int rc = bdlu_gpu_refactor(ctx, id, r.vals.data(), BDLU_MEM_HOST, prec, &fs);
    uint64_t use = id;
    bool tmp_plan = false;
    std::unique_ptr<bdlu::HostPlan> tmp;
    if (rc == BDLU_E_REPIVOT) {
      // This request's values fail the fixed pivot order: re-analyse a private
      // copy with its values (the registered plan stays as it is).
      tmp = std::make_unique<bdlu::HostPlan>(*pt->plan);
      bdlu::Canonical c2 = pt->canon;
      c2.val = r.vals;
      bdlu::PlanStatus ps;
      try {
        ps = bdlu::reanalyze_sparse_blocks(c2, *tmp);
      } catch (const std::exception&) {
        ps.ok = false;
      }
      if (!ps.ok) {
        r.res.error = SolverError::FactorizationFailed;
        return;
      }
      if (upload(ctx, *tmp, resident, lru, &use) != BDLU_OK) {
        r.res.error = SolverError::BackendUnavailable;
        return;
      }
      tmp_plan = true;
      r.res.repivots = 1;
      rc = bdlu_gpu_refactor(ctx, use, r.vals.data(), BDLU_MEM_HOST, prec, &fs);
    }
    r.res.fp64_fallback_blocks = fs.fp64_fallback_blocks;
    if (rc == BDLU_OK) {
      bdlu_ir_opts ir{};
      ir.max_iters = refine_cap(precision_);
      ir.stop_rule = ir.max_iters > 0 ? 0 : -1;  // -1: report the residual without refining
      ir.tol = 1e-12;
      bdlu_solve_stats ss{};
      rc = bdlu_gpu_solve(ctx, use, r.B.data(), r.nrhs, pt->n, BDLU_MEM_HOST, &ir, &ss);
      r.res.ir_iterations = ss.iterations;
      r.res.rel_residual = ss.rel_residual;
      r.res.backward_error = ss.backward_error;
      r.res.componentwise_error = ss.componentwise_error;
      if (rc != BDLU_OK) r.res.error = SolverError::SolveFailed;
    } else {
      r.res.error = rc == BDLU_E_BACKEND ? SolverError::BackendUnavailable : SolverError::FactorizationFailed;
    }
    if (tmp_plan) bdlu_gpu_plan_release(ctx, use);
  }

  void worker(int w) {
    bdlu_ctx* ctx = nullptr;
    if (bdlu_gpu_create(device_, budget_ / static_cast<uint64_t>(slots_), &ctx) != BDLU_OK) ctx = nullptr;
    std::unordered_map<int32_t, uint64_t> resident;
    std::list<int32_t> lru;
    for (;;) {
      Request* r = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
        if (queue_.empty()) break;  // stop requested and nothing left
        r = queue_.front();
        queue_.pop_front();
        r->res.t_start_ms = now_ms();
      }
      r->res.slot = w;
      if (ctx) run(ctx, resident, lru, *r);
      else r->res.error = SolverError::BackendUnavailable;
      {
        std::lock_guard<std::mutex> g(mu_);
        r->res.t_done_ms = now_ms();
        r->done = true;
      }
      done_cv_.notify_all();
    }
    if (ctx) {
      for (auto& kv : resident) bdlu_gpu_plan_release(ctx, kv.second);
      bdlu_gpu_destroy(ctx);
    }
  }

  const int device_;
  const int slots_;
  const uint64_t budget_;
  const Precision precision_;
  const Clock::time_point t0_;
  mutable std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::vector<std::unique_ptr<Pattern>> patterns_;
  std::deque<Request*> queue_;
  std::unordered_map<int64_t, std::unique_ptr<Request>> requests_;
  int64_t next_id_ = 0;
  bool stop_ = false;
  SolverError last_error_ = SolverError::None;
  std::vector<std::thread> workers_;
};

Server::Server(int device, int slots, uint64_t hbm_budget_bytes, Precision precision)
    : engine_(new ServeEngine(device, slots, hbm_budget_bytes, precision)) {}
Server::~Server() = default;

int32_t Server::add_pattern(const CooMatrix& m, int32_t dimension) { return engine_->add_pattern(m, dimension); }
int32_t Server::pattern_n(int32_t pattern) const { return engine_->pattern_n(pattern); }
int64_t Server::submit(int32_t pattern, const std::vector<double>& values, const std::vector<double>& B,
                       int32_t nrhs) {
  return engine_->submit(pattern, values.data(), static_cast<int64_t>(values.size()), B.data(), nrhs);
}
bool Server::wait(int64_t request, std::vector<double>& X, ServeResult& result) {
  return engine_->wait(request, nullptr, &X, result);
}
SolverError Server::last_error() const { return engine_->last_error(); }

// C-ABI helpers (capi.cpp) go straight to the engine so they avoid copies.
ServeEngine* serve_engine_create(int device, int slots, uint64_t budget, int precision) {
  return new ServeEngine(device, slots, budget, precision == 1 ? Precision::Fp32Ir : Precision::Fp64);
}
void serve_engine_destroy(ServeEngine* e) { delete e; }
int32_t serve_engine_add_pattern(ServeEngine* e, const CooMatrix& m, int32_t dim) { return e->add_pattern(m, dim); }
int32_t serve_engine_pattern_n(ServeEngine* e, int32_t p) { return e->pattern_n(p); }
int64_t serve_engine_submit(ServeEngine* e, int32_t p, const double* v, int64_t nv, const double* B, int32_t k) {
  return e->submit(p, v, nv, B, k);
}
bool serve_engine_wait(ServeEngine* e, int64_t id, double* X, ServeResult& res) { return e->wait(id, X, nullptr, res); }
SolverError serve_engine_last_error(ServeEngine* e) { return e->last_error(); }

}  // namespace blockdaglu
