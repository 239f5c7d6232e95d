// Serving loop: registered patterns, device plans resident per slot under one
// HBM budget for all slots, near-FIFO admission of requests onto slots (one
// engine context and HIP stream per slot, so requests in different slots
// overlap on the device). No counterpart in the reference; each request runs
// the same refactor + solve (+ refinement) as Solver does.
#include "blockdaglu/serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <deque>
#include <map>
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
  // Input entry -> canonical entry, or -1 when its (row, col) group summed to
  // exactly zero at registration (then zgroup names the group).
  std::vector<int64_t> slot;
  std::vector<int32_t> zgroup;
  int32_t nzgroups = 0;
  // Canonical entries the plan left out under the drop rule (|v| < drop_abs_lt).
  std::vector<uint8_t> dropped;
  double drop_abs_lt = 0.0;
  bdlu::Canonical canon;
  std::unique_ptr<bdlu::HostPlan> plan;
  bool singular = false;
  uint64_t bytes = 0;  // device bytes of one resident plan, known after its first upload
};

struct Request {
  int32_t pattern = 0;
  std::vector<double> vals;  // canonical order
  std::vector<double> B;     // right-hand sides in, solutions out (column-major n x nrhs)
  // Device-resident request (submit_device): canonical values, right-hand
  // sides and the solution buffer in device memory, owned by the caller.
  const double* d_vals = nullptr;
  const double* d_B = nullptr;
  double* d_X = nullptr;
  int32_t nrhs = 1;
  ServeResult res;
  bool done = false;
};

// One pattern's plan on one slot.
struct Resident {
  uint64_t id = 0;
  uint64_t bytes = 0;
  bool in_use = false;
  uint64_t last_use = 0;
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
    // Contexts are created here, one after another, before any worker starts.
    ctx_.assign(slots_, nullptr);
    // One stream per slot unless BLOCKDAGLU_GPU_STREAMS asks otherwise: the
    // slots already overlap requests, and side streams of concurrent contexts
    // would share the hardware queues (measured: 15-20% more requests per
    // second without them).
    // The dense-block solve's persistent grid is capped at 64 workgroups per
    // slot unless BLOCKDAGLU_GPU_SEG_GRID_MAX says otherwise (measured with 4
    // slots: +10% requests/s with FP32 factors and refinement, +8% FP64 with
    // refinement, against the 256-1024 a context alone picks).
    // Refinement stop tests stay on the host unless BLOCKDAGLU_GPU_IR_DEVICE
    // is set: while one slot waits on its test, the others keep the GPU busy,
    // and the device form's steps past the stop cost GPU time (measured with
    // 4 slots: about 1% more requests/s with host tests, FP32+IR and FP64+IR).
    const bool side = std::getenv("BLOCKDAGLU_GPU_STREAMS") != nullptr;
    const bool grid_env = std::getenv("BLOCKDAGLU_GPU_SEG_GRID_MAX") != nullptr;
    const bool ir_env = std::getenv("BLOCKDAGLU_GPU_IR_DEVICE") != nullptr;
    for (int w = 0; w < slots_; ++w) {
      if (bdlu_gpu_create(device_, 0, &ctx_[w]) != BDLU_OK) {
        ctx_[w] = nullptr;
        continue;
      }
      if (!side) bdlu_gpu_set_side_streams(ctx_[w], 0);
      if (!grid_env) bdlu_gpu_set_solve_grid_max(ctx_[w], 64);
      if (!ir_env) bdlu_gpu_set_device_refinement(ctx_[w], 0);
    }
    resident_.resize(slots_);
    for (int w = 0; w < slots_; ++w) workers_.emplace_back([this, w] { worker(w); });
  }
  ~ServeEngine() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
    for (int w = 0; w < slots_; ++w) {
      if (!ctx_[w]) continue;
      for (auto& kv : resident_[w]) bdlu_gpu_plan_release(ctx_[w], kv.second.id);
      bdlu_gpu_destroy(ctx_[w]);
    }
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
    // (row, col) group mapped to the canonical entry with that key if it exists;
    // groups whose registration sum was exactly zero get a group id instead.
    std::vector<int64_t> ord(m.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
      return std::get<0>(m[a]) != std::get<0>(m[b]) ? std::get<0>(m[a]) < std::get<0>(m[b])
                                                     : std::get<1>(m[a]) < std::get<1>(m[b]);
    });
    pt->slot.assign(m.size(), -1);
    pt->zgroup.assign(m.size(), -1);
    std::size_t ci = 0;
    for (std::size_t k = 0; k < ord.size(); ++k) {
      const auto& e = m[ord[k]];
      while (ci < pt->canon.nnz() && (pt->canon.row[ci] < std::get<0>(e) ||
                                      (pt->canon.row[ci] == std::get<0>(e) && pt->canon.col[ci] < std::get<1>(e))))
        ++ci;
      if (ci < pt->canon.nnz() && pt->canon.row[ci] == std::get<0>(e) && pt->canon.col[ci] == std::get<1>(e)) {
        pt->slot[ord[k]] = static_cast<int64_t>(ci);
      } else {
        const bool same_group = k > 0 && pt->slot[ord[k - 1]] < 0 && std::get<0>(m[ord[k - 1]]) == std::get<0>(e) &&
                                std::get<1>(m[ord[k - 1]]) == std::get<1>(e);
        pt->zgroup[ord[k]] = same_group ? pt->zgroup[ord[k - 1]] : pt->nzgroups++;
      }
    }
    pt->plan = std::make_unique<bdlu::HostPlan>();
    try {
      const bdlu::PlanStatus ps = bdlu::build_host_plan(pt->canon, *pt->plan);
      pt->singular = !ps.ok;
    } catch (const std::exception&) {
      last_error_ = SolverError::AnalysisFailed;
      return -1;
    }
    pt->drop_abs_lt = pt->plan->dec.drop_abs_lt;
    pt->dropped.assign(pt->canon.nnz(), 0);
    for (std::size_t e = 0; e < pt->plan->vmap.size() && e < pt->dropped.size(); ++e)
      pt->dropped[e] = pt->plan->vmap[e] < 0 ? 1 : 0;
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
    for (int64_t k = 0; k < nvals; ++k)
      if (!std::isfinite(values[k])) {
        last_error_ = SolverError::NonFiniteValue;
        return -1;
      }
    auto r = std::make_unique<Request>();
    r->pattern = pattern;
    r->nrhs = nrhs;
    // Canonical values: duplicates summed in long double (as the canonical form).
    std::vector<long double> acc(pt->canon.nnz(), 0.0L);
    std::vector<long double> zacc(static_cast<std::size_t>(pt->nzgroups), 0.0L);
    for (int64_t k = 0; k < nvals; ++k) {
      if (pt->slot[k] >= 0) acc[pt->slot[k]] += static_cast<long double>(values[k]);
      else zacc[pt->zgroup[k]] += static_cast<long double>(values[k]);
    }
    // Values where the registered pattern has no entry make another pattern.
    bool outside = false;
    for (long double z : zacc) outside = outside || z != 0.0L;
    r->vals.resize(acc.size());
    for (std::size_t k = 0; k < acc.size(); ++k) {
      r->vals[k] = static_cast<double>(acc[k]);
      if (pt->dropped[k] && !(std::abs(r->vals[k]) < pt->drop_abs_lt)) outside = true;
    }
    if (outside) {
      last_error_ = SolverError::AnalysisFailed;
      return -1;
    }
    r->B.assign(B, B + static_cast<int64_t>(pt->n) * nrhs);
    for (double v : r->B)
      if (!std::isfinite(v)) {
        last_error_ = SolverError::NonFiniteValue;
        return -1;
      }
    std::lock_guard<std::mutex> g(mu_);
    r->res.t_submit_ms = now_ms();
    const int64_t id = next_id_++;
    queue_.push_back(r.get());
    requests_.emplace(id, std::move(r));
    cv_.notify_one();
    return id;
  }

  // A request whose inputs are already in device memory: d_vals in canonical
  // order (canon.nnz values), d_B and d_X n x nrhs column-major; the solution
  // is written to d_X. The buffers must stay valid until wait() returns.
  int64_t submit_device(int32_t pattern, const double* d_vals, const double* d_B, double* d_X, int32_t nrhs) {
    {
      std::lock_guard<std::mutex> g(mu_);
      if (pattern < 0 || pattern >= static_cast<int32_t>(patterns_.size())) {
        last_error_ = SolverError::InvalidIndex;
        return -1;
      }
    }
    if (!d_vals || !d_B || !d_X || nrhs < 1) {
      last_error_ = SolverError::InvalidRightHandSide;
      return -1;
    }
    auto r = std::make_unique<Request>();
    r->pattern = pattern;
    r->nrhs = nrhs;
    r->d_vals = d_vals;
    r->d_B = d_B;
    r->d_X = d_X;
    std::lock_guard<std::mutex> g(mu_);
    r->res.t_submit_ms = now_ms();
    const int64_t id = next_id_++;
    queue_.push_back(r.get());
    requests_.emplace(id, std::move(r));
    cv_.notify_one();
    return id;
  }

  // Blocks until request `id` is done; its solutions go to X (raw pointer, n*nrhs)
  // or Xv (resized), whichever is given (device requests: already in d_X).
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
    if (!r->d_X) {
      if (X) std::copy(r->B.begin(), r->B.end(), X);
      if (Xv) *Xv = std::move(r->B);
    }
    requests_.erase(it);
    return out.error == SolverError::None;
  }

  SolverError last_error() const { return last_error_; }

 private:
  double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0_).count(); }

  // Picks the least recently used idle plan of any slot, removes it from the
  // books and returns it (slot, id); false if no plan is idle. Caller holds mu_.
  bool take_victim(int* slot, uint64_t* id) {
    int vs = -1;
    int32_t vp = -1;
    uint64_t oldest = ~uint64_t(0);
    for (int w = 0; w < slots_; ++w)
      for (auto& kv : resident_[w])
        if (!kv.second.in_use && kv.second.last_use < oldest) {
          oldest = kv.second.last_use;
          vs = w;
          vp = kv.first;
        }
    if (vs < 0) return false;
    *slot = vs;
    *id = resident_[vs][vp].id;
    used_ -= resident_[vs][vp].bytes;
    resident_[vs].erase(vp);
    return true;
  }

  bool any_in_use() const {
    for (const auto& m : resident_)
      for (const auto& kv : m)
        if (kv.second.in_use) return true;
    return false;
  }

  // Releases a plan taken by take_victim (without mu_: it waits for that slot's
  // context, which may be finishing another request).
  void release_plan(int slot, uint64_t id) {
    bdlu_gpu_plan_release(ctx_[slot], id);
    std::lock_guard<std::mutex> g(mu_);
    mem_cv_.notify_all();
  }

  // The resident plan of pattern p on slot w, uploaded if needed. While the
  // budget cannot take it, idle plans (any slot, least recently used first) are
  // evicted, and if none is idle the request waits for requests in flight to
  // finish. Returns a SolverError; on success the plan is marked in use.
  SolverError acquire(int w, int32_t p, uint64_t* id) {
    bdlu_ctx* ctx = ctx_[w];
    std::unique_lock<std::mutex> lk(mu_);
    auto it = resident_[w].find(p);
    if (it != resident_[w].end()) {
      it->second.in_use = true;
      it->second.last_use = ++tick_;
      *id = it->second.id;
      return SolverError::None;
    }
    Pattern& pt = *patterns_[p];
    for (;;) {
      const uint64_t need = pt.bytes;  // 0 until the pattern's first upload measured it
      if (budget_ == 0 || used_ + need <= budget_) break;
      int vs;
      uint64_t vid;
      if (take_victim(&vs, &vid)) {
        lk.unlock();
        release_plan(vs, vid);
        lk.lock();
        continue;
      }
      if (!any_in_use()) return SolverError::BackendUnavailable;  // larger than the whole budget
      mem_cv_.wait(lk);
    }
    used_ += pt.bytes;
    const uint64_t reserved = pt.bytes;
    lk.unlock();
    uint64_t bytes = 0;
    int rc = bdlu_gpu_plan_upload(ctx, &pt.plan->desc, id);
    if (rc == BDLU_OK) rc = bdlu_gpu_plan_bytes(ctx, *id, &bytes);
    lk.lock();
    used_ -= reserved;
    if (rc != BDLU_OK) {
      mem_cv_.notify_all();
      return SolverError::BackendUnavailable;
    }
    pt.bytes = bytes;
    if (budget_ != 0 && bytes > budget_) {
      // Larger than the whole budget: this pattern cannot be served here.
      lk.unlock();
      bdlu_gpu_plan_release(ctx, *id);
      lk.lock();
      mem_cv_.notify_all();
      return SolverError::BackendUnavailable;
    }
    used_ += bytes;
    resident_[w][p] = Resident{*id, bytes, true, ++tick_};
    // A pattern's first upload is measured only after the fact: make room now.
    while (budget_ != 0 && used_ > budget_) {
      int vs;
      uint64_t vid;
      if (!take_victim(&vs, &vid)) break;
      lk.unlock();
      release_plan(vs, vid);
      lk.lock();
    }
    return SolverError::None;
  }

  void unuse(int w, int32_t p) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = resident_[w].find(p);
    if (it != resident_[w].end()) it->second.in_use = false;
    mem_cv_.notify_all();
  }

  void run(int w, Request& r) {
    bdlu_ctx* ctx = ctx_[w];
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
    const SolverError ae = acquire(w, r.pattern, &id);
    {
      std::lock_guard<std::mutex> g(mu_);
      r.res.t_start_ms = now_ms();
    }
    if (ae != SolverError::None) {
      r.res.error = ae;
      return;
    }
    const int prec = precision_ == Precision::Fp32Ir ? BDLU_FP32 : BDLU_FP64;
    bdlu_factor_stats fs{};
    const bool dev = r.d_vals != nullptr;
    int rc = dev ? bdlu_gpu_refactor(ctx, id, r.d_vals, BDLU_MEM_DEVICE, prec, &fs)
                 : bdlu_gpu_refactor(ctx, id, r.vals.data(), BDLU_MEM_HOST, prec, &fs);
    uint64_t use = id;
    uint64_t tmp_bytes = 0;
    if (rc == BDLU_E_REPIVOT && dev) {
      // The re-analysis runs on the host: fetch this request's values.
      r.vals.resize(pt->canon.nnz());
      if (bdlu_gpu_plan_values(ctx, id, r.vals.data()) != BDLU_OK) {
        unuse(w, r.pattern);
        r.res.error = SolverError::BackendUnavailable;
        return;
      }
    }
    if (rc == BDLU_E_REPIVOT) {
      // This request's values fail the fixed pivot order: re-analyse a private
      // copy with its values (the registered plan stays as it is).
      auto tmp = std::make_unique<bdlu::HostPlan>(*pt->plan);
      bdlu::Canonical c2 = pt->canon;
      c2.val = r.vals;
      bdlu::PlanStatus ps;
      try {
        ps = bdlu::reanalyze_sparse_blocks(c2, *tmp);
      } catch (const std::exception&) {
        ps.ok = false;
      }
      if (!ps.ok) {
        unuse(w, r.pattern);
        r.res.error = SolverError::FactorizationFailed;
        return;
      }
      if (bdlu_gpu_plan_upload(ctx, &tmp->desc, &use) != BDLU_OK) {
        unuse(w, r.pattern);
        r.res.error = SolverError::BackendUnavailable;
        return;
      }
      bdlu_gpu_plan_bytes(ctx, use, &tmp_bytes);
      {
        std::lock_guard<std::mutex> g(mu_);
        used_ += tmp_bytes;
      }
      r.res.repivots = 1;
      rc = bdlu_gpu_refactor(ctx, use, r.vals.data(), BDLU_MEM_HOST, prec, &fs);
    }
    r.res.fp64_fallback_blocks = fs.fp64_fallback_blocks;
    if (rc == BDLU_OK) {
      bdlu_ir_opts ir{};
      ir.max_iters = refine_cap(precision_);
      ir.stop_rule = ir.max_iters > 0 ? 0 : -1;  // -1: report the residual without refining
      ir.tol = 1e-12;
      ir.fp64_fallback = precision_ == Precision::Fp32Ir ? 1 : 0;
      bdlu_solve_stats ss{};
      if (dev) {
        rc = bdlu_gpu_copy(ctx, r.d_X, r.d_B, sizeof(double) * static_cast<uint64_t>(pt->n) * r.nrhs);
        if (rc == BDLU_OK) rc = bdlu_gpu_solve(ctx, use, r.d_X, r.nrhs, pt->n, BDLU_MEM_DEVICE, &ir, &ss);
      } else {
        rc = bdlu_gpu_solve(ctx, use, r.B.data(), r.nrhs, pt->n, BDLU_MEM_HOST, &ir, &ss);
      }
      r.res.ir_iterations = ss.iterations;
      r.res.ir_fp64_refactor = ss.fp64_refactor;
      r.res.rel_residual = ss.rel_residual;
      r.res.backward_error = ss.backward_error;
      r.res.componentwise_error = ss.componentwise_error;
      if (rc != BDLU_OK) r.res.error = SolverError::SolveFailed;
    } else {
      r.res.error = rc == BDLU_E_BACKEND ? SolverError::BackendUnavailable : SolverError::FactorizationFailed;
    }
    if (use != id) {
      bdlu_gpu_plan_release(ctx, use);
      std::lock_guard<std::mutex> g(mu_);
      used_ -= tmp_bytes;
    }
    unuse(w, r.pattern);
  }

  void worker(int w) {
    for (;;) {
      Request* r = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
        if (queue_.empty()) break;  // stop requested and nothing left
        // Oldest request among the next few whose plan this slot holds, else the oldest.
        std::size_t pick = 0;
        const std::size_t window = std::min<std::size_t>(queue_.size(), 2 * static_cast<std::size_t>(slots_));
        for (std::size_t i = 0; i < window; ++i)
          if (resident_[w].count(queue_[i]->pattern)) {
            pick = i;
            break;
          }
        r = queue_[pick];
        queue_.erase(queue_.begin() + static_cast<std::ptrdiff_t>(pick));
      }
      r->res.slot = w;
      if (ctx_[w]) {
        run(w, *r);
      } else {
        r->res.error = SolverError::BackendUnavailable;
      }
      {
        std::lock_guard<std::mutex> g(mu_);
        r->res.t_done_ms = now_ms();
        r->done = true;
      }
      done_cv_.notify_all();
    }
  }

  const int device_;
  const int slots_;
  const uint64_t budget_;
  const Precision precision_;
  const Clock::time_point t0_;
  mutable std::mutex mu_;
  std::condition_variable cv_, done_cv_, mem_cv_;
  std::vector<bdlu_ctx*> ctx_;
  std::vector<std::map<int32_t, Resident>> resident_;  // per slot: pattern -> plan
  uint64_t used_ = 0;  // device bytes of all resident plans (plus private re-analysed ones in flight)
  uint64_t tick_ = 0;
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
int64_t Server::submit_device(int32_t pattern, const double* d_values, const double* d_B, double* d_X,
                              int32_t nrhs) {
  return engine_->submit_device(pattern, d_values, d_B, d_X, nrhs);
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
int64_t serve_engine_submit_device(ServeEngine* e, int32_t p, const double* d_v, const double* d_B, double* d_X,
                                   int32_t k) {
  return e->submit_device(p, d_v, d_B, d_X, k);
}
bool serve_engine_wait(ServeEngine* e, int64_t id, double* X, ServeResult& res) { return e->wait(id, X, nullptr, res); }
SolverError serve_engine_last_error(ServeEngine* e) { return e->last_error(); }

}  // namespace blockdaglu
