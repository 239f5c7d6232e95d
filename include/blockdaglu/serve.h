#pragma once
// Serving loop for the BLOCK engine (additive API; the reference has no
// counterpart). Sparsity patterns are registered once: their host analysis
// (canonical order, UBTF decomposition, per-block symbolic analysis) is kept.
// Requests (new values for a registered pattern plus k right-hand sides) are
// queued in arrival order and run on slots. Every slot owns an engine context
// and HIP stream, so requests that arrive while others run overlap on the GPU
// (continuous batching). A slot keeps the device plans (symbolic arrays plus
// its own factor storage) of the patterns it served resident in HBM; all
// resident plans of all slots together stay within one HBM budget. A request
// is admitted when a slot is free and its plan is resident or fits: idle plans
// are evicted least recently used first, in any slot, and a request waits while
// the memory it needs is held by requests in flight. A free slot prefers the
// oldest queued request whose plan it already holds (among the next few in
// line), so a pattern is not copied into every slot. A request's result does
// not depend on what else is in flight: each runs on its own buffers with
// deterministic kernels.
#include <cstdint>
#include <memory>
#include <vector>

#include "blockdaglu/solver.h"

namespace blockdaglu {

struct ServeResult {
  SolverError error = SolverError::None;
  int32_t slot = -1;
  int32_t ir_iterations = 0;
  int32_t fp64_fallback_blocks = 0;
  int32_t repivots = 0;
  int32_t ir_fp64_refactor = 0;  // FP32 refinement did not converge; refactored in FP64
  double rel_residual = -1.0;    // max over the request's right-hand sides
  double backward_error = -1.0;  // max over the request's right-hand sides
  double componentwise_error = -1.0;  // max over the request's right-hand sides
  double t_submit_ms = 0.0;      // server clock (ms since the server started)
  double t_start_ms = 0.0;       // admitted onto a slot
  double t_done_ms = 0.0;
};

class ServeEngine;

class Server {
 public:
  // slots: concurrent requests on the device. hbm_budget_bytes: resident plans
  // of all slots together (0: no limit). precision: Fp64 (plain) or Fp32Ir.
  Server(int device, int slots, uint64_t hbm_budget_bytes, Precision precision);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // Registers a pattern (entries in any order, duplicates allowed; the values
  // choose the first pivots of blocks larger than 64). Returns the pattern id
  // or -1 (see last_error()).
  int32_t add_pattern(const CooMatrix& m, int32_t dimension = -1);
  int32_t pattern_n(int32_t pattern) const;

  // Queues a request: values in the registration's entry order, B column-major
  // n x nrhs. Returns a request id (>= 0) or -1. A request whose values are
  // nonzero where the registered pattern has no entry (a duplicate group that
  // summed to zero at registration, or an entry the 1e-20 drop rule removed) is
  // a different sparsity pattern and is rejected with AnalysisFailed.
  int64_t submit(int32_t pattern, const std::vector<double>& values, const std::vector<double>& B, int32_t nrhs);
  // The same with inputs already in device memory (HBM): d_values in canonical
  // order (sorted by (row, col), duplicates summed: the registered pattern's
  // canonical entries), d_B the right-hand sides and d_X the solutions, n x nrhs
  // column-major. The caller keeps the three buffers valid until wait() returns;
  // wait()'s X is left empty.
  int64_t submit_device(int32_t pattern, const double* d_values, const double* d_B, double* d_X, int32_t nrhs);
  // Blocks until the request is done; X gets n x nrhs (column-major).
  bool wait(int64_t request, std::vector<double>& X, ServeResult& result);

  SolverError last_error() const;

 private:
  std::unique_ptr<ServeEngine> engine_;
};

}  // namespace blockdaglu
