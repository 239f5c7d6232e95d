#pragma once
// blockdaglu::Solver — the reference's public API, unchanged
// (blockdaglu/include/blockdaglu/solver.h:19-126), backed by the MI355X engine:
//   set_matrix      storage only (reference src/solver.cpp:745-800)
//   analyze         host: validate/canonicalize, signatures, DM/BTF, merge,
//                   per-block ordering and symbolic analysis, device plan upload
//   factor          device: value scatter + numeric refactorization of every
//                   diagonal block (HIP kernels for gfx950)
//   solve_in_place  device: block-DAG solve with fused off-diagonal updates,
//                   optional FP64 refinement (reference src/solver.cpp:1449-1536)
//
// Additions below the reference API are marked "additive"; they have no
// counterpart in the reference and never change the meaning of the calls above.

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "blockdaglu/types.h"

struct bdlu_factor_stats;  // include/bdlu_gpu.h

namespace blockdaglu {

// Same enumerators, same order as the reference (solver.h:19-31).
enum class SolverError {
  None,
  NotPrepared,
  InvalidDimension,
  InvalidIndex,
  NonFiniteValue,
  InvalidRightHandSide,
  BackendUnavailable,
  AnalysisFailed,
  FactorizationFailed,
  SolveFailed,
  ExportFailed,
};

// Additive: arithmetic used by factor()/solve_in_place().
//   Fp64   — plain FP64 factors and solves (default); USERALGO_REFINE_ITERS
//            steps of FP64 refinement as in the reference.
//   Fp32Ir — FP32 factors; FP64 residuals and FP64 correction accumulation
//            until ||b - A x||_2 / ||b||_2 < 1e-12 (BLOCKDAGLU_GPU_IR_STOP
//            selects the normwise or componentwise backward error instead),
//            at most BLOCKDAGLU_GPU_IR_MAX (30) steps. If the stop rule is not
//            met, or a step fails to halve the stop measure, the matrix is
//            refactored in FP64 and solved again (report().ir_fp64_refactor).
enum class Precision { Fp64 = 0, Fp32Ir = 1 };

// Additive: what the last factor()/solve_in_place() did.
struct EngineReport {
  int32_t nblocks = 0;
  int32_t dense_blocks = 0;
  int32_t sparse_blocks = 0;
  int32_t max_block = 0;
  int32_t fp64_fallback_blocks = 0;  // FP32 blocks refactored in FP64 after a pivot check
  int32_t repivot_blocks = 0;        // blocks re-analysed because fixed pivots failed the check
  int32_t ir_iterations = 0;
  bool ir_fp64_refactor = false;     // FP32 refinement did not converge; refactored in FP64
  double rel_residual = -1.0;        // ||b - A x||_2 / ||b||_2 (FP64, on device)
  double backward_error = -1.0;      // ||r||_inf / (||A||_inf ||x||_inf + ||b||_inf)
  double componentwise_error = -1.0; // max_i |r_i| / (|A| |x| + |b|)_i
  bool plan_reused = false;          // analyze() found a resident plan for this pattern
};

struct SolverImpl;

class Solver final {
 public:
  Solver();
  ~Solver();
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;

  bool set_matrix(CooMatrix m);
  bool set_matrix(CooMatrix m, int32_t dimension);
  bool analyze();
  bool factor();
  bool solve_in_place(std::vector<double>& b);

  SolverError last_error() const { return last_error_; }
  const std::string& last_error_message() const { return last_error_message_; }

  CooMatrix get_rearranged_matrix();
  CooMatrix getQ();
  CooMatrix getP();
  CooMatrix getL();
  CooMatrix getU();

  int32_t n() const { return n_; }
  // "KLU" when the last analyze() selected the global-KLU route, else "BLOCK".
  const char* selected_mode() const;

  // ---- additive ----------------------------------------------------------
  // k right-hand sides, column-major with leading dimension n (B.size() == n*k).
  bool solve_multi_in_place(std::vector<double>& B, int32_t nrhs);
  void set_precision(Precision p) { precision_ = p; }
  Precision precision() const { return precision_; }
  const EngineReport& report() const { return report_; }
  // Values in canonical order (sorted by (row, col), duplicates summed) and
  // right-hand sides already resident in device memory (HBM); same semantics
  // as factor() / solve_multi_in_place().
  bool factor_device(const double* d_values);
  bool solve_device(double* d_B, int32_t nrhs, int64_t ldb);
  // The engine context (bdlu_ctx*) and its HIP stream, for callers that order
  // or time device work around these calls.
  void* device_context() const;
  void* device_stream() const;
  // Id of the resident device plan of the last analyze() in that context (0 if none).
  uint64_t device_plan() const;
  // Structural permutations of the BLOCK decomposition (perm index -> original).
  const std::vector<int32_t>& block_row_perm() const;
  const std::vector<int32_t>& block_col_perm() const;
  const std::vector<int32_t>& block_bounds() const;
  // Column order and pivot rows (local indices) of diagonal block `block` if it
  // is above the dense threshold: q[k] = column at position k, prow[k] = row
  // pivoted there (the sparse part's first, then the dense tail's). Returns the
  // number of sparse leading columns k0, or -1 if the block is dense or unknown.
  int32_t sparse_block_order(int32_t block, std::vector<int32_t>& q, std::vector<int32_t>& prow) const;

 private:
  bool clear_error_ok();
  bool validate();
  bool ensure_export();
  bool analyze_klu();
  int refactor_plan(const double* values, bool on_device, int precision, bdlu_factor_stats& stats);
  bool factor_values(const double* values, bool on_device);
  bool fail(SolverError error, std::string message);

  int32_t n_ = 0;
  Precision precision_ = Precision::Fp64;
  EngineReport report_{};
  SolverError last_error_ = SolverError::None;
  std::string last_error_message_;
  std::unique_ptr<SolverImpl> impl_;
};

}  // namespace blockdaglu
