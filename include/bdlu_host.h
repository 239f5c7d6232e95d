#pragma once
// C-ABI of the host side (libblockdaglu.so): the blockdaglu::Solver
// pipeline over plain pointers, read-only queries of the host analysis, and the
// seeded workload generator. All functions are extern "C", never throw, and
// report failure through return values.
//
// Reference interfaces replaced or mirrored:
//   bdlu_solver_*   blockdaglu::Solver (blockdaglu/include/blockdaglu/solver.h:43-65);
//                   the reference exposes it to the organizer tool through
//                   CreateAlgorithm/DestroyAlgorithm (src/exports.cpp:14-24) and
//                   the ALInterface stage mapping (src/UserAlgorithmWrapper.cpp:8-22).
//   bdlu_host_*     build_decomposition_btf + per-block KLU symbolic analysis
//                   (src/ubtf_block_solver.h:131-136, ubtf_block_solver.cpp:1749-1825),
//                   as read-only queries so callers can check permutations.
//   bdlu_gen_*      no reference counterpart: the seeded synthetic inputs the
//                   benchmarks and tests run on.

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// ---- Solver --------------------------------------------------------------------
// Error codes equal the blockdaglu::SolverError enumerators (solver.h:19-31).
enum {
  BDLU_ERR_NONE = 0,
  BDLU_ERR_NOT_PREPARED = 1,
  BDLU_ERR_INVALID_DIMENSION = 2,
  BDLU_ERR_INVALID_INDEX = 3,
  BDLU_ERR_NON_FINITE = 4,
  BDLU_ERR_INVALID_RHS = 5,
  BDLU_ERR_BACKEND_UNAVAILABLE = 6,
  BDLU_ERR_ANALYSIS_FAILED = 7,
  BDLU_ERR_FACTORIZATION_FAILED = 8,
  BDLU_ERR_SOLVE_FAILED = 9,
  BDLU_ERR_EXPORT_FAILED = 10,
};

typedef struct bdlu_report {
  int32_t nblocks, dense_blocks, sparse_blocks, max_block;
  int32_t fp64_fallback_blocks, repivot_blocks, ir_iterations;
  int32_t plan_reused;
  double rel_residual, backward_error, componentwise_error;
  int32_t ir_fp64_refactor;  // 1: FP32 refinement did not converge; refactored in FP64 and solved again
} bdlu_report;

void* bdlu_solver_create(void);
void bdlu_solver_destroy(void* s);
// Returns 1 on success, 0 on failure (then see last_error / last_error_message).
int bdlu_solver_set_matrix(void* s, int64_t nnz, const int32_t* rows, const int32_t* cols, const double* vals,
                           int32_t dimension /* -1: infer */);
int bdlu_solver_analyze(void* s);
int bdlu_solver_factor(void* s);
int bdlu_solver_solve(void* s, double* b, int64_t len);
int bdlu_solver_solve_multi(void* s, double* B, int64_t n, int32_t nrhs);
int bdlu_solver_last_error(void* s);
const char* bdlu_solver_last_error_message(void* s);
int32_t bdlu_solver_n(void* s);
const char* bdlu_solver_selected_mode(void* s);
void bdlu_solver_set_precision(void* s, int precision /* 0 FP64, 1 FP32+IR */);
void bdlu_solver_report(void* s, bdlu_report* out);
// Device-resident variants (additive): values in canonical order and B in device
// memory (HBM); same return convention.
int bdlu_solver_factor_device(void* s, const double* d_values);
int bdlu_solver_solve_device(void* s, double* d_B, int32_t nrhs, int64_t ldb);
// Engine context (bdlu_ctx*, see bdlu_gpu.h), its HIP stream, and the id of the
// resident device plan of the last analyze() (0 if none), for callers of the
// device C-ABI.
void* bdlu_solver_device_context(void* s);
uint64_t bdlu_solver_device_plan(void* s);
void* bdlu_solver_device_stream(void* s);
// Structural permutations of the last analyze(); arrays of n (n+1 / nblocks+1 for R).
int32_t bdlu_solver_nblocks(void* s);
void bdlu_solver_copy_perm(void* s, int32_t* P, int32_t* Q, int32_t* R);
// Column order and pivot rows (size arrays of the block size) of block b of the
// last analyze() when it is above the dense threshold; returns the number of
// sparse leading columns k0 (the rest is the dense tail), or -1.
int32_t bdlu_solver_sparse_order(void* s, int32_t b, int32_t* q, int32_t* prow);
// Exports (reference Solver::getP/getQ/getL/getU/get_rearranged_matrix). Two-call
// protocol: pass nullptr arrays to get the entry count, then arrays of that size.
// which: 0 rearranged, 1 P, 2 Q, 3 L, 4 U. Returns entry count or -1 on failure.
int64_t bdlu_solver_export(void* s, int which, int32_t* rows, int32_t* cols, double* vals);

// ---- serving loop (blockdaglu/serve.h) -----------------------------------------
// No counterpart in the reference: for callers that serve many refactor+solve
// requests on registered patterns (Newton-like loops, circuit simulators). Slots
// run concurrently on the device, each with its own stream and resident plans;
// requests are admitted near arrival order as slots free up, and the resident
// plans of all slots together stay within one HBM budget (blockdaglu/serve.h).
typedef struct bdlu_serve_result {
  int32_t error;                 // SolverError value (0 = success)
  int32_t slot;
  int32_t ir_iterations;
  int32_t fp64_fallback_blocks;
  int32_t repivots;
  double rel_residual;           // worst right-hand side of the request
  double backward_error;         // worst right-hand side of the request
  double componentwise_error;    // worst right-hand side of the request
  double t_submit_ms, t_start_ms, t_done_ms;  // server clock
  int32_t ir_fp64_refactor;      // 1: FP32 refinement did not converge; refactored in FP64
} bdlu_serve_result;
void* bdlu_serve_create(int device, int slots, uint64_t hbm_budget_bytes, int precision /* 0 FP64, 1 FP32+IR */);
void bdlu_serve_destroy(void* s);
// Returns the pattern id (>= 0) or -(SolverError).
int32_t bdlu_serve_add_pattern(void* s, int64_t nnz, const int32_t* rows, const int32_t* cols, const double* vals,
                               int32_t dimension);
int32_t bdlu_serve_pattern_n(void* s, int32_t pattern);
// values: nnz entries in the registration's entry order; B: n x nrhs column-major.
// Returns a request id (>= 0) or -(SolverError).
int64_t bdlu_serve_submit(void* s, int32_t pattern, const double* vals, int64_t nnz, const double* B, int32_t nrhs);
// The same with inputs already in device memory: d_vals the registered
// pattern's canonical values (sorted by (row, col), duplicates summed), d_B the
// right-hand sides and d_X the solutions (n x nrhs column-major, device
// memory, kept valid until bdlu_serve_wait returns; its X may be NULL).
int64_t bdlu_serve_submit_device(void* s, int32_t pattern, const double* d_vals, const double* d_B, double* d_X,
                                 int32_t nrhs);
// Blocks until done; X (n x nrhs) receives the solutions. Returns a SolverError value.
int bdlu_serve_wait(void* s, int64_t request, double* X, bdlu_serve_result* out);

// ---- host analysis queries (CPU only) ---------------------------------------
void* bdlu_host_analyze(int64_t nnz, const int32_t* rows, const int32_t* cols, const double* vals,
                        int32_t dimension, int* error_code);
// The same for the global-KLU route with KLU's settings (btf 0/1, scale
// <= 0 / 1 / 2, pivot tolerance): KLU's BTF or one block, no merging, blocks of
// 2..64 in their COLAMD order (baked into the permutations), whole-row
// scaling (src/solver.cpp:878-968).
void* bdlu_host_analyze_klu(int64_t nnz, const int32_t* rows, const int32_t* cols, const double* vals,
                            int32_t dimension, int btf, int scale, double tol, int* error_code);
// Route Solver::analyze would take for this COO matrix under the current
// environment (src/solver.cpp:820-860, :432-565): out4 = {1 global KLU / 0
// BLOCK, KLU btf, KLU scale, KLU ordering}, *tol = KLU pivot tolerance.
// Returns 0, or an error code if the matrix fails validation.
int bdlu_route_for(int64_t nnz, const int32_t* rows, const int32_t* cols, const double* vals, int32_t dimension,
                   int32_t* out4, double* tol);
void bdlu_host_free(void* h);
// n, raw/merged block counts, nmatch, nnz (canonical), off-diagonal nnz, levels.
void bdlu_host_info(void* h, int32_t* n, int32_t* nblocks_raw, int32_t* nblocks, int32_t* nmatch, int64_t* nnz,
                    int64_t* nnz_off, int32_t* nlevels);
void bdlu_host_copy_perm(void* h, int32_t* P, int32_t* Q, int32_t* R);
// Per block: 0 dense, 1 sparse; size.
void bdlu_host_block_kinds(void* h, int32_t* kind, int32_t* size);
// Sparse block statistics for block b (returns 0 if b is not a sparse block):
// stats = {s, k0, m, levels, offdiag_pivots, scale}, fstats = {nnz_l, nnz_u, updates, flops}.
int bdlu_host_sparse_stats(void* h, int32_t b, int32_t* stats6, double* fstats4);
// Column order (s) and pivot rows (s: the sparse part's, then the dense tail's
// in its pivot order) of sparse block b (local indices).
int bdlu_host_sparse_order(void* h, int32_t b, int32_t* q, int32_t* prow);
// Column order (n) that the analysis gives an n x n block with this CSC pattern
// (the COLAMD role of klu_analyze, ordering = 1): order[k] = column at position k.
void bdlu_host_colamd(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t* order);
// Signatures (reference src/solver.cpp:90-136) of the canonical matrix.
void bdlu_host_signatures(void* h, uint64_t* sig, uint64_t* sig2, uint64_t* value_sig);

// ---- workload generator -------------------------------------------------------
enum { BDLU_GEN_GEOMETRIC = 0, BDLU_GEN_UNIFORM = 1, BDLU_GEN_LOGUNIFORM = 2 };
// large_fill kinds; BDLU_GEN_BERNOULLI also names Bernoulli(rate) coupling.
enum { BDLU_GEN_POISSON4 = 0, BDLU_GEN_ZIPF = 1, BDLU_GEN_BERNOULLI = 2 };
enum { BDLU_GEN_POISSON = 0 /* coupling: Poisson(rate) entries per row */ };

typedef struct bdlu_gen_spec {
  uint64_t seed;
  int32_t n_small_blocks;
  int32_t size_dist;      // BDLU_GEN_GEOMETRIC (size_param = p) | UNIFORM (size_lo..size_param) | LOGUNIFORM (exp(U(0, ln size_param)))
  double size_param;
  int32_t size_lo, size_hi;  // clamp
  int64_t n_small_total;  // > 0: sizes adjusted by +-1 to sum to this
  int32_t n_large_blocks;
  int32_t large_size_lo, large_size_hi;
  int32_t large_fill;     // BDLU_GEN_ZIPF | BDLU_GEN_BERNOULLI (large_fill_param) | BDLU_GEN_POISSON4
  double large_fill_param;
  double extra_per_row;   // Poisson mean of in-block extra entries per row
  int32_t coupling_kind;  // BDLU_GEN_POISSON | BDLU_GEN_BERNOULLI
  double coupling_rate;
  double diag_factor;
  int32_t nrhs;
  double scale_exp;       // > 0: A <- D_r A D_c with d ~ 10^U(-scale_exp, scale_exp)
  uint64_t scale_seed;
  int32_t keep_order;     // 1: skip the random row/column permutation that hides the block structure
} bdlu_gen_spec;

void bdlu_gen_config_spec(int cfg, int pattern, bdlu_gen_spec* spec);
void* bdlu_gen_create(const bdlu_gen_spec* spec);
void bdlu_gen_info(const void* h, int32_t* n, int64_t* nnz, int32_t* nrhs);
void bdlu_gen_copy(const void* h, int32_t* rows, int32_t* cols, double* vals, double* rhs);
void bdlu_gen_free(void* h);
void bdlu_gen_perturb(double* vals, int64_t nnz, uint64_t seed, double rel);

#ifdef __cplusplus
}
#endif
