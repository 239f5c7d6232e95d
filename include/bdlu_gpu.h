#pragma once
// C-ABI of the MI355X numeric engine (libbdlu_gpu.so, HIP for gfx950).
//
// It replaces the numeric half of the reference BLOCK route, the internal seam
// declared at blockdaglu/src/ubtf_block_solver.h:145-152:
//   factor_blocks_parallel(dec, threads)  -> bdlu_gpu_refactor
//        (called at src/solver.cpp:1252; dense branch ubtf:1441-1498, KLU
//         branch ubtf:1749-1825)
//   free_factors(fac)                     -> bdlu_gpu_plan_release
//        (called at src/solver.cpp:728, :756, :1245, :1262)
//   solve_ubtf(dec, fac, b, threads)      -> bdlu_gpu_solve
//        (called at src/solver.cpp:1481 and :1529; body ubtf:2220-2555; the
//         refinement loop of src/solver.cpp:1486-1536 runs inside when asked)
// build_decomposition_btf (ubtf:570) stays on the host; its output, plus the
// per-block symbolic analysis, crosses this boundary once per sparsity
// pattern as a plan (plain arrays, borrowed for the duration of the call).
//
// Conventions: every function is extern "C", never throws, returns 0 on
// success or a status code below, and records a message for
// bdlu_gpu_last_error. A context owns one HIP stream; calls on one context are
// serialized by an internal mutex and complete before they return.

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bdlu_ctx bdlu_ctx;

// Status codes (the nonzero ones equal blockdaglu::SolverError values where one
// applies, so the host maps them 1:1).
enum {
  BDLU_OK = 0,
  BDLU_E_INVALID = 5,             // bad argument / size (InvalidRightHandSide role)
  BDLU_E_BACKEND = 6,             // no device, HIP error, out of the HBM budget
  BDLU_E_ANALYSIS = 7,            // inconsistent plan
  BDLU_E_FACTOR = 8,              // a block has an exactly-zero pivot
  BDLU_E_SOLVE = 9,               // solve failed (zero diagonal, non-finite result)
  BDLU_E_REPIVOT = 20,            // a fixed-pivot check failed: re-analyse with the current values
  BDLU_E_BUDGET = 21,             // plan does not fit the context's HBM budget
};

enum { BDLU_MEM_HOST = 0, BDLU_MEM_DEVICE = 1 };
enum { BDLU_FP64 = 0, BDLU_FP32 = 1 };

// One diagonal block larger than the dense threshold (positions are the pivot
// order: position k < k0 pivots on local row prow[k] in column q[k]; positions
// k0 + t pivot on local row trow[tperm[t]] in column q[k0 + t]).
typedef struct bdlu_sparse_desc {
  int32_t block;
  int32_t s, k0, m;
  int32_t scale;          // 1: rows scaled by the sum of |a_ij| inside the block (KLU scale=1)
  double tol;             // KLU pivot tolerance the host analysis used
  int64_t a_val0;         // value-arena offset of the block's entries, in a_ptr order
  const int32_t* a_ptr;   // s+1: entries of column position j
  const int32_t* a_row;   // local row of each entry
  const int32_t* q;       // s: position -> local column
  const int32_t* prow;    // k0
  const int32_t* l_ptr;   // k0+1
  const int32_t* l_row;   // local rows of L(:,k), pivot row excluded
  const int32_t* u_ptr;   // s+1
  const int32_t* u_idx;   // positions k < min(j, k0), ascending per column
  const int32_t* lev_ptr; // nlev+1
  const int32_t* lev_col; // k0: sparse columns grouped by refactor level
  int32_t nlev;
  const int32_t* trow;    // m: tail-local row -> local row
  const int32_t* tperm;   // m: tail position -> tail-local row
} bdlu_sparse_desc;

typedef struct bdlu_plan_desc {
  int32_t n;
  int64_t nnz;              // canonical entries = length of the values array of refactor
  const int64_t* vmap;      // nnz: value-arena slot of each canonical entry, -1 if dropped
  int64_t arena_len;        // doubles
  const int64_t* csr_ptr;   // n+1: canonical entries by original row (canonical order is row-major)
  const int32_t* csr_col;   // nnz: original column
  const int32_t* P;         // n: permuted row k -> original row
  const int32_t* Q;         // n: permuted col k -> original col
  int32_t nblocks;
  const int32_t* R;         // nblocks+1
  const int32_t* off_ptr;   // n+1: strictly-upper inter-block entries by permuted row
  const int32_t* off_col;   // permuted column
  int64_t off_val0;         // arena offset of the off-diagonal values (CSR order)
  const int32_t* dep_ptr;   // nblocks+1: blocks each block's rows couple to (all > b)
  const int32_t* dep_blk;
  int32_t ndense;
  const int32_t* dense_block;  // block id of each dense block
  const int64_t* dense_val0;   // arena offset of its s*s column-major values
  int32_t nsparse;
  const bdlu_sparse_desc* sparse;
  // Pivot rule of the dense blocks: 0 partial pivoting, first largest |a| (the
  // BLOCK route's dense branch, ubtf:1465-1477); 1 KLU's threshold rule (the
  // global-KLU route, src/solver.cpp:1198): the diagonal row is kept when
  // |a_kk| >= dense_tol * max_i |a_ik|, else the largest, ties to the smaller
  // dense_key (n values, one per permuted row; NULL with rule 0).
  int32_t dense_rule;
  double dense_tol;
  const int32_t* dense_key;
  // 0: none; 1 / 2: every row of A is divided by the sum / maximum of its |a_ij|
  // (KLU scale 1 / 2 over whole rows) before the blocks are factored, and the
  // right-hand side by the same factors before the block solve (klu_solve).
  int32_t row_scale;
} bdlu_plan_desc;

typedef struct bdlu_factor_stats {
  int32_t failed_blocks;        // exactly-zero pivot (dense: partial pivoting found none)
  int32_t repivot_blocks;       // fixed-pivot check failed in FP64
  int32_t fp64_fallback_blocks; // FP32 check failed, block refactored in FP64
  int32_t first_bad_block;      // a block id needing re-analysis, or -1
  double max_abs_l;             // largest |L| seen by the checks (sparse blocks)
} bdlu_factor_stats;

typedef struct bdlu_ir_opts {
  int32_t max_iters;        // 0: plain solve. >0: refinement steps at most
  int32_t stop_rule;        // 0: reference rule ||r||_2/||b||_2 < tol (src/solver.cpp:1508-1521)
                            // 1: normwise backward error <= tol
                            // 2: componentwise backward error <= tol
  double tol;
  int32_t fp64_fallback;    // 1: when the factors in use include FP32 ones and the stop rule is not met
                            // within max_iters (or a step fails to halve the worst stop measure), the
                            // plan is refactored in FP64 from the same values and the columns are
                            // solved and refined again (bdlu_solve_stats.fp64_refactor = 1)
} bdlu_ir_opts;

typedef struct bdlu_solve_stats {
  int32_t iterations;
  double rel_residual;      // ||b - A x||_2 / ||b||_2 (FP64)
  double backward_error;    // ||r||_inf / (||A||_inf ||x||_inf + ||b||_inf)
  double componentwise_error;  // max_i |r_i| / (|A| |x| + |b|)_i (Oettli-Prager)
  int32_t fp64_refactor;    // 1: FP32 factors did not converge and the plan was refactored in FP64
} bdlu_solve_stats;

int bdlu_gpu_device_count(int* count);
// Source commit and target of this build, e.g. "1a2b3c4d5e6f gfx950" (no
// counterpart in the reference).
const char* bdlu_gpu_build_info(void);
int bdlu_gpu_create(int device, uint64_t hbm_budget_bytes, bdlu_ctx** out);
void bdlu_gpu_destroy(bdlu_ctx* ctx);
const char* bdlu_gpu_last_error(bdlu_ctx* ctx);

int bdlu_gpu_plan_upload(bdlu_ctx* ctx, const bdlu_plan_desc* plan, uint64_t* plan_id);
int bdlu_gpu_plan_release(bdlu_ctx* ctx, uint64_t plan_id);
int bdlu_gpu_plan_bytes(bdlu_ctx* ctx, uint64_t plan_id, uint64_t* bytes);
// Launch schedule the autotuner picked for a plan (no counterpart in the
// reference): out[0] trailing updates below this size use 64x64 tiles,
// out[1] dense blocks launched largest first (0/1), out[2] persistent
// workgroups of the dense-block solve, out[3] bit 0/1: refactor tuned for
// FP64/FP32, bit 2: solve tuned. Tuning happens on a plan's first refactor
// (per precision) and first solve unless BLOCKDAGLU_GPU_AUTOTUNE=0; every
// candidate gives bit-identical results.
int bdlu_gpu_plan_schedule(bdlu_ctx* ctx, uint64_t plan_id, int32_t* out4);

// New values for the plan's pattern, canonical order (values == NULL: refactor
// the values of the plan's previous refactor again, e.g. in the other
// precision). precision: BDLU_FP64 or BDLU_FP32. Sparse blocks refactor on the analysis' pivot order and check the
// growth of L: an FP32 block with |L| above BLOCKDAGLU_GPU_FP32_GROWTH (1e4)
// is refactored in FP64; an FP64 block with |L| above
// BLOCKDAGLU_GPU_REPIVOT_GROWTH (1e8), a zero or a non-finite pivot returns
// BDLU_E_REPIVOT so the host re-analyses it with the current values.
int bdlu_gpu_refactor(bdlu_ctx* ctx, uint64_t plan_id, const double* values, int where, int precision,
                      bdlu_factor_stats* stats);
// B: n x nrhs column-major (leading dimension ldb), overwritten by X.
// With refinement and graphs on, the solve and the refinement steps (as many
// as the fewest of the plan's last four solves needed) replay as one graph
// whose stop tests run on the device; a loop still running at its end
// continues with one host stop test per step. Either way the same kernels run
// and the results are the same bit for bit.
int bdlu_gpu_solve(bdlu_ctx* ctx, uint64_t plan_id, double* B, int32_t nrhs, int64_t ldb, int where,
                   const bdlu_ir_opts* ir, bdlu_solve_stats* stats);

// No counterpart in the reference. For callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid allocation):
// bdlu_gpu_solve with B in device memory and no allocation of any kind. The
// solve's per-call vectors and counters come from a static device array in the
// library (128 MiB per GPU, kept for the life of the process) instead of the
// plan, launches are issued directly (no graph is instantiated, no autotuning),
// and FP32 factors that do not reach the stop rule return BDLU_E_SOLVE rather
// than being refactored in FP64. Limit: 4 vectors of n x k doubles (7 with
// refinement), k = nrhs rounded up to 1, 2, 4 or 8, plus a few KiB, must fit
// in 16Mi doubles; larger requests return BDLU_E_INVALID. Results equal
// bdlu_gpu_solve's bit for bit. Calls on one GPU share the array: calls that
// overlap in time on one GPU (from any context) race.
int bdlu_gpu_solve_noalloc(bdlu_ctx* ctx, uint64_t plan_id, double* d_B, int32_t nrhs, int64_t ldb,
                           const bdlu_ir_opts* ir, bdlu_solve_stats* stats);

// Device-to-device copy of `bytes` on the context's stream, complete on return
// (no counterpart in the reference; the serving loop uses it to place a
// device-resident request's right-hand sides in its solution buffer).
int bdlu_gpu_copy(bdlu_ctx* ctx, void* dst, const void* src, uint64_t bytes);

// Copies the plan's current canonical-order values (last refactor) to host memory.
int bdlu_gpu_plan_values(bdlu_ctx* ctx, uint64_t plan_id, double* host_values);

// Side streams of a refactor (no counterpart in the reference): with enable
// = 1 (the default) the dense-block LU and the dense tails of different sparse
// blocks run on the context's side streams, joined back into its stream; 0
// keeps every launch on the one stream (the serving loop runs one context per
// slot, and side streams of concurrent contexts would share hardware queues).
// Results are identical either way.
int bdlu_gpu_set_side_streams(bdlu_ctx* ctx, int enable);

// Largest persistent grid of the dense-block solve (no counterpart in the
// reference; default 1024, or BLOCKDAGLU_GPU_SEG_GRID_MAX). The solve's
// workgroups wait on one another's results, so where several contexts share
// the GPU (the serving loop) a small grid leaves compute units to the other
// contexts' work; alone, the autotuner picks among grids up to this size.
// Results are identical for every grid.
int bdlu_gpu_set_solve_grid_max(bdlu_ctx* ctx, int32_t max_workgroups);

// Refinement stop tests on the device, the solve and its steps replayed as one
// graph (enable = 1, the default, or BLOCKDAGLU_GPU_IR_DEVICE), or one host
// stop test per step (0). No counterpart in the reference. The device form
// saves a host round trip per step, which matters to a lone caller waiting on
// its solve; where several contexts share the GPU (the serving loop) other
// contexts fill those gaps anyway. Results are identical either way.
int bdlu_gpu_set_device_refinement(bdlu_ctx* ctx, int enable);

// The context's HIP stream (a hipStream_t), so callers can order or time their
// own work around engine calls.
void* bdlu_gpu_stream(bdlu_ctx* ctx);

// Per-launch timing with HIP events on the engine stream, by kernel family.
// Enabling resets the counters. flops/bytes are algorithmic (from the plan's
// shapes), accumulated over the timed launches.
enum {
  BDLU_PROF_TAIL_GEMM = 0,   // trailing MFMA update of the dense Schur tails (with lookahead: the fused
                             // launch that also factors the next diagonal block)
  BDLU_PROF_TAIL_PANEL = 1,  // diagonal LU + both triangular solves of a tail panel
  BDLU_PROF_DENSE_LU = 2,    // batched dense LU of the blocks <= 64
  BDLU_PROF_SPARSE = 3,      // sparse leading columns + Schur columns
  BDLU_PROF_SOLVE = 4,       // one block-DAG solve
  BDLU_PROF_KINDS = 5
};
int bdlu_gpu_profile(bdlu_ctx* ctx, int enable);
int bdlu_gpu_profile_read(bdlu_ctx* ctx, int kind, double* total_ms, int64_t* launches, double* flops,
                          double* bytes);

// Row pivots of the dense blocks from the last refactor (test hook, no
// counterpart in the reference): host_piv[i] (n values) = the block-local row
// that the pivoting brought to permuted row i (i inside a dense block).
int bdlu_gpu_plan_dense_pivots(bdlu_ctx* ctx, uint64_t plan_id, int32_t* host_piv);

// Test hook: no-pivot dense LU of an m x m column-major matrix on the device
// through the same kernels the tail factorization uses; A is overwritten by L\U.
// lookahead selects the launch structure (0 separate diagonal LU launches,
// 1 diagonal LU fused into each bulk update, 2 paired panels with K = 128 bulk
// updates); trailing updates below small_tile_below rows use 64x64 tiles.
int bdlu_gpu_debug_dense_lu(bdlu_ctx* ctx, int32_t m, double* A, int precision, int lookahead,
                            int32_t small_tile_below);

#ifdef __cplusplus
}
#endif
