#pragma once
// Symbolic analysis of one diagonal block larger than the dense threshold, the
// host half of the reference's KLU branch (ubtf:1749-1825): column ordering
// (COLAMD role), optional row scaling (KLU scale=1 "sum" for blocks of 4096 or
// more, ubtf:556-568), and a first left-looking factorization with KLU's
// threshold partial pivoting (tol 0.001, diagonal preference with diagonal
// tracking) that fixes the pivot rows and the L/U patterns of the sparse
// leading columns. When the active submatrix becomes dense the factorization
// stops at column k0; the trailing m = s - k0 columns form a dense Schur
// complement that the device factors with the same pivot rule.
//
// Only patterns, pivot rows and the tail's starting "diagonal" bookkeeping leave
// this analysis; every numeric value used by solves is computed on the device.

#include <cstdint>
#include <vector>

namespace bdlu {

struct SparseBlockPlan {
  int32_t s = 0;
  int32_t k0 = 0;   // sparse leading columns
  int32_t m = 0;    // dense tail size
  bool scale = false;
  std::vector<int32_t> q;       // position -> local column
  // Block entries grouped by pivot-column position: local row and the index of
  // the entry in the block's local entry list.
  std::vector<int32_t> a_ptr, a_row, a_ent;
  std::vector<int32_t> prow;    // k0: pivot row (local) of position k
  std::vector<int32_t> l_ptr, l_row;   // k0+1: rows of L(:,k) without the pivot row, ascending
  std::vector<int32_t> u_ptr, u_idx;   // s+1: U(k,j) positions k < min(j,k0), ascending
  std::vector<int32_t> lev_ptr, lev_col;  // refactor levels of the sparse columns
  std::vector<int32_t> trow;    // m tail rows (local), ascending
  std::vector<int32_t> tdiag;   // m: tail-local diagonal row of tail column t, or -1
  std::vector<int32_t> tperm;   // m: tail-local row pivoted at tail position t
  // statistics
  int64_t nnz_l = 0, nnz_u = 0, updates = 0;
  double flops = 0.0;
  int32_t levels = 0;
  int32_t offdiag_pivots = 0;
};

struct SymbolicOptions {
  double tol = 0.001;            // KLU default pivot tolerance (sparse leading columns)
  // Pivot tolerance inside the dense Schur tail. 1.0 (default) = partial
  // pivoting (the diagonal wins only ties), which bounds |L| by 1: the device
  // factors the tail with explicit 64x64 triangular inverses in its MFMA panel
  // products, which amplify |L| growth (with KLU's tol 0.001 the plain FP64
  // residual of the circuit-like system grows from 7e-11 to 1.8e-6).
  // tol (BLOCKDAGLU_GPU_TAIL_TOL=0.001) gives klu_factor's pivot rows over the
  // whole block; <= 0 means the same as tol.
  double tail_tol = 1.0;
  double tail_density = 0.2;     // switch to the dense tail when |L(:,j)| > density*(s-j)
  int32_t tail_min = 32;         // ... and at least this many columns remain
  int32_t scale_min_n = 4096;    // KLU scaling for blocks at least this large
  int32_t scale_kind = 1;        // 1: row sums (KLU scale=1), 2: row maxima (KLU scale=2)
};

// Numeric values of the first factorization, for callers that export factors
// (analyze_sparse_block otherwise keeps patterns and pivot rows only).
struct SparseFactorValues {
  std::vector<double> rs;     // s: row scale factors by local row (1 when unscaled)
  std::vector<double> lx;     // L(:,k) below the pivot, parallel to l_row
  std::vector<double> ux;     // U(k,j) above the diagonal, parallel to u_idx
  std::vector<double> udiag;  // k0: pivots
};
SymbolicOptions symbolic_options_from_env();

// rows/cols/vals: the block's local entries (canonical order). Returns false if
// a column has no admissible pivot (numerically or structurally singular).
bool analyze_sparse_block(int32_t s, const std::vector<int32_t>& rows, const std::vector<int32_t>& cols,
                          const std::vector<double>& vals, const SymbolicOptions& opt, SparseBlockPlan& out,
                          SparseFactorValues* values = nullptr);

// Threshold partial pivoting with diagonal preference and KLU's diagonal
// tracking on a dense m x m column-major matrix (factored in place, rows
// physically swapped). diag[t]: starting diagonal row (original row id) of
// column t or -1. perm[i] out: original row id at position i. Ties in |a| go to
// the smaller original row id. Returns false if a column has no nonzero.
bool dense_threshold_lu(int32_t m, double* a, const std::vector<int32_t>& diag, double tol,
                        std::vector<int32_t>& perm);

}  // namespace bdlu
