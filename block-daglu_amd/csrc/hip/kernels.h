#pragma once
// Device kernels of the BLOCK-route engine (gfx950) and their launchers.
// Launchers enqueue on the given stream and never synchronize.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace bdk {

// Per-sparse-block device view (pointers into device memory).
struct SpView {
  int32_t s, k0, m, scale;
  double inv_tol;          // largest |L| the fixed-pivot check accepts
  int64_t a_val0;
  const int32_t* a_ptr;    // s+1
  const int32_t* a_row;
  const int32_t* prow;     // k0
  const int32_t* l_ptr;    // k0+1
  const int32_t* l_row;
  const int32_t* u_ptr;    // s+1
  const int32_t* u_idx;
  const int32_t* trow;     // m
  const int32_t* tpos;     // m: tail-local row -> tail position
  const int32_t* rs_ptr;   // s+1: block entries by local row (for the row scaling)
  const int32_t* rs_idx;
  const int32_t* rpos;     // s: local row -> pivot position
};

// Status bits written by the factor kernels into a per-block int.
enum : int32_t { kStZeroPivot = 1, kStCheck = 2, kStNonFinite = 4 };

void launch_scatter(hipStream_t st, int64_t nnz, const double* vals, const int64_t* vmap, double* arena);
void launch_fill_i32(hipStream_t st, int32_t* p, int64_t n, int32_t v);
void launch_fill_f64(hipStream_t st, double* p, int64_t n, double v);
void launch_fill_f32(hipStream_t st, float* p, int64_t n, float v);
// Fills with the solve's "not yet published" pattern (see kernels.hip).
void launch_fill_sentinel(hipStream_t st, double* p, int64_t n);
void launch_fill_sentinel_cols(hipStream_t st, double* p, int32_t m, int32_t k, int64_t ld);  // k columns, stride ld

// key == nullptr: partial pivoting (BLOCK route); else KLU's threshold rule
// with tolerance tol and per-row tie keys (global-KLU route).
template <typename T>
void launch_dense_lu(hipStream_t st, int32_t nd, const int32_t* sizes, const int32_t* starts, const int64_t* val0,
                     const int64_t* lu0, const double* arena, T* lu, int32_t* piv, int32_t* pinv, int32_t* status,
                     const int32_t* block_of, const int32_t* order = nullptr, double tol = 1.0,
                     const int32_t* key = nullptr);
// KLU whole-row scaling: rs[i] = sum (kind 1) or max (kind 2) of |vals| in CSR
// row i (1 for a row without a nonzero), svals[e] = vals[e] / rs[row of e].
void launch_row_scale(hipStream_t st, int32_t n, const int64_t* ptr, const double* vals, int32_t kind, double* rs,
                      double* svals);

void launch_rowscale(hipStream_t st, const SpView& v, const double* arena, double* rs);
// Columns of one refactor level whose reach is long, done row-oriented by
// level (one workgroup each) after the others: is_long (k0) marks them,
// long_j lists this level's, reach/lev hold their reach positions grouped by
// forward level (nlev levels), lr_*: L by position rows.
struct SpColsLong {
  int32_t nlong = 0, nlev = 0;
  const int32_t* is_long = nullptr;
  const int32_t* long_j = nullptr;
  const int32_t* reach = nullptr;
  const int32_t* lev = nullptr;
  const int32_t* lr_ptr = nullptr;
  const int32_t* lr_col = nullptr;
  const int32_t* lr_vidx = nullptr;
};
template <typename T>
void launch_sp_cols(hipStream_t st, const SpView& v, const double* arena, const double* rs, int32_t ncols,
                    const int32_t* cols, T* xs, int32_t nslots, T* Lx, T* Ux, T* Ud, int32_t* status,
                    const SpColsLong* lg = nullptr);
// Tail columns whose sparse reach is long, done row-oriented by level (one
// workgroup each) after the others: is_long (m) marks them, long_t lists them,
// reach/lev hold their reach positions grouped by forward level (nlev levels),
// lr_*: L by position rows.
struct SpTailLong {
  int32_t nlong = 0, nlev = 0;
  const int32_t* is_long = nullptr;
  const int32_t* long_t = nullptr;
  const int32_t* reach = nullptr;
  const int32_t* lev = nullptr;
  const int32_t* lr_ptr = nullptr;
  const int32_t* lr_col = nullptr;
  const int32_t* lr_vidx = nullptr;
};
template <typename T>
void launch_sp_tail(hipStream_t st, const SpView& v, const double* arena, const double* rs, T* xs, int32_t nslots,
                    const T* Lx, T* Ux, T* S, int64_t lds, const SpTailLong* lg = nullptr);
// Optional per-launch event timing of the tail kernels (profiling runs only).
struct TailTimer {
  std::vector<hipEvent_t>* gemm_events = nullptr;   // pairs (start, stop)
  std::vector<double>* gemm_flops = nullptr;
  std::vector<double>* gemm_bytes = nullptr;
  std::vector<hipEvent_t>* panel_events = nullptr;  // pairs around diag + trsm
};

// No-pivot blocked LU of the m x m column-major S (the tail after its rows were
// placed in the fixed pivot order). Check: |L| <= limit and nonzero pivots.
// inv: 2*64*64 per 64-column panel (ceil(m/64) panels): the inverses of the
// diagonal blocks, L^{-1} then U^{-1}, row-major, kept for the tail solve.
// lookahead 1: each step's next diagonal LU runs as one workgroup of the bulk
// trailing-update launch; 2: in addition, panels go in pairs and the bulk
// update applies both at once (K = 128). Results are identical in every mode.
// With a timer, events bracket every panel launch (and the strip updates
// between paired panels) and every bulk trailing-update launch (the fused one
// with lookahead) on the same launch structure.
template <typename T>
void launch_tail_lu(hipStream_t st, int32_t m, T* S, int64_t lds, double limit, int32_t* status, T* inv,
                    TailTimer* timer = nullptr, int32_t small_tile_below = 0, int lookahead = 0);

// Explicit inverse M = U^{-1} L^{-1} of each dense block's factors, s x s
// column-major at the block's factor offset (x = M (P b) solves the block).
template <typename T>
void launch_dense_inv(hipStream_t st, int32_t nd, const int32_t* sizes, const int64_t* lu0, const T* lu, T* minv);
// ---- solve ----
// Vector sets hold kc columns (right-hand sides), column c at base + c * ld.
// Kernels with per-column registers are instantiated for 1, 2, 4 and 8 columns
// (kc is rounded up by the caller, padding with zero columns).
void launch_permute(hipStream_t st, int32_t n, int32_t k, const int32_t* P, const double* src, int64_t lds,
                    double* dst, int64_t ldd, const double* rs = nullptr);  // dst[c][i] = src[c][P[i]] (/ rs[P[i]])
void launch_unpermute(hipStream_t st, int32_t n, int32_t k, const int32_t* Q, const double* src, int64_t lds,
                      double* dst, int64_t ldd);  // dst[c][Q[i]] = src[c][i]
// Dense-block solve of one segment (persistent, sync-free). ticket starts at 0;
// xperm entries of this segment's blocks start as the unpublished pattern.
void launch_dense_seg_solve(hipStream_t st, int32_t kc, int32_t ntasks, const int32_t* tasks, const int32_t* R,
                            const int64_t* lu0_of_block, const double* minv64, const float* minv32,
                            const int32_t* prec_of_block,
                            const int32_t* gpinv, const int32_t* off_ptr, const int32_t* off_col,
                            const double* off_val, int32_t* ticket, const double* bperm, double* xperm, int64_t ldv,
                            int32_t* err, int32_t grid = 1024);
// Sparse block solve pieces; y is an s x kc work set in position order.
void launch_sp_gather(hipStream_t st, int32_t s, int32_t start, int32_t kc, const int32_t* rowperm, const double* rs,
                      const int32_t* off_ptr, const int32_t* off_col, const double* off_val, const double* bperm,
                      const double* xperm, int64_t ldv, double* y, int64_t ldy);
template <typename T>
void launch_sp_rows(hipStream_t st, int32_t kc, int32_t nrows, const int32_t* rows, const int32_t* ptr,
                    const int32_t* col, const int32_t* vidx, const T* vals, const T* diag, double* y, int64_t ldy);
// Couplings of the dense tail's 64-row bands to their neighbours in solve
// order (D_w^{-1} times the sub- or superdiagonal tile), for launch_tail_trsv;
// g has the layout of inv (2*64*64 per band).
template <typename T>
void launch_tail_coupling(hipStream_t st, int32_t m, const T* S, int64_t lds, const T* inv, T* g);
// Dense tail triangular solve over 64-row bands; inv holds the diagonal tiles'
// inverses as launch_tail_lu leaves them, g the couplings. ticket starts at 0;
// yout starts as the unpublished pattern and does not alias yin.
template <typename T>
void launch_tail_trsv(hipStream_t st, int32_t kc, int32_t m, const T* S, int64_t lds, const T* inv, const T* g,
                      const double* yin, int64_t ldi, double* yout, int64_t ldo, bool upper, int32_t* ticket,
                      int32_t* err);
void launch_sp_scatter_x(hipStream_t st, int32_t s, int32_t start, int32_t kc, const int32_t* q, const double* x,
                         int64_t ldx, double* xperm, int64_t ldv);

// ---- residual / norms / refinement ----
// r = b - A x; cw_i = |r_i| / (|A||x| + |b|)_i (componentwise backward-error ratio).
// Rows flagged in is_long (nlong of them, listed in long_rows) are taken one
// wavefront each; with nlong == 0 the flags are not read.
void launch_residual(hipStream_t st, int32_t n, int32_t kc, const int64_t* csr_ptr, const int32_t* csr_col,
                     const double* vals, const double* x, const double* b, double* r, double* cw, int64_t ld,
                     const uint8_t* is_long, int32_t nlong, const int32_t* long_rows);
// Per column c: out[c*9+0]=sum r^2, [1]=max|r|, [2]=max|x|, [3]=sum b^2, [4]=max|b|,
// [5]=max cw. partials: 512 * 6 * kc doubles.
void launch_norms(hipStream_t st, int32_t n, int32_t kc, const double* r, const double* x, const double* b,
                  const double* cw, int64_t ld, double* partials, double* out);
// out[0] = max_i sum_j |a_ij|
void launch_anorm(hipStream_t st, int32_t n, const int64_t* csr_ptr, const double* vals, double* partials,
                  double* out);
void launch_axpy(hipStream_t st, int32_t n, int32_t kc, const double* d, double* x, int64_t ld);  // x += d
// Refinement state of a loop run on the device (doubles): still running, steps
// taken, converged, worst measure of the last evaluation, ended by a stall.
constexpr int kIrActive = 0, kIrIters = 1, kIrConverged = 2, kIrWorst = 3, kIrStalled = 4, kIrStateLen = 8;
void launch_axpy_gated(hipStream_t st, int32_t n, int32_t kc, const double* d, double* x, int64_t ld,
                       const double* state);  // x += d while state[kIrActive] != 0
void launch_ir_check(hipStream_t st, const double* norms, int32_t kc, int32_t rule, double tol, const double* anorm,
                     int32_t max_iters, bool watch_stall, bool first, double* state);

}  // namespace bdk
