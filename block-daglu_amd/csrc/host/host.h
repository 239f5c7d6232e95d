#pragma once
// Host-side integer work of the BLOCK route: canonical COO, signatures, DM/BTF,
// block merge, UBTF decomposition, block-DAG levels, per-block ordering and the
// symbolic analysis that produces the device plan.
//
// Everything here is deterministic and bit-exact with the reference semantics it
// cites; floating point appears only where the reference itself computes with
// values during analysis (duplicate summation, the |v| < 1e-20 drop rule, and
// the pivot choice of the first factorization of blocks larger than 64).

#include <cstdint>
#include <string>
#include <tuple>
#include <vector>

namespace bdlu {

// Canonical COO: entries sorted by (row, col), duplicates summed, exact zeros
// dropped (reference src/solver.cpp:642-722). Structure of arrays.
struct Canonical {
  int32_t n = 0;
  std::vector<int32_t> row;
  std::vector<int32_t> col;
  std::vector<double> val;
  std::size_t nnz() const { return row.size(); }
};

enum class CanonStatus { Ok, InvalidDimension, InvalidIndex, NonFinite, DuplicateOverflow };

// Validates and canonicalizes `coo` (reference src/solver.cpp:642-722).
// On success `out` holds the canonical matrix and `coo` is rewritten to it too
// (the reference replaces the stored matrix by its canonical form).
CanonStatus canonicalize(std::vector<std::tuple<int, int, double>>& coo, int32_t dimension_hint,
                         Canonical& out, std::string& message);

// 64-bit signatures (reference src/solver.cpp:90-136); same formulas, so the same
// matrix gives the same numbers as the reference.
uint64_t matrix_sig(int32_t n, const Canonical& a);
uint64_t matrix_sig2(int32_t n, const Canonical& a);
uint64_t matrix_value_sig(const Canonical& a);
uint64_t cache_config_sig(std::size_t build_threads);

// ---- DM/BTF ------------------------------------------------------------------
constexpr int32_t kEmpty = -1;
inline int32_t btf_flip(int32_t j) { return -j - 2; }
inline int32_t btf_unflip(int32_t j) { return j < kEmpty ? btf_flip(j) : j; }

// Maximum transversal (depth-first augmenting paths with a cheap-assignment
// pass), SuiteSparse BTF's btf_maxtrans semantics. Match[i] = column matched to
// row i or kEmpty. Returns the number of matches.
int32_t btf_maxtrans(int32_t nrow, int32_t ncol, const int32_t* Ap, const int32_t* Ai, double maxwork,
                     double* work, int32_t* Match);
// Tarjan SCC of A(:,Q); P/R out, Q replaced by Q*P. Returns number of blocks.
int32_t btf_strongcomp(int32_t n, const int32_t* Ap, const int32_t* Ai, int32_t* Q, int32_t* P, int32_t* R);
// btf_order: maxtrans, completion of a structurally singular matching with
// flipped entries, then strongcomp. Called by the reference at ubtf:661.
int32_t btf_order(int32_t n, const int32_t* Ap, const int32_t* Ai, double maxwork, double* work, int32_t* P,
                  int32_t* Q, int32_t* R, int32_t* nmatch);

// ---- UBTF decomposition (reference ubtf:570-938) -------------------------------
struct Decomposition {
  int32_t n = 0;
  int32_t nblocks_raw = 0;
  int32_t nmatch = 0;
  int32_t nblocks = 0;
  double drop_abs_lt = 0.0;
  std::vector<int32_t> P, Q, invP, invQ, R;  // perm index -> original; R: block bounds
  // Strictly-upper inter-block entries in permuted coordinates (CSR by permuted
  // row, canonical-COO order within a row). off_src[p] = canonical entry index.
  std::vector<int32_t> off_ptr, off_col;
  std::vector<int64_t> off_src;
  // Diagonal-block entries, local coordinates, canonical-COO order per block.
  std::vector<int64_t> dblk_ptr;  // nblocks+1
  std::vector<int32_t> dblk_r, dblk_c;
  std::vector<int64_t> dblk_src;
  // Block DAG levels (ubtf:857-909): level(b) = 1 + max level(deps(b)).
  std::vector<int32_t> level;
  int32_t max_level = 0;
  std::vector<int32_t> level_ptr, level_blocks;
  // With DecompOptions::order_small_max: for each permuted position, the
  // block-local index it had before its block's column order was applied (the
  // row id KLU's pivot search breaks ties on); empty otherwise.
  std::vector<int32_t> tie_key;
};

struct DecompOptions {
  double drop_abs_lt = -1.0;  // <0: reference auto rule (1e-20 iff n >= 50000)
  double btf_maxwork = 0.0;
  bool merge = true;
  int32_t merge_max = 64;
  bool build_levels = true;
  // Global-KLU settings (reference src/solver.cpp:878-968 through klu_analyze):
  // btf=false keeps the whole matrix as one block (P = Q = identity), and
  // blocks of 2..order_small_max get their COLAMD column order applied
  // symmetrically to P and Q (KLU combines its block ordering with the BTF
  // permutations that way); larger blocks are ordered by their own analysis.
  bool btf = true;
  int32_t order_small_max = 0;
};
DecompOptions decomp_options_from_env();

// Throws std::runtime_error on failure (as the reference does).
Decomposition build_decomposition(const Canonical& a, const DecompOptions& opt);

std::vector<int32_t> merge_block_boundaries(const std::vector<int32_t>& R, int32_t n, int32_t merge_max);

// ---- route selection (reference src/solver.cpp:182-556, :820-860) ------------------
enum class RouteMode { Auto, Klu, Block };
RouteMode route_mode_from_env();
struct ProfileLite {
  double avg_deg = 0.0, diag_cov = 0.0;
  int32_t row_max = 0, col_max = 0, row_p95 = 0, col_p95 = 0;
};
ProfileLite profile_lite(const Canonical& a);
// true: BLOCK route, false: global KLU (AUTO mode).
bool auto_choose_block(const Canonical& a, uint64_t sig, uint64_t sig2, uint64_t value_sig, uint64_t config);
void note_block_analysis(const Canonical& a, uint64_t sig, uint64_t sig2, uint64_t value_sig, uint64_t config,
                         double build_ms, int32_t max_block);
int klu_btf_for(int32_t n, std::size_t nnz);
int klu_scale_for(int32_t n, std::size_t nnz, int btf);
double klu_tol_for(int32_t n, std::size_t nnz, int btf);
int klu_ordering();

// ---- per-block ordering ---------------------------------------------------------
// Column approximate-minimum-degree ordering of an s x s pattern in CSC
// (the role SuiteSparse COLAMD plays inside klu_analyze for ordering=1).
std::vector<int32_t> colamd_order(int32_t s, const int32_t* Ap, const int32_t* Ai);

}  // namespace bdlu
