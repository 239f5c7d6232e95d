// C-ABI of the device engine (declared in include/bdlu_gpu.h): plan upload,
// refactorization, block-DAG solve and FP64 refinement on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/bdlu_gpu.h"
#include "kernels.h"

// Right-hand sides solved together per pass (work sets are n x kMaxRhs), and the
// slot of ||A||_inf after the per-column norms.
constexpr int32_t kMaxRhs = 8;
constexpr int kAnormSlot = 9 * kMaxRhs;

namespace {

struct HipError {
  std::string msg;
};

#define HIPCHECK(expr)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) throw HipError{std::string(#expr) + ": " + hipGetErrorString(e_)}; \
  } while (0)

// Device allocations of one plan (or of one call). Copies from host memory are
// asynchronous on the caller's stream: a raw host pointer must stay valid until
// that stream is synchronized (the borrowed plan arrays do, since the upload
// synchronizes before returning); a host vector may be a temporary, so its
// bytes are first copied into `staged`, which lives until drop_staged() after
// the synchronization.
struct Alloc {
  std::vector<void*> ptrs;
  std::vector<std::vector<unsigned char>> staged;
  uint64_t bytes = 0;
  template <typename T>
  T* get(std::size_t count) {
    if (count == 0) count = 1;
    void* p = nullptr;
    HIPCHECK(hipMalloc(&p, count * sizeof(T)));
    ptrs.push_back(p);
    bytes += count * sizeof(T);
    return static_cast<T*>(p);
  }
  template <typename T>
  T* copy(const T* host, std::size_t count, hipStream_t st) {
    T* d = get<T>(count);
    if (count && host) HIPCHECK(hipMemcpyAsync(d, host, count * sizeof(T), hipMemcpyHostToDevice, st));
    return d;
  }
  template <typename T>
  T* copy(const std::vector<T>& v, hipStream_t st) {
    if (v.empty()) return get<T>(0);
    const auto* src = reinterpret_cast<const unsigned char*>(v.data());
    staged.emplace_back(src, src + v.size() * sizeof(T));
    return copy(reinterpret_cast<const T*>(staged.back().data()), v.size(), st);
  }
  // Call after the stream that carried the copies has been synchronized.
  void drop_staged() { staged.clear(); }
  void release() {
    for (void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
    staged.clear();
    bytes = 0;
  }
};

struct SparseDev {
  int32_t block = 0, start = 0, s = 0, k0 = 0, m = 0, scale = 0;
  double tol = 0.001;
  bdk::SpView view{};
  std::vector<int32_t> lev_ptr;  // refactor levels (host)
  int32_t* lev_col = nullptr;
  int32_t* rowperm = nullptr;
  int32_t* q = nullptr;
  // solve structures
  std::vector<int32_t> fl_ptr, bl_ptr;  // level pointers (host)
  int32_t *fl_rows = nullptr, *tl_rows = nullptr, *bl_rows = nullptr;
  int32_t *lr_ptr = nullptr, *lr_col = nullptr, *lr_vidx = nullptr;
  int32_t *ur_ptr = nullptr, *ur_col = nullptr, *ur_vidx = nullptr;
  int64_t nL = 0, nU = 0;
  double *Lx64 = nullptr, *Ux64 = nullptr, *Ud64 = nullptr, *S64 = nullptr;
  float *Lx32 = nullptr, *Ux32 = nullptr, *Ud32 = nullptr, *S32 = nullptr;
  // Column stride of the dense tail S (m rounded up to 64 elements, so every
  // column starts on a cache-line boundary in both precisions).
  int64_t ldt = 1;
  double* rs = nullptr;
  double* y = nullptr;
  double* inv64 = nullptr;  // 2*64*64 per panel: the tail's diagonal-block inverses
  float* inv32 = nullptr;
  double* cpl64 = nullptr;  // same layout: the band couplings for the tail solve
  float* cpl32 = nullptr;
  double* ty = nullptr;  // m: tail forward-solve output (published values)
  bdk::SpTailLong tail_long;  // tail columns with a long sparse reach
  std::vector<bdk::SpColsLong> cols_long;  // per refactor level: columns with a long reach
  int32_t* status = nullptr;  // device int
  int prec = 0;
};

// Per-call vectors of a solve (n x kMaxRhs sets, column stride n) and its
// counters. A plan owns one set; the no-allocation entry point carves one out
// of the library's static scratch array instead.
struct Work {
  double *b = nullptr, *x = nullptr, *r = nullptr, *d = nullptr, *bperm = nullptr, *xperm = nullptr;
  double* cw = nullptr;        // per-row componentwise backward-error ratios of the last residual
  double* partials = nullptr;  // 512 * 6 * kMaxRhs
  double* norms = nullptr;     // per column c: [c*9 + 0..5] residual norms
  int32_t* tickets = nullptr;  // the plan's nticket counters
  int32_t* err = nullptr;
  double* irs = nullptr;  // refinement state of a loop run on the device (bdk::kIrStateLen)
};

struct Phase {
  bool sparse = false;
  int32_t index = 0;   // sparse block index or segment id
  int32_t ntasks = 0;
  int32_t* tasks = nullptr;
};

struct Plan {
  Alloc mem;
  int32_t n = 0, nblocks = 0;
  int64_t nnz = 0, arena_len = 0;
  int64_t* vmap = nullptr;
  double* arena = nullptr;
  double* vals = nullptr;
  int64_t* csr_ptr = nullptr;
  int32_t* csr_col = nullptr;
  // Rows with more than kLongRow entries, for the wavefront-per-row residual.
  uint8_t* res_is_long = nullptr;
  int32_t* res_long_rows = nullptr;
  int32_t res_nlong = 0;
  int32_t *P = nullptr, *Q = nullptr, *R = nullptr;
  int32_t *off_ptr = nullptr, *off_col = nullptr;
  double* off_val = nullptr;
  // dense blocks
  int32_t ndense = 0;
  std::vector<int32_t> h_dense_block, h_dense_size, h_dense_start;
  std::vector<int64_t> h_dense_val0, h_dense_lu0;
  int32_t *d_size = nullptr, *d_start = nullptr, *d_block = nullptr;
  int64_t *d_val0 = nullptr, *d_lu0 = nullptr;
  int64_t* lu0_of_block = nullptr;
  int64_t lu_len = 0;
  double* lu64 = nullptr;
  float* lu32 = nullptr;
  // Explicit inverses M = U^{-1} L^{-1} of the dense blocks, at the factor offsets.
  double* minv64 = nullptr;
  float* minv32 = nullptr;
  // Pivot rule of the dense blocks (0 partial pivoting; 1 KLU threshold rule
  // with tolerance dense_tol and per-row tie keys dense_key) and KLU's whole-row
  // scaling (row_scale 1 sums / 2 maxima; rs per original row, svals the
  // scaled canonical values the refactor scatters).
  int32_t dense_rule = 0;
  double dense_tol = 1.0;
  int32_t* dense_key = nullptr;
  int32_t row_scale = 0;
  double* rs = nullptr;
  double* svals = nullptr;
  int32_t* piv = nullptr;  // per row: block-local source row of the pivot gather
  int32_t* pinv = nullptr; // per row: block-local position the gather moves it to (inverse of piv)
  int32_t* block_status = nullptr;  // per block (dense status)
  int32_t* prec_of_block = nullptr; // per block: 0 fp64 factors, 1 fp32
  std::vector<int32_t> h_prec;
  // sparse
  std::vector<SparseDev> sp;
  std::vector<int32_t> sp_order;  // sparse blocks by decreasing tail size (refactor order)
  std::vector<int32_t> sp_of_block;
  void* xscratch = nullptr;
  int32_t nslots = 0;
  // solve
  std::vector<Phase> phases;
  int32_t nticket = 0;
  Work work;                // the plan's own solve vectors (work.norms aliases norms)
  double* norms = nullptr;  // [kAnormSlot]: ||A||_inf of the current values
  bool factored = false;
  bool has_values = false;  // P.vals holds the values of a previous refactor
  // Launch schedule, picked per plan by the autotuner (every choice gives the
  // same results bit for bit; only the time differs).
  struct Schedule {
    int32_t small_tile_below = 0;  // trailing updates of size < this use 64x64 tiles
    bool dense_lpt = false;        // dense blocks launched largest first
    int32_t seg_grid = 1024;       // persistent workgroups of the dense-block solve
    bool factor_tuned[2] = {false, false};
    bool solve_tuned = false;
  } sched;
  int32_t* d_lpt = nullptr;  // dense blocks by size, largest first
  // Refinement steps of the last four solves (-1: none yet); their minimum is
  // the length of the next device-side loop, so that steps it runs past the
  // stop (each a whole correction solve, gated off) stay rare.
  int32_t ir_steps[4] = {-1, -1, -1, -1};
  int32_t ir_next = 0;
  // Captured launch sequences, replayed per call: the refactor per precision,
  // and the solve per (column count, vector pair). Lazily instantiated.
  hipGraphExec_t refactor_graph[2] = {nullptr, nullptr};
  std::map<std::vector<int32_t>, hipGraphExec_t> solve_graph;
  void drop_graphs() {
    for (auto& e : refactor_graph)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
    for (auto& kv : solve_graph) (void)hipGraphExecDestroy(kv.second);
    solve_graph.clear();
  }
  ~Plan() { drop_graphs(); }
};

}  // namespace

// Scratch of the no-allocation solve (bdlu_gpu_solve_noalloc): the solve's
// per-call vectors and counters, carved from one array in the code object's
// device memory, so that entry point calls no allocator. Every GPU the library
// runs on holds its own copy for the life of the process; host code reaches it
// only through the per-device address hipGetSymbolAddress returns. Calls that
// use it at the same time on one GPU race.
constexpr int64_t kScratchDoubles = int64_t(16) << 20;  // 128 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_addr[kMaxDevices] = {};

// The device address of `device`'s scratch array, looked up once per device
// with that device current. Returns nullptr if the lookup fails.
double* scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    g_scratch_addr[device] = static_cast<double*>(ptr);
  }
  return g_scratch_addr[device];
}

}  // namespace

struct bdlu_ctx {
  // Device address of this device's no-allocation scratch (nullptr if unavailable).
  double* scratch = nullptr;
  // Largest |L| a refactorization on the fixed pivot order may produce before
  // the block is refactored in FP64 (FP32 runs) or re-pivoted (FP64 runs).
  double fp32_growth = 1e4;
  double fp64_growth = 1e8;
  bool prof = false;
  bool graphs = true;  // capture launch sequences into HIP graphs (BLOCKDAGLU_GPU_GRAPH=0 disables)
  bool device_ir = true;  // refinement stop tests on the device (BLOCKDAGLU_GPU_IR_DEVICE=0 disables)
  bool autotune = true;  // time schedule candidates on a plan's first calls (BLOCKDAGLU_GPU_AUTOTUNE=0 disables)
  double prof_ms[BDLU_PROF_KINDS] = {0};
  int64_t prof_n[BDLU_PROF_KINDS] = {0};
  double prof_flops[BDLU_PROF_KINDS] = {0};
  double prof_bytes[BDLU_PROF_KINDS] = {0};
  int device = 0;
  uint64_t budget = 0;
  uint64_t used = 0;
  hipStream_t stream = nullptr;
  // Side streams (default priority) for independent parts of a refactor: the
  // dense-block LU and the dense tails of different sparse blocks. Joined back
  // into `stream` with events, so callers see one ordered stream.
  static constexpr int kSide = 4;
  hipStream_t side[kSide] = {};
  bool fork = true;  // BLOCKDAGLU_GPU_STREAMS=1 keeps every launch on one stream
  // Dense-tail launch structure (BLOCKDAGLU_GPU_LOOKAHEAD): 2 (default) paired
  // panels with K = 128 bulk updates and the next diagonal LU fused into them,
  // 1 one bulk update per panel, 0 the next diagonal LU in its own launch.
  int lookahead = 2;
  // Largest persistent grid of the dense-block solve (BLOCKDAGLU_GPU_SEG_GRID_MAX):
  // its workgroups wait for one another, so where several contexts share the
  // GPU (serving) a smaller grid leaves room for the others' work.
  int32_t seg_grid_max = 1024;
  std::vector<hipEvent_t> events;  // fork/join points of a refactor
  hipEvent_t event(std::size_t i) {
    while (events.size() <= i) {
      hipEvent_t e = nullptr;
      HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      events.push_back(e);
    }
    return events[i];
  }
  std::mutex mu;
  std::string last_error;
  uint64_t next_id = 1;
  std::map<uint64_t, std::unique_ptr<Plan>> plans;
};

namespace {

// Event pairs recorded on the engine stream while profiling.
struct ProfEvents {
  std::vector<hipEvent_t> ev[BDLU_PROF_KINDS];
  std::vector<double> gemm_flops, gemm_bytes;
  void mark(int kind, hipStream_t st) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) == hipSuccess) {
      (void)hipEventRecord(e, st);
      ev[kind].push_back(e);
    }
  }
  // Call after the stream is synchronized.
  void collect(bdlu_ctx* c) {
    for (int k = 0; k < BDLU_PROF_KINDS; ++k) {
      for (std::size_t i = 0; i + 1 < ev[k].size(); i += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[k][i], ev[k][i + 1]) == hipSuccess) {
          c->prof_ms[k] += ms;
          c->prof_n[k] += 1;
        }
      }
      for (hipEvent_t e : ev[k]) (void)hipEventDestroy(e);
      ev[k].clear();
    }
    for (double f : gemm_flops) c->prof_flops[BDLU_PROF_TAIL_GEMM] += f;
    for (double b : gemm_bytes) c->prof_bytes[BDLU_PROF_TAIL_GEMM] += b;
    gemm_flops.clear();
    gemm_bytes.clear();
  }
};

int set_err(bdlu_ctx* c, int code, const std::string& m) {
  if (c) c->last_error = m;
  return code;
}

// Dense triangular-solve structures of one sparse block, from its patterns.
void build_sparse_solve(const bdlu_sparse_desc& d, std::vector<int32_t>& rowperm, std::vector<int32_t>& pos,
                        std::vector<int32_t>& lr_ptr, std::vector<int32_t>& lr_col, std::vector<int32_t>& lr_vidx,
                        std::vector<int32_t>& fl_ptr, std::vector<int32_t>& fl_rows,
                        std::vector<int32_t>& ur_ptr, std::vector<int32_t>& ur_col,
                        std::vector<int32_t>& ur_vidx, std::vector<int32_t>& bl_ptr,
                        std::vector<int32_t>& bl_rows) {
  const int32_t s = d.s, k0 = d.k0, m = d.m;
  rowperm.resize(s);
  pos.assign(s, -1);
  for (int32_t k = 0; k < k0; ++k) rowperm[k] = d.prow[k];
  for (int32_t t = 0; t < m; ++t) rowperm[k0 + t] = d.trow[d.tperm[t]];
  for (int32_t k = 0; k < s; ++k) pos[rowperm[k]] = k;
  // L by position rows: entries (column k', value index p).
  std::vector<int32_t> cnt(s + 1, 0);
  const int32_t nL = d.l_ptr[k0];
  for (int32_t p = 0; p < nL; ++p) cnt[pos[d.l_row[p]] + 1]++;
  for (int32_t i = 0; i < s; ++i) cnt[i + 1] += cnt[i];
  lr_ptr = cnt;
  lr_col.resize(nL);
  lr_vidx.resize(nL);
  {
    std::vector<int32_t> next(cnt.begin(), cnt.end() - 1);
    for (int32_t kc = 0; kc < k0; ++kc)
      for (int32_t p = d.l_ptr[kc]; p < d.l_ptr[kc + 1]; ++p) {
        const int32_t w = next[pos[d.l_row[p]]]++;
        lr_col[w] = kc;
        lr_vidx[w] = p;
      }
  }
  // Forward levels over positions < k0.
  std::vector<int32_t> lev(k0, 0);
  int32_t maxl = -1;
  for (int32_t k = 0; k < k0; ++k) {
    int32_t l = 0;
    for (int32_t p = lr_ptr[k]; p < lr_ptr[k + 1]; ++p) l = std::max(l, lev[lr_col[p]] + 1);
    lev[k] = l;
    maxl = std::max(maxl, l);
  }
  fl_ptr.assign(maxl + 2, 0);
  for (int32_t k = 0; k < k0; ++k) fl_ptr[lev[k] + 1]++;
  for (int32_t l = 0; l <= maxl; ++l) fl_ptr[l + 1] += fl_ptr[l];
  fl_rows.resize(k0);
  {
    std::vector<int32_t> next(fl_ptr.begin(), fl_ptr.end() - 1);
    for (int32_t k = 0; k < k0; ++k) fl_rows[next[lev[k]]++] = k;
  }
  // U rows k < k0: entries (column position j, value index p).
  const int32_t nU = d.u_ptr[s];
  std::vector<int32_t> uc(k0 + 1, 0);
  for (int32_t p = 0; p < nU; ++p) uc[d.u_idx[p] + 1]++;
  for (int32_t k = 0; k < k0; ++k) uc[k + 1] += uc[k];
  ur_ptr = uc;
  ur_col.resize(nU);
  ur_vidx.resize(nU);
  {
    std::vector<int32_t> next(uc.begin(), uc.end() - 1);
    for (int32_t j = 0; j < s; ++j)
      for (int32_t p = d.u_ptr[j]; p < d.u_ptr[j + 1]; ++p) {
        const int32_t w = next[d.u_idx[p]]++;
        ur_col[w] = j;
        ur_vidx[w] = p;
      }
  }
  std::vector<int32_t> blev(k0, 0);
  maxl = -1;
  for (int32_t k = k0 - 1; k >= 0; --k) {
    int32_t l = 0;
    for (int32_t p = ur_ptr[k]; p < ur_ptr[k + 1]; ++p)
      if (ur_col[p] < k0) l = std::max(l, blev[ur_col[p]] + 1);
    blev[k] = l;
    maxl = std::max(maxl, l);
  }
  bl_ptr.assign(maxl + 2, 0);
  for (int32_t k = 0; k < k0; ++k) bl_ptr[blev[k] + 1]++;
  for (int32_t l = 0; l <= maxl; ++l) bl_ptr[l + 1] += bl_ptr[l];
  bl_rows.resize(k0);
  {
    std::vector<int32_t> next(bl_ptr.begin(), bl_ptr.end() - 1);
    for (int32_t k = 0; k < k0; ++k) bl_rows[next[blev[k]]++] = k;
  }
}

void upload_into(bdlu_ctx* c, const bdlu_plan_desc& d, Plan& P) {
  hipStream_t st = c->stream;
  Alloc& A = P.mem;
  P.n = d.n;
  P.nnz = d.nnz;
  P.nblocks = d.nblocks;
  P.arena_len = d.arena_len;
  const int32_t n = d.n, nb = d.nblocks;
  P.vmap = A.copy(d.vmap, d.nnz, st);
  P.arena = A.get<double>(d.arena_len);
  HIPCHECK(hipMemsetAsync(P.arena, 0, sizeof(double) * std::max<int64_t>(d.arena_len, 1), st));
  P.vals = A.get<double>(d.nnz);
  P.csr_ptr = A.copy(d.csr_ptr, n + 1, st);
  P.csr_col = A.copy(d.csr_col, d.nnz, st);
  // (vector copies are staged by Alloc, so temporaries may go out of scope)
  std::vector<uint8_t> res_flag(static_cast<std::size_t>(n), 0);
  std::vector<int32_t> res_rows;
  {
    constexpr int64_t kLongRow = 64;
    for (int32_t i = 0; i < n; ++i)
      if (d.csr_ptr[i + 1] - d.csr_ptr[i] > kLongRow) {
        res_flag[i] = 1;
        res_rows.push_back(i);
      }
    P.res_nlong = static_cast<int32_t>(res_rows.size());
    if (P.res_nlong > 0) {
      P.res_is_long = A.copy(res_flag, st);
      P.res_long_rows = A.copy(res_rows, st);
    }
  }
  P.P = A.copy(d.P, n, st);
  P.Q = A.copy(d.Q, n, st);
  P.R = A.copy(d.R, nb + 1, st);
  P.off_ptr = A.copy(d.off_ptr, n + 1, st);
  const int64_t noff = d.off_ptr[n];
  P.off_col = A.copy(d.off_col, noff, st);
  P.off_val = P.arena + d.off_val0;
  if (d.dense_rule != 0 && d.dense_rule != 1) throw HipError{"unknown dense pivot rule"};
  if (d.row_scale < 0 || d.row_scale > 2) throw HipError{"unknown row scaling"};
  P.dense_rule = d.dense_rule;
  P.dense_tol = d.dense_tol;
  if (d.dense_rule == 1) {
    if (!d.dense_key) throw HipError{"KLU pivot rule without tie keys"};
    P.dense_key = A.copy(d.dense_key, n, st);
  }
  P.row_scale = d.row_scale;
  if (d.row_scale > 0) {
    P.rs = A.get<double>(n);
    P.svals = A.get<double>(d.nnz);
  }

  // Dense blocks.
  P.ndense = d.ndense;
  P.h_dense_block.assign(d.dense_block, d.dense_block + d.ndense);
  P.h_dense_val0.assign(d.dense_val0, d.dense_val0 + d.ndense);
  std::vector<int64_t> lu0_of_block(nb, 0);
  int64_t lu = 0;
  for (int32_t t = 0; t < d.ndense; ++t) {
    const int32_t b = d.dense_block[t];
    const int32_t s = d.R[b + 1] - d.R[b];
    if (s < 1 || s > 64) throw HipError{"dense block larger than 64"};
    P.h_dense_size.push_back(s);
    P.h_dense_start.push_back(d.R[b]);
    P.h_dense_lu0.push_back(lu);
    lu0_of_block[b] = lu;
    lu += static_cast<int64_t>(s) * s;
  }
  P.lu_len = lu;
  P.d_size = A.copy(P.h_dense_size, st);
  P.d_start = A.copy(P.h_dense_start, st);
  P.d_block = A.copy(P.h_dense_block, st);
  {
    std::vector<int32_t> lpt(P.h_dense_size.size());
    std::iota(lpt.begin(), lpt.end(), 0);
    std::stable_sort(lpt.begin(), lpt.end(),
                     [&](int32_t a, int32_t b) { return P.h_dense_size[a] > P.h_dense_size[b]; });
    P.d_lpt = A.copy(lpt, st);
  }
  P.d_val0 = A.copy(P.h_dense_val0, st);
  P.d_lu0 = A.copy(P.h_dense_lu0, st);
  P.lu0_of_block = A.copy(lu0_of_block, st);
  P.lu64 = A.get<double>(lu);
  P.lu32 = A.get<float>(lu);
  P.minv64 = A.get<double>(lu);
  P.minv32 = A.get<float>(lu);
  P.piv = A.get<int32_t>(n);
  P.pinv = A.get<int32_t>(n);
  P.block_status = A.get<int32_t>(nb);
  P.prec_of_block = A.get<int32_t>(nb);
  P.h_prec.assign(nb, 0);

  // Sparse blocks.
  P.sp_of_block.assign(nb, -1);
  int64_t max_s = 0;
  int32_t max_width = 1;
  for (int32_t i = 0; i < d.nsparse; ++i) {
    const bdlu_sparse_desc& sd = d.sparse[i];
    SparseDev S;
    S.block = sd.block;
    S.start = d.R[sd.block];
    S.s = sd.s;
    S.k0 = sd.k0;
    S.m = sd.m;
    S.ldt = std::max<int64_t>((int64_t(sd.m) + 63) / 64 * 64, 1);
    S.scale = sd.scale;
    S.tol = sd.tol;
    if (sd.s != d.R[sd.block + 1] - d.R[sd.block] || sd.k0 + sd.m != sd.s) throw HipError{"sparse block sizes"};
    P.sp_of_block[sd.block] = i;
    max_s = std::max<int64_t>(max_s, sd.s);
    for (int32_t l = 0; l < sd.nlev; ++l) max_width = std::max(max_width, sd.lev_ptr[l + 1] - sd.lev_ptr[l]);
    max_width = std::max(max_width, sd.m);
    const int32_t ne = sd.a_ptr[sd.s];
    S.nL = sd.l_ptr[sd.k0];
    S.nU = sd.u_ptr[sd.s];
    // Entries by local row for the row scaling, each row's entries in
    // ascending local column order (the order klu_scale sums them in).
    std::vector<int32_t> rs_ptr(sd.s + 1, 0), rs_idx(ne);
    for (int32_t p = 0; p < ne; ++p) rs_ptr[sd.a_row[p] + 1]++;
    for (int32_t r = 0; r < sd.s; ++r) rs_ptr[r + 1] += rs_ptr[r];
    {
      std::vector<int32_t> next(rs_ptr.begin(), rs_ptr.end() - 1), pos_of_col(sd.s);
      for (int32_t j = 0; j < sd.s; ++j) pos_of_col[sd.q[j]] = j;
      for (int32_t c = 0; c < sd.s; ++c) {
        const int32_t j = pos_of_col[c];
        for (int32_t p = sd.a_ptr[j]; p < sd.a_ptr[j + 1]; ++p) rs_idx[next[sd.a_row[p]]++] = p;
      }
    }
    std::vector<int32_t> tpos(sd.m);
    for (int32_t t = 0; t < sd.m; ++t) tpos[sd.tperm[t]] = t;
    std::vector<int32_t> rowperm, pos, lr_ptr, lr_col, lr_vidx, fl_ptr, fl_rows, ur_ptr, ur_col, ur_vidx, bl_ptr,
        bl_rows;
    build_sparse_solve(sd, rowperm, pos, lr_ptr, lr_col, lr_vidx, fl_ptr, fl_rows, ur_ptr, ur_col, ur_vidx, bl_ptr,
                       bl_rows);
    bdk::SpView& v = S.view;
    v.s = sd.s;
    v.k0 = sd.k0;
    v.m = sd.m;
    v.scale = sd.scale;
    v.inv_tol = sd.tol > 0 ? 1.0 / sd.tol : 1e300;
    v.a_val0 = sd.a_val0;
    v.a_ptr = A.copy(sd.a_ptr, sd.s + 1, st);
    v.a_row = A.copy(sd.a_row, ne, st);
    v.prow = A.copy(sd.prow, sd.k0, st);
    v.l_ptr = A.copy(sd.l_ptr, sd.k0 + 1, st);
    v.l_row = A.copy(sd.l_row, S.nL, st);
    v.u_ptr = A.copy(sd.u_ptr, sd.s + 1, st);
    v.u_idx = A.copy(sd.u_idx, S.nU, st);
    v.trow = A.copy(sd.trow, sd.m, st);
    v.tpos = A.copy(tpos, st);
    v.rs_ptr = A.copy(rs_ptr, st);
    v.rs_idx = A.copy(rs_idx, st);
    v.rpos = A.copy(pos, st);
    S.lev_ptr.assign(sd.lev_ptr, sd.lev_ptr + sd.nlev + 1);
    S.lev_col = A.copy(sd.lev_col, sd.k0, st);
    S.rowperm = A.copy(rowperm, st);
    S.q = A.copy(sd.q, sd.s, st);
    S.fl_ptr = fl_ptr;
    S.bl_ptr = bl_ptr;
    S.fl_rows = A.copy(fl_rows, st);
    std::vector<int32_t> tl_rows(sd.m);
    for (int32_t t = 0; t < sd.m; ++t) tl_rows[t] = sd.k0 + t;
    S.tl_rows = A.copy(tl_rows, st);
    {
      // Tail columns whose reach in the sparse part exceeds kLongReach: their
      // reach positions grouped by forward level, for k_sp_tail_long.
      constexpr int32_t kLongReach = 64;
      const int32_t nlev = static_cast<int32_t>(fl_ptr.size()) - 1;
      std::vector<int32_t> lev_of(std::max(sd.k0, 1), 0);
      for (int32_t l = 0; l < nlev; ++l)
        for (int32_t e = fl_ptr[l]; e < fl_ptr[l + 1]; ++e) lev_of[fl_rows[e]] = l;
      std::vector<int32_t> is_long(std::max(sd.m, 1), 0), long_t, reach, levp;
      for (int32_t t = 0; t < sd.m && nlev > 0; ++t) {
        const int32_t j = sd.k0 + t;
        const int32_t p0 = sd.u_ptr[j], p1 = sd.u_ptr[j + 1];
        if (p1 - p0 <= kLongReach) continue;
        is_long[t] = 1;
        long_t.push_back(t);
        std::vector<int32_t> cnt(nlev + 1, 0);
        for (int32_t p = p0; p < p1; ++p) cnt[lev_of[sd.u_idx[p]] + 1]++;
        for (int32_t l = 0; l < nlev; ++l) cnt[l + 1] += cnt[l];
        const int32_t base = static_cast<int32_t>(reach.size());
        for (int32_t l = 0; l <= nlev; ++l) levp.push_back(base + cnt[l]);
        reach.resize(base + (p1 - p0));
        for (int32_t p = p0; p < p1; ++p) reach[base + cnt[lev_of[sd.u_idx[p]]]++] = sd.u_idx[p];
      }
      if (!long_t.empty()) {
        bdk::SpTailLong& lg = S.tail_long;
        lg.nlong = static_cast<int32_t>(long_t.size());
        lg.nlev = nlev;
        lg.is_long = A.copy(is_long, st);
        lg.long_t = A.copy(long_t, st);
        lg.reach = A.copy(reach, st);
        lg.lev = A.copy(levp, st);
      }
      // The same for the sparse columns, per refactor level.
      std::vector<int32_t> col_long(std::max(sd.k0, 1), 0);
      S.cols_long.assign(sd.nlev, bdk::SpColsLong{});
      bool any = false;
      for (int32_t l = 0; l < sd.nlev && nlev > 0; ++l) {
        std::vector<int32_t> lj, rch, lp;
        for (int32_t e = sd.lev_ptr[l]; e < sd.lev_ptr[l + 1]; ++e) {
          const int32_t j = sd.lev_col[e];
          const int32_t p0 = sd.u_ptr[j], p1 = sd.u_ptr[j + 1];
          if (p1 - p0 <= kLongReach) continue;
          col_long[j] = 1;
          lj.push_back(j);
          std::vector<int32_t> cnt(nlev + 1, 0);
          for (int32_t p = p0; p < p1; ++p) cnt[lev_of[sd.u_idx[p]] + 1]++;
          for (int32_t q = 0; q < nlev; ++q) cnt[q + 1] += cnt[q];
          const int32_t base = static_cast<int32_t>(rch.size());
          for (int32_t q = 0; q <= nlev; ++q) lp.push_back(base + cnt[q]);
          rch.resize(base + (p1 - p0));
          for (int32_t p = p0; p < p1; ++p) rch[base + cnt[lev_of[sd.u_idx[p]]]++] = sd.u_idx[p];
        }
        if (lj.empty()) continue;
        any = true;
        bdk::SpColsLong& g = S.cols_long[l];
        g.nlong = static_cast<int32_t>(lj.size());
        g.nlev = nlev;
        g.long_j = A.copy(lj, st);
        g.reach = A.copy(rch, st);
        g.lev = A.copy(lp, st);
      }
      if (any) {
        const int32_t* d_col_long = A.copy(col_long, st);
        for (bdk::SpColsLong& g : S.cols_long) g.is_long = d_col_long;
      }
    }
    S.bl_rows = A.copy(bl_rows, st);
    S.lr_ptr = A.copy(lr_ptr, st);
    S.lr_col = A.copy(lr_col, st);
    S.lr_vidx = A.copy(lr_vidx, st);
    S.tail_long.lr_ptr = S.lr_ptr;
    S.tail_long.lr_col = S.lr_col;
    S.tail_long.lr_vidx = S.lr_vidx;
    for (bdk::SpColsLong& g : S.cols_long) {
      g.lr_ptr = S.lr_ptr;
      g.lr_col = S.lr_col;
      g.lr_vidx = S.lr_vidx;
    }
    S.ur_ptr = A.copy(ur_ptr, st);
    S.ur_col = A.copy(ur_col, st);
    S.ur_vidx = A.copy(ur_vidx, st);
    S.Lx64 = A.get<double>(S.nL);
    S.Ux64 = A.get<double>(S.nU);
    S.Ud64 = A.get<double>(sd.k0);
    S.S64 = A.get<double>(static_cast<std::size_t>(S.ldt) * sd.m);
    S.Lx32 = A.get<float>(S.nL);
    S.Ux32 = A.get<float>(S.nU);
    S.Ud32 = A.get<float>(sd.k0);
    S.S32 = A.get<float>(static_cast<std::size_t>(S.ldt) * sd.m);
    S.rs = A.get<double>(sd.s);
    S.y = A.get<double>(static_cast<std::size_t>(sd.s) * kMaxRhs);
    const std::size_t npan = (static_cast<std::size_t>(sd.m) + 63) / 64;
    S.inv64 = A.get<double>(2 * 64 * 64 * std::max<std::size_t>(npan, 1));
    S.inv32 = A.get<float>(2 * 64 * 64 * std::max<std::size_t>(npan, 1));
    S.cpl64 = A.get<double>(2 * 64 * 64 * std::max<std::size_t>(npan, 1));
    S.cpl32 = A.get<float>(2 * 64 * 64 * std::max<std::size_t>(npan, 1));
    S.ty = A.get<double>(static_cast<std::size_t>(std::max(sd.m, 1)) * kMaxRhs);
    S.status = A.get<int32_t>(1);
    P.sp.push_back(std::move(S));
  }
  P.sp_order.resize(P.sp.size());
  for (std::size_t i = 0; i < P.sp.size(); ++i) P.sp_order[i] = static_cast<int32_t>(i);
  std::stable_sort(P.sp_order.begin(), P.sp_order.end(),
                   [&](int32_t x, int32_t y) { return P.sp[x].m > P.sp[y].m; });
  if (!P.sp.empty()) {
    // Column slots of the sparse phases: each sparse or tail column is a serial
    // chain of dependent memory steps, so the phase time falls with the number
    // of columns in flight. One slot holds one s-sized work vector; slots are
    // sized from a 768 MB scratch budget, at least 1024 and at most 8192.
    constexpr int64_t kScratchBudget = int64_t(768) << 20;
    const int64_t by_budget = std::max<int64_t>(1024, kScratchBudget / (int64_t(sizeof(double)) * max_s));
    P.nslots = static_cast<int32_t>(std::min<int64_t>(std::min<int64_t>(8192, by_budget), max_width));
    const std::size_t xbytes = sizeof(double) * static_cast<std::size_t>(P.nslots) * max_s;
    P.xscratch = A.get<char>(xbytes);
    HIPCHECK(hipMemsetAsync(P.xscratch, 0, xbytes, st));
  }

  // Solve phases: blocks in descending order; runs of dense blocks form one
  // persistent-kernel segment, each sparse block is its own phase. Inside a
  // segment the tickets follow the solve levels (level(b) = 1 + max level of
  // the blocks b reads): a topological order in which blocks that are ready
  // come before blocks still waiting. Within a level, blocks with the longest
  // chain of dependents after them (height) come first, so the critical path
  // is claimed as soon as it can run; ties keep descending block id.
  std::vector<int32_t> lev(static_cast<std::size_t>(nb), 0);
  for (int32_t b = nb - 1; b >= 0; --b)
    for (int32_t e = d.dep_ptr[b]; e < d.dep_ptr[b + 1]; ++e) lev[b] = std::max(lev[b], lev[d.dep_blk[e]] + 1);
  std::vector<int32_t> height(static_cast<std::size_t>(nb), 0);  // blocks read only blocks of higher id
  for (int32_t c = 0; c < nb; ++c)
    for (int32_t e = d.dep_ptr[c]; e < d.dep_ptr[c + 1]; ++e)
      height[d.dep_blk[e]] = std::max(height[d.dep_blk[e]], height[c] + 1);
  std::vector<int32_t> cur;
  auto flush = [&]() {
    if (cur.empty()) return;
    std::stable_sort(cur.begin(), cur.end(), [&](int32_t x, int32_t y) {
      return lev[x] != lev[y] ? lev[x] < lev[y] : height[x] > height[y];
    });
    Phase ph;
    ph.sparse = false;
    ph.index = static_cast<int32_t>(P.phases.size());
    ph.ntasks = static_cast<int32_t>(cur.size());
    ph.tasks = A.copy(cur, st);
    P.phases.push_back(ph);
    cur.clear();
  };
  for (int32_t b = nb - 1; b >= 0; --b) {
    if (P.sp_of_block[b] >= 0) {
      flush();
      Phase ph;
      ph.sparse = true;
      ph.index = P.sp_of_block[b];
      P.phases.push_back(ph);
    } else {
      cur.push_back(b);
    }
  }
  flush();
  P.sched.seg_grid = std::min(P.sched.seg_grid, c->seg_grid_max);
  P.nticket = static_cast<int32_t>(P.phases.size()) + 2 * d.nsparse;
  Work& W = P.work;
  W.tickets = A.get<int32_t>(P.nticket);
  W.err = A.get<int32_t>(1);
  const std::size_t nk = static_cast<std::size_t>(n) * kMaxRhs;
  W.b = A.get<double>(nk);
  W.x = A.get<double>(nk);
  W.r = A.get<double>(nk);
  W.d = A.get<double>(nk);
  W.bperm = A.get<double>(nk);
  W.xperm = A.get<double>(nk);
  W.cw = A.get<double>(nk);
  W.partials = A.get<double>(512 * 6 * kMaxRhs);
  W.irs = A.get<double>(bdk::kIrStateLen);
  P.norms = A.get<double>(kAnormSlot + 8);
  W.norms = P.norms;
  HIPCHECK(hipStreamSynchronize(st));
}

std::unique_ptr<Plan> upload(bdlu_ctx* c, const bdlu_plan_desc& d) {
  auto pl = std::make_unique<Plan>();
  try {
    upload_into(c, d, *pl);
  } catch (...) {
    // Copies already queued still read the staged host bytes: wait for them
    // before those and the device buffers go away.
    (void)hipStreamSynchronize(c->stream);
    pl->mem.release();
    throw;
  }
  pl->mem.drop_staged();
  return pl;
}

// Solve A x = b for the resident factors; b and x are device vectors of length n.
// Runs `body` (a fixed sequence of kernel launches on `st`): the first time it is
// captured into a graph and instantiated into `exec`, then the graph is
// replayed. With use == false the launches are issued directly.
template <class F>
void run_captured(hipStream_t st, hipGraphExec_t& exec, bool use, F&& body) {
  if (!use) {
    body();
    return;
  }
  if (!exec) {
    HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipGraph_t graph = nullptr;
    try {
      body();
    } catch (...) {
      (void)hipStreamEndCapture(st, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    HIPCHECK(hipStreamEndCapture(st, &graph));
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      exec = nullptr;
      throw HipError{std::string("hipGraphInstantiate: ") + hipGetErrorString(e)};
    }
  }
  HIPCHECK(hipGraphLaunch(exec, st));
}

// Solves kt columns at once: b and x are n x kt sets (column stride n); W
// supplies the permuted vectors and the counters.
void block_solve_launches(Plan& P, const Work& W, hipStream_t st, const double* b, double* x, int32_t kt);
// With graphs (the plan's own work set only), the launches are captured once
// per key and replayed.
void block_solve(Plan& P, hipStream_t st, const double* b, double* x, int32_t kt, bool graphs) {
  // The launch sequence depends on the column count, the vector pair and the
  // precision each sparse block was last factored in (its factor arrays).
  std::vector<int32_t> key{kt, b == P.work.b ? 0 : 1};
  for (const SparseDev& S : P.sp) key.push_back(S.prec);
  hipGraphExec_t& exec = P.solve_graph[key];
  run_captured(st, exec, graphs, [&] { block_solve_launches(P, P.work, st, b, x, kt); });
}

void block_solve_launches(Plan& P, const Work& W, hipStream_t st, const double* b, double* x, int32_t kt) {
  const int64_t ld = P.n;
  bdk::launch_fill_i32(st, W.tickets, P.nticket, 0);
  bdk::launch_fill_sentinel(st, W.xperm, int64_t(P.n) * kt);
  bdk::launch_permute(st, P.n, kt, P.P, b, ld, W.bperm, ld, P.row_scale ? P.rs : nullptr);
  int32_t tix = 0;
  for (const Phase& ph : P.phases) {
    if (!ph.sparse) {
      // Each dense block's factors are read in the precision it was factored in.
      bdk::launch_dense_seg_solve(st, kt, ph.ntasks, ph.tasks, P.R, P.lu0_of_block, P.minv64, P.minv32,
                                  P.prec_of_block, P.pinv, P.off_ptr, P.off_col, P.off_val, W.tickets + tix,
                                  W.bperm, W.xperm, ld, W.err, P.sched.seg_grid);
      ++tix;
      continue;
    }
    SparseDev& S = P.sp[ph.index];
    ++tix;
    const int64_t ldy = S.s, ldt = std::max(S.m, 1);
    bdk::launch_sp_gather(st, S.s, S.start, kt, S.rowperm, S.scale ? S.rs : nullptr, P.off_ptr, P.off_col, P.off_val,
                          W.bperm, W.xperm, ld, S.y, ldy);
    auto run = [&](auto* Lx, auto* Ux, auto* Ud, auto* Sm, auto* inv, auto* cpl) {
      using T = std::remove_pointer_t<decltype(Lx)>;
      for (std::size_t l = 0; l + 1 < S.fl_ptr.size(); ++l)
        bdk::launch_sp_rows<T>(st, kt, S.fl_ptr[l + 1] - S.fl_ptr[l], S.fl_rows + S.fl_ptr[l], S.lr_ptr, S.lr_col,
                               S.lr_vidx, Lx, nullptr, S.y, ldy);
      bdk::launch_sp_rows<T>(st, kt, S.m, S.tl_rows, S.lr_ptr, S.lr_col, S.lr_vidx, Lx, nullptr, S.y, ldy);
      // Tail: forward into ty, backward back into y (each output starts unpublished).
      bdk::launch_fill_sentinel_cols(st, S.ty, S.m, kt, ldt);
      bdk::launch_tail_trsv<T>(st, kt, S.m, Sm, S.ldt, inv, cpl, S.y + S.k0, ldy, S.ty, ldt, false, W.tickets + tix,
                               W.err);
      bdk::launch_fill_sentinel_cols(st, S.y + S.k0, S.m, kt, ldy);
      bdk::launch_tail_trsv<T>(st, kt, S.m, Sm, S.ldt, inv, cpl, S.ty, ldt, S.y + S.k0, ldy, true, W.tickets + tix + 1,
                               W.err);
      for (std::size_t l = 0; l + 1 < S.bl_ptr.size(); ++l)
        bdk::launch_sp_rows<T>(st, kt, S.bl_ptr[l + 1] - S.bl_ptr[l], S.bl_rows + S.bl_ptr[l], S.ur_ptr, S.ur_col,
                               S.ur_vidx, Ux, Ud, S.y, ldy);
    };
    if (S.prec == 0) run(S.Lx64, S.Ux64, S.Ud64, S.S64, S.inv64, S.cpl64);
    else run(S.Lx32, S.Ux32, S.Ud32, S.S32, S.inv32, S.cpl32);
    tix += 2;
    bdk::launch_sp_scatter_x(st, S.s, S.start, kt, S.q, S.y, ldy, W.xperm, ld);
  }
  bdk::launch_unpermute(st, P.n, kt, P.Q, W.xperm, ld, x, ld);
}

// Autotuner: times each candidate schedule once (after one warm-up run) with
// events on the engine stream and keeps the fastest. Candidates only change
// tile sizes, workgroup order and grid size, never the arithmetic.
float time_launches(hipStream_t st, const std::function<void()>& f) {
  hipEvent_t a, b;
  HIPCHECK(hipEventCreate(&a));
  HIPCHECK(hipEventCreate(&b));
  HIPCHECK(hipEventRecord(a, st));
  f();
  HIPCHECK(hipEventRecord(b, st));
  HIPCHECK(hipEventSynchronize(b));
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms;
}

void autotune_factor(Plan& P, hipStream_t st, const std::function<void()>& launches) {
  int32_t max_m = 0;
  for (const SparseDev& S : P.sp) max_m = std::max(max_m, S.m);
  std::vector<int32_t> tiles{0};
  if (max_m > 256) tiles = {0, 768, 2048, 1 << 30};
  std::vector<bool> lpt{false};
  if (P.ndense > 1) lpt = {false, true};
  launches();  // warm-up
  float best = 1e30f;
  Plan::Schedule pick = P.sched;
  for (int32_t t : tiles)
    for (bool l : lpt) {
      P.sched.small_tile_below = t;
      P.sched.dense_lpt = l;
      const float ms = time_launches(st, launches);
      if (ms < best) {
        best = ms;
        pick = P.sched;
      }
    }
  P.sched.small_tile_below = pick.small_tile_below;
  P.sched.dense_lpt = pick.dense_lpt;
}

void autotune_solve(Plan& P, hipStream_t st, int32_t kt, int32_t grid_max) {
  std::vector<int32_t> grids;
  for (int32_t g : {256, 512, 1024})
    if (g <= grid_max) grids.push_back(g);
  if (grids.empty()) grids.push_back(grid_max);
  auto run = [&] { block_solve_launches(P, P.work, st, P.work.b, P.work.x, kt); };
  run();  // warm-up
  float best = 1e30f;
  int32_t pick = P.sched.seg_grid;
  for (int32_t g : grids) {
    P.sched.seg_grid = g;
    const float ms = time_launches(st, run);
    if (ms < best) {
      best = ms;
      pick = g;
    }
  }
  P.sched.seg_grid = pick;
}

// Sparse leading columns and the assembly of the dense tail (these share the
// plan's column scratch, so sparse blocks take this phase one at a time).
template <typename T>
void factor_sparse_cols(Plan& P, SparseDev& S, hipStream_t st, T* Lx, T* Ux, T* Ud, T* Sm, ProfEvents* pe,
                        double growth_limit) {
  bdk::SpView view = S.view;
  view.inv_tol = growth_limit;
  bdk::launch_fill_i32(st, S.status, 1, 0);
  if (pe) pe->mark(BDLU_PROF_SPARSE, st);
  if (S.scale) bdk::launch_rowscale(st, view, P.arena, S.rs);
  T* xs = static_cast<T*>(P.xscratch);
  for (std::size_t l = 0; l + 1 < S.lev_ptr.size(); ++l)
    bdk::launch_sp_cols<T>(st, view, P.arena, S.rs, S.lev_ptr[l + 1] - S.lev_ptr[l], S.lev_col + S.lev_ptr[l], xs,
                           P.nslots, Lx, Ux, Ud, S.status,
                           l < S.cols_long.size() && S.cols_long[l].nlong > 0 ? &S.cols_long[l] : nullptr);
  if (sizeof(T) == 8) bdk::launch_fill_f64(st, reinterpret_cast<double*>(Sm), S.ldt * S.m, 0.0);
  else bdk::launch_fill_f32(st, reinterpret_cast<float*>(Sm), S.ldt * S.m, 0.0f);
  bdk::launch_sp_tail<T>(st, view, P.arena, S.rs, xs, P.nslots, Lx, Ux, Sm, S.ldt,
                         S.tail_long.nlong > 0 ? &S.tail_long : nullptr);
  if (pe) pe->mark(BDLU_PROF_SPARSE, st);
}

// LU of the dense tail, then its band couplings for the solve (touches only the
// block's own tail, inverses, couplings and status, so tails of different
// sparse blocks may run concurrently).
template <typename T>
void factor_sparse_tail(Plan& P, SparseDev& S, hipStream_t st, T* Sm, ProfEvents* pe, double growth_limit,
                        int lookahead = 0) {
  bdk::TailTimer tt;
  if (pe) {
    tt.gemm_events = &pe->ev[BDLU_PROF_TAIL_GEMM];
    tt.gemm_flops = &pe->gemm_flops;
    tt.gemm_bytes = &pe->gemm_bytes;
    tt.panel_events = &pe->ev[BDLU_PROF_TAIL_PANEL];
  }
  T* inv = sizeof(T) == 8 ? reinterpret_cast<T*>(S.inv64) : reinterpret_cast<T*>(S.inv32);
  bdk::launch_tail_lu<T>(st, S.m, Sm, S.ldt, growth_limit, S.status, inv, pe ? &tt : nullptr,
                         P.sched.small_tile_below, lookahead);
  T* cpl = sizeof(T) == 8 ? reinterpret_cast<T*>(S.cpl64) : reinterpret_cast<T*>(S.cpl32);
  bdk::launch_tail_coupling<T>(st, S.m, Sm, S.ldt, inv, cpl);
}

// Refactorization of a resident plan: the body of bdlu_gpu_refactor, run with
// the context lock held. values == nullptr refactors the values of the
// previous call again (the refinement uses it to redo FP32 factors in FP64).
int refactor_impl(bdlu_ctx* c, Plan& P, const double* values, int where, int precision, bdlu_factor_stats& fs) {
  hipStream_t st = c->stream;
  fs = bdlu_factor_stats{0, 0, 0, -1, 0.0};
  if (values) {
    HIPCHECK(hipMemcpyAsync(P.vals, values, sizeof(double) * P.nnz,
                            where == BDLU_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    P.has_values = true;
  } else if (!P.has_values) {
    return set_err(c, BDLU_E_INVALID, "no values to refactor");
  }
  const bool fp32 = precision == BDLU_FP32;
  ProfEvents pev;
  ProfEvents* pe = c->prof ? &pev : nullptr;
  for (SparseDev& S : P.sp) S.prec = fp32 ? 1 : 0;
  auto launches = [&] {
    if (P.row_scale) {
      bdk::launch_row_scale(st, P.n, P.csr_ptr, P.vals, P.row_scale, P.rs, P.svals);
      bdk::launch_scatter(st, P.nnz, P.svals, P.vmap, P.arena);
    } else {
      bdk::launch_scatter(st, P.nnz, P.vals, P.vmap, P.arena);
    }
    bdk::launch_anorm(st, P.n, P.csr_ptr, P.vals, P.work.partials, P.norms + kAnormSlot);
    // Dense blocks (dynamic partial pivoting, reference semantics), on a side
    // stream unless profiling (the profile events time one stream).
    const bool fork = c->fork && !pe;
    hipStream_t ds = st;
    if (fork) {
      HIPCHECK(hipEventRecord(c->event(0), st));
      HIPCHECK(hipStreamWaitEvent(c->side[0], c->event(0), 0));
      ds = c->side[0];
    }
    if (pe) pe->mark(BDLU_PROF_DENSE_LU, st);
    const int32_t* order = P.sched.dense_lpt ? P.d_lpt : nullptr;
    if (fp32) {
      bdk::launch_dense_lu<float>(ds, P.ndense, P.d_size, P.d_start, P.d_val0, P.d_lu0, P.arena, P.lu32, P.piv, P.pinv,
                                  P.block_status, P.d_block, order, P.dense_tol, P.dense_key);
      bdk::launch_dense_inv<float>(ds, P.ndense, P.d_size, P.d_lu0, P.lu32, P.minv32);
    } else {
      bdk::launch_dense_lu<double>(ds, P.ndense, P.d_size, P.d_start, P.d_val0, P.d_lu0, P.arena, P.lu64, P.piv, P.pinv,
                                   P.block_status, P.d_block, order, P.dense_tol, P.dense_key);
      bdk::launch_dense_inv<double>(ds, P.ndense, P.d_size, P.d_lu0, P.lu64, P.minv64);
    }
    if (pe) pe->mark(BDLU_PROF_DENSE_LU, st);
    if (fork) HIPCHECK(hipEventRecord(c->event(1), ds));
    // Sparse blocks, largest tail first. With several blocks, each tail goes
    // to the least loaded side stream (load: tail size cubed) once its
    // columns are done; the column phases stay on the main stream.
    const bool spread = fork && P.sp.size() > 1;
    double load[bdlu_ctx::kSide] = {0};
    bool used[bdlu_ctx::kSide] = {false};
    for (std::size_t q = 0; q < P.sp_order.size(); ++q) {
      SparseDev& S = P.sp[P.sp_order[q]];
      if (fp32) factor_sparse_cols<float>(P, S, st, S.Lx32, S.Ux32, S.Ud32, S.S32, pe, c->fp32_growth);
      else factor_sparse_cols<double>(P, S, st, S.Lx64, S.Ux64, S.Ud64, S.S64, pe, c->fp64_growth);
      hipStream_t ts = st;
      if (spread && S.m > 0) {
        int pick = 1;
        for (int k = 2; k < bdlu_ctx::kSide; ++k)
          if (load[k] < load[pick]) pick = k;
        load[pick] += double(S.m) * S.m * S.m;
        used[pick] = true;
        HIPCHECK(hipEventRecord(c->event(2 + q), st));
        HIPCHECK(hipStreamWaitEvent(c->side[pick], c->event(2 + q), 0));
        ts = c->side[pick];
      }
      if (fp32) factor_sparse_tail<float>(P, S, ts, S.S32, pe, c->fp32_growth, c->lookahead);
      else factor_sparse_tail<double>(P, S, ts, S.S64, pe, c->fp64_growth, c->lookahead);
    }
    if (fork) {
      // Join: the main stream waits for the dense-block LU and every side stream used.
      HIPCHECK(hipStreamWaitEvent(st, c->event(1), 0));
      for (int k = 1; k < bdlu_ctx::kSide; ++k) {
        if (!used[k]) continue;
        const std::size_t e = 2 + P.sp_order.size() + k;
        HIPCHECK(hipEventRecord(c->event(e), c->side[k]));
        HIPCHECK(hipStreamWaitEvent(st, c->event(e), 0));
      }
    }
  };
  const int pi = fp32 ? 1 : 0;
  if (c->autotune && !pe && !P.sched.factor_tuned[pi]) {
    autotune_factor(P, st, launches);
    P.sched.factor_tuned[pi] = true;
    P.drop_graphs();  // recapture with the chosen schedule
  }
  run_captured(st, P.refactor_graph[pi], c->graphs && !pe, launches);
  // Collect statuses.
  std::vector<int32_t> bst(P.nblocks, 0);
  if (P.ndense) HIPCHECK(hipMemcpyAsync(bst.data(), P.block_status, sizeof(int32_t) * P.nblocks,
                                        hipMemcpyDeviceToHost, st));
  std::vector<int32_t> sst(P.sp.size(), 0);
  for (std::size_t i = 0; i < P.sp.size(); ++i)
    HIPCHECK(hipMemcpyAsync(&sst[i], P.sp[i].status, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  if (pe) pe->collect(c);
  std::fill(P.h_prec.begin(), P.h_prec.end(), fp32 ? 1 : 0);
  // FP32 dense blocks that failed the small-pivot check: redo in FP64.
  std::vector<int32_t> redo;
  for (int32_t t = 0; t < P.ndense; ++t) {
    const int32_t b = P.h_dense_block[t];
    if (bst[b] & bdk::kStZeroPivot) {
      fs.failed_blocks++;
      if (fs.first_bad_block < 0) fs.first_bad_block = b;
    } else if (fp32 && (bst[b] & bdk::kStCheck)) {
      redo.push_back(t);
    }
  }
  if (!redo.empty()) {
    std::vector<int32_t> sz, sta, blk;
    std::vector<int64_t> v0, l0;
    for (int32_t t : redo) {
      sz.push_back(P.h_dense_size[t]);
      sta.push_back(P.h_dense_start[t]);
      blk.push_back(P.h_dense_block[t]);
      v0.push_back(P.h_dense_val0[t]);
      l0.push_back(P.h_dense_lu0[t]);
      P.h_prec[P.h_dense_block[t]] = 0;
    }
    Alloc tmp;
    int32_t* dsz = tmp.copy(sz, st);
    int32_t* dst_ = tmp.copy(sta, st);
    int32_t* dbl = tmp.copy(blk, st);
    int64_t* dv0 = tmp.copy(v0, st);
    int64_t* dl0 = tmp.copy(l0, st);
    bdk::launch_dense_lu<double>(st, static_cast<int32_t>(redo.size()), dsz, dst_, dv0, dl0, P.arena, P.lu64,
                                 P.piv, P.pinv, P.block_status, dbl, nullptr, P.dense_tol, P.dense_key);
    bdk::launch_dense_inv<double>(st, static_cast<int32_t>(redo.size()), dsz, dl0, P.lu64, P.minv64);
    // The FP64 redo writes its own statuses: a block that is singular in
    // FP64 (exact zero pivot) fails like any other dense block.
    HIPCHECK(hipMemcpyAsync(bst.data(), P.block_status, sizeof(int32_t) * P.nblocks, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    tmp.release();
    fs.fp64_fallback_blocks += static_cast<int32_t>(redo.size());
    for (int32_t t : redo) {
      const int32_t b = P.h_dense_block[t];
      if (bst[b] & bdk::kStZeroPivot) {
        fs.failed_blocks++;
        if (fs.first_bad_block < 0) fs.first_bad_block = b;
      }
    }
  }
  for (std::size_t i = 0; i < P.sp.size(); ++i) {
    SparseDev& S = P.sp[i];
    int32_t s0 = sst[i];
    if (fp32 && s0 != 0) {
      // FP32 refactor failed its check: redo this block in FP64.
      factor_sparse_cols<double>(P, S, st, S.Lx64, S.Ux64, S.Ud64, S.S64, nullptr, c->fp64_growth);
      factor_sparse_tail<double>(P, S, st, S.S64, nullptr, c->fp64_growth);
      S.prec = 0;
      HIPCHECK(hipMemcpyAsync(&s0, S.status, sizeof(int32_t), hipMemcpyDeviceToHost, st));
      HIPCHECK(hipStreamSynchronize(st));
      fs.fp64_fallback_blocks++;
      P.h_prec[S.block] = 0;
    }
    if (s0 & (bdk::kStCheck | bdk::kStZeroPivot | bdk::kStNonFinite)) {
      fs.repivot_blocks++;
      if (fs.first_bad_block < 0) fs.first_bad_block = S.block;
    }
  }
  HIPCHECK(hipMemcpyAsync(P.prec_of_block, P.h_prec.data(), sizeof(int32_t) * P.nblocks, hipMemcpyHostToDevice, st));
  HIPCHECK(hipStreamSynchronize(st));
  P.factored = fs.failed_blocks == 0 && fs.repivot_blocks == 0;
  if (fs.failed_blocks) return set_err(c, BDLU_E_FACTOR, "block factorization failed (zero pivot)");
  if (fs.repivot_blocks) return set_err(c, BDLU_E_REPIVOT, "fixed pivot order failed the pivot check");
  return BDLU_OK;
}

// Solves the chunk of right-hand sides in W.b (kt columns, the first kc real)
// into W.x and, when asked, refines it in FP64. cs receives the chunk's
// statistics. Returns true if the stop rule was met or no refinement was
// asked. With watch_stall, refinement also ends early when a step fails to
// halve the worst stop measure (FP32 factors too inaccurate for this matrix).
// The plan's own work set replays captured graphs; any other set (the
// no-allocation entry point's) is launched directly, without profiling.
// kc columns of n doubles between column-major arrays with leading dimensions
// ldd and lds: one copy when both are contiguous, else one per column.
void copy_columns(double* dst, int64_t ldd, const double* src, int64_t lds, int64_t n, int32_t kc, hipMemcpyKind kind,
                  hipStream_t st) {
  if (ldd == n && lds == n) {
    HIPCHECK(hipMemcpyAsync(dst, src, sizeof(double) * n * kc, kind, st));
    return;
  }
  for (int32_t k = 0; k < kc; ++k) HIPCHECK(hipMemcpyAsync(dst + k * ldd, src + k * lds, sizeof(double) * n, kind, st));
}

// anorm is read only after the stream has been synchronized (the callers queue
// its copy from the device just before).
bool solve_chunk(bdlu_ctx* c, Plan& P, const Work& W, int32_t kt, int32_t kc, const bdlu_ir_opts* ir,
                 const double& anorm, bool watch_stall, bdlu_solve_stats& cs) {
  hipStream_t st = c->stream;
  const bool own = &W == &P.work;
  const bool prof = own && c->prof;
  auto solve = [&](const double* b, double* x) {
    if (own) block_solve(P, st, b, x, kt, c->graphs && !c->prof);
    else block_solve_launches(P, W, st, b, x, kt);
  };
  const int64_t n = P.n;
  const int32_t max_iters = ir ? ir->max_iters : 0;
  const bool want_stats = max_iters > 0 || (ir && ir->stop_rule < 0);
  cs = bdlu_solve_stats{0, -1.0, -1.0, -1.0, 0};
  double h[9 * kMaxRhs];
  auto residual = [&]() {
    bdk::launch_residual(st, P.n, kt, P.csr_ptr, P.csr_col, P.vals, W.x, W.b, W.r, W.cw, n, P.res_is_long,
                         P.res_nlong, P.res_long_rows);
    bdk::launch_norms(st, P.n, kt, W.r, W.x, W.b, W.cw, n, W.partials, W.norms);
  };
  auto eval = [&]() {
    residual();
    HIPCHECK(hipMemcpyAsync(h, W.norms, sizeof(double) * 9 * kt, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
  };
  int32_t iters = 0;
  bool converged = false, ended = false;
  double prev = 0.0;
  const bool device_loop = own && c->graphs && c->device_ir && !prof && max_iters > 0;
  if (device_loop) {
    // The solve, the first evaluation and as many refinement steps as the
    // fewest of the plan's recent solves took, as one graph with the stop
    // decisions taken on the device: steps after the loop has ended leave x
    // alone (the gated update) and re-evaluate the same x, so the result is
    // the host loop's. A loop still running at the graph's end goes on below.
    int32_t steps = max_iters;
    bool seen = false;
    for (int32_t v : P.ir_steps)
      if (v >= 0) steps = std::min(steps, v), seen = true;
    if (!seen) steps = 0;
    const int32_t rule = ir->stop_rule;
    int32_t tol_bits[2];
    std::memcpy(tol_bits, &ir->tol, sizeof(tol_bits));
    // The stop test runs over all kt columns: the padding columns have b = 0
    // and count as converged, as they do for the host's test over kc.
    std::vector<int32_t> key{kt, 2, steps, rule, max_iters, watch_stall ? 1 : 0, tol_bits[0], tol_bits[1]};
    for (const SparseDev& S : P.sp) key.push_back(S.prec);
    hipGraphExec_t& exec = P.solve_graph[key];
    const double* anorm_d = P.norms + kAnormSlot;
    run_captured(st, exec, true, [&] {
      block_solve_launches(P, W, st, W.b, W.x, kt);
      residual();
      bdk::launch_ir_check(st, W.norms, kt, rule, ir->tol, anorm_d, max_iters, watch_stall, true, W.irs);
      for (int32_t i = 0; i < steps; ++i) {
        block_solve_launches(P, W, st, W.r, W.d, kt);
        bdk::launch_axpy_gated(st, P.n, kt, W.d, W.x, n, W.irs);
        residual();
        bdk::launch_ir_check(st, W.norms, kt, rule, ir->tol, anorm_d, max_iters, watch_stall, false, W.irs);
      }
    });
    double s[bdk::kIrStateLen];
    HIPCHECK(hipMemcpyAsync(h, W.norms, sizeof(double) * 9 * kt, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(s, W.irs, sizeof(s), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    iters = static_cast<int32_t>(s[bdk::kIrIters]);
    converged = s[bdk::kIrConverged] != 0.0;
    prev = s[bdk::kIrWorst];
    ended = s[bdk::kIrActive] == 0.0;
  } else {
    ProfEvents pev;
    if (prof) pev.mark(BDLU_PROF_SOLVE, st);
    solve(W.b, W.x);
    if (prof) {
      pev.mark(BDLU_PROF_SOLVE, st);
      HIPCHECK(hipStreamSynchronize(st));
      pev.collect(c);
    }
    if (!want_stats) return true;
  }
  auto rel = [&](int k) {
    const double bn2 = std::sqrt(h[k * 9 + 3]);
    return bn2 > 0 ? std::sqrt(h[k * 9 + 0]) / bn2 : 0.0;
  };
  auto berr = [&](int k) {
    const double den = anorm * h[k * 9 + 2] + h[k * 9 + 4];
    return den > 0 ? h[k * 9 + 1] / den : 0.0;
  };
  const int rule = ir->stop_rule;
  auto measure = [&](int k) { return rule == 1 ? berr(k) : rule == 2 ? h[k * 9 + 5] : rel(k); };
  // Columns with b = 0 have x = 0 and count as converged.
  auto worst = [&]() {
    double w = 0.0;
    for (int k = 0; k < kc; ++k)
      if (h[k * 9 + 3] > 0.0) w = std::max(w, measure(k));
    return w;
  };
  auto met = [&]() {
    for (int k = 0; k < kc; ++k) {
      const bool stop = rule == 0 ? rel(k) < ir->tol : measure(k) <= ir->tol;
      if (!stop && h[k * 9 + 3] > 0.0) return false;
    }
    return true;
  };
  if (!ended && iters == 0) {
    if (!device_loop) eval();
    converged = max_iters <= 0 || met();
    prev = worst();
  }
  while (!ended && !converged && iters < max_iters) {
    solve(W.r, W.d);
    bdk::launch_axpy(st, P.n, kt, W.d, W.x, n);
    ++iters;
    eval();
    converged = met();
    const double now = worst();
    if (!converged && watch_stall && !(now <= 0.5 * prev)) break;
    prev = now;
  }
  if (own && max_iters > 0) {
    P.ir_steps[P.ir_next] = iters;
    P.ir_next = (P.ir_next + 1) % 4;
  }
  // Worst right-hand side of the chunk.
  for (int k = 0; k < kc; ++k) {
    cs.rel_residual = std::max(cs.rel_residual, rel(k));
    cs.backward_error = std::max(cs.backward_error, berr(k));
    cs.componentwise_error = std::max(cs.componentwise_error, h[k * 9 + 5]);
  }
  cs.iterations = iters;
  return converged;
}

}  // namespace

extern "C" {


#ifndef BDLU_BUILD_INFO
#define BDLU_BUILD_INFO "unknown"
#endif
const char* bdlu_gpu_build_info(void) { return BDLU_BUILD_INFO; }

int bdlu_gpu_device_count(int* count) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  if (count) *count = c;
  return c > 0 ? BDLU_OK : BDLU_E_BACKEND;
}

int bdlu_gpu_create(int device, uint64_t hbm_budget_bytes, bdlu_ctx** out) {
  if (!out) return BDLU_E_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return BDLU_E_BACKEND;
  if (device < 0 || device >= count) return BDLU_E_INVALID;
  auto* c = new (std::nothrow) bdlu_ctx;
  if (!c) return BDLU_E_BACKEND;
  c->device = device;
  c->budget = hbm_budget_bytes;
  if (const char* e = std::getenv("BLOCKDAGLU_GPU_FP32_GROWTH")) c->fp32_growth = std::max(1.0, std::atof(e));
  if (const char* e = std::getenv("BLOCKDAGLU_GPU_REPIVOT_GROWTH")) c->fp64_growth = std::max(1.0, std::atof(e));
  if (const char* e = std::getenv("BLOCKDAGLU_GPU_GRAPH")) c->graphs = std::atoi(e) != 0;
  if (const char* e = std::getenv("BLOCKDAGLU_GPU_IR_DEVICE")) c->device_ir = std::atoi(e) != 0;
  if (const char* e = std::getenv("BLOCKDAGLU_GPU_AUTOTUNE")) c->autotune = std::atoi(e) != 0;
  if (const char* e = std::getenv("BLOCKDAGLU_GPU_STREAMS")) c->fork = std::atoi(e) > 1;
  if (const char* e = std::getenv("BLOCKDAGLU_GPU_LOOKAHEAD")) c->lookahead = std::max(0, std::min(2, std::atoi(e)));
  if (const char* e = std::getenv("BLOCKDAGLU_GPU_SEG_GRID_MAX")) c->seg_grid_max = std::max(1, std::atoi(e));
  bool ok = hipSetDevice(device) == hipSuccess &&
            hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
  if (ok) c->scratch = scratch_address(device);
  for (int i = 0; ok && i < bdlu_ctx::kSide; ++i)
    ok = hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking) == hipSuccess;
  if (!ok) {
    for (hipStream_t q : c->side)
      if (q) (void)hipStreamDestroy(q);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return BDLU_E_BACKEND;
  }
  *out = c;
  return BDLU_OK;
}

void bdlu_gpu_destroy(bdlu_ctx* c) {
  if (!c) return;
  {
    std::lock_guard<std::mutex> g(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (hipStream_t q : c->side) (void)hipStreamSynchronize(q);
    for (auto& kv : c->plans) kv.second->mem.release();
    c->plans.clear();
    for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
    for (hipStream_t q : c->side) (void)hipStreamDestroy(q);
    (void)hipStreamDestroy(c->stream);
  }
  delete c;
}

const char* bdlu_gpu_last_error(bdlu_ctx* c) { return c ? c->last_error.c_str() : "null context"; }

int bdlu_gpu_plan_upload(bdlu_ctx* c, const bdlu_plan_desc* d, uint64_t* plan_id) {
  if (!c || !d || !plan_id) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  if (d->n <= 0 || d->nblocks <= 0 || d->nnz < 0) return set_err(c, BDLU_E_ANALYSIS, "empty plan");
  try {
    HIPCHECK(hipSetDevice(c->device));
    std::unique_ptr<Plan> pl = upload(c, *d);
    if (c->budget && c->used + pl->mem.bytes > c->budget) {
      pl->mem.release();
      return set_err(c, BDLU_E_BUDGET, "plan exceeds the HBM budget");
    }
    c->used += pl->mem.bytes;
    const uint64_t id = c->next_id++;
    c->plans[id] = std::move(pl);
    *plan_id = id;
    return BDLU_OK;
  } catch (const HipError& e) {
    return set_err(c, BDLU_E_BACKEND, e.msg);
  } catch (const std::exception& e) {
    return set_err(c, BDLU_E_BACKEND, e.what());
  }
}

int bdlu_gpu_plan_release(bdlu_ctx* c, uint64_t id) {
  if (!c) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  auto it = c->plans.find(id);
  if (it == c->plans.end()) return set_err(c, BDLU_E_INVALID, "unknown plan");
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->used -= it->second->mem.bytes;
  it->second->mem.release();
  c->plans.erase(it);
  return BDLU_OK;
}

int bdlu_gpu_plan_schedule(bdlu_ctx* c, uint64_t id, int32_t* out4) {
  if (!c || !out4) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  auto it = c->plans.find(id);
  if (it == c->plans.end()) return set_err(c, BDLU_E_INVALID, "unknown plan");
  const Plan::Schedule& sc = it->second->sched;
  out4[0] = sc.small_tile_below;
  out4[1] = sc.dense_lpt ? 1 : 0;
  out4[2] = sc.seg_grid;
  out4[3] = (sc.factor_tuned[0] ? 1 : 0) | (sc.factor_tuned[1] ? 2 : 0) | (sc.solve_tuned ? 4 : 0);
  return BDLU_OK;
}

int bdlu_gpu_plan_bytes(bdlu_ctx* c, uint64_t id, uint64_t* bytes) {
  if (!c || !bytes) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  auto it = c->plans.find(id);
  if (it == c->plans.end()) return set_err(c, BDLU_E_INVALID, "unknown plan");
  *bytes = it->second->mem.bytes;
  return BDLU_OK;
}

int bdlu_gpu_refactor(bdlu_ctx* c, uint64_t id, const double* values, int where, int precision,
                      bdlu_factor_stats* stats) {
  if (!c) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  auto it = c->plans.find(id);
  if (it == c->plans.end()) return set_err(c, BDLU_E_INVALID, "unknown plan");
  bdlu_factor_stats fs{0, 0, 0, -1, 0.0};
  try {
    HIPCHECK(hipSetDevice(c->device));
    const int rc = refactor_impl(c, *it->second, values, where, precision, fs);
    if (stats) *stats = fs;
    return rc;
  } catch (const HipError& e) {
    return set_err(c, BDLU_E_BACKEND, e.msg);
  }
}

int bdlu_gpu_solve(bdlu_ctx* c, uint64_t id, double* B, int32_t nrhs, int64_t ldb, int where, const bdlu_ir_opts* ir,
                   bdlu_solve_stats* stats) {
  if (!c || !B || nrhs < 1) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  auto it = c->plans.find(id);
  if (it == c->plans.end()) return set_err(c, BDLU_E_INVALID, "unknown plan");
  Plan& P = *it->second;
  if (!P.factored) return set_err(c, BDLU_E_SOLVE, "plan has no valid factors");
  if (ldb < P.n) return set_err(c, BDLU_E_INVALID, "ldb < n");
  hipStream_t st = c->stream;
  const int32_t max_iters = ir ? ir->max_iters : 0;
  const bool want_stats = max_iters > 0 || (ir && ir->stop_rule < 0);
  bdlu_solve_stats ss{0, -1.0, -1.0, -1.0, 0};
  const int64_t n = P.n;
  try {
    HIPCHECK(hipSetDevice(c->device));
    const hipMemcpyKind in = where == BDLU_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind out = where == BDLU_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    const Work& W = P.work;
    bdk::launch_fill_i32(st, W.err, 1, 0);
    double anorm = 0.0;
    if (want_stats) HIPCHECK(hipMemcpyAsync(&anorm, P.norms + kAnormSlot, sizeof(double), hipMemcpyDeviceToHost, st));
    // Columns in chunks of up to kMaxRhs, each solved (and refined) together.
    for (int32_t c0 = 0; c0 < nrhs; c0 += kMaxRhs) {
      const int32_t kc = std::min(kMaxRhs, nrhs - c0);
      const int32_t kt = kc <= 1 ? 1 : kc <= 2 ? 2 : kc <= 4 ? 4 : 8;
      copy_columns(W.b, n, B + c0 * ldb, ldb, n, kc, in, st);
      if (kt > kc) bdk::launch_fill_f64(st, W.b + kc * n, (kt - kc) * n, 0.0);
      if (c->autotune && !c->prof && !P.sched.solve_tuned) {
        autotune_solve(P, st, kt, c->seg_grid_max);
        P.sched.solve_tuned = true;
        P.drop_graphs();
      }
      const bool fp32_factors = std::any_of(P.h_prec.begin(), P.h_prec.end(), [](int32_t p) { return p != 0; });
      const bool fallback = ir && ir->fp64_fallback && max_iters > 0 && fp32_factors;
      bdlu_solve_stats cs{};
      if (!solve_chunk(c, P, W, kt, kc, ir, anorm, fallback, cs) && fallback) {
        // FP32 factors did not reach the stop rule: refactor the same values in
        // FP64 (W.b still holds this chunk) and solve the chunk again.
        bdlu_factor_stats fs{};
        const int rc = refactor_impl(c, P, nullptr, BDLU_MEM_DEVICE, BDLU_FP64, fs);
        if (rc != BDLU_OK) return rc;
        ss.fp64_refactor = 1;
        const int32_t first_iters = cs.iterations;
        solve_chunk(c, P, W, kt, kc, ir, anorm, false, cs);
        cs.iterations += first_iters;
      }
      ss.rel_residual = std::max(ss.rel_residual, cs.rel_residual);
      ss.backward_error = std::max(ss.backward_error, cs.backward_error);
      ss.componentwise_error = std::max(ss.componentwise_error, cs.componentwise_error);
      ss.iterations = std::max(ss.iterations, cs.iterations);
      copy_columns(B + c0 * ldb, ldb, W.x, n, n, kc, out, st);
    }
    int32_t herr = 0;
    HIPCHECK(hipMemcpyAsync(&herr, W.err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (stats) *stats = ss;
    if (herr) return set_err(c, BDLU_E_SOLVE, "solve wait bound exceeded");
    return BDLU_OK;
  } catch (const HipError& e) {
    return set_err(c, BDLU_E_BACKEND, e.msg);
  }
}

int bdlu_gpu_solve_noalloc(bdlu_ctx* c, uint64_t id, double* d_B, int32_t nrhs, int64_t ldb,
                           const bdlu_ir_opts* ir, bdlu_solve_stats* stats) {
  if (!c || !d_B || nrhs < 1) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  auto it = c->plans.find(id);
  if (it == c->plans.end()) return set_err(c, BDLU_E_INVALID, "unknown plan");
  Plan& P = *it->second;
  if (!P.factored) return set_err(c, BDLU_E_SOLVE, "plan has no valid factors");
  if (ldb < P.n) return set_err(c, BDLU_E_INVALID, "ldb < n");
  if (!c->scratch) return set_err(c, BDLU_E_BACKEND, "no scratch address on this device");
  hipStream_t st = c->stream;
  const int32_t max_iters = ir ? ir->max_iters : 0;
  const bool want_stats = max_iters > 0 || (ir && ir->stop_rule < 0);
  const int64_t n = P.n;
  // Work set in the scratch array: b, x, bperm, xperm (+ r, d, cw with
  // refinement) of n x kt each, the norm partials, the norms and the counters,
  // each piece starting on a 256-byte boundary.
  const int32_t kt_max = nrhs <= 1 ? 1 : nrhs <= 2 ? 2 : nrhs <= 4 ? 4 : kMaxRhs;
  auto round32 = [](int64_t v) { return (v + 31) / 32 * 32; };
  const int64_t vec = round32(n * kt_max);
  const int64_t nvec = want_stats ? 7 : 4;
  const int64_t need = nvec * vec + round32(512 * 6 * kMaxRhs) + round32(9 * kMaxRhs) +
                       round32((static_cast<int64_t>(P.nticket) + 1 + 1) / 2);
  if (need > kScratchDoubles)
    return set_err(c, BDLU_E_INVALID, "solve needs " + std::to_string(need) + " doubles of scratch, more than the " +
                                          std::to_string(kScratchDoubles) + " of the no-allocation entry point");
  Work W;
  double* p = c->scratch;
  W.b = p, p += vec;
  W.x = p, p += vec;
  W.bperm = p, p += vec;
  W.xperm = p, p += vec;
  if (want_stats) {
    W.r = p, p += vec;
    W.d = p, p += vec;
    W.cw = p, p += vec;
  }
  W.partials = p, p += round32(512 * 6 * kMaxRhs);
  W.norms = p, p += round32(9 * kMaxRhs);
  W.tickets = reinterpret_cast<int32_t*>(p);
  W.err = W.tickets + P.nticket;
  bdlu_solve_stats ss{0, -1.0, -1.0, -1.0, 0};
  bool all_converged = true;
  try {
    HIPCHECK(hipSetDevice(c->device));
    bdk::launch_fill_i32(st, W.err, 1, 0);
    double anorm = 0.0;
    if (want_stats) HIPCHECK(hipMemcpyAsync(&anorm, P.norms + kAnormSlot, sizeof(double), hipMemcpyDeviceToHost, st));
    for (int32_t c0 = 0; c0 < nrhs; c0 += kMaxRhs) {
      const int32_t kc = std::min(kMaxRhs, nrhs - c0);
      const int32_t kt = kc <= 1 ? 1 : kc <= 2 ? 2 : kc <= 4 ? 4 : 8;
      copy_columns(W.b, n, d_B + c0 * ldb, ldb, n, kc, hipMemcpyDeviceToDevice, st);
      if (kt > kc) bdk::launch_fill_f64(st, W.b + kc * n, (kt - kc) * n, 0.0);
      bdlu_solve_stats cs{};
      all_converged = solve_chunk(c, P, W, kt, kc, ir, anorm, false, cs) && all_converged;
      ss.rel_residual = std::max(ss.rel_residual, cs.rel_residual);
      ss.backward_error = std::max(ss.backward_error, cs.backward_error);
      ss.componentwise_error = std::max(ss.componentwise_error, cs.componentwise_error);
      ss.iterations = std::max(ss.iterations, cs.iterations);
      copy_columns(d_B + c0 * ldb, ldb, W.x, n, n, kc, hipMemcpyDeviceToDevice, st);
    }
    int32_t herr = 0;
    HIPCHECK(hipMemcpyAsync(&herr, W.err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (stats) *stats = ss;
    if (herr) return set_err(c, BDLU_E_SOLVE, "solve wait bound exceeded");
    const bool fp32_factors = std::any_of(P.h_prec.begin(), P.h_prec.end(), [](int32_t q) { return q != 0; });
    if (!all_converged && ir && ir->fp64_fallback && fp32_factors)
      return set_err(c, BDLU_E_SOLVE, "refinement on FP32 factors did not reach the stop rule (this entry point "
                                      "does not refactor in FP64)");
    return BDLU_OK;
  } catch (const HipError& e) {
    return set_err(c, BDLU_E_BACKEND, e.msg);
  }
}

int bdlu_gpu_plan_values(bdlu_ctx* c, uint64_t id, double* host_values) {
  if (!c || !host_values) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  auto it = c->plans.find(id);
  if (it == c->plans.end()) return set_err(c, BDLU_E_INVALID, "unknown plan");
  Plan& P = *it->second;
  if (hipSetDevice(c->device) != hipSuccess ||
      hipMemcpyAsync(host_values, P.vals, sizeof(double) * P.nnz, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return set_err(c, BDLU_E_BACKEND, "value copy failed");
  return BDLU_OK;
}

int bdlu_gpu_plan_dense_pivots(bdlu_ctx* c, uint64_t id, int32_t* host_piv) {
  if (!c || !host_piv) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  auto it = c->plans.find(id);
  if (it == c->plans.end()) return set_err(c, BDLU_E_INVALID, "unknown plan");
  Plan& P = *it->second;
  try {
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipMemcpyAsync(host_piv, P.piv, sizeof(int32_t) * P.n, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
  } catch (const HipError& e) {
    return set_err(c, BDLU_E_BACKEND, e.msg);
  }
  return BDLU_OK;
}

int bdlu_gpu_copy(bdlu_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c || (bytes && (!dst || !src))) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  try {
    HIPCHECK(hipSetDevice(c->device));
    if (bytes) HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
  } catch (const HipError& e) {
    return set_err(c, BDLU_E_BACKEND, e.msg);
  }
  return BDLU_OK;
}

int bdlu_gpu_set_side_streams(bdlu_ctx* c, int enable) {
  if (!c) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  c->fork = enable != 0;
  for (auto& kv : c->plans) kv.second->drop_graphs();  // recapture with the new launch structure
  return BDLU_OK;
}

int bdlu_gpu_set_solve_grid_max(bdlu_ctx* c, int32_t max_workgroups) {
  if (!c || max_workgroups < 1) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  c->seg_grid_max = max_workgroups;
  for (auto& kv : c->plans) {
    Plan& P = *kv.second;
    P.sched.seg_grid = std::min(P.sched.seg_grid, max_workgroups);
    P.drop_graphs();  // recapture with the new grid
  }
  return BDLU_OK;
}

int bdlu_gpu_set_device_refinement(bdlu_ctx* c, int enable) {
  if (!c) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  c->device_ir = enable != 0;
  return BDLU_OK;
}

void* bdlu_gpu_stream(bdlu_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

int bdlu_gpu_profile(bdlu_ctx* c, int enable) {
  if (!c) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  c->prof = enable != 0;
  for (int k = 0; k < BDLU_PROF_KINDS; ++k) {
    c->prof_ms[k] = c->prof_flops[k] = c->prof_bytes[k] = 0.0;
    c->prof_n[k] = 0;
  }
  return BDLU_OK;
}

int bdlu_gpu_profile_read(bdlu_ctx* c, int kind, double* total_ms, int64_t* launches, double* flops, double* bytes) {
  if (!c || kind < 0 || kind >= BDLU_PROF_KINDS) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  if (total_ms) *total_ms = c->prof_ms[kind];
  if (launches) *launches = c->prof_n[kind];
  if (flops) *flops = c->prof_flops[kind];
  if (bytes) *bytes = c->prof_bytes[kind];
  return BDLU_OK;
}

int bdlu_gpu_debug_dense_lu(bdlu_ctx* c, int32_t m, double* A, int precision, int lookahead,
                            int32_t small_tile_below) {
  if (!c || !A || m <= 0 || lookahead < 0 || lookahead > 2) return BDLU_E_INVALID;
  std::lock_guard<std::mutex> g(c->mu);
  hipStream_t st = c->stream;
  Alloc tmp;
  try {
    HIPCHECK(hipSetDevice(c->device));
    const std::size_t mm = static_cast<std::size_t>(m) * m;
    int32_t* status = tmp.get<int32_t>(1);
    bdk::launch_fill_i32(st, status, 1, 0);
    if (precision == BDLU_FP32) {
      std::vector<float> h(mm);
      for (std::size_t i = 0; i < mm; ++i) h[i] = static_cast<float>(A[i]);
      float* d = tmp.copy(h, st);
      float* inv = tmp.get<float>(2 * 64 * 64 * ((static_cast<std::size_t>(m) + 63) / 64));
      bdk::launch_tail_lu<float>(st, m, d, m, 1e300, status, inv, nullptr, small_tile_below, lookahead);
      HIPCHECK(hipMemcpyAsync(h.data(), d, sizeof(float) * mm, hipMemcpyDeviceToHost, st));
      HIPCHECK(hipStreamSynchronize(st));
      for (std::size_t i = 0; i < mm; ++i) A[i] = h[i];
    } else {
      double* d = tmp.copy(A, mm, st);
      double* inv = tmp.get<double>(2 * 64 * 64 * ((static_cast<std::size_t>(m) + 63) / 64));
      bdk::launch_tail_lu<double>(st, m, d, m, 1e300, status, inv, nullptr, small_tile_below, lookahead);
      HIPCHECK(hipMemcpyAsync(A, d, sizeof(double) * mm, hipMemcpyDeviceToHost, st));
      HIPCHECK(hipStreamSynchronize(st));
    }
    tmp.release();
    return BDLU_OK;
  } catch (const HipError& e) {
    tmp.release();
    return set_err(c, BDLU_E_BACKEND, e.msg);
  }
}

}  // extern "C"
