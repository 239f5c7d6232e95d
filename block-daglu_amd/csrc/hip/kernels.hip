// HIP kernels for gfx950 (CDNA4, 64-wide wavefronts, MFMA, 160 KB LDS/CU).
//
// Built with -ffp-contract=off: every multiply-subtract below rounds twice,
// in the same order as the reference's loops (ubtf:1465-1494 for the dense
// LU), which keeps the FP64 dense factors bitwise equal to the CPU restatement
// and the pivot sequence identical. The tail GEMM uses MFMA (fused, its own
// accumulation order); only its results are tolerance-checked.
#include "kernels.h"

#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

namespace bdk {

namespace {

constexpr int kWave = 64;
constexpr int kLdsPad = 65;  // odd leading dimension for 64x64 tiles in LDS

__device__ __forceinline__ void wave_sync_mem() {
  // Orders this wave's global-memory accesses: every lane's earlier loads and
  // stores complete before any later one issues (all lanes share the CU's L1).
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

template <typename T>
__device__ __forceinline__ T tabs(T x) {
  return x < T(0) ? -x : x;
}
template <typename T>
__device__ __forceinline__ T tmax(T a, T b) {
  return a < b ? b : a;
}
__device__ __forceinline__ float tfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double tfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

__device__ __forceinline__ int32_t load_acquire(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_release(int32_t* p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}


// Wave-uniform copy of lane 0's value. Loop exits and trip counts must come
// from uniform values: a single-wave workgroup's barrier is not a convergence
// point for the compiler, and a loop whose exit depends on a per-lane value
// (even one that is equal in every lane) may be split so that lane 0 leaves
// the loop while the others iterate on, which deadlocks a ticket loop.
__device__ __forceinline__ int32_t uniform_from_lane0(int32_t v) {
  return __builtin_amdgcn_readlane(v, 0);  // lane 0's register, into a scalar register
}

// Cross-lane moves through DPP (the VALU's lane network, no LDS round trip as
// with __shfl): kCtrl 0xb1 / 0x4e quad_perm [1,0,3,2] / [2,3,0,1], 0x124 /
// 0x128 row_ror 4 / 8, 0x142 / 0x143 row_bcast 15 / 31. A lane whose source is
// invalid (row 0 for row_bcast 15, rows 0-1 for row_bcast 31) keeps its value.
template <int kCtrl>
__device__ __forceinline__ int32_t dpp_i32(int32_t v) {
  return __builtin_amdgcn_update_dpp(v, v, kCtrl, 0xF, 0xF, false);
}
template <int kCtrl>
__device__ __forceinline__ float dpp_t(float v) {
  return __builtin_bit_cast(float, dpp_i32<kCtrl>(__builtin_bit_cast(int32_t, v)));
}
template <int kCtrl>
__device__ __forceinline__ double dpp_t(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = static_cast<uint32_t>(dpp_i32<kCtrl>(static_cast<int32_t>(static_cast<uint32_t>(u))));
  const uint32_t hi = static_cast<uint32_t>(dpp_i32<kCtrl>(static_cast<int32_t>(static_cast<uint32_t>(u >> 32))));
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ __forceinline__ float lane63(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int32_t, v), 63));
}
__device__ __forceinline__ double lane63(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int32_t>(static_cast<uint32_t>(u)), 63);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int32_t>(static_cast<uint32_t>(u >> 32)), 63);
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}
// fmax over the wave's 64 lanes (NaN lanes ignored, as fmax does), returned
// wave-uniform. Every pairing order gives the same value: max is exact.
template <typename T>
__device__ __forceinline__ T wave_fmax(T v) {
  v = fmax(v, dpp_t<0xb1>(v));
  v = fmax(v, dpp_t<0x4e>(v));
  v = fmax(v, dpp_t<0x124>(v));
  v = fmax(v, dpp_t<0x128>(v));
  v = fmax(v, dpp_t<0x142>(v));
  v = fmax(v, dpp_t<0x143>(v));
  return lane63(v);
}
__device__ __forceinline__ int32_t wave_min_i32(int32_t v) {
  v = min(v, dpp_i32<0xb1>(v));
  v = min(v, dpp_i32<0x4e>(v));
  v = min(v, dpp_i32<0x124>(v));
  v = min(v, dpp_i32<0x128>(v));
  v = min(v, dpp_i32<0x142>(v));
  v = min(v, dpp_i32<0x143>(v));
  return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ int32_t load_relaxed(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Published-value protocol of the solves: an output vector starts filled with
// kXSentinel (a signaling-NaN pattern, which no arithmetic produces: results
// are quiet NaNs), each value is written once with a relaxed agent-scope
// atomic store, and a consumer polls the very values it needs until they
// differ from the sentinel. The value is its own flag, so no fences, flags or
// dependency lists sit on the critical path.
constexpr uint64_t kXSentinel = 0x7FF4DEADBEEF0001ull;

__device__ __forceinline__ uint64_t ld_x_relaxed(const double* p) {
  return __hip_atomic_load(reinterpret_cast<const uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_x_relaxed(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<uint64_t*>(p), __builtin_bit_cast(uint64_t, v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
// Per-lane bounded wait for a published value (first poll already done: u).
// Valid schedules never come near the bound; expiry sets `late`.
__device__ __forceinline__ double finish_wait_x(const double* p, uint64_t u, bool& late) {
  for (int32_t it = 0; u == kXSentinel && it < (1 << 22); ++it) {
    __builtin_amdgcn_s_sleep(2);
    u = ld_x_relaxed(p);
  }
  if (u == kXSentinel) late = true;
  return __builtin_bit_cast(double, u);
}
__device__ __forceinline__ double wait_x(const double* p, bool& late) { return finish_wait_x(p, ld_x_relaxed(p), late); }

inline int grid_for(int64_t n, int block, int cap = 65535) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<int>(g);
}

}  // namespace

// ---------------------------------------------------------------------------
// value routing
__global__ void k_scatter(int64_t nnz, const double* __restrict__ vals, const int64_t* __restrict__ vmap,
                          double* __restrict__ arena) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < nnz; e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t d = vmap[e];
    if (d >= 0) arena[d] = vals[e];
  }
}
void launch_scatter(hipStream_t st, int64_t nnz, const double* vals, const int64_t* vmap, double* arena) {
  if (nnz <= 0) return;
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(nnz, 256, 8192)), dim3(256), 0, st, nnz, vals, vmap, arena);
}

template <typename T>
__global__ void k_fill(T* p, int64_t n, T v) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) p[i] = v;
}
void launch_fill_i32(hipStream_t st, int32_t* p, int64_t n, int32_t v) {
  if (n > 0) hipLaunchKernelGGL(k_fill<int32_t>, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, p, n, v);
}
void launch_fill_f64(hipStream_t st, double* p, int64_t n, double v) {
  if (n > 0) hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, p, n, v);
}
void launch_fill_f32(hipStream_t st, float* p, int64_t n, float v) {
  if (n > 0) hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, p, n, v);
}
// The unpublished mark in rows [0, m) of k columns p + c * ld (one launch).
__global__ void k_fill_sentinel_cols(uint64_t* p, int32_t m, int32_t k, int64_t ld) {
  const int64_t n = int64_t(m) * k;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t c = e / m;
    p[c * ld + (e - c * m)] = kXSentinel;
  }
}
void launch_fill_sentinel_cols(hipStream_t st, double* p, int32_t m, int32_t k, int64_t ld) {
  if (m > 0 && k > 0)
    hipLaunchKernelGGL(k_fill_sentinel_cols, dim3(grid_for(int64_t(m) * k, 256, 4096)), dim3(256), 0, st,
                       reinterpret_cast<uint64_t*>(p), m, k, ld);
}
void launch_fill_sentinel(hipStream_t st, double* p, int64_t n) {
  if (n > 0)
    hipLaunchKernelGGL(k_fill<uint64_t>, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st,
                       reinterpret_cast<uint64_t*>(p), n, kXSentinel);
}

// ---------------------------------------------------------------------------
// K1: batched dense LU with row pivoting, one wavefront per block (s <= 64).
// Same arithmetic as the reference dense branch (ubtf:1465-1494): fail iff the
// largest candidate |a_ik| is 0, full row swap, l_ik = a_ik / a_kk,
// a_ij -= l_ik * a_kj. Pivot choice:
//   key == nullptr: partial pivoting, the first row with the largest |a_ik|
//     (ties -> smaller row position), as the reference's dense branch;
//   key != nullptr: KLU's threshold rule (global-KLU route): the row at
//     position k (the column's diagonal; the physical swaps carry KLU's
//     diagonal tracking) is kept when |a_kk| >= tol * max_i |a_ik|, else the
//     largest, ties to the smaller key (each row's id before the block's column
//     ordering, key[starts + position], moved with its row).
// piv out: for each row position of the block, the block-local row that the
// swaps bring there.
template <typename T>
__global__ __launch_bounds__(64) void k_dense_lu(int32_t nd, const int32_t* __restrict__ sizes,
                                                 const int32_t* __restrict__ starts, const int64_t* __restrict__ val0,
                                                 const int64_t* __restrict__ lu0, const double* __restrict__ arena,
                                                 T* __restrict__ lu, int32_t* __restrict__ piv,
                                                 int32_t* __restrict__ pinv, int32_t* __restrict__ status, const int32_t* __restrict__ block_of,
                                                 const int32_t* __restrict__ order, double tol,
                                                 const int32_t* __restrict__ key) {
  __shared__ T a[64 * kLdsPad];
  const int lane = threadIdx.x;
  // order: which dense block each workgroup takes (e.g. largest first), or none.
  const int32_t t = order ? order[blockIdx.x] : static_cast<int32_t>(blockIdx.x);
  if (t >= nd) return;
  const int s = __builtin_amdgcn_readfirstlane(sizes[t]);
  const double* src = arena + val0[t];
  for (int idx = lane; idx < s * s; idx += kWave) {
    const int j = idx / s, i = idx - j * s;
    a[j * kLdsPad + i] = static_cast<T>(src[idx]);
  }
  __syncthreads();
  int32_t mypiv = lane;
  int32_t st = 0;
  const bool klu = key != nullptr;
  int32_t mykey = (klu && lane < s) ? key[starts[t] + lane] : lane;
  for (int k = 0; k < s; ++k) {
    const T v = (lane >= k && lane < s) ? tabs(a[k * kLdsPad + lane]) : T(-1);
    // The largest candidate (a max reduction over the wave through DPP; lanes
    // outside k..s-1 hold -1), then the first lane holding it (ballot), or
    // with KLU's rule the smallest key among the lanes holding it. Rows are
    // lanes (physical swaps), so the first lane is the first row position, as
    // in the reference.
    const T vm = wave_fmax(v);
    if (uniform_from_lane0(!(vm > T(0)) ? 1 : 0)) {  // best == 0 (or NaN): the reference marks the block failed
      st |= kStZeroPivot;
      break;
    }
    const bool top = v == vm;
    int p;
    if (!klu) {
      p = __builtin_ctzll(__ballot(top));
    } else {
      const int kk = wave_min_i32(top ? mykey : INT_MAX);
      p = __builtin_ctzll(__ballot(top && mykey == kk));
      // Diagonal preference (KLU's test: |diag| >= tol * largest).
      if (tabs(a[k * kLdsPad + k]) >= static_cast<T>(tol) * vm) p = k;
    }
    p = uniform_from_lane0(p);
    if (klu) {
      const int32_t kk_k = __builtin_amdgcn_readlane(mykey, k), kk_p = __builtin_amdgcn_readlane(mykey, p);
      mykey = (lane == k) ? kk_p : ((lane == p) ? kk_k : mykey);
    }
    mypiv = (lane == k) ? p : mypiv;
    if (p != k) {
      if (lane < s) {
        const T x = a[lane * kLdsPad + k];
        a[lane * kLdsPad + k] = a[lane * kLdsPad + p];
        a[lane * kLdsPad + p] = x;
      }
      __syncthreads();
    }
    const T akk = a[k * kLdsPad + k];
    if (lane > k && lane < s) {
      const T l = a[k * kLdsPad + lane] / akk;
      a[k * kLdsPad + lane] = l;
      // Columns in groups of 8: a group's LDS reads issue together, then its
      // updates and stores (one column at a time, each read waited for the
      // previous store). Per element the same a - l * u.
      int j = k + 1;
      for (; j + 8 <= s; j += 8) {
        T u[8], v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          u[q] = a[(j + q) * kLdsPad + k];
          v[q] = a[(j + q) * kLdsPad + lane];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) a[(j + q) * kLdsPad + lane] = v[q] - l * u[q];
      }
      for (; j < s; ++j) a[j * kLdsPad + lane] = a[j * kLdsPad + lane] - l * a[j * kLdsPad + k];
    }
    __syncthreads();
  }
  if (sizeof(T) == 4 && st == 0) {
    // FP32 small-pivot check, invariant under row and column scaling: pivot k
    // was computed as a_kk - sum_i l_ki u_ik, so its rounding error is about
    // u * (|L||U|)_kk (Higham, Accuracy and Stability of Numerical Algorithms,
    // Thm 9.3). A pivot below 2^-12 of that sum has lost more than half of
    // FP32's 24 bits to cancellation; the block is then redone in FP64.
    int bad = 0;
    if (lane < s) {
      const T ukk = tabs(a[lane * kLdsPad + lane]);
      T mass = ukk;
      for (int i = 0; i < lane; ++i) mass += tabs(a[i * kLdsPad + lane]) * tabs(a[lane * kLdsPad + i]);
      bad = (ukk * T(4096) < mass) ? 1 : 0;
    }
    if (__ballot(bad) != 0ull) st |= kStCheck;
  }
  T* dst = lu + lu0[t];
  for (int idx = lane; idx < s * s; idx += kWave) {
    const int j = idx / s, i = idx - j * s;
    dst[idx] = a[j * kLdsPad + i];
  }
  // Gather form of the row swaps (row piv_k <-> row k for k = 0..s-1, in order):
  // lane i follows the swaps backwards to the original row landing at i.
  int32_t srcrow = lane;
  for (int k = s - 1; k >= 0; --k) {
    const int32_t pk = __builtin_amdgcn_readlane(mypiv, k);
    srcrow = (srcrow == k) ? pk : ((srcrow == pk) ? k : srcrow);
  }
  if (lane < s) {
    piv[starts[t] + lane] = srcrow;
    pinv[starts[t] + srcrow] = lane;  // the inverse: where the solve's gather puts row `srcrow`
  }
  if (lane == 0) status[block_of[t]] = st;
}

template <typename T>
void launch_dense_lu(hipStream_t st, int32_t nd, const int32_t* sizes, const int32_t* starts, const int64_t* val0,
                     const int64_t* lu0, const double* arena, T* lu, int32_t* piv, int32_t* pinv, int32_t* status,
                     const int32_t* block_of, const int32_t* order, double tol, const int32_t* key) {
  if (nd <= 0) return;
  hipLaunchKernelGGL(k_dense_lu<T>, dim3(nd), dim3(64), 0, st, nd, sizes, starts, val0, lu0, arena, lu, piv, pinv,
                     status, block_of, order, tol, key);
}
template void launch_dense_lu<double>(hipStream_t, int32_t, const int32_t*, const int32_t*, const int64_t*,
                                      const int64_t*, const double*, double*, int32_t*, int32_t*, int32_t*, const int32_t*,
                                      const int32_t*, double, const int32_t*);
template void launch_dense_lu<float>(hipStream_t, int32_t, const int32_t*, const int32_t*, const int64_t*,
                                     const int64_t*, const double*, float*, int32_t*, int32_t*, int32_t*, const int32_t*,
                                     const int32_t*, double, const int32_t*);

// KLU row scaling over whole rows (klu_scale, scale 1 sums / 2 maxima of |a_ij|
// in column order; a row without a nonzero gets 1), and the scaled values
// a_ij / rs_i that the blocks are factored from. One thread per row of the
// canonical CSR.
__global__ void k_row_scale(int32_t n, const int64_t* __restrict__ ptr, const double* __restrict__ vals, int32_t kind,
                            double* __restrict__ rs, double* __restrict__ svals) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int64_t e0 = ptr[i], e1 = ptr[i + 1];
    double r = 0.0;
    for (int64_t e = e0; e < e1; ++e) {
      const double v = vals[e] < 0.0 ? -vals[e] : vals[e];
      r = kind == 1 ? r + v : (r < v ? v : r);
    }
    if (r == 0.0) r = 1.0;
    rs[i] = r;
    for (int64_t e = e0; e < e1; ++e) svals[e] = vals[e] / r;
  }
}
void launch_row_scale(hipStream_t st, int32_t n, const int64_t* ptr, const double* vals, int32_t kind, double* rs,
                      double* svals) {
  if (n > 0)
    hipLaunchKernelGGL(k_row_scale, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, n, ptr, vals, kind, rs, svals);
}

// The explicit inverse of each dense block's factors, for the solve:
// M = U^{-1} L^{-1} (so x = M (P b) solves the block), s x s column-major at the
// block's factor offset. One workgroup of two wavefronts per block: wave 0
// forms the columns of L^{-1}, wave 1 those of U^{-1} (lane j column j,
// right-looking, 63-k independent FMAs per step, over a column-major LDS copy
// of the factor read at wave-uniform addresses); wave 1 leaves U^{-1} in LDS
// and wave 0 multiplies it into its column: M(i,j) = sum_{k>=i} U^{-1}(i,k)
// L^{-1}(k,j), k ascending, in place for i ascending (row i of the product
// reads only entries k >= i). Applying M in the solve replaces two chains of s
// dependent substitution steps per block by one matrix-vector product.
// Two waves per SIMD: with FP64 the column in registers (x[64]) would
// otherwise take ~280 registers and one wave per SIMD; capped at 256 a few
// values spill, and the kernel still runs 1.8x faster on many small blocks.
template <typename T>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) void k_dense_inv(
    int32_t nd, const int32_t* __restrict__ sizes, const int64_t* __restrict__ lu0, const T* __restrict__ lu,
    T* __restrict__ minv) {
  constexpr int LC = 66;
  // The factors; after the two substitutions U^{-1} replaces U on and above
  // the diagonal (fc[k*66 + i] = U^{-1}(i, k), i <= k): wave 0 reads only the
  // strictly lower part, and wave 1 has finished its reads of the upper part
  // when it writes.
  __shared__ T fc[64 * LC];
  const int32_t t = static_cast<int32_t>(blockIdx.x);
  if (t >= nd) return;
  const int s = __builtin_amdgcn_readfirstlane(sizes[t]);
  const T* src = lu + lu0[t];
  for (int idx = threadIdx.x; idx < s * s; idx += 128) {
    const int j = idx / s, i = idx - j * s;
    fc[j * LC + i] = src[idx];
  }
  // Rows and columns past s: identity, so every value the loops read is finite.
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 128) {
    const int j = idx >> 6, i = idx & 63;
    if (i >= s || j >= s) fc[j * LC + i] = (i == j) ? T(1) : T(0);
  }
  __syncthreads();
  const int j = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  T x[64];
#pragma unroll
  for (int i = 0; i < 64; ++i) x[i] = (i == j) ? T(1) : T(0);
  if (w == 0) {
#pragma unroll
    for (int k = 0; k < 63; ++k) {
      if (k + 1 < s) {
#pragma unroll
        for (int i = k + 1; i < 64; ++i) x[i] = tfma(-fc[k * LC + i], x[k], x[i]);
      }
    }
  } else {
#pragma unroll
    for (int k = 63; k >= 0; --k) {
      if (k < s) {
        x[k] = x[k] / fc[k * LC + k];
#pragma unroll
        for (int i = 0; i < k; ++i) x[i] = tfma(-fc[k * LC + i], x[k], x[i]);
      }
    }
#pragma unroll
    for (int i = 0; i < 64; ++i)
      if (i <= j) fc[j * LC + i] = x[i];
  }
  __syncthreads();
  if (w == 0) {
    // M(i, j) = sum over k >= i of U^{-1}(i, k) x[k], k ascending for every i,
    // eight rows at a time with k outermost, so eight sums advance together
    // (the same operations in the same order per entry). In place: a group's
    // rows are read by no later group.
#pragma unroll
    for (int i0 = 0; i0 < 64; i0 += 8) {
      if (i0 < s) {
        T acc[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] = T(0);
#pragma unroll
        for (int k = i0; k < 64; ++k) {
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (i0 + u <= k) acc[u] = tfma(fc[k * LC + i0 + u], x[k], acc[u]);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (i0 + u < s) x[i0 + u] = acc[u];
      }
    }
    if (j < s) {
      T* dst = minv + lu0[t] + int64_t(j) * s;
#pragma unroll
      for (int i = 0; i < 64; ++i)
        if (i < s) dst[i] = x[i];
    }
  }
}

template <typename T>
void launch_dense_inv(hipStream_t st, int32_t nd, const int32_t* sizes, const int64_t* lu0, const T* lu, T* minv) {
  if (nd <= 0) return;
  hipLaunchKernelGGL(k_dense_inv<T>, dim3(nd), dim3(128), 0, st, nd, sizes, lu0, lu, minv);
}
template void launch_dense_inv<double>(hipStream_t, int32_t, const int32_t*, const int64_t*, const double*, double*);
template void launch_dense_inv<float>(hipStream_t, int32_t, const int32_t*, const int64_t*, const float*, float*);

// ---------------------------------------------------------------------------
// Sparse blocks: KLU scale=1 row sums, in a fixed order per row.
__global__ void k_rowscale(SpView v, const double* __restrict__ arena, double* __restrict__ rs) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < v.s; i += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int32_t p = v.rs_ptr[i]; p < v.rs_ptr[i + 1]; ++p) acc += fabs(arena[v.a_val0 + v.rs_idx[p]]);
    rs[i] = acc == 0.0 ? 1.0 : acc;
  }
}
void launch_rowscale(hipStream_t st, const SpView& v, const double* arena, double* rs) {
  hipLaunchKernelGGL(k_rowscale, dim3(grid_for(v.s, 256, 1024)), dim3(256), 0, st, v, arena, rs);
}

// Left-looking refactorization of the sparse leading columns on the fixed
// pattern, one wavefront per column, one level per launch. x is the wave's
// dense scratch column (length s, all zero between columns). Updates follow
// ascending pivot position, the order the host analysis used.
template <typename T>
__global__ __launch_bounds__(64) void k_sp_cols(SpView v, const double* __restrict__ arena,
                                                const double* __restrict__ rs, int32_t ncols,
                                                const int32_t* __restrict__ cols, T* __restrict__ xs,
                                                T* __restrict__ Lx, T* __restrict__ Ux, T* __restrict__ Ud,
                                                int32_t* __restrict__ status, const int32_t* __restrict__ is_long) {
  const int lane = threadIdx.x;
  T* x = xs + int64_t(blockIdx.x) * v.s;
  int32_t st = 0;
  T lmax_all = T(0);
  for (int32_t c = blockIdx.x; c < ncols; c += gridDim.x) {
    const int32_t j = cols[c];
    if (is_long && is_long[j]) continue;  // done by k_sp_cols_long
    for (int32_t p = v.a_ptr[j] + lane; p < v.a_ptr[j + 1]; p += kWave) {
      const int32_t r = v.a_row[p];
      const double a = arena[v.a_val0 + p];
      x[r] = static_cast<T>(v.scale ? a / rs[r] : a);
    }
    wave_sync_mem();
    for (int32_t up = v.u_ptr[j]; up < v.u_ptr[j + 1]; ++up) {
      const int32_t k = v.u_idx[up];
      const T xk = x[v.prow[k]];
      for (int32_t p = v.l_ptr[k] + lane; p < v.l_ptr[k + 1]; p += kWave) {
        const int32_t r = v.l_row[p];
        x[r] = x[r] - Lx[p] * xk;
      }
      wave_sync_mem();
    }
    // x[prow[k]] for k in U(:,j) is final once k was applied (later updates
    // only touch rows not pivotal before them), so U is read out here.
    const int32_t pr = v.prow[j];
    const T piv = x[pr];
    Ud[j] = piv;  // every lane stores the same value
    if (piv == T(0)) st |= kStZeroPivot;
    for (int32_t p = v.l_ptr[j] + lane; p < v.l_ptr[j + 1]; p += kWave) {
      const int32_t r = v.l_row[p];
      const T l = x[r] / piv;
      Lx[p] = l;
      const T al = tabs(l);
      if (!(al <= lmax_all)) lmax_all = al;  // NaN propagates as "large"
      x[r] = T(0);
    }
    for (int32_t up = v.u_ptr[j] + lane; up < v.u_ptr[j + 1]; up += kWave) {
      const int32_t r = v.prow[v.u_idx[up]];
      Ux[up] = x[r];
      x[r] = T(0);
    }
    wave_sync_mem();
    x[pr] = T(0);  // every lane stores the same value
    wave_sync_mem();
  }
  if (!(lmax_all <= static_cast<T>(v.inv_tol))) st |= kStCheck;
  for (int off = 32; off > 0; off >>= 1) st |= __shfl_xor(st, off);
  if (lane == 0 && st) atomicOr(status, st);
}

// The same for the few columns of a level whose reach is long (the column
// chain above would set the level's time): one 256-thread workgroup per
// column, row-oriented over the forward levels of the reach, as
// k_sp_tail_long; then the pivot row and the rows of L(:,j) each pull their
// row of L over the columns before j (later columns are not yet refactored).
template <typename T>
__global__ __launch_bounds__(256) void k_sp_cols_long(SpView v, const double* __restrict__ arena,
                                                      const double* __restrict__ rs, T* __restrict__ xs,
                                                      T* __restrict__ Lx, T* __restrict__ Ux, T* __restrict__ Ud,
                                                      int32_t* __restrict__ status, int32_t nlong,
                                                      const int32_t* __restrict__ long_j,
                                                      const int32_t* __restrict__ reach,
                                                      const int32_t* __restrict__ lev, int32_t nlev,
                                                      const int32_t* __restrict__ lr_ptr,
                                                      const int32_t* __restrict__ lr_col,
                                                      const int32_t* __restrict__ lr_vidx) {
  __shared__ int32_t st_sh;
  __shared__ T piv_sh;
  const int tid = threadIdx.x;
  T* x = xs + int64_t(blockIdx.x) * v.s;
  int32_t st = 0;
  T lmax_all = T(0);
  for (int32_t c = blockIdx.x; c < nlong; c += gridDim.x) {
    const int32_t j = long_j[c];
    for (int32_t p = v.a_ptr[j] + tid; p < v.a_ptr[j + 1]; p += 256) {
      const int32_t r = v.a_row[p];
      const double a = arena[v.a_val0 + p];
      x[r] = static_cast<T>(v.scale ? a / rs[r] : a);
    }
    __syncthreads();
    const int32_t* lv = lev + int64_t(c) * (nlev + 1);
    for (int32_t l = 0; l < nlev; ++l) {
      for (int32_t e = lv[l] + tid; e < lv[l + 1]; e += 256) {
        const int32_t k = reach[e];
        const int32_t r = v.prow[k];
        T acc = x[r];
        for (int32_t p = lr_ptr[k]; p < lr_ptr[k + 1]; ++p) {
          const T xv = x[v.prow[lr_col[p]]];
          acc = xv != T(0) ? acc - Lx[lr_vidx[p]] * xv : acc;
        }
        x[r] = acc;
      }
      __syncthreads();
    }
    // Rows at positions >= j: the pivot row (e = -1) and the rows of L(:,j).
    // Each reads only x of reach positions (< j), never another row of this
    // set, so every row is written as soon as it is summed.
    const int32_t nl = v.l_ptr[j + 1] - v.l_ptr[j];
    for (int32_t e = tid - 1; e < nl; e += 256) {
      const int32_t r = e < 0 ? v.prow[j] : v.l_row[v.l_ptr[j] + e];
      const int32_t kr = v.rpos[r];
      T acc = x[r];
      // Columns before j outside its reach have x = 0 here but need not be
      // refactored yet (same or later level): their L values are skipped.
      for (int32_t p = lr_ptr[kr]; p < lr_ptr[kr + 1] && lr_col[p] < j; ++p) {
        const T xv = x[v.prow[lr_col[p]]];
        acc = xv != T(0) ? acc - Lx[lr_vidx[p]] * xv : acc;
      }
      x[r] = acc;
      if (e < 0) piv_sh = acc;
    }
    __syncthreads();
    const T piv = piv_sh;
    if (tid == 0) {
      Ud[j] = piv;
      if (piv == T(0)) st |= kStZeroPivot;
    }
    for (int32_t p = v.l_ptr[j] + tid; p < v.l_ptr[j + 1]; p += 256) {
      const int32_t r = v.l_row[p];
      const T l = x[r] / piv;
      Lx[p] = l;
      const T al = tabs(l);
      if (!(al <= lmax_all)) lmax_all = al;
    }
    for (int32_t up = v.u_ptr[j] + tid; up < v.u_ptr[j + 1]; up += 256) Ux[up] = x[v.prow[v.u_idx[up]]];
    __syncthreads();
    for (int32_t p = v.l_ptr[j] + tid; p < v.l_ptr[j + 1]; p += 256) x[v.l_row[p]] = T(0);
    for (int32_t up = v.u_ptr[j] + tid; up < v.u_ptr[j + 1]; up += 256) x[v.prow[v.u_idx[up]]] = T(0);
    if (tid == 0) x[v.prow[j]] = T(0);
    __syncthreads();
  }
  if (!(lmax_all <= static_cast<T>(v.inv_tol))) st |= kStCheck;
  if (tid == 0) st_sh = 0;
  __syncthreads();
  if (st) atomicOr(&st_sh, st);
  __syncthreads();
  if (tid == 0 && st_sh) atomicOr(status, st_sh);
}

template <typename T>
void launch_sp_cols(hipStream_t st, const SpView& v, const double* arena, const double* rs, int32_t ncols,
                    const int32_t* cols, T* xs, int32_t nslots, T* Lx, T* Ux, T* Ud, int32_t* status,
                    const SpColsLong* lg) {
  if (ncols <= 0) return;
  const int g = ncols < nslots ? ncols : nslots;
  hipLaunchKernelGGL(k_sp_cols<T>, dim3(g), dim3(64), 0, st, v, arena, rs, ncols, cols, xs, Lx, Ux, Ud, status,
                     lg ? lg->is_long : nullptr);
  if (lg && lg->nlong > 0) {
    const int gl = lg->nlong < nslots ? lg->nlong : nslots;
    hipLaunchKernelGGL(k_sp_cols_long<T>, dim3(gl), dim3(256), 0, st, v, arena, rs, xs, Lx, Ux, Ud, status,
                       lg->nlong, lg->long_j, lg->reach, lg->lev, lg->nlev, lg->lr_ptr, lg->lr_col, lg->lr_vidx);
  }
}
template void launch_sp_cols<double>(hipStream_t, const SpView&, const double*, const double*, int32_t,
                                     const int32_t*, double*, int32_t, double*, double*, double*, int32_t*,
                                     const SpColsLong*);
template void launch_sp_cols<float>(hipStream_t, const SpView&, const double*, const double*, int32_t,
                                    const int32_t*, float*, int32_t, float*, float*, float*, int32_t*,
                                    const SpColsLong*);

// Tail columns: U(0:k0, j) and the Schur column S(:, t) = A(tail, j) -
// L(tail, 0:k0) U(0:k0, j). S arrives zero-filled, in the tail's fixed
// pivot-row order (column-major, leading dimension lds); rows pivotal in the
// sparse part live in the wave's scratch x, tail rows are updated in place in S
// (rpos maps a local row to its pivot position).
template <typename T>
__global__ __launch_bounds__(64) void k_sp_tail(SpView v, const double* __restrict__ arena,
                                                const double* __restrict__ rs, T* __restrict__ xs,
                                                const T* __restrict__ Lx, T* __restrict__ Ux, T* __restrict__ S,
                                                int64_t lds, const int32_t* __restrict__ is_long) {
  const int lane = threadIdx.x;
  T* x = xs + int64_t(blockIdx.x) * v.s;
  for (int32_t t = blockIdx.x; t < v.m; t += gridDim.x) {
    if (is_long && is_long[t]) continue;  // done by k_sp_tail_long
    const int32_t j = v.k0 + t;
    T* col = S + int64_t(t) * lds;
    for (int32_t p = v.a_ptr[j] + lane; p < v.a_ptr[j + 1]; p += kWave) {
      const int32_t r = v.a_row[p];
      const double a = arena[v.a_val0 + p];
      const T val = static_cast<T>(v.scale ? a / rs[r] : a);
      const int32_t q = v.rpos[r];
      if (q < v.k0) x[r] = val;
      else col[q - v.k0] = val;
    }
    wave_sync_mem();
    for (int32_t up = v.u_ptr[j]; up < v.u_ptr[j + 1]; ++up) {
      const int32_t k = v.u_idx[up];
      const T xk = x[v.prow[k]];
      for (int32_t p = v.l_ptr[k] + lane; p < v.l_ptr[k + 1]; p += kWave) {
        const int32_t r = v.l_row[p];
        const T d = Lx[p] * xk;
        const int32_t q = v.rpos[r];
        if (q < v.k0) x[r] = x[r] - d;
        else col[q - v.k0] = col[q - v.k0] - d;
      }
      wave_sync_mem();
    }
    for (int32_t up = v.u_ptr[j] + lane; up < v.u_ptr[j + 1]; up += kWave) {
      const int32_t r = v.prow[v.u_idx[up]];
      Ux[up] = x[r];
      x[r] = T(0);
    }
    wave_sync_mem();
  }
}
// The same tail columns for the few whose sparse reach is long (thousands of
// entries make the column-oriented chain above the longest path of the phase).
// One 256-thread workgroup per column, row-oriented: the reach positions come
// grouped by forward level, and within a level every position k computes
// U(k,j) = a(k,j) - sum_k' L(k,k') U(k',j) over its row of L (rows of one level
// depend only on lower levels, so there are no write conflicts and the sums
// have a fixed order). Then every tail row pulls its L21 products into S(:,t).
// reach/lev: the column's reach positions by level, lev[c*(nlev+1) + l] the
// start of level l in `reach` for long column c; lr_*: L by position rows.
template <typename T>
__global__ __launch_bounds__(256) void k_sp_tail_long(SpView v, const double* __restrict__ arena,
                                                      const double* __restrict__ rs, T* __restrict__ xs,
                                                      const T* __restrict__ Lx, T* __restrict__ Ux,
                                                      T* __restrict__ S, int64_t lds, int32_t nlong,
                                                      const int32_t* __restrict__ long_t,
                                                      const int32_t* __restrict__ reach,
                                                      const int32_t* __restrict__ lev, int32_t nlev,
                                                      const int32_t* __restrict__ lr_ptr,
                                                      const int32_t* __restrict__ lr_col,
                                                      const int32_t* __restrict__ lr_vidx) {
  const int tid = threadIdx.x;
  T* x = xs + int64_t(blockIdx.x) * v.s;  // this workgroup's work slot
  for (int32_t c = blockIdx.x; c < nlong; c += gridDim.x) {
  const int32_t t = long_t[c];
  const int32_t j = v.k0 + t;
  T* col = S + int64_t(t) * lds;
  for (int32_t p = v.a_ptr[j] + tid; p < v.a_ptr[j + 1]; p += 256) {
    const int32_t r = v.a_row[p];
    const double a = arena[v.a_val0 + p];
    const T val = static_cast<T>(v.scale ? a / rs[r] : a);
    const int32_t q = v.rpos[r];
    if (q < v.k0) x[r] = val;
    else col[q - v.k0] = val;
  }
  __syncthreads();
  const int32_t* lv = lev + int64_t(c) * (nlev + 1);
  for (int32_t l = 0; l < nlev; ++l) {
    for (int32_t e = lv[l] + tid; e < lv[l + 1]; e += 256) {
      const int32_t k = reach[e];
      const int32_t r = v.prow[k];
      T acc = x[r];
      for (int32_t p = lr_ptr[k]; p < lr_ptr[k + 1]; ++p) acc = acc - Lx[lr_vidx[p]] * x[v.prow[lr_col[p]]];
      x[r] = acc;
    }
    __syncthreads();
  }
  for (int32_t tt = tid; tt < v.m; tt += 256) {
    const int32_t k = v.k0 + tt;
    const int32_t p0 = lr_ptr[k], p1 = lr_ptr[k + 1];
    if (p0 == p1) continue;
    T acc = col[tt];
    for (int32_t p = p0; p < p1; ++p) acc = acc - Lx[lr_vidx[p]] * x[v.prow[lr_col[p]]];
    col[tt] = acc;
  }
  __syncthreads();
  for (int32_t up = v.u_ptr[j] + tid; up < v.u_ptr[j + 1]; up += 256) Ux[up] = x[v.prow[v.u_idx[up]]];
  __syncthreads();
  for (int32_t up = v.u_ptr[j] + tid; up < v.u_ptr[j + 1]; up += 256) x[v.prow[v.u_idx[up]]] = T(0);
  __syncthreads();
  }
}

template <typename T>
void launch_sp_tail(hipStream_t st, const SpView& v, const double* arena, const double* rs, T* xs, int32_t nslots,
                    const T* Lx, T* Ux, T* S, int64_t lds, const SpTailLong* lg) {
  if (v.m <= 0) return;
  const int g = v.m < nslots ? v.m : nslots;
  hipLaunchKernelGGL(k_sp_tail<T>, dim3(g), dim3(64), 0, st, v, arena, rs, xs, Lx, Ux, S, lds,
                     lg ? lg->is_long : nullptr);
  if (lg && lg->nlong > 0) {
    const int gl = lg->nlong < nslots ? lg->nlong : nslots;
    hipLaunchKernelGGL(k_sp_tail_long<T>, dim3(gl), dim3(256), 0, st, v, arena, rs, xs, Lx, Ux, S, lds, lg->nlong,
                       lg->long_t, lg->reach, lg->lev, lg->nlev, lg->lr_ptr, lg->lr_col, lg->lr_vidx);
  }
}
template void launch_sp_tail<double>(hipStream_t, const SpView&, const double*, const double*, double*, int32_t,
                                     const double*, double*, double*, int64_t, const SpTailLong*);
template void launch_sp_tail<float>(hipStream_t, const SpView&, const double*, const double*, float*, int32_t,
                                    const float*, float*, float*, int64_t, const SpTailLong*);

// ---------------------------------------------------------------------------
// Dense tail: no-pivot blocked right-looking LU, panel width 64. Per panel:
//   1. k_diag_lu_inv: LU of the 64x64 diagonal block (one wavefront, rows
//      in registers), with L11^{-1} and U11^{-1} formed alongside (one
//      wavefront each);
//   2. k_tail_panel: L21 = A21 U11^{-1} and U12 = L11^{-1} A12 as MFMA tile
//      products (one launch, both jobs), with the |L| growth check;
//   3. k_tail_update: trailing update A22 -= L21 U12 on MFMA.
constexpr int kNb = 64;

// inv[0..4095]: L11^{-1} row-major (unit lower); inv[4096..8191]: U11^{-1}
// row-major (upper). Rows/columns at or past b are an identity pad. Each panel
// keeps its pair (the tail solve applies them to its diagonal tiles).
//
// Blocked LU of the 64x64 block in LDS, four 16-column steps, one workgroup of
// four wavefronts. Step K (columns c0 = 16K .. c0+15):
//   wave 0: the column panel's unblocked LU over all rows from c0 down (lane i
//     keeps row i's 16 entries in registers; at step k every lane takes pivot
//     row c0+k from lane c0+k with readlane and updates its own row), so the
//     16x16 diagonal factors and L_IK = A_IK U_KK^{-1} for every block below
//     come out of the same 16 dependent steps; column k of the panel is
//     published to LDS after step k;
//   wave 1 (lanes j >= c0+16): the row panel U_KJ = L_KK^{-1} A_KJ by forward
//     substitution down column j, from the published columns;
//   wave 3 (lanes j < 16): column j of the diagonal block's L_KK^{-1}, the same way;
//   waves 1 and 3 run up to kDiagStep steps behind wave 0: every wave passes a
//     workgroup barrier every kDiagStep steps (wave 0 after its steps, the
//     others before theirs), so every wave executes the same barriers;
//   wave 2, between the same barriers: U_{K-1,K-1}^{-1} (column by column,
//     backward substitution) and whichever off-diagonal blocks of the inverses
//     have become computable (below);
//   then the trailing update A_IJ -= L_IK U_KJ for I, J > K on MFMA, 16x16x16
//     block products spread over the four waves.
// Pivot reciprocals use the hardware reciprocal with Newton steps.
// The inverses' off-diagonal blocks follow from the diagonal ones:
// X_IJ = -Linv_II sum_{M=J}^{I-1} L_IM X_MJ (lower) and Y_IJ = -Uinv_II
// sum_{M=I+1}^{J} U_IM Y_MJ (upper), one MFMA chain each. Block row I of
// L^{-1} and block column J of U^{-1} are computable once panel I (J) is done,
// so wave 2 forms them during the next panel; only L^{-1}'s last block row and
// U^{-1}'s last block column remain after the LU.
constexpr int kDiagThreads = 4 * kNb;
constexpr int kDiagLF = kNb + 1;  // LDS column length of F and G (column-major, padded)
constexpr int kDiagStep = 4;      // panel steps between barriers
template <typename T>
constexpr std::size_t diag_lds_elems() {
  return 2 * kNb * kDiagLF + 16 + 16 / sizeof(T);  // F, G, a 16-element sink, one int32 (status)
}

__device__ __forceinline__ float readlane_t(float v, int src) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), src));
}
__device__ __forceinline__ double readlane_t(double v, int src) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(u)), src);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(u >> 32)), src);
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// Inputs of both forms: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16].
// Output element r of lane l is D[row(l / 16, r)][l % 16]; the f64 form
// interleaves its rows (measured on gfx950: row = 4 r + l/16), the f32 form
// keeps them in blocks of four (row = 4 (l/16) + r).
template <typename T>
struct MfmaOps;
template <>
struct MfmaOps<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 mma(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int lk, int r) { return 4 * r + lk; }
};
template <>
struct MfmaOps<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int lk, int r) { return 4 * lk + r; }
};

// 1/x from the hardware reciprocal refined by Newton steps (one for FP32, two
// for FP64): a few dependent instructions instead of the IEEE division
// sequence, within an ulp or so of 1/x; 1/0 stays infinite.
__device__ __forceinline__ float fast_rcp(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
}
__device__ __forceinline__ double fast_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
  return __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
}

// 16x16 block products on column-major LDS arrays (column length kDiagLF):
// acc (the output layout of MfmaOps) += sign * X Y with X at rows xr, columns
// xc of X0 and Y at rows yr, columns yc of Y0. MX / MY read a diagonal block
// of G as one triangle: 0 the whole block, 1 the unit lower triangle (L^{-1}),
// 2 the upper triangle with the diagonal (U^{-1}).
template <typename T>
__device__ __forceinline__ T tri_pick(T v, int r, int c, int mode) {
  return mode == 0 ? v : mode == 1 ? (r > c ? v : (r == c ? T(1) : T(0))) : (r <= c ? v : T(0));
}
template <typename T, int MX = 0, int MY = 0>
__device__ __forceinline__ void blk_mma(const T* X0, int xr, int xc, const T* Y0, int yr, int yc, T sign,
                                        typename MfmaOps<T>::v4& acc, int li, int lk) {
#pragma unroll
  for (int k0 = 0; k0 < 16; k0 += 4) {
    const T a = sign * tri_pick(X0[(xc + k0 + lk) * kDiagLF + xr + li], li, k0 + lk, MX);
    const T b = tri_pick(Y0[(yc + li) * kDiagLF + yr + k0 + lk], k0 + lk, li, MY);
    acc = MfmaOps<T>::mma(a, b, acc);
  }
}
template <typename T>
__device__ __forceinline__ void blk_load(const T* C0, int r0, int c0, typename MfmaOps<T>::v4& acc, int li, int lk) {
#pragma unroll
  for (int r = 0; r < 4; ++r) acc[r] = C0[(c0 + li) * kDiagLF + r0 + MfmaOps<T>::row(lk, r)];
}
template <typename T>
__device__ __forceinline__ void blk_store(T* C0, int r0, int c0, const typename MfmaOps<T>::v4& acc, int li, int lk) {
#pragma unroll
  for (int r = 0; r < 4; ++r) C0[(c0 + li) * kDiagLF + r0 + MfmaOps<T>::row(lk, r)] = acc[r];
}

// Body of the diagonal LU + inverses for a workgroup of 256 threads. fc: LDS
// of diag_lds_elems<T>() elements, 16-byte aligned. kPreloaded: the block is
// already in fc (F, identity past b, as the load below writes it), put there
// by the tile update that produced it, so it is not read back from S.
template <typename T, bool kPreloaded = false>
__device__ __forceinline__ void diag_lu_inv_body(T* __restrict__ S, int64_t lds, int32_t p, int32_t b, double limit,
                                                 int32_t* __restrict__ status, T* __restrict__ inv, T* fc) {
  using V = typename MfmaOps<T>::v4;
  constexpr int LF = kDiagLF;
  T* F = fc;                   // F[j*LF + i] = A(i, j), then the factors
  // G: the inverses in one array, L^{-1} below the diagonal and U^{-1} on and
  // above it (in the diagonal blocks too: L^{-1}'s unit diagonal is implied).
  T* G = fc + kNb * LF;
  T* sink = fc + 2 * kNb * LF;  // where lanes without a column of their own store theirs
  int32_t* st_sh = reinterpret_cast<int32_t*>(fc + 2 * kNb * LF + 16);
  const int tid = threadIdx.x;
  const int lane = tid & (kNb - 1);
  const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(tid >> 6));
  const int li = lane & 15, lk = lane >> 4;
  if (tid == 0) *st_sh = 0;
  if constexpr (!kPreloaded) {
    __syncthreads();
    // The block into LDS, identity past b (coalesced down columns; clamped
    // addresses, unconditional loads).
#pragma unroll
    for (int u = 0; u < kNb * kNb / kDiagThreads; ++u) {
      const int e = u * kDiagThreads + tid;
      const int i = e & (kNb - 1), j = e >> 6;
      const T v = S[int64_t(p + (j < b ? j : 0)) * lds + p + (i < b ? i : 0)];
      F[j * LF + i] = (i < b && j < b) ? v : (i == j ? T(1) : T(0));
    }
  }
  __syncthreads();
  // Pieces of the inverses (column j of U_bb^{-1}; block (I,J) of L^{-1} or of
  // U^{-1} off the diagonal), run by wave 2 between the panel steps' barriers
  // as soon as their inputs are final, the rest after the last panel.
  // Column j (lane j < 16) of U_bb^{-1} by backward substitution, column
  // oriented (each step one multiply on the critical path), in two halves
  // (steps 15..8, then 7..0 and the store) so that it can span two of the
  // panel's barrier intervals; every lane stores unconditionally, as in the
  // panel.
  T ux[16];
  auto u_leaf_a = [&](int bb) {
    const int b0 = 16 * bb;
#pragma unroll
    for (int q = 0; q < 16; ++q) ux[q] = q == lane ? T(1) : T(0);
#pragma unroll
    for (int c = 15; c >= 8; --c) {
      ux[c] = ux[c] * fast_rcp(F[(b0 + c) * LF + b0 + c]);
#pragma unroll
      for (int r = 0; r < c; ++r) ux[r] = tfma(-F[(b0 + c) * LF + b0 + r], ux[c], ux[r]);
      if (c % 2 == 0) __builtin_amdgcn_sched_barrier(0);  // LDS reads of 2 steps in flight at a time (registers)
    }
  };
  auto u_leaf_b = [&](int bb) {
    const int b0 = 16 * bb;
#pragma unroll
    for (int c = 7; c >= 0; --c) {
      ux[c] = ux[c] * fast_rcp(F[(b0 + c) * LF + b0 + c]);
#pragma unroll
      for (int r = 0; r < c; ++r) ux[r] = tfma(-F[(b0 + c) * LF + b0 + r], ux[c], ux[r]);
      if (c % 2 == 0) __builtin_amdgcn_sched_barrier(0);
    }
    const int j = lane;
    T* col = j < 16 ? G + (b0 + j) * LF + b0 : sink;
#pragma unroll
    for (int q = 0; q < 16; ++q) (q > j ? sink : col)[q] = ux[q];  // the upper triangle, with the diagonal
  };
  auto l_item = [&](int I, int J) {  // X_IJ = -Linv_II sum_{M=J}^{I-1} L_IM X_MJ
    V acc = V{T(0), T(0), T(0), T(0)};
    blk_mma<T, 0, 1>(F, 16 * I, 16 * J, G, 16 * J, 16 * J, T(1), acc, li, lk);
    for (int M = J + 1; M < I; ++M) blk_mma<T>(F, 16 * I, 16 * M, G, 16 * M, 16 * J, T(1), acc, li, lk);
    blk_store<T>(G, 16 * I, 16 * J, acc, li, lk);
    acc = V{T(0), T(0), T(0), T(0)};
    blk_mma<T, 1, 0>(G, 16 * I, 16 * I, G, 16 * I, 16 * J, T(-1), acc, li, lk);
    blk_store<T>(G, 16 * I, 16 * J, acc, li, lk);
  };
  auto u_item = [&](int I, int J) {  // Y_IJ = -Uinv_II sum_{M=I+1}^{J} U_IM Y_MJ
    V acc = V{T(0), T(0), T(0), T(0)};
    for (int M = I + 1; M < J; ++M) blk_mma<T>(F, 16 * I, 16 * M, G, 16 * M, 16 * J, T(1), acc, li, lk);
    blk_mma<T, 0, 2>(F, 16 * I, 16 * J, G, 16 * J, 16 * J, T(1), acc, li, lk);
    blk_store<T>(G, 16 * I, 16 * J, acc, li, lk);
    acc = V{T(0), T(0), T(0), T(0)};
    blk_mma<T, 2, 0>(G, 16 * I, 16 * I, G, 16 * I, 16 * J, T(-1), acc, li, lk);
    blk_store<T>(G, 16 * I, 16 * J, acc, li, lk);
  };
  int32_t st = 0;
#pragma unroll 1
  for (int K = 0; K < 4; ++K) {
    const int c0 = 16 * K;
    if (w == 0) {
      // Column panel: rows c0..63 of columns c0..c0+15 (lanes above c0 only
      // follow along). The next pivot's reciprocal starts as soon as that
      // pivot is final, so the division overlaps the rest of the step.
      T a[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) a[q] = F[(c0 + q) * LF + lane];
      T akk = readlane_t(a[0], c0);
      T rp = fast_rcp(akk);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (akk == T(0)) st |= kStZeroPivot;
        const bool below = lane > c0 + k;
        const T lik = a[k] * rp;
        const T m = below ? -lik : T(0);
        if (k + 1 < 16) {
          a[k + 1] = tfma(m, readlane_t(a[k + 1], c0 + k), a[k + 1]);
          akk = readlane_t(a[k + 1], c0 + k + 1);
          rp = fast_rcp(akk);
        }
#pragma unroll
        for (int j = k + 2; j < 16; ++j) a[j] = tfma(m, readlane_t(a[j], c0 + k), a[j]);
        a[k] = below ? lik : a[k];
        if (lane >= c0) F[(c0 + k) * LF + lane] = a[k];  // column k: final in both factors
        if ((k + 1) % kDiagStep == 0) __syncthreads();
      }
    } else if (w == 1 || w == 3) {
      // Wave 1, lane j >= c0+16: column j of the row panel; wave 3, lane j <
      // 16: column j of L_KK^{-1}, from e_j (the identity is written to its
      // place first, so both waves run the same code on different columns).
      // Every lane stores its result unconditionally (lanes without a column
      // into the sink): a store on some lanes only lets the compiler sink the
      // arithmetic into that branch while every step's LDS reads stay above
      // it, all live at once.
      const int j = lane;
      const bool mine = w == 1 ? j >= c0 + 16 : j < 16;
      T* col = !mine ? sink : w == 1 ? F + j * LF + c0 : G + (c0 + j) * LF + c0;
      T x[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const T v = F[(w == 1 ? j : c0) * LF + c0 + q];  // wave 3: any in-bounds address
        x[q] = w == 1 ? v : (q == j ? T(1) : T(0));
      }
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (k % kDiagStep == 0) __syncthreads();
#pragma unroll
        for (int q = k + 1; q < 16; ++q) x[q] = tfma(-F[(c0 + k) * LF + c0 + q], x[k], x[q]);
      }
      // Wave 3 keeps the strictly lower entries (q > j); the rest of the
      // column of G belongs to U_KK^{-1}.
#pragma unroll
      for (int q = 0; q < 16; ++q) (w == 3 && q <= j ? sink : col)[q] = x[q];
    } else {
      // Wave 2, during panel K: U_{K-1}^{-1}, then the blocks of L^{-1}'s row
      // K-1 and of U^{-1}'s column K-1 (their inputs were final when panel K
      // started; panel K writes none of what they read or write). One piece
      // per interval between the panel's barriers, the first before the first
      // barrier; FP64 spreads U_{K-1}^{-1} over two intervals (its steps
      // outlast one: -4% on the diagonal LU). In FP32 that split raised the
      // fused kernel's registers and made it 8% slower, so U_{K-1}^{-1} stays
      // whole there.
#pragma unroll 1
      for (int t = 0; t < 16 / kDiagStep; ++t) {
        if constexpr (sizeof(T) == 8) {
          if (K >= 1 && t == 0) u_leaf_a(K - 1);
          if (K >= 1 && t == 1) u_leaf_b(K - 1);
          if (K >= 2 && t == 2)
            for (int J = 0; J < K - 1; ++J) l_item(K - 1, J);
          if (K >= 2 && t == 3)
            for (int I = K - 2; I >= 0; --I) u_item(I, K - 1);
          __syncthreads();
        } else {
          if (K >= 1 && t == 0) {
            u_leaf_a(K - 1);
            u_leaf_b(K - 1);
          }
          if (K >= 2 && t == 1)
            for (int J = 0; J < K - 1; ++J) l_item(K - 1, J);
          if (K >= 2 && t == 2)
            for (int I = K - 2; I >= 0; --I) u_item(I, K - 1);
          __syncthreads();
        }
      }
    }
    __syncthreads();
    // Trailing update A_IJ -= L_IK U_KJ, I, J > K.
    const int nr = 3 - K;
    for (int pr = w; pr < nr * nr; pr += 4) {
      const int I = K + 1 + pr % nr, J = K + 1 + pr / nr;
      V acc;
      blk_load<T>(F, 16 * I, 16 * J, acc, li, lk);
      blk_mma<T>(F, 16 * I, c0, F, c0, 16 * J, T(-1), acc, li, lk);
      blk_store<T>(F, 16 * I, 16 * J, acc, li, lk);
    }
    __syncthreads();
  }
  // L^{-1}'s last block row (waves 0, 1, 3) and U^{-1}'s last block column
  // (wave 2, after U_33^{-1}).
  if (w == 2) {
    u_leaf_a(3);
    u_leaf_b(3);
    for (int I = 2; I >= 0; --I) u_item(I, 3);
  } else {
    l_item(3, w == 3 ? 2 : w);
  }
  // |L| check on the finished factor (a NaN fails it) and the status, then the
  // factor back to S (coalesced down columns) and the inverses, row-major, to
  // inv, with no barrier after the stores.
  T lmax = T(0);
#pragma unroll
  for (int u = 0; u < kNb * kNb / kDiagThreads; ++u) {
    const int e = u * kDiagThreads + tid;
    const int i = e & (kNb - 1), j = e >> 6;
    const T al = i > j ? tabs(F[j * LF + i]) : T(0);
    lmax = (al <= lmax || lmax != lmax) ? lmax : al;
  }
  if (!(lmax <= static_cast<T>(limit))) st |= kStCheck;
  if (st) atomicOr(st_sh, st);
  __syncthreads();
  if (tid == 0 && *st_sh) atomicOr(status, *st_sh);
#pragma unroll
  for (int u = 0; u < kNb * kNb / kDiagThreads; ++u) {
    const int e = u * kDiagThreads + tid;
    const int i = e & (kNb - 1), j = e >> 6;
    if (i < b && j < b) S[int64_t(p + j) * lds + p + i] = F[j * LF + i];
  }
#pragma unroll
  for (int u = 0; u < kNb * kNb / kDiagThreads; ++u) {
    const int e = u * kDiagThreads + tid;
    const int i = e >> 6, j = e & (kNb - 1);
    const T g = G[j * LF + i];
    inv[e] = i > j ? g : (i == j ? T(1) : T(0));
    inv[kNb * kNb + e] = i <= j ? g : T(0);
  }
}

template <typename T>
__global__ __launch_bounds__(kDiagThreads) void k_diag_lu_inv(T* __restrict__ S, int64_t lds, int32_t p, int32_t b,
                                                              double limit, int32_t* __restrict__ status,
                                                              T* __restrict__ inv) {
  __shared__ __attribute__((aligned(16))) T fc[diag_lds_elems<T>()];
  diag_lu_inv_body<T>(S, lds, p, b, limit, status, inv, fc);
}

// One tile product job: C(MxN) op= A(MxK) B(KxN), K <= 64, element (i,j) of A
// at A[i*a_r + j*a_c] (B likewise); C is column-major (c_r must be 1, column
// stride c_c). op: 0 C -= AB, 1 C = AB, 2 C = AB and check |C| <= limit.
template <typename T>
struct TileJob {
  int32_t M, N, K, op, tiles_m, tiles_n;
  const T* A;
  int64_t a_r, a_c;
  const T* B;
  int64_t b_r, b_c;
  T* C;
  int64_t c_r, c_c;
};

// Panel products (ops 1 and 2). 128x128 output tile per 256-thread workgroup
// (each wavefront a 64x64 quarter, 4x4 MFMA 16x16 blocks), K <= 64 staged
// whole in LDS (zero padded), so C may alias A (a tile covers all K columns of
// its rows) or B (all K rows of its columns). Each MFMA forms a block of C^T
// (B^T as the row operand, A^T as the column operand): the output's fast lane
// index then runs down a column of C, so every C access is a 16-element
// contiguous run. All global loads are unconditional with clamped indices (a
// conditional load turns into a branch that waits for it). Workgroups below
// j0.tiles_m*j0.tiles_n run job 0, the rest job 1.
// kTile: output tile edge (128, or 64 for small trailing matrices, where 128
// leaves most CUs idle); each wavefront computes a kTile/2 square.
template <int kTile, typename T>
constexpr std::size_t tile_lds_elems() { return 2 * kNb * (kTile + 1); }

template <typename T, int kTile>
__device__ __forceinline__ void tile_gemm(const TileJob<T>& j0, const TileJob<T>& j1, double limit,
                                          int32_t* __restrict__ status, T* lds, int32_t bid) {
  using V = typename MfmaOps<T>::v4;
  constexpr int kGT = kTile;
  constexpr int kGLd = kGT + 1;    // LDS row length (padded against bank conflicts)
  constexpr int kWT = kGT / 2;     // wavefront sub-tile edge
  constexpr int kMB = kGT / 32;    // 16x16 MFMA blocks per wavefront and dimension
  constexpr int kRounds = kGT / 64;  // rounds of 16 staged elements per thread and operand
  T* As = lds;                // As[k*kGLd + i] = A(i0 + i, k)
  T* Bs = lds + kNb * kGLd;   // Bs[k*kGLd + j] = +-B(k, c0 + j)
  const int32_t n0 = j0.tiles_m * j0.tiles_n;
  const bool second = bid >= n0;
  const TileJob<T>& J = second ? j1 : j0;
  const int32_t tb = second ? bid - n0 : bid;
  const int32_t i0 = (tb % J.tiles_m) * kGT, c0 = (tb / J.tiles_m) * kGT;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * kWT, wc = (wave & 1) * kWT;
  const int li = lane & 15, lk = lane >> 4;
  const int32_t M = J.M, N = J.N, K = J.K;
  V acc[kMB][kMB];  // acc[c][a]: C^T block (j-block c, i-block a)
#pragma unroll
  for (int c = 0; c < kMB; ++c)
#pragma unroll
    for (int a = 0; a < kMB; ++a) acc[c][a] = V{T(0), T(0), T(0), T(0)};
  // A tile: 128 x 64, B tile: 64 x 128, 32 elements each per thread, loaded in
  // two rounds of 16 with the unit-stride index running across threads.
#pragma unroll
  for (int half = 0; half < kRounds; ++half) {
    T v[16];
    int ii[16], kk[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int e = (half * 16 + u) * 256 + tid;
      if (J.a_r == 1) {
        ii[u] = e & (kGT - 1);
        kk[u] = e / kGT;
      } else {
        kk[u] = e & (kNb - 1);
        ii[u] = e >> 6;
      }
      const int32_t gi = min(i0 + ii[u], M - 1), gk = min(kk[u], K - 1);
      v[u] = J.A[int64_t(gi) * J.a_r + int64_t(gk) * J.a_c];
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) As[kk[u] * kGLd + ii[u]] = (i0 + ii[u] < M && kk[u] < K) ? v[u] : T(0);
  }
#pragma unroll
  for (int half = 0; half < kRounds; ++half) {
    T v[16];
    int jj[16], kk[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int e = (half * 16 + u) * 256 + tid;
      if (J.b_r == 1) {
        kk[u] = e & (kNb - 1);
        jj[u] = e >> 6;
      } else {
        jj[u] = e & (kGT - 1);
        kk[u] = e / kGT;
      }
      const int32_t gj = min(c0 + jj[u], N - 1), gk = min(kk[u], K - 1);
      v[u] = J.B[int64_t(gk) * J.b_r + int64_t(gj) * J.b_c];
    }
#pragma unroll
    for (int u = 0; u < 16; ++u)
      Bs[kk[u] * kGLd + jj[u]] = (c0 + jj[u] < N && kk[u] < K) ? v[u] : T(0);
  }
  __syncthreads();
  for (int k = 0; k < kNb; k += 4) {
    T af[kMB], bf[kMB];
#pragma unroll
    for (int a = 0; a < kMB; ++a) af[a] = As[(k + lk) * kGLd + wr + a * 16 + li];
#pragma unroll
    for (int c = 0; c < kMB; ++c) bf[c] = Bs[(k + lk) * kGLd + wc + c * 16 + li];
#pragma unroll
    for (int c = 0; c < kMB; ++c)
#pragma unroll
      for (int a = 0; a < kMB; ++a) acc[c][a] = MfmaOps<T>::mma(bf[c], af[a], acc[c][a]);
  }
  T lmax = T(0);
#pragma unroll
  for (int c = 0; c < kMB; ++c)
#pragma unroll
    for (int a = 0; a < kMB; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int32_t i = i0 + wr + a * 16 + li;
        const int32_t j = c0 + wc + c * 16 + MfmaOps<T>::row(lk, r);
        const T v = acc[c][a][r];
        if (i < M && j < N) J.C[i + int64_t(j) * J.c_c] = v;
        const T av = (i < M && j < N) ? tabs(v) : T(0);
        lmax = (av <= lmax || lmax != lmax) ? lmax : av;
      }
  if (J.op == 2) {
    int32_t st = !(lmax <= static_cast<T>(limit)) ? kStCheck : 0;
    for (int off = 32; off > 0; off >>= 1) st |= __shfl_xor(st, off);
    if (lane == 0 && st) atomicOr(status, st);
  }
}
// Trailing update C -= A B (op 0), A (M x K) and B (K x N) column-major
// (a_r = b_r = 1), same tiles, waves and MFMA blocks as the panel product. The
// accumulators start as C and B is negated, so the MFMA chain yields C - AB.
// K runs in chunks of 16 through a double-buffered LDS ring: the next chunk's
// global loads are in flight while the MFMAs take the current one, and the C
// loads overlap the first chunk's staging. Per element the MFMA sequence is
// that of a single staged pass over K = 0..63. C does not alias A or B.
constexpr int kUpdKc = 32;
template <int kTile, typename T>
constexpr std::size_t update_lds_elems() { return 2 * 2 * kUpdKc * (kTile + 1); }

// kK: the largest K the tile takes (64, or 128 for two panels applied at once;
// the chunks past J.K stage zeros).
// kToDiag (one 64x64 tile at the job's origin): the result goes to lds in the
// diagonal body's layout (column-major, leading dimension kDiagLF, identity
// past M x N) instead of to C.
template <typename T, int kTile, bool kAllChunks = false, int kK = kNb, bool kToDiag = false>
__device__ __forceinline__ void tile_update(const TileJob<T>& j0, const TileJob<T>& j1, T* lds, int32_t bid) {
  using V = typename MfmaOps<T>::v4;
  constexpr int kLd = kTile + 1;                // LDS row length (padded against bank conflicts)
  constexpr int kWT = kTile / 2;                // wavefront sub-tile edge
  constexpr int kMB = kTile / 32;               // 16x16 MFMA blocks per wavefront and dimension
  constexpr int kPer = kTile * kUpdKc / 256;    // staged elements per thread, operand and chunk
  constexpr int kChunks = kK / kUpdKc;
  static_assert(!kAllChunks || kK == kNb, "one register set per chunk only for K = 64");
  // Ring slot s: As = lds + s * 2 * kUpdKc * kLd (As[k*kLd + i] = A(i0 + i, k0 + k)),
  // Bs = As + kUpdKc * kLd (Bs[k*kLd + j] = -B(k0 + k, c0 + j)).
  const int32_t n0 = j0.tiles_m * j0.tiles_n;
  const bool second = bid >= n0;
  const TileJob<T>& J = second ? j1 : j0;
  const int32_t tb = second ? bid - n0 : bid;
  const int32_t i0 = (tb % J.tiles_m) * kTile, c0 = (tb / J.tiles_m) * kTile;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * kWT, wc = (wave & 1) * kWT;
  const int li = lane & 15, lk = lane >> 4;
  const int32_t M = J.M, N = J.N, K = J.K;
  V acc[kMB][kMB];  // acc[c][a]: C^T block (j-block c, i-block a)
#pragma unroll
  for (int c = 0; c < kMB; ++c)
#pragma unroll
    for (int a = 0; a < kMB; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int32_t i = min(i0 + wr + a * 16 + li, M - 1);
        const int32_t j = min(c0 + wc + c * 16 + MfmaOps<T>::row(lk, r), N - 1);
        acc[c][a][r] = J.C[i + int64_t(j) * J.c_c];
      }
  // A chunk: kTile rows x 16 columns with the row index running across
  // threads; B chunk: 16 rows x kTile columns with the row (k) index running
  // across threads (both unit stride in memory). Register sets: one (the next
  // chunk's loads in flight while the MFMAs take the current one), or with
  // kAllChunks one per chunk, all loads issued up front, for kernels whose
  // register count allows few waves per SIMD, so each wave keeps more of its
  // own loads in flight.
  constexpr int kSets = kAllChunks ? kChunks : 1;
  T ra[kSets][kPer], rb[kSets][kPer];
  auto gload = [&](int set, int k0) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = u * 256 + tid;
      const int ii = e % kTile, ka = e / kTile;
      ra[set][u] = J.A[int64_t(min(i0 + ii, M - 1)) + int64_t(min(k0 + ka, K - 1)) * J.a_c];
      const int kb = e % kUpdKc, jj = e / kUpdKc;
      rb[set][u] = J.B[int64_t(min(k0 + kb, K - 1)) + int64_t(min(c0 + jj, N - 1)) * J.b_c];
    }
  };
  auto lstore = [&](int slot, int set, int k0) {
    T* As = lds + slot * 2 * kUpdKc * kLd;
    T* Bs = As + kUpdKc * kLd;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = u * 256 + tid;
      const int ii = e % kTile, ka = e / kTile;
      As[ka * kLd + ii] = (i0 + ii < M && k0 + ka < K) ? ra[set][u] : T(0);
      const int kb = e % kUpdKc, jj = e / kUpdKc;
      Bs[kb * kLd + jj] = (c0 + jj < N && k0 + kb < K) ? -rb[set][u] : T(0);
    }
  };
  if constexpr (kSets == 1) {
    gload(0, 0);
  } else {
#pragma unroll
    for (int c = 0; c < kChunks; ++c) gload(c, c * kUpdKc);
  }
  lstore(0, 0, 0);
  __syncthreads();
#pragma unroll
  for (int ch = 0; ch < kChunks; ++ch) {
    if (kSets == 1 && ch + 1 < kChunks) gload(0, (ch + 1) * kUpdKc);
    const T* As = lds + (ch & 1) * 2 * kUpdKc * kLd;
    const T* Bs = As + kUpdKc * kLd;
#pragma unroll
    for (int k = 0; k < kUpdKc; k += 4) {
      T af[kMB], bf[kMB];
#pragma unroll
      for (int a = 0; a < kMB; ++a) af[a] = As[(k + lk) * kLd + wr + a * 16 + li];
#pragma unroll
      for (int c = 0; c < kMB; ++c) bf[c] = Bs[(k + lk) * kLd + wc + c * 16 + li];
#pragma unroll
      for (int c = 0; c < kMB; ++c)
#pragma unroll
        for (int a = 0; a < kMB; ++a) acc[c][a] = MfmaOps<T>::mma(bf[c], af[a], acc[c][a]);
    }
    // The other slot was last read before the previous barrier.
    if (ch + 1 < kChunks) lstore((ch + 1) & 1, kSets == 1 ? 0 : ch + 1, (ch + 1) * kUpdKc);
    __syncthreads();
  }
  if constexpr (kToDiag) {
    static_assert(kTile == kNb, "the diagonal block is one 64x64 tile");
    // After the loop's last barrier every wave is done with the staging ring.
#pragma unroll
    for (int c = 0; c < kMB; ++c)
#pragma unroll
      for (int a = 0; a < kMB; ++a)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = wr + a * 16 + li;
          const int j = wc + c * 16 + MfmaOps<T>::row(lk, r);
          lds[j * kDiagLF + i] = (i < M && j < N) ? acc[c][a][r] : (i == j ? T(1) : T(0));
        }
    return;
  }
#pragma unroll
  for (int c = 0; c < kMB; ++c)
#pragma unroll
    for (int a = 0; a < kMB; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int32_t i = i0 + wr + a * 16 + li;
        const int32_t j = c0 + wc + c * 16 + MfmaOps<T>::row(lk, r);
        if (i < M && j < N) J.C[i + int64_t(j) * J.c_c] = acc[c][a][r];
      }
}

// Trailing update and the panel pair: separate kernels so profiles and
// counters separate them.
template <typename T, int kTile>
__global__ __launch_bounds__(256) void k_tail_update(TileJob<T> j0, TileJob<T> j1, double limit,
                                                     int32_t* __restrict__ status) {
  __shared__ T lds[update_lds_elems<kTile, T>()];
  tile_update<T, kTile>(j0, j1, lds, static_cast<int32_t>(blockIdx.x));
}
// A trailing update with the next step's diagonal block folded in: workgroup 0
// first applies the update to the next diagonal block (job `next`, one 64x64
// tile, the same tile code), then factors it and forms its inverses (three of
// its four waves); workgroups from 1 on update the rest (jobs j0, j1), so the
// diagonal work runs beside the bulk of the update from the start. The tile
// and diagonal code share one LDS buffer.
// FP32: three waves per SIMD with 64x64 tiles and two with 128x128 (the
// diagonal path would otherwise raise the register count of the whole kernel;
// three with 128x128 tiles would spill).
// kK = 128: the jobs apply two consecutive panels at once (K = 128), which
// reads and writes the trailing matrix once per two panels; each element still
// gets the MFMA sequence of the two single-panel updates in turn.
template <typename T, int kTile, int kK = kNb>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? (kTile == 64 ? 3 : 2) : 1)))
void k_tail_update_diag(TileJob<T> j0, TileJob<T> j1, TileJob<T> next, double limit,
                        int32_t* __restrict__ status, T* __restrict__ S, int64_t lds_s, int32_t q, int32_t nb,
                        T* __restrict__ inv) {
  constexpr std::size_t kTileElems = update_lds_elems<kTile, T>();
  constexpr std::size_t kDiagElems = diag_lds_elems<T>();
  __shared__ __attribute__((aligned(16))) T lds[kTileElems > kDiagElems ? kTileElems : kDiagElems];
  if (blockIdx.x == 0) {
    // The updated block goes straight into the diagonal body's LDS (the tile
    // writes it after its last barrier; the body's first barrier makes it
    // visible), not through S.
    tile_update<T, 64, sizeof(T) == 8 && kK == kNb, kK, true>(next, next, lds, 0);
    diag_lu_inv_body<T, true>(S, lds_s, q, nb, limit, status, inv, lds);
    return;
  }
  // FP64 64x64 tiles with K = 64: all K chunks' loads up front (two waves per
  // SIMD here, set by the diagonal path's registers).
  tile_update<T, kTile, sizeof(T) == 8 && kTile == 64 && kK == kNb, kK>(j0, j1, lds,
                                                                      static_cast<int32_t>(blockIdx.x) - 1);
}
template <typename T, int kTile>
__global__ __launch_bounds__(256) void k_tail_panel(TileJob<T> j0, TileJob<T> j1, double limit,
                                                    int32_t* __restrict__ status) {
  __shared__ T lds[tile_lds_elems<kTile, T>()];
  tile_gemm<T, kTile>(j0, j1, limit, status, lds, static_cast<int32_t>(blockIdx.x));
}

namespace {
hipEvent_t timer_event(std::vector<hipEvent_t>* pool, hipStream_t st) {
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  (void)hipEventRecord(e, st);
  pool->push_back(e);
  return e;
}
}  // namespace

// Panel jobs and trailing update of one panel step with kTile x kTile tiles.
template <typename T, int kTile>
void tail_step(hipStream_t st, T* S, int64_t lds, int32_t p, int32_t b, int32_t rest, double limit,
               int32_t* status, T* inv, TailTimer* timer) {
  // inv: this panel's pair of diagonal-block inverses.
  const int32_t tr = (rest + kTile - 1) / kTile;
  TileJob<T> l21{rest, b, b, 2, tr, 1, S + int64_t(p) * lds + p + b, 1, lds, inv + kNb * kNb, kNb, 1,
                 S + int64_t(p) * lds + p + b, 1, lds};
  TileJob<T> u12{b, rest, b, 1, 1, tr, inv, kNb, 1, S + int64_t(p + b) * lds + p, 1, lds,
                 S + int64_t(p + b) * lds + p, 1, lds};
  hipLaunchKernelGGL((k_tail_panel<T, kTile>), dim3(2 * tr), dim3(256), 0, st, l21, u12, limit, status);
  if (timer && timer->panel_events) timer_event(timer->panel_events, st);
  TileJob<T> upd{rest, rest, b, 0, tr, tr, S + int64_t(p) * lds + p + b, 1, lds,
                 S + int64_t(p + b) * lds + p, 1, lds, S + int64_t(p + b) * lds + p + b, 1, lds};
  if (timer && timer->gemm_events) timer_event(timer->gemm_events, st);
  hipLaunchKernelGGL((k_tail_update<T, kTile>), dim3(tr * tr), dim3(256), 0, st, upd, upd, limit, status);
  if (timer && timer->gemm_events) {
    timer_event(timer->gemm_events, st);
    const double r = rest;
    timer->gemm_flops->push_back(2.0 * r * r * b);
    timer->gemm_bytes->push_back(sizeof(T) * (2.0 * r * r + 2.0 * r * b));
  }
}

// One step with lookahead: the panel pair of step p, then one launch for the
// whole trailing update: workgroup 0 updates the next diagonal block and
// factors it (the next step's diagonal LU), the others update the rest of the
// trailing matrix (two tile jobs: everything right of the next block column,
// and the block column below the next diagonal block). Every element gets the
// same MFMA sequence as in the single update, so results are identical.
template <typename T, int kTile>
void tail_step_lookahead(hipStream_t st, T* S, int64_t lds, int32_t p, int32_t b, int32_t rest, double limit,
                         int32_t* status, T* inv, TailTimer* timer) {
  // inv: this panel's pair of diagonal-block inverses; the next panel's follows.
  const int32_t tr = (rest + kTile - 1) / kTile;
  TileJob<T> l21{rest, b, b, 2, tr, 1, S + int64_t(p) * lds + p + b, 1, lds, inv + kNb * kNb, kNb, 1,
                 S + int64_t(p) * lds + p + b, 1, lds};
  TileJob<T> u12{b, rest, b, 1, 1, tr, inv, kNb, 1, S + int64_t(p + b) * lds + p, 1, lds,
                 S + int64_t(p + b) * lds + p, 1, lds};
  if (timer && timer->panel_events) timer_event(timer->panel_events, st);
  hipLaunchKernelGGL((k_tail_panel<T, kTile>), dim3(2 * tr), dim3(256), 0, st, l21, u12, limit, status);
  if (timer && timer->panel_events) timer_event(timer->panel_events, st);
  const int32_t q = p + b;                      // next diagonal block
  const int32_t nb = rest < kNb ? rest : kNb;   // its size
  const T* L21 = S + int64_t(p) * lds + q;      // rest x b, column stride lds
  const T* U12 = S + int64_t(q) * lds + p;      // b x rest, column stride lds
  TileJob<T> a{nb, nb, b, 0, 1, 1, L21, 1, lds, U12, 1, lds, S + int64_t(q) * lds + q, 1, lds};
  const int32_t n0 = rest - nb;
  // Job 0: all trailing rows, columns right of the next block column.
  TileJob<T> right{rest, n0, b, 0, tr, (n0 + kTile - 1) / kTile, L21, 1, lds, U12 + int64_t(nb) * lds, 1, lds,
                   S + int64_t(q + nb) * lds + q, 1, lds};
  // Job 1: the next block column below its diagonal block.
  TileJob<T> below{n0, nb, b, 0, (n0 + kTile - 1) / kTile, 1, L21 + nb, 1, lds, U12, 1, lds,
                   S + int64_t(q) * lds + q + nb, 1, lds};
  if (n0 <= 0) right.tiles_m = right.tiles_n = below.tiles_m = below.tiles_n = 0;
  const int32_t g = 1 + right.tiles_m * right.tiles_n + below.tiles_m * below.tiles_n;
  if (timer && timer->gemm_events) timer_event(timer->gemm_events, st);
  hipLaunchKernelGGL((k_tail_update_diag<T, kTile>), dim3(g), dim3(256), 0, st, right, below, a, limit, status, S,
                     lds, q, nb, inv + 2 * kNb * kNb);
  if (timer && timer->gemm_events) {
    timer_event(timer->gemm_events, st);
    // Algorithmic work of the launch: the whole rest x rest trailing update
    // (C read and written once, the two panels read once) plus the next
    // diagonal block's LU and its two triangular inverses.
    const double r = rest, k = b, d = nb;
    timer->gemm_flops->push_back(2.0 * r * r * k + (2.0 / 3.0 + 2.0 / 3.0) * d * d * d);
    timer->gemm_bytes->push_back(sizeof(T) * (2.0 * r * r + 2.0 * r * k));
  }
}

// Panel products of the panel at p (L21 = A21 U11^{-1} over the rest rows
// below it, U12 = L11^{-1} A12 over the rest columns right of it).
template <typename T, int kTile>
void tail_panel(hipStream_t st, T* S, int64_t lds, int32_t p, int32_t b, int32_t rest, double limit,
                int32_t* status, const T* inv, TailTimer* timer) {
  const int32_t tr = (rest + kTile - 1) / kTile;
  TileJob<T> l21{rest, b, b, 2, tr, 1, S + int64_t(p) * lds + p + b, 1, lds, inv + kNb * kNb, kNb, 1,
                 S + int64_t(p) * lds + p + b, 1, lds};
  TileJob<T> u12{b, rest, b, 1, 1, tr, inv, kNb, 1, S + int64_t(p + b) * lds + p, 1, lds,
                 S + int64_t(p + b) * lds + p, 1, lds};
  if (timer && timer->panel_events) timer_event(timer->panel_events, st);
  hipLaunchKernelGGL((k_tail_panel<T, kTile>), dim3(2 * tr), dim3(256), 0, st, l21, u12, limit, status);
  if (timer && timer->panel_events) timer_event(timer->panel_events, st);
}

// Two panels per trailing update. Panel s (at p, width 64, rest > 64 rows
// below it): its products, then one launch that applies it only to what panel
// s+1 needs: workgroup 0 updates the next diagonal block and factors it, the
// others the block column below it and the block row right of it (64-wide
// strips). Panel s+1 (at q = p + 64): its products, then one launch that
// applies both panels (K = 128) to the rest of the trailing matrix, with
// workgroup 0 updating and factoring the diagonal block after them. Every
// element gets the same MFMA sequence as with one update per panel, so the
// factors are identical; the trailing matrix is read and written once per two
// panels. The panel products use 64x64 tiles at every size: one side of each
// product is 64 wide, so a 128x128 tile would leave half its MFMAs idle and
// halve the workgroups (the per-element MFMA sequence is the same either way).
template <typename T>
void tail_pair(hipStream_t st, T* S, int64_t lds, int32_t p, int32_t rest, double limit, int32_t* status, T* inv,
               TailTimer* timer, int32_t small_tile_below) {
  const int64_t kInvPair = 2 * kNb * kNb;
  const int32_t b = kNb;
  tail_panel<T, 64>(st, S, lds, p, b, rest, limit, status, inv, timer);
  const int32_t q = p + b;   // panel s+1 (64 columns, since rest > 64)
  const int32_t n0 = rest - kNb;
  {
    const T* L21 = S + int64_t(p) * lds + q;  // rest x 64
    const T* U12 = S + int64_t(q) * lds + p;  // 64 x rest
    TileJob<T> a{kNb, kNb, b, 0, 1, 1, L21, 1, lds, U12, 1, lds, S + int64_t(q) * lds + q, 1, lds};
    const int32_t t0 = (n0 + kNb - 1) / kNb;
    TileJob<T> row{kNb, n0, b, 0, 1, t0, L21, 1, lds, U12 + int64_t(kNb) * lds, 1, lds,
                   S + int64_t(q + kNb) * lds + q, 1, lds};
    TileJob<T> col{n0, kNb, b, 0, t0, 1, L21 + kNb, 1, lds, U12, 1, lds, S + int64_t(q) * lds + q + kNb, 1, lds};
    if (timer && timer->panel_events) timer_event(timer->panel_events, st);
    hipLaunchKernelGGL((k_tail_update_diag<T, 64, kNb>), dim3(1 + 2 * t0), dim3(256), 0, st, row, col, a, limit,
                       status, S, lds, q, kNb, inv + kInvPair);
    if (timer && timer->panel_events) timer_event(timer->panel_events, st);
  }
  // Panel s+1.
  const int32_t rest1 = n0;
  tail_panel<T, 64>(st, S, lds, q, kNb, rest1, limit, status, inv + kInvPair, timer);
  const int32_t q2 = q + kNb;                        // next diagonal block
  const int32_t nb2 = rest1 < kNb ? rest1 : kNb;     // its size
  const int32_t n02 = rest1 - nb2;
  const int32_t K = 2 * kNb;
  const T* L21 = S + int64_t(p) * lds + q2;          // rest1 x 128 (panels s and s+1)
  const T* U12 = S + int64_t(q2) * lds + p;          // 128 x rest1
  TileJob<T> a{nb2, nb2, K, 0, 1, 1, L21, 1, lds, U12, 1, lds, S + int64_t(q2) * lds + q2, 1, lds};
  auto launch = [&](auto tile) {
    constexpr int kTile = decltype(tile)::value;
    const int32_t tr = (rest1 + kTile - 1) / kTile;
    TileJob<T> right{rest1, n02, K, 0, tr, (n02 + kTile - 1) / kTile, L21, 1, lds, U12 + int64_t(nb2) * lds, 1, lds,
                     S + int64_t(q2 + nb2) * lds + q2, 1, lds};
    TileJob<T> below{n02, nb2, K, 0, (n02 + kTile - 1) / kTile, 1, L21 + nb2, 1, lds, U12, 1, lds,
                     S + int64_t(q2) * lds + q2 + nb2, 1, lds};
    if (n02 <= 0) right.tiles_m = right.tiles_n = below.tiles_m = below.tiles_n = 0;
    const int32_t g = 1 + right.tiles_m * right.tiles_n + below.tiles_m * below.tiles_n;
    if (timer && timer->gemm_events) timer_event(timer->gemm_events, st);
    hipLaunchKernelGGL((k_tail_update_diag<T, kTile, 2 * kNb>), dim3(g), dim3(256), 0, st, right, below, a, limit,
                       status, S, lds, q2, nb2, inv + 2 * kInvPair);
  };
  if (rest1 < small_tile_below) launch(std::integral_constant<int, 64>{});
  else launch(std::integral_constant<int, 128>{});
  if (timer && timer->gemm_events) {
    timer_event(timer->gemm_events, st);
    // Algorithmic work of the launch: the rest1 x rest1 trailing update with
    // K = 128 (C read and written once, both panels read once) plus the next
    // diagonal block's LU and its two triangular inverses.
    const double r = rest1, k = K, d = nb2;
    timer->gemm_flops->push_back(2.0 * r * r * k + (2.0 / 3.0 + 2.0 / 3.0) * d * d * d);
    timer->gemm_bytes->push_back(sizeof(T) * (2.0 * r * r + 2.0 * r * k));
  }
}

template <typename T>
void launch_tail_lu(hipStream_t st, int32_t m, T* S, int64_t lds, double limit, int32_t* status, T* inv,
                    TailTimer* timer, int32_t small_tile_below, int lookahead) {
  if (lookahead) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_diag_lu_inv<T>, dim3(1), dim3(kDiagThreads), 0, st, S, lds, 0, m < kNb ? m : kNb, limit,
                       status, inv);
    const int64_t kInvPair = 2 * kNb * kNb;
    for (int32_t p = 0; p < m;) {
      const int32_t b = (m - p) < kNb ? (m - p) : kNb;
      const int32_t rest = m - p - b;
      if (rest <= 0) break;
      if (lookahead >= 2 && rest > kNb) {
        tail_pair<T>(st, S, lds, p, rest, limit, status, inv + p / kNb * kInvPair, timer, small_tile_below);
        p += 2 * kNb;
        continue;
      }
      if (rest < small_tile_below)
        tail_step_lookahead<T, 64>(st, S, lds, p, b, rest, limit, status, inv + p / kNb * kInvPair, timer);
      else
        tail_step_lookahead<T, 128>(st, S, lds, p, b, rest, limit, status, inv + p / kNb * kInvPair, timer);
      p += b;
    }
    return;
  }
  for (int32_t p = 0; p < m; p += kNb) {
    const int32_t b = (m - p) < kNb ? (m - p) : kNb;
    const int32_t rest = m - p - b;
    if (timer && timer->panel_events) timer_event(timer->panel_events, st);
    T* pinv = inv + int64_t(p / kNb) * 2 * kNb * kNb;
    hipLaunchKernelGGL(k_diag_lu_inv<T>, dim3(1), dim3(kDiagThreads), 0, st, S, lds, p, b, limit, status, pinv);
    if (rest <= 0) {
      if (timer && timer->panel_events) timer_event(timer->panel_events, st);
      break;
    }
    if (rest < small_tile_below)
      tail_step<T, 64>(st, S, lds, p, b, rest, limit, status, pinv, timer);
    else
      tail_step<T, 128>(st, S, lds, p, b, rest, limit, status, pinv, timer);
  }
}
template void launch_tail_lu<double>(hipStream_t, int32_t, double*, int64_t, double, int32_t*, double*, TailTimer*,
                                     int32_t, int);
template void launch_tail_lu<float>(hipStream_t, int32_t, float*, int64_t, double, int32_t*, float*, TailTimer*,
                                    int32_t, int);

// ---------------------------------------------------------------------------
// Solve.
// Multi-column vectors: column c of a vector set lives at base + c * ld.
// dst(i) = src(P[i]), divided by rs[P[i]] when row scale factors are given.
__global__ void k_permute(int32_t n, int32_t k, const int32_t* __restrict__ P, const double* __restrict__ src,
                          int64_t lds, double* __restrict__ dst, int64_t ldd, const double* __restrict__ rs) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int32_t pi = P[i];
    if (rs) {
      const double r = rs[pi];
      for (int32_t c = 0; c < k; ++c) dst[c * ldd + i] = src[c * lds + pi] / r;
    } else {
      for (int32_t c = 0; c < k; ++c) dst[c * ldd + i] = src[c * lds + pi];
    }
  }
}
__global__ void k_unpermute(int32_t n, int32_t k, const int32_t* __restrict__ Q, const double* __restrict__ src,
                            int64_t lds, double* __restrict__ dst, int64_t ldd) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int32_t qi = Q[i];
    for (int32_t c = 0; c < k; ++c) dst[c * ldd + qi] = src[c * lds + i];
  }
}
void launch_permute(hipStream_t st, int32_t n, int32_t k, const int32_t* P, const double* src, int64_t lds,
                    double* dst, int64_t ldd, const double* rs) {
  hipLaunchKernelGGL(k_permute, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, n, k, P, src, lds, dst, ldd, rs);
}
void launch_unpermute(hipStream_t st, int32_t n, int32_t k, const int32_t* Q, const double* src, int64_t lds,
                      double* dst, int64_t ldd) {
  hipLaunchKernelGGL(k_unpermute, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, n, k, Q, src, lds, dst, ldd);
}

__device__ __forceinline__ double readlane_f64(double v, int src) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(u)), src);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(u >> 32)), src);
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// Persistent solve of one segment of dense diagonal blocks (a segment is a
// maximal run of dense blocks between sparse blocks in the solve order). Each
// workgroup (one wavefront) claims block tickets in order (a topological order
// of the block DAG: solve level, then the longest dependent chain first) and
// solves block b: rhs = bperm - A_off * xperm (ubtf:2269-2286), then the pivot
// gather and the unit-L and U solves (ubtf:2071-2097), applied as one product
// with the block's explicit inverse M = U^{-1} L^{-1} (minv64/minv32,
// column-major s x s at the block's factor offset). x values are published
// values (kXSentinel protocol): everything that does not need other blocks' x
// (the rows of M into registers, row gather, the first off-diagonal entries,
// the next ticket) is issued first, then the block polls exactly the x entries
// its rows use.
// Deadlock-free: the lowest unfinished ticket is always the current block of
// some workgroup (a claimed, unstarted ticket is higher than its holder's
// current one), and every x entry it reads belongs to a lower ticket.
// x = M v over the first N columns (N >= s; entries of M past s are zero): the
// rows of M in registers, the input vector broadcast from LDS at wave-uniform
// addresses. No masks and no per-step branches.
template <int N, int K>
__device__ __forceinline__ void block_inverse_product(const double (&rowm)[64], double* __restrict__ vb,
                                                      double (&v)[K], int at) {
#pragma unroll
  for (int c = 0; c < K; ++c) vb[c * 64 + at] = v[c];
  // vb is this wavefront's own slice: a wave-level barrier orders the writes
  // before the broadcast reads (the other columns' waves need not wait).
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int c = 0; c < K; ++c) v[c] = 0.0;
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int c = 0; c < K; ++c) v[c] = __builtin_fma(rowm[j], vb[c * 64 + j], v[c]);
}

// NW right-hand sides, one wavefront each: the workgroup claims a ticket (thread
// 0, shared through LDS), and wave c solves the block for column c, so a block
// costs what it costs for one column however many there are.
template <int NW>
__global__ __launch_bounds__(64 * NW) void k_dense_seg_solve(
    int32_t ntasks, const int32_t* __restrict__ tasks, const int32_t* __restrict__ R,
    const int64_t* __restrict__ lu0_of_block, const double* __restrict__ minv64, const float* __restrict__ minv32,
    const int32_t* __restrict__ prec_of_block, const int32_t* __restrict__ gpinv, const int32_t* __restrict__ off_ptr,
    const int32_t* __restrict__ off_col, const double* __restrict__ off_val, int32_t* __restrict__ ticket,
    const double* __restrict__ bperm, double* __restrict__ xperm, int64_t ldv, int32_t* __restrict__ err) {
  __shared__ double vb[NW * 64];
  __shared__ int32_t tk_sh[2];  // ticket slots, alternating so one barrier per block suffices
  constexpr int kPre = 4;
  const int lane = threadIdx.x & 63;
  const int col = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const double* bcol = bperm + col * ldv;
  double* xcol = xperm + col * ldv;
  double* vbw = vb + col * 64;
  if (threadIdx.x == 0) tk_sh[0] = atomicAdd(ticket, 1);
  __syncthreads();
  int32_t t = __builtin_amdgcn_readfirstlane(tk_sh[0]);
  int slot = 1;
  bool late = false;
  while (t < ntasks) {
    if (threadIdx.x == 0) tk_sh[slot] = atomicAdd(ticket, 1);  // the next ticket, claimed ahead
    const int32_t b = __builtin_amdgcn_readfirstlane(tasks[t]);
    const int32_t r0 = __builtin_amdgcn_readfirstlane(R[b]);
    const int s = __builtin_amdgcn_readfirstlane(R[b + 1]) - r0;
    const bool act = lane < s;
    const int32_t r = r0 + lane;
    // --- independent of other blocks ---
    const int32_t at = act ? gpinv[r] : lane;  // this row's position after the pivot gather
    double acc[1] = {act ? bcol[r] : 0.0};
    const int32_t p0 = act ? off_ptr[r] : 0, p1 = act ? off_ptr[r + 1] : 0;
    int32_t oc[kPre];
    double ov[kPre];
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
      const bool ok = p0 + u < p1;
      oc[u] = ok ? off_col[p0 + u] : -1;
      ov[u] = ok ? off_val[p0 + u] : 0.0;
    }
    // Each lane's row of M straight into registers (clamped, unconditional
    // loads; zero past s and on lanes past s), read in the block's precision.
    const int li = lane < s ? lane : s - 1;
    const int64_t o = lu0_of_block[b];
    double rowm[64];
    if (__builtin_amdgcn_readfirstlane(prec_of_block[b]) != 0) {
#pragma unroll
      for (int j = 0; j < 64; ++j) {
        const double m = static_cast<double>(minv32[o + (j < s ? j : s - 1) * s + li]);
        rowm[j] = (j < s && act) ? m : 0.0;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 64; ++j) {
        const double m = minv64[o + (j < s ? j : s - 1) * s + li];
        rowm[j] = (j < s && act) ? m : 0.0;
      }
    }
    // --- published x of other blocks: all first polls in flight together, then
    // the rest in groups of kPre ---
    for (int32_t pg = p0;;) {
      uint64_t xu[kPre];
#pragma unroll
      for (int u = 0; u < kPre; ++u) xu[u] = oc[u] >= 0 ? ld_x_relaxed(xcol + oc[u]) : 0ull;
#pragma unroll
      for (int u = 0; u < kPre; ++u)
        if (oc[u] >= 0) acc[0] = acc[0] - ov[u] * finish_wait_x(xcol + oc[u], xu[u], late);
      pg += kPre;
      if (pg >= p1) break;
#pragma unroll
      for (int u = 0; u < kPre; ++u) {
        const bool ok = pg + u < p1;
        oc[u] = ok ? off_col[pg + u] : -1;
        ov[u] = ok ? off_val[pg + u] : 0.0;
      }
    }
    // Row gather (the block's pivoting), then x = M (P rhs) as a matrix-vector
    // product: every step reads the finished input vector, so there is no
    // dependency chain across the block.
    if (!act) acc[0] = 0.0;
    if (s <= 16) block_inverse_product<16, 1>(rowm, vbw, acc, at);
    else if (s <= 32) block_inverse_product<32, 1>(rowm, vbw, acc, at);
    else block_inverse_product<64, 1>(rowm, vbw, acc, at);
    if (act) st_x_relaxed(xcol + r, acc[0]);
    __syncthreads();  // vb is rewritten by the next block; the next ticket is in its slot
    t = __builtin_amdgcn_readfirstlane(tk_sh[slot]);
    slot ^= 1;
  }
  if (late) atomicOr(err, 1);
}

void launch_dense_seg_solve(hipStream_t st, int32_t kc, int32_t ntasks, const int32_t* tasks, const int32_t* R,
                            const int64_t* lu0_of_block, const double* minv64, const float* minv32,
                            const int32_t* prec_of_block,
                            const int32_t* gpinv, const int32_t* off_ptr, const int32_t* off_col,
                            const double* off_val, int32_t* ticket, const double* bperm,
                            double* xperm, int64_t ldv, int32_t* err, int32_t grid) {
  if (ntasks <= 0) return;
  const int g = ntasks < grid ? ntasks : grid;  // persistent workgroups, one wavefront per column
#define BDLU_SEG(KK)                                                                                               \
  hipLaunchKernelGGL(k_dense_seg_solve<KK>, dim3(g), dim3(64 * KK), 0, st, ntasks, tasks, R, lu0_of_block, minv64, \
                     minv32, prec_of_block, gpinv, off_ptr, off_col, off_val, ticket, bperm, xperm, ldv, err)
  switch (kc) {
    case 1: BDLU_SEG(1); break;
    case 2: BDLU_SEG(2); break;
    case 4: BDLU_SEG(4); break;
    default: BDLU_SEG(8); break;
  }
#undef BDLU_SEG
}

// Sparse block: y[k] = (bperm[r] - A_off(r,:) xperm) / Rs(row), r = start + rowperm[k],
// for each of the kc columns (bperm/xperm column stride ldv, y column stride ldy).
__global__ void k_sp_gather(int32_t s, int32_t start, int32_t kc, const int32_t* __restrict__ rowperm,
                            const double* __restrict__ rs, const int32_t* __restrict__ off_ptr,
                            const int32_t* __restrict__ off_col, const double* __restrict__ off_val,
                            const double* __restrict__ bperm, const double* __restrict__ xperm, int64_t ldv,
                            double* __restrict__ y, int64_t ldy) {
  for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < s; k += gridDim.x * blockDim.x) {
    const int32_t lr = rowperm[k];
    const int32_t r = start + lr;
    const double scale = rs ? rs[lr] : 1.0;
    for (int32_t c = 0; c < kc; ++c) {
      double acc = bperm[c * ldv + r];
      for (int32_t p = off_ptr[r]; p < off_ptr[r + 1]; ++p) acc = acc - off_val[p] * xperm[c * ldv + off_col[p]];
      y[c * ldy + k] = rs ? acc / scale : acc;
    }
  }
}
void launch_sp_gather(hipStream_t st, int32_t s, int32_t start, int32_t kc, const int32_t* rowperm, const double* rs,
                      const int32_t* off_ptr, const int32_t* off_col, const double* off_val, const double* bperm,
                      const double* xperm, int64_t ldv, double* y, int64_t ldy) {
  hipLaunchKernelGGL(k_sp_gather, dim3(grid_for(s, 256, 4096)), dim3(256), 0, st, s, start, kc, rowperm, rs, off_ptr,
                     off_col, off_val, bperm, xperm, ldv, y, ldy);
}

// Row-oriented sparse triangular step over one level: for each listed row k and
// each of the K columns, y[k] = (y[k] - sum vals[vidx[p]] * y[col[p]]) [/ diag[k]].
// A wavefront per row; the row's entries are read once for all K columns.
template <typename T, int K>
__global__ __launch_bounds__(64) void k_sp_rows(int32_t nrows, const int32_t* __restrict__ rows,
                                                const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                                                const int32_t* __restrict__ vidx, const T* __restrict__ vals,
                                                const T* __restrict__ diag, double* __restrict__ y, int64_t ldy) {
  const int lane = threadIdx.x;
  for (int32_t t = blockIdx.x; t < nrows; t += gridDim.x) {
    const int32_t k = rows[t];
    double acc[K];
#pragma unroll
    for (int c = 0; c < K; ++c) acc[c] = 0.0;
    for (int32_t p = ptr[k] + lane; p < ptr[k + 1]; p += kWave) {
      const double v = static_cast<double>(vals[vidx[p]]);
      const int32_t j = col[p];
#pragma unroll
      for (int c = 0; c < K; ++c) acc[c] += v * y[c * ldy + j];
    }
#pragma unroll
    for (int c = 0; c < K; ++c)
      for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off);
    // Every lane holds the same sums and stores the same values (no lane-0-only
    // region inside the loop).
#pragma unroll
    for (int c = 0; c < K; ++c) {
      double v = y[c * ldy + k] - acc[c];
      if (diag) v = v / static_cast<double>(diag[k]);
      y[c * ldy + k] = v;
    }
  }
}
template <typename T>
void launch_sp_rows(hipStream_t st, int32_t kc, int32_t nrows, const int32_t* rows, const int32_t* ptr,
                    const int32_t* col, const int32_t* vidx, const T* vals, const T* diag, double* y, int64_t ldy) {
  if (nrows <= 0) return;
  const dim3 g(grid_for(nrows, 1, 8192)), b(64);
  switch (kc) {
    case 1: hipLaunchKernelGGL((k_sp_rows<T, 1>), g, b, 0, st, nrows, rows, ptr, col, vidx, vals, diag, y, ldy); break;
    case 2: hipLaunchKernelGGL((k_sp_rows<T, 2>), g, b, 0, st, nrows, rows, ptr, col, vidx, vals, diag, y, ldy); break;
    case 4: hipLaunchKernelGGL((k_sp_rows<T, 4>), g, b, 0, st, nrows, rows, ptr, col, vidx, vals, diag, y, ldy); break;
    default: hipLaunchKernelGGL((k_sp_rows<T, 8>), g, b, 0, st, nrows, rows, ptr, col, vidx, vals, diag, y, ldy); break;
  }
}
template void launch_sp_rows<double>(hipStream_t, int32_t, int32_t, const int32_t*, const int32_t*, const int32_t*,
                                     const int32_t*, const double*, const double*, double*, int64_t);
template void launch_sp_rows<float>(hipStream_t, int32_t, int32_t, const int32_t*, const int32_t*, const int32_t*,
                                    const int32_t*, const float*, const float*, double*, int64_t);

// Coupling of each 64-row band of the dense tail to its neighbour in solve
// order, for the tail solve: G_w = L_ww^{-1} L_{w,w-1} (forward, bands w >= 1)
// and H_w = U_ww^{-1} U_{w,w+1} (backward, bands w < np-1), 64 x 64 row-major
// at g + (2 w + kind) * 64 * 64 (kind 0 forward, 1 backward), from the
// diagonal inverses the tail LU leaves in inv (same layout). One 256-thread
// workgroup per product: thread (i, quarter) forms row i's 16 columns of the
// quarter, k ascending; pads (rows or columns past the matrix) are zero.
template <typename T>
__global__ __launch_bounds__(256) void k_tail_coupling(int32_t m, const T* __restrict__ S, int64_t lds,
                                                       const T* __restrict__ inv, T* __restrict__ g) {
  constexpr int kL = kNb + 1;
  __shared__ T Ds[kNb * kL];  // Ds[i*65 + k] = D^{-1}(i, k)
  __shared__ T Ts[kNb * kL];  // Ts[j*65 + k] = T(k, j)
  const int32_t np = (m + kNb - 1) / kNb;
  const int kind = static_cast<int>(blockIdx.x) / np;
  const int32_t w = static_cast<int32_t>(blockIdx.x) % np;
  if (kind == 0 ? w == 0 : w == np - 1) return;  // no neighbour
  const int32_t r0 = w * kNb;
  const int32_t rows = (m - r0) < kNb ? (m - r0) : kNb;
  const int32_t c0 = (kind == 0 ? w - 1 : w + 1) * kNb;
  const int32_t cols = (m - c0) < kNb ? (m - c0) : kNb;
  const T* dinv = inv + (int64_t(w) * 2 + kind) * kNb * kNb;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < kNb * kNb; idx += 256) {
    Ds[(idx >> 6) * kL + (idx & 63)] = dinv[idx];
    const int j = idx >> 6, k = idx & 63;  // T(k, j) = S(r0 + k, c0 + j), column-major in S
    const T v = S[int64_t(c0 + (j < cols ? j : cols - 1)) * lds + r0 + (k < rows ? k : rows - 1)];
    Ts[j * kL + k] = (j < cols && k < rows) ? v : T(0);
  }
  __syncthreads();
  const int i = tid & 63, jq = tid >> 6;
  T acc[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) acc[u] = T(0);
#pragma unroll 8
  for (int k = 0; k < kNb; ++k) {
    const T a = Ds[i * kL + k];
#pragma unroll
    for (int u = 0; u < 16; ++u) acc[u] = tfma(a, Ts[(jq * 16 + u) * kL + k], acc[u]);
  }
  T* dst = g + (int64_t(w) * 2 + kind) * kNb * kNb + i * kNb + jq * 16;
#pragma unroll
  for (int u = 0; u < 16; ++u) dst[u] = acc[u];
}
template <typename T>
void launch_tail_coupling(hipStream_t st, int32_t m, const T* S, int64_t lds, const T* inv, T* g) {
  if (m <= kNb) return;  // one band: nothing to couple
  const int32_t np = (m + kNb - 1) / kNb;
  hipLaunchKernelGGL(k_tail_coupling<T>, dim3(2 * np), dim3(256), 0, st, m, S, lds, inv, g);
}
template void launch_tail_coupling<double>(hipStream_t, int32_t, const double*, int64_t, const double*, double*);
template void launch_tail_coupling<float>(hipStream_t, int32_t, const float*, int64_t, const float*, float*);

// Dense tail triangular solve (unit-lower forward or upper backward) on the
// column-major m x m factor, the inverses D_w^{-1} of its 64 x 64 diagonal
// tiles (inv) and the band couplings G_w (g, from launch_tail_coupling), one
// workgroup of four wavefronts per 64-row band and right-hand side, claimed by
// ticket in dependency order (ticket t: band t / kc in solve order, column
// t % kc; each column is its own solve). For band w with neighbour band v (w-1 forward, w+1
// backward): x_w = D_w^{-1} (y_w - sum over the older bands) - G_w x_v, so only
// the last product waits for the neighbour's values. Waves 1-3 take the older
// bands' tiles round-robin (each keeping the next tile piece in flight, polling
// the 64 values of the tile's band as they are published, kXSentinel protocol
// on yout, and broadcasting them through LDS); their partial sums are combined
// in a fixed order; wave 0 applies D_w^{-1} (staged in LDS) and then, once the
// neighbour's values are visible, G_w (its row of G_w in registers, loaded at
// the start) and publishes the band's 64 values. yin and yout must not alias;
// yout starts filled with kXSentinel.
template <typename T>
__global__ __launch_bounds__(256) void k_tail_trsv(int32_t m, int32_t kc, const T* __restrict__ S, int64_t lds,
                                                   const T* __restrict__ inv, const T* __restrict__ g,
                                                   const double* __restrict__ yin_all, int64_t ldi,
                                                   double* __restrict__ yout_all, int64_t ldo, int upper,
                                                   int32_t* __restrict__ ticket, int32_t* __restrict__ err) {
  constexpr int K = 1;
  constexpr int kW = 4;         // wavefronts per band; waves 1..kW-1 take the older tiles
  constexpr int kDL = kNb + 1;  // LDS row length of the diagonal inverse (padded against bank conflicts)
  __shared__ double xs[kW][K][kNb];
  __shared__ double part[kW][K][kNb];
  __shared__ T dt[kNb * kDL];   // inverse of the diagonal tile, row-major
  __shared__ int32_t tk_sh;
  const int tid = threadIdx.x;
  const int lane = tid & (kNb - 1);
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (tid == 0) tk_sh = atomicAdd(ticket, 1);
  __syncthreads();
  const int32_t tk = __builtin_amdgcn_readfirstlane(tk_sh);
  const int32_t np = (m + kNb - 1) / kNb;
  if (tk >= np * kc) return;  // uniform over the workgroup
  const int32_t t = tk / kc;
  const double* __restrict__ yin = yin_all + (tk % kc) * ldi;
  double* __restrict__ yout = yout_all + (tk % kc) * ldo;
  const int32_t w = upper ? (np - 1 - t) : t;
  const int32_t r0 = w * kNb;
  const int32_t rows = (m - r0) < kNb ? (m - r0) : kNb;
  const int32_t row = r0 + lane;
  const bool live = lane < rows;
  // Loads are unconditional with clamped indices, so they issue back to back
  // (a per-element condition becomes a branch that waits on each load). Lanes
  // past the band never store; columns past the matrix meet x = 0.
  const int32_t rowc = live ? row : r0;
  const int kind = upper ? 1 : 0;
  const T* tinv = inv + (int64_t(w) * 2 + kind) * kNb * kNb;
  for (int idx = tid; idx < kNb * kNb; idx += kW * kNb) dt[(idx >> 6) * kDL + (idx & (kNb - 1))] = tinv[idx];
  const int32_t nd = upper ? (np - 1 - w) : w;  // earlier bands
  const int32_t nold = nd > 0 ? nd - 1 : 0;     // all but the neighbour
  bool late = false;
  double acc[K];
#pragma unroll
  for (int c = 0; c < K; ++c) acc[c] = 0.0;
  if (wv != 0) {
    // Tiles are loaded in pieces of kP columns (fewer with more right-hand
    // sides, for registers), two pieces in flight.
    constexpr int kP = K >= 4 ? 16 : 32;
    constexpr int kPieces = kNb / kP;
    T buf[2][kP];
    auto load_piece = [&](T(&v)[kP], int32_t q, int h) {
      const int32_t j = upper ? (np - 1 - q) : q;
      const int32_t cols = (m - j * kNb) < kNb ? (m - j * kNb) : kNb;
#pragma unroll
      for (int c = 0; c < kP; ++c) {
        const int cc = h * kP + c;
        v[c] = S[int64_t(j * kNb + (cc < cols ? cc : cols - 1)) * lds + rowc];
      }
    };
    int32_t q = wv - 1;
    if (q < nold) load_piece(buf[0], q, 0);
    while (q < nold) {
      const int32_t j = upper ? (np - 1 - q) : q;
      const int32_t cols = (m - j * kNb) < kNb ? (m - j * kNb) : kNb;
      const int32_t qn = q + (kW - 1);
      load_piece(buf[1], q, 1);
#pragma unroll
      for (int k = 0; k < K; ++k)
        xs[wv][k][lane] = lane < cols ? wait_x(yout + k * ldo + j * kNb + lane, late) : 0.0;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int h = 0; h < kPieces; ++h) {
#pragma unroll
        for (int c = 0; c < kP; ++c)
#pragma unroll
          for (int k = 0; k < K; ++k)
            acc[k] = __builtin_fma(static_cast<double>(buf[h & 1][c]), xs[wv][k][h * kP + c], acc[k]);
        // The buffer just consumed takes the piece after the one in flight
        // (unconditionally: after the last tile it reloads this one, which keeps
        // the loop free of branches around loads and their register copies).
        if (h + 2 < kPieces)
          load_piece(buf[h & 1], q, h + 2);
        else
          load_piece(buf[h & 1], qn < nold ? qn : q, h + 2 - kPieces);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      q = qn;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) part[wv][k][lane] = acc[k];
  }
  // Wave 0: its row of G_w (the neighbour coupling) into registers meanwhile.
  T grow[kNb];
  if (wv == 0 && nd > 0) {
    const T* gr = g + (int64_t(w) * 2 + kind) * kNb * kNb + (live ? lane : 0) * kNb;
#pragma unroll
    for (int c = 0; c < kNb; ++c) grow[c] = gr[c];
  }
  __syncthreads();
  if (wv != 0) {
    if (late) atomicOr(err, 2);
    return;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double sum = (part[1][k][lane] + part[2][k][lane]) + part[3][k][lane];
    xs[0][k][lane] = live ? yin[k * ldi + row] - sum : 0.0;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // D_w^{-1} (yin - sum): row `lane` of the inverse at dt[lane * kDL + c], the
  // right-hand side broadcast through LDS. Rows past the band take a zero
  // right-hand side (the inverse's pad is the identity).
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
#pragma unroll
  for (int c = 0; c < kNb; ++c) {
    const double a = static_cast<double>(dt[lane * kDL + c]);
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = __builtin_fma(a, xs[0][k][c], acc[k]);
  }
  if (nd > 0) {
    // - G_w x_v, once the neighbour's values are visible.
    const int32_t j = upper ? (w + 1) : (w - 1);
    const int32_t cols = (m - j * kNb) < kNb ? (m - j * kNb) : kNb;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int k = 0; k < K; ++k) xs[0][k][lane] = lane < cols ? wait_x(yout + k * ldo + j * kNb + lane, late) : 0.0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int c = 0; c < kNb; ++c) {
      const double gc = static_cast<double>(grow[c]);
#pragma unroll
      for (int k = 0; k < K; ++k) acc[k] = __builtin_fma(-gc, xs[0][k][c], acc[k]);
    }
  }
  if (live) {
#pragma unroll
    for (int k = 0; k < K; ++k) st_x_relaxed(yout + k * ldo + row, acc[k]);
  }
  if (late) atomicOr(err, 2);
}
template <typename T>
void launch_tail_trsv(hipStream_t st, int32_t kc, int32_t m, const T* S, int64_t lds, const T* inv, const T* g,
                      const double* yin, int64_t ldi, double* yout, int64_t ldo, bool upper, int32_t* ticket,
                      int32_t* err) {
  if (m <= 0 || kc <= 0) return;
  const int32_t np = (m + kNb - 1) / kNb;
  hipLaunchKernelGGL(k_tail_trsv<T>, dim3(np * kc), dim3(256), 0, st, m, kc, S, lds, inv, g, yin, ldi, yout, ldo,
                     upper ? 1 : 0, ticket, err);
}
template void launch_tail_trsv<double>(hipStream_t, int32_t, int32_t, const double*, int64_t, const double*,
                                       const double*, const double*, int64_t, double*, int64_t, bool, int32_t*,
                                       int32_t*);
template void launch_tail_trsv<float>(hipStream_t, int32_t, int32_t, const float*, int64_t, const float*,
                                      const float*, const double*, int64_t, double*, int64_t, bool, int32_t*,
                                      int32_t*);

__global__ void k_sp_scatter_x(int32_t s, int32_t start, int32_t kc, const int32_t* __restrict__ q,
                               const double* __restrict__ x, int64_t ldx, double* __restrict__ xperm, int64_t ldv) {
  for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < s; k += gridDim.x * blockDim.x)
    for (int32_t c = 0; c < kc; ++c) xperm[c * ldv + start + q[k]] = x[c * ldx + k];
}
void launch_sp_scatter_x(hipStream_t st, int32_t s, int32_t start, int32_t kc, const int32_t* q, const double* x,
                         int64_t ldx, double* xperm, int64_t ldv) {
  hipLaunchKernelGGL(k_sp_scatter_x, dim3(grid_for(s, 256, 4096)), dim3(256), 0, st, s, start, kc, q, x, ldx, xperm,
                     ldv);
}

// ---------------------------------------------------------------------------
// Residual r = b - A x over the canonical matrix, row by row in canonical
// order (the reference's refinement loop, src/solver.cpp:1497-1506).
// r = b - A x and, per row, the componentwise backward-error ratio
// |r_i| / (|A| |x| + |b|)_i (Oettli-Prager) into cw.
template <int K>
__global__ void k_residual(int32_t n, const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                           const double* __restrict__ vals, const double* __restrict__ x,
                           const double* __restrict__ b, double* __restrict__ r, double* __restrict__ cw,
                           int64_t ld, const uint8_t* __restrict__ is_long) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (is_long && is_long[i]) continue;  // k_residual_long takes it
    double acc[K], den[K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
      acc[c] = b[c * ld + i];
      den[c] = fabs(acc[c]);
    }
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) {
      const double v = vals[p];
      const int32_t j = col[p];
#pragma unroll
      for (int c = 0; c < K; ++c) {
        const double xj = x[c * ld + j];
        acc[c] = acc[c] - v * xj;
        den[c] += fabs(v * xj);
      }
    }
#pragma unroll
    for (int c = 0; c < K; ++c) {
      r[c * ld + i] = acc[c];
      const double a = fabs(acc[c]);
      cw[c * ld + i] = den[c] > 0.0 ? a / den[c] : (a == 0.0 ? 0.0 : INFINITY);
    }
  }
}
// Rows with many entries (dense-ish blocks), one wavefront each: lane l takes
// entries l, l+64, ... in order, then a fixed xor-butterfly sums the lanes, so
// the result does not depend on scheduling. Same outputs as k_residual.
template <int K>
__global__ __launch_bounds__(256) void k_residual_long(int32_t nlong, const int32_t* __restrict__ rows,
                                                       const int64_t* __restrict__ ptr,
                                                       const int32_t* __restrict__ col,
                                                       const double* __restrict__ vals,
                                                       const double* __restrict__ x, const double* __restrict__ b,
                                                       double* __restrict__ r, double* __restrict__ cw, int64_t ld) {
  const int lane = threadIdx.x & 63;
  const int32_t wpb = blockDim.x >> 6;
  for (int32_t t = blockIdx.x * wpb + (threadIdx.x >> 6); t < nlong; t += gridDim.x * wpb) {
    const int32_t i = rows[t];
    double s[K], d[K];
#pragma unroll
    for (int c = 0; c < K; ++c) s[c] = d[c] = 0.0;
    for (int64_t p = ptr[i] + lane; p < ptr[i + 1]; p += 64) {
      const double v = vals[p];
      const int32_t j = col[p];
#pragma unroll
      for (int c = 0; c < K; ++c) {
        const double t2 = v * x[c * ld + j];
        s[c] += t2;
        d[c] += fabs(t2);
      }
    }
#pragma unroll
    for (int c = 0; c < K; ++c)
      for (int off = 32; off > 0; off >>= 1) {
        s[c] += __shfl_xor(s[c], off);
        d[c] += __shfl_xor(d[c], off);
      }
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < K; ++c) {
        const double bi = b[c * ld + i];
        const double acc = bi - s[c];
        const double den = fabs(bi) + d[c];
        r[c * ld + i] = acc;
        const double a = fabs(acc);
        cw[c * ld + i] = den > 0.0 ? a / den : (a == 0.0 ? 0.0 : INFINITY);
      }
    }
  }
}
void launch_residual(hipStream_t st, int32_t n, int32_t kc, const int64_t* csr_ptr, const int32_t* csr_col,
                     const double* vals, const double* x, const double* b, double* r, double* cw, int64_t ld,
                     const uint8_t* is_long, int32_t nlong, const int32_t* long_rows) {
  const dim3 g(grid_for(n, 256, 8192)), bl(256);
#define BDLU_RES(KK)                                                                                             \
  do {                                                                                                           \
    hipLaunchKernelGGL(k_residual<KK>, g, bl, 0, st, n, csr_ptr, csr_col, vals, x, b, r, cw, ld,                \
                       nlong > 0 ? is_long : nullptr);                                                           \
    if (nlong > 0)                                                                                               \
      hipLaunchKernelGGL(k_residual_long<KK>, dim3(grid_for(nlong, 4, 4096)), dim3(256), 0, st, nlong,          \
                         long_rows, csr_ptr, csr_col, vals, x, b, r, cw, ld);                                    \
  } while (0)
  switch (kc) {
    case 1: BDLU_RES(1); break;
    case 2: BDLU_RES(2); break;
    case 4: BDLU_RES(4); break;
    default: BDLU_RES(8); break;
  }
#undef BDLU_RES
}

// Deterministic reductions: fixed grid, fixed per-block order, then a fixed
// lane tree. Per column c: out[c*9 + 0] = sum r^2, [1] max|r|, [2] max|x|,
// [3] sum b^2, [4] max|b|, [5] max componentwise backward-error ratio.
// 128 partial blocks: the norms are a small part of a refinement step, and a
// narrower launch leaves more of the GPU to other contexts' solves (serving:
// about +1% requests/s against 512).
constexpr int kRedBlocks = 128;
constexpr int kNormQ = 6;
__global__ __launch_bounds__(256) void k_norms_partial(int32_t n, int32_t kc, const double* __restrict__ r,
                                                       const double* __restrict__ x, const double* __restrict__ b,
                                                       const double* __restrict__ cw, int64_t ld,
                                                       double* __restrict__ part) {
  __shared__ double sh[kNormQ][256];
  for (int32_t c = 0; c < kc; ++c) {
    double v[kNormQ] = {0, 0, 0, 0, 0, 0};
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      const double ri = r[c * ld + i], xi = x[c * ld + i], bi = b[c * ld + i];
      v[0] += ri * ri;
      v[1] = fmax(v[1], fabs(ri));
      v[2] = fmax(v[2], fabs(xi));
      v[3] += bi * bi;
      v[4] = fmax(v[4], fabs(bi));
      v[5] = fmax(v[5], cw[c * ld + i]);
    }
    for (int q = 0; q < kNormQ; ++q) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (threadIdx.x < h) {
        sh[0][threadIdx.x] += sh[0][threadIdx.x + h];
        sh[3][threadIdx.x] += sh[3][threadIdx.x + h];
        for (int q : {1, 2, 4, 5}) sh[q][threadIdx.x] = fmax(sh[q][threadIdx.x], sh[q][threadIdx.x + h]);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0)
      for (int q = 0; q < kNormQ; ++q) part[(c * kRedBlocks + blockIdx.x) * kNormQ + q] = sh[q][0];
    __syncthreads();
  }
}
// One wavefront per column: lane l folds partials l, l+64, ... in order, then a
// fixed butterfly.
__global__ __launch_bounds__(64) void k_norms_final(int32_t nb, const double* __restrict__ part,
                                                    double* __restrict__ out) {
  const int lane = threadIdx.x, c = blockIdx.x;
  double v[kNormQ] = {0, 0, 0, 0, 0, 0};
  for (int32_t i = lane; i < nb; i += kWave) {
    const double* p = part + (c * nb + i) * kNormQ;
    v[0] += p[0];
    v[3] += p[3];
    for (int q : {1, 2, 4, 5}) v[q] = fmax(v[q], p[q]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    v[0] += __shfl_xor(v[0], off);
    v[3] += __shfl_xor(v[3], off);
    for (int q : {1, 2, 4, 5}) v[q] = fmax(v[q], __shfl_xor(v[q], off));
  }
  if (lane == 0)
    for (int q = 0; q < kNormQ; ++q) out[c * 9 + q] = v[q];
}
void launch_norms(hipStream_t st, int32_t n, int32_t kc, const double* r, const double* x, const double* b,
                  const double* cw, int64_t ld, double* partials, double* out) {
  hipLaunchKernelGGL(k_norms_partial, dim3(kRedBlocks), dim3(256), 0, st, n, kc, r, x, b, cw, ld, partials);
  hipLaunchKernelGGL(k_norms_final, dim3(kc), dim3(64), 0, st, kRedBlocks, partials, out);
}

__global__ __launch_bounds__(256) void k_anorm_partial(int32_t n, const int64_t* __restrict__ ptr,
                                                       const double* __restrict__ vals, double* __restrict__ part) {
  __shared__ double sh[256];
  double v = 0.0;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) s += fabs(vals[p]);
    v = fmax(v, s);
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
// One wave: strided maxima per lane, then a butterfly (max is exact and
// order-free, so the result does not depend on the reduction order).
__global__ __launch_bounds__(64) void k_max_final(int32_t nb, const double* __restrict__ part,
                                                  double* __restrict__ out) {
  const int lane = threadIdx.x;
  double v = 0.0;
  for (int32_t i = lane; i < nb; i += kWave) v = fmax(v, part[i]);
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  if (lane == 0) out[0] = v;
}
void launch_anorm(hipStream_t st, int32_t n, const int64_t* csr_ptr, const double* vals, double* partials,
                  double* out) {
  hipLaunchKernelGGL(k_anorm_partial, dim3(kRedBlocks), dim3(256), 0, st, n, csr_ptr, vals, partials);
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(64), 0, st, kRedBlocks, partials, out);
}

__global__ void k_axpy(int32_t n, int32_t kc, const double* __restrict__ d, double* __restrict__ x, int64_t ld) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int32_t c = 0; c < kc; ++c) x[c * ld + i] = x[c * ld + i] + d[c * ld + i];
}
void launch_axpy(hipStream_t st, int32_t n, int32_t kc, const double* d, double* x, int64_t ld) {
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, n, kc, d, x, ld);
}

// x += d while the refinement state says the loop is still running.
__global__ void k_axpy_gated(int32_t n, int32_t kc, const double* __restrict__ d, double* __restrict__ x, int64_t ld,
                             const double* __restrict__ state) {
  if (state[kIrActive] == 0.0) return;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int32_t c = 0; c < kc; ++c) x[c * ld + i] = x[c * ld + i] + d[c * ld + i];
}
void launch_axpy_gated(hipStream_t st, int32_t n, int32_t kc, const double* d, double* x, int64_t ld,
                       const double* state) {
  hipLaunchKernelGGL(k_axpy_gated, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, n, kc, d, x, ld, state);
}

// One stop decision of iterative refinement, taken on the device from the
// norms of the last residual (h[c*9 + 0..5] per column, as k_norms_final
// writes them) so that a whole refinement loop can run without a host round
// trip. The measures, the stop rule (0: ||r||_2/||b||_2 < tol; 1: normwise
// backward error ||r||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= tol; 2:
// componentwise backward error <= tol), columns with b = 0 counting as
// converged, and the stall rule (a step that fails to halve the worst measure
// ends the loop) are the host loop's, in the same double operations
// (correctly rounded sqrt and division, no contraction), so both take the
// same decisions. `first` evaluates the initial solution; later calls do
// nothing once the loop has ended.
__global__ void k_ir_check(const double* __restrict__ h, int32_t kc, int32_t rule, double tol,
                           const double* __restrict__ anorm, int32_t max_iters, int32_t watch_stall, int32_t first,
                           double* __restrict__ state) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  if (!first && state[kIrActive] == 0.0) return;
  const double an = *anorm;
  bool met = true;
  double worst = 0.0;
  for (int32_t k = 0; k < kc; ++k) {
    const double* q = h + k * 9;
    const double bn2 = sqrt(q[3]);
    const double rel = bn2 > 0 ? sqrt(q[0]) / bn2 : 0.0;
    const double den = an * q[2] + q[4];
    const double berr = den > 0 ? q[1] / den : 0.0;
    const double m = rule == 1 ? berr : rule == 2 ? q[5] : rel;
    const bool stop = rule == 0 ? rel < tol : m <= tol;
    if (q[3] > 0.0) {
      if (!stop) met = false;
      worst = worst < m ? m : worst;
    }
  }
  if (first) {
    const bool conv = max_iters <= 0 || met;
    state[kIrIters] = 0.0;
    state[kIrConverged] = conv ? 1.0 : 0.0;
    state[kIrWorst] = worst;
    state[kIrStalled] = 0.0;
    state[kIrActive] = conv ? 0.0 : 1.0;
    return;
  }
  const double iters = state[kIrIters] + 1.0;
  state[kIrIters] = iters;
  state[kIrConverged] = met ? 1.0 : 0.0;
  if (!met && watch_stall && !(worst <= 0.5 * state[kIrWorst])) {
    state[kIrStalled] = 1.0;
    state[kIrActive] = 0.0;
    return;
  }
  state[kIrWorst] = worst;
  state[kIrActive] = (!met && iters < double(max_iters)) ? 1.0 : 0.0;
}
void launch_ir_check(hipStream_t st, const double* norms, int32_t kc, int32_t rule, double tol, const double* anorm,
                     int32_t max_iters, bool watch_stall, bool first, double* state) {
  hipLaunchKernelGGL(k_ir_check, dim3(1), dim3(64), 0, st, norms, kc, rule, tol, anorm, max_iters,
                     watch_stall ? 1 : 0, first ? 1 : 0, state);
}

}  // namespace bdk
