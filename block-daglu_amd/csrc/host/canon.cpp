// Validation, canonicalization and signatures of the input COO.
//   canonicalize  — reference src/solver.cpp:642-722
//   signatures    — reference src/solver.cpp:90-165
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "blockdaglu/env.h"
#include "host.h"

namespace bdlu {

CanonStatus canonicalize(std::vector<std::tuple<int, int, double>>& coo, int32_t dimension_hint, Canonical& out,
                         std::string& message) {
  // Checks in input order, same precedence as the reference loop (:650-670):
  // negative index, then non-finite value, then outside the declared dimension.
  if (dimension_hint < -1) {
    message = "matrix dimension must be non-negative";
    return CanonStatus::InvalidDimension;
  }
  int32_t max_index = -1;
  for (const auto& e : coo) {
    const int r = std::get<0>(e);
    const int c = std::get<1>(e);
    const double v = std::get<2>(e);
    if (r < 0 || c < 0) {
      message = "matrix contains a negative row or column index";
      return CanonStatus::InvalidIndex;
    }
    if (!std::isfinite(v)) {
      message = "matrix contains NaN or infinity";
      return CanonStatus::NonFinite;
    }
    if (dimension_hint >= 0 && (r >= dimension_hint || c >= dimension_hint)) {
      message = "matrix entry lies outside the declared dimension";
      return CanonStatus::InvalidIndex;
    }
    max_index = std::max<int32_t>(max_index, std::max<int32_t>(r, c));
  }
  int32_t n = 0;
  if (dimension_hint >= 0) {
    n = dimension_hint;
  } else {
    if (max_index == std::numeric_limits<int32_t>::max()) {
      message = "inferred matrix dimension overflows int32";
      return CanonStatus::InvalidDimension;
    }
    n = max_index + 1;
  }
  if (n <= 0) {
    message = "matrix dimension must be positive";
    return CanonStatus::InvalidDimension;
  }

  // Order by (row, col): a stable counting pass by row, then a stable insertion
  // sort by column inside each row (rows are short in the sparse inputs this
  // path serves, and stability fixes the summation order of duplicates to the
  // input order; the reference's std::sort leaves that order unspecified, which
  // only matters when a long double duplicate sum is inexact).
  const std::size_t nnz = coo.size();
  std::vector<int64_t> rptr(static_cast<std::size_t>(n) + 1, 0);
  for (const auto& e : coo) rptr[static_cast<std::size_t>(std::get<0>(e)) + 1]++;
  for (int32_t i = 0; i < n; ++i) rptr[i + 1] += rptr[i];
  std::vector<int32_t> scol(nnz);
  std::vector<double> sval(nnz);
  {
    std::vector<int64_t> next(rptr.begin(), rptr.end() - 1);
    for (const auto& e : coo) {
      const int64_t p = next[static_cast<std::size_t>(std::get<0>(e))]++;
      scol[p] = std::get<1>(e);
      sval[p] = std::get<2>(e);
    }
  }
  for (int32_t i = 0; i < n; ++i) {
    const int64_t b = rptr[i], e = rptr[i + 1];
    if (e - b <= 32) {
      for (int64_t p = b + 1; p < e; ++p) {
        const int32_t c = scol[p];
        const double v = sval[p];
        int64_t q = p;
        while (q > b && scol[q - 1] > c) {
          scol[q] = scol[q - 1];
          sval[q] = sval[q - 1];
          --q;
        }
        scol[q] = c;
        sval[q] = v;
      }
    } else {
      std::vector<int64_t> idx(static_cast<std::size_t>(e - b));
      for (int64_t p = b; p < e; ++p) idx[p - b] = p;
      std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) { return scol[x] < scol[y]; });
      std::vector<int32_t> tc(idx.size());
      std::vector<double> tv(idx.size());
      for (std::size_t t = 0; t < idx.size(); ++t) {
        tc[t] = scol[idx[t]];
        tv[t] = sval[idx[t]];
      }
      std::copy(tc.begin(), tc.end(), scol.begin() + b);
      std::copy(tv.begin(), tv.end(), sval.begin() + b);
    }
  }

  out.n = n;
  out.row.clear();
  out.col.clear();
  out.val.clear();
  out.row.reserve(nnz);
  out.col.reserve(nnz);
  out.val.reserve(nnz);
  for (int32_t i = 0; i < n; ++i) {
    int64_t p = rptr[i];
    const int64_t e = rptr[i + 1];
    while (p < e) {
      const int32_t c = scol[p];
      long double sum = 0.0L;
      int64_t q = p;
      while (q < e && scol[q] == c) {
        sum += static_cast<long double>(sval[q]);
        ++q;
      }
      const double v = static_cast<double>(sum);
      if (!std::isfinite(v)) {
        message = "combining duplicate matrix entries overflowed";
        return CanonStatus::DuplicateOverflow;
      }
      if (v != 0.0) {
        out.row.push_back(i);
        out.col.push_back(c);
        out.val.push_back(v);
      }
      p = q;
    }
  }
  coo.resize(out.nnz());
  for (std::size_t k = 0; k < out.nnz(); ++k) coo[k] = std::make_tuple(out.row[k], out.col[k], out.val[k]);
  message.clear();
  return CanonStatus::Ok;
}

namespace {
inline uint64_t fnv_step(uint64_t h, uint64_t x) {
  h ^= x;
  return h * 1099511628211ull;
}
}  // namespace

uint64_t matrix_sig(int32_t n, const Canonical& a) {
  uint64_t h = 1469598103934665603ull;
  h = fnv_step(h, static_cast<uint64_t>(n));
  h = fnv_step(h, static_cast<uint64_t>(a.nnz()));
  for (std::size_t k = 0; k < a.nnz(); ++k) {
    const uint64_t r = static_cast<uint32_t>(a.row[k]);
    const uint64_t c = static_cast<uint32_t>(a.col[k]);
    h = fnv_step(h, (r << 32) ^ c);
  }
  return h;
}

uint64_t matrix_sig2(int32_t n, const Canonical& a) {
  uint64_t h = 0x6a09e667f3bcc909ull ^ static_cast<uint64_t>(n);
  for (std::size_t k = 0; k < a.nnz(); ++k) {
    const uint64_t r = static_cast<uint32_t>(a.row[k]);
    const uint64_t c = static_cast<uint32_t>(a.col[k]);
    uint64_t z = ((r << 32) ^ c) ^ 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    h ^= z + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
  }
  return h ^ static_cast<uint64_t>(a.nnz());
}

uint64_t matrix_value_sig(const Canonical& a) {
  uint64_t h = 0xbb67ae8584caa73bull;
  for (double v : a.val) {
    uint64_t bits;
    std::memcpy(&bits, &v, sizeof(bits));
    h = fnv_step(h, bits);
  }
  return h ^ static_cast<uint64_t>(a.nnz());
}

uint64_t cache_config_sig(std::size_t build_threads) {
  static const char* const kKeys[] = {
      "USERALGO_BTF_IMPL",     "USERALGO_BTF_KLU_ORDERING",     "USERALGO_BTF_MAXWORK",
      "USERALGO_DROP_ABS_LT",  "USERALGO_BLOCK_MERGE",          "USERALGO_BLOCK_MERGE_MAX_SIZE",
      "USERALGO_DENSE_BLOCK_MAX", "USERALGO_CSR_MODE",          "USERALGO_BLOCK_SCHED",
  };
  uint64_t h = fnv_step(1469598103934665603ull, static_cast<uint64_t>(build_threads));
  for (const char* key : kKeys) {
    for (const char* p = key; *p; ++p) h = fnv_step(h, static_cast<unsigned char>(*p));
    const char* v = blockdaglu_getenv_compat(key);
    if (v == nullptr) {
      h = fnv_step(h, 0xff);
      continue;
    }
    for (const char* p = v; *p; ++p) h = fnv_step(h, static_cast<unsigned char>(*p));
  }
  return h;
}

}  // namespace bdlu
