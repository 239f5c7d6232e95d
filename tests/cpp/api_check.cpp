// Drop-in check of the unchanged C++ API: a program written only against
// include/blockdaglu/solver.h and linked to libblockdaglu.so, as a caller of the
// reference library would be. It runs the reference's own known answers
// (blockdaglu/apps/tests/main.cpp): the validation error codes (:93-109), the
// 3x3 system with duplicate entries and its solution (:111-128), the n = 1100
// cache matrices with solution x = 1, for two patterns and two value sets
// (:130-168), the 4x4 export identity P A Q = L U (:170-195) and two solvers
// used from two threads (:197-204).
// Usage: bdlu_api_check [--host-only]   (host-only: checks that need no GPU)
// Prints "api_check: all N checks passed" and exits 0, or lists failures and exits 1.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "blockdaglu/solver.h"

using blockdaglu::CooMatrix;
using blockdaglu::Solver;
using blockdaglu::SolverError;

namespace {

std::atomic<int> g_checks{0}, g_failed{0};

void expect(bool ok, const std::string& what) {
  ++g_checks;
  if (!ok) {
    ++g_failed;
    std::fprintf(stderr, "FAIL: %s\n", what.c_str());
  }
}

bool near(double a, double b, double tol = 1e-9) {
  const double scale = std::max(1.0, std::max(std::fabs(a), std::fabs(b)));
  return std::fabs(a - b) <= tol * scale;
}

std::vector<double> apply(const CooMatrix& a, int n, const std::vector<double>& x) {
  std::vector<double> y(n, 0.0);
  for (const auto& [i, j, v] : a) y[i] += v * x[j];
  return y;
}

std::vector<double> to_dense(const CooMatrix& a, int n) {
  std::vector<double> d(static_cast<std::size_t>(n) * n, 0.0);
  for (const auto& [i, j, v] : a) d[static_cast<std::size_t>(i) * n + j] += v;
  return d;
}

void error_codes() {
  Solver neg;
  expect(neg.set_matrix({{-1, 0, 1.0}}), "negative index: set_matrix stores");
  expect(!neg.analyze() && neg.last_error() == SolverError::InvalidIndex, "negative index -> InvalidIndex");
  Solver inf;
  expect(inf.set_matrix({{0, 0, std::numeric_limits<double>::infinity()}}), "infinity: set_matrix stores");
  expect(!inf.analyze() && inf.last_error() == SolverError::NonFiniteValue, "infinity -> NonFiniteValue");
  Solver dim;
  expect(dim.set_matrix({{2, 2, 1.0}}, 2), "explicit dimension: set_matrix stores");
  expect(!dim.analyze() && dim.last_error() == SolverError::InvalidIndex, "entry outside dimension -> InvalidIndex");
  Solver fresh;
  expect(!fresh.analyze() && fresh.last_error() == SolverError::NotPrepared, "analyze before set_matrix");
  expect(!fresh.factor() && fresh.last_error() == SolverError::NotPrepared, "factor before analyze");
}

void export_identity() {
  Solver s;
  const CooMatrix a{{0, 0, 100.0}, {0, 1, 20.0}, {0, 2, 7.0}, {1, 0, 30.0}, {1, 1, 80.0}, {1, 3, 5.0},
                    {2, 2, 0.25},  {2, 3, 0.05}, {3, 2, 0.10}, {3, 3, 0.50}};
  expect(s.set_matrix(a, 4), "export: set_matrix");
  const CooMatrix pa = s.get_rearranged_matrix(), l = s.getL(), u = s.getU();
  expect(!pa.empty() && !l.empty() && !u.empty(), "export: non-empty rearranged, L and U");
  const std::vector<double> A = to_dense(pa, 4), L = to_dense(l, 4), U = to_dense(u, 4);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double lu = 0.0;
      for (int k = 0; k < 4; ++k) lu += L[i * 4 + k] * U[k * 4 + j];
      expect(near(A[i * 4 + j], lu), "export: P A Q == L U at (" + std::to_string(i) + "," + std::to_string(j) + ")");
    }
  const CooMatrix p = s.getP(), q = s.getQ();
  expect(p.size() == 4 && q.size() == 4, "export: P and Q have one entry per row");
}

void three_by_three() {
  Solver s;
  CooMatrix a{{0, 0, 4.0}, {0, 1, 1.0}, {1, 0, 1.0}, {1, 1, 3.0}, {1, 2, 1.0}, {2, 1, 1.0}, {2, 2, 2.0},
              {0, 0, 1.0}, {0, 0, -1.0}};
  expect(s.set_matrix(std::move(a), 3), "3x3: set_matrix");
  expect(s.analyze(), "3x3: analyze: " + s.last_error_message());
  expect(s.factor(), "3x3: factor: " + s.last_error_message());
  std::vector<double> b{1.0, 2.0, 3.0};
  expect(s.solve_in_place(b), "3x3: solve: " + s.last_error_message());
  const double x[3] = {0.2222222222222222, 0.1111111111111111, 1.4444444444444444};
  for (int i = 0; i < 3; ++i) expect(near(b[i], x[i]), "3x3: x[" + std::to_string(i) + "]");
  std::vector<double> short_b{1.0, 2.0};
  expect(!s.solve_in_place(short_b) && s.last_error() == SolverError::InvalidRightHandSide,
         "3x3: wrong right-hand-side size -> InvalidRightHandSide");
  expect(std::string(s.selected_mode()) == "BLOCK" || std::string(s.selected_mode()) == "KLU", "3x3: mode name");
}

// Diagonal 4 plus one 0.25 entry per row (at row n/2 in column `changed`), the
// last two rows with a duplicate diagonal 0.5 instead.
CooMatrix cache_matrix(int n, int changed, double scale = 1.0) {
  CooMatrix a;
  for (int i = 0; i < n; ++i) {
    a.emplace_back(i, i, 4.0 * scale);
    if (i + 2 < n) a.emplace_back(i, i == n / 2 ? changed : i + 1, 0.25 * scale);
    else a.emplace_back(i, i, 0.5 * scale);
  }
  return a;
}

void solve_to_ones(CooMatrix a, int n, const std::string& label) {
  std::vector<double> b = apply(a, n, std::vector<double>(n, 1.0));
  Solver s;
  expect(s.set_matrix(std::move(a), n), label + ": set_matrix");
  expect(s.analyze(), label + ": analyze: " + s.last_error_message());
  expect(s.factor(), label + ": factor: " + s.last_error_message());
  expect(s.solve_in_place(b), label + ": solve: " + s.last_error_message());
  bool ones = b.size() == static_cast<std::size_t>(n);
  for (double v : b) ones = ones && near(v, 1.0);
  expect(ones, label + ": x == 1");
}

}  // namespace

int main(int argc, char** argv) {
  const bool host_only = argc > 1 && std::strcmp(argv[1], "--host-only") == 0;
  error_codes();
  export_identity();
  if (!host_only) {
    three_by_three();
    const int n = 1100;
    solve_to_ones(cache_matrix(n, n / 2 + 1), n, "cache pattern A");
    solve_to_ones(cache_matrix(n, n / 2 + 2), n, "cache pattern B");
    solve_to_ones(cache_matrix(n, n / 2 + 1), n, "cache values A");
    solve_to_ones(cache_matrix(n, n / 2 + 1, 2.0), n, "cache values B");
    std::thread t1([] { solve_to_ones(cache_matrix(256, 129), 256, "thread A"); });
    std::thread t2([] { solve_to_ones(cache_matrix(256, 129), 256, "thread B"); });
    t1.join();
    t2.join();
  }
  if (g_failed) {
    std::fprintf(stderr, "api_check: %d of %d checks failed\n", g_failed.load(), g_checks.load());
    return 1;
  }
  std::printf("api_check: all %d checks passed\n", g_checks.load());
  return 0;
}
