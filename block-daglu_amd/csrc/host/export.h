#pragma once
// Factor export of the reference Solver (getP/getQ/getL/getU/
// get_rearranged_matrix, src/solver.cpp:1541-1725): one global sparse LU of the
// whole matrix with KLU's settings for that export (btf = 0, COLAMD ordering,
// threshold partial pivoting, the reference's scale and tol pickers,
// src/solver.cpp:392-556), so that P (R \ A) Q = L U with no off-diagonal part.
// Host only; not on the refactor/solve path.
#include <cstdint>
#include <string>
#include <tuple>
#include <vector>

#include "host.h"

namespace bdlu {

struct ExportFactors {
  int32_t n = 0;
  std::vector<int32_t> P;   // P[k]: original row at pivot position k
  std::vector<int32_t> Q;   // Q[k]: original column at position k
  std::vector<double> Rs;   // row scale factors by original row (1 when unscaled)
  std::vector<std::tuple<int32_t, int32_t, double>> L;  // unit lower, positions
  std::vector<std::tuple<int32_t, int32_t, double>> U;  // upper with diagonal, positions
};

// Returns false (message set) if the matrix has a column with no admissible pivot.
bool export_factors(const Canonical& a, ExportFactors& out, std::string& message);

}  // namespace bdlu
