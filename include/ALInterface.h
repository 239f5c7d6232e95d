#pragma once
// The competition's algorithm interface, which the reference implements with
// UserAlgorithmWrapper (blockdaglu/include/UserAlgorithmWrapper.h:16-35) and
// exports through CreateAlgorithm/DestroyAlgorithm (src/exports.cpp:14-24).
// The organizer's header is not part of the reference; this declaration is
// restated from those uses: the same COO type, the same stages in the same
// order (prepareData -> rearrangeMatrix -> decompositionLU -> solveLinearSystem,
// apps/benchmark/main.cpp:165-183) and the same exports. Its destructor is not
// virtual, as the reference notes of the organizer's, so an instance is
// destroyed through DestroyAlgorithm only.
#include <tuple>
#include <vector>

using COOMatrix = std::vector<std::tuple<int, int, double>>;

class ALInterface {
 public:
  virtual bool prepareData(COOMatrix m) = 0;            // storage only
  virtual bool rearrangeMatrix() = 0;                   // ordering / analysis
  virtual bool decompositionLU() = 0;                   // numeric factorization
  virtual bool solveLinearSystem(std::vector<double>& b) = 0;  // b <- x
  virtual COOMatrix getRearrangedMatrix() = 0;
  virtual COOMatrix getQ() = 0;
  virtual COOMatrix getP() = 0;
  virtual COOMatrix getL() = 0;
  virtual COOMatrix getU() = 0;

 protected:
  ~ALInterface() = default;
};
