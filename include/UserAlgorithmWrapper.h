#pragma once
// The competition adapter over blockdaglu::Solver (reference
// blockdaglu/include/UserAlgorithmWrapper.h:16-35, src/UserAlgorithmWrapper.cpp:8-42):
// each ALInterface stage is one Solver call, so the harness's stage timings are
// the Solver's stages. Built into libblockdaglu.so, which exports
//   extern "C" ALInterface* CreateAlgorithm();
//   extern "C" void DestroyAlgorithm(ALInterface*);
// for harnesses that load the library at run time (dlopen + dlsym, as the
// reference's dll_benchmark does with LoadLibraryA + GetProcAddress).
#include <memory>
#include <vector>

#include "ALInterface.h"
#include "blockdaglu/solver.h"

class UserAlgorithmWrapper final : public ALInterface {
 public:
  UserAlgorithmWrapper();
  ~UserAlgorithmWrapper();

  bool prepareData(COOMatrix m) override;
  bool rearrangeMatrix() override;
  bool decompositionLU() override;
  bool solveLinearSystem(std::vector<double>& b) override;

  COOMatrix getRearrangedMatrix() override;
  COOMatrix getQ() override;
  COOMatrix getP() override;
  COOMatrix getL() override;
  COOMatrix getU() override;

  const char* selectedMode() const;
  // The Solver behind the stages (for its last error and message).
  blockdaglu::Solver& solver() { return *solver_; }

 private:
  std::unique_ptr<blockdaglu::Solver> solver_;
};

extern "C" {
ALInterface* CreateAlgorithm();
void DestroyAlgorithm(ALInterface* algorithm);
// Mode the last rearrangeMatrix() of this instance selected ("BLOCK" | "KLU"),
// for harnesses that cannot see the C++ class (no counterpart in the reference).
const char* AlgorithmSelectedMode(ALInterface* algorithm);
}
