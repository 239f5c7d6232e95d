// The competition plugin: UserAlgorithmWrapper maps each ALInterface stage onto
// one blockdaglu::Solver call (reference src/UserAlgorithmWrapper.cpp:8-42),
// and the factory functions are exported with C linkage (src/exports.cpp:14-24).
#include "UserAlgorithmWrapper.h"

UserAlgorithmWrapper::UserAlgorithmWrapper() : solver_(std::make_unique<blockdaglu::Solver>()) {}
UserAlgorithmWrapper::~UserAlgorithmWrapper() = default;

bool UserAlgorithmWrapper::prepareData(COOMatrix m) { return solver_->set_matrix(std::move(m)); }
bool UserAlgorithmWrapper::rearrangeMatrix() { return solver_->analyze(); }
bool UserAlgorithmWrapper::decompositionLU() { return solver_->factor(); }
bool UserAlgorithmWrapper::solveLinearSystem(std::vector<double>& b) { return solver_->solve_in_place(b); }
COOMatrix UserAlgorithmWrapper::getRearrangedMatrix() { return solver_->get_rearranged_matrix(); }
COOMatrix UserAlgorithmWrapper::getQ() { return solver_->getQ(); }
COOMatrix UserAlgorithmWrapper::getP() { return solver_->getP(); }
COOMatrix UserAlgorithmWrapper::getL() { return solver_->getL(); }
COOMatrix UserAlgorithmWrapper::getU() { return solver_->getU(); }
const char* UserAlgorithmWrapper::selectedMode() const { return solver_->selected_mode(); }

extern "C" {

__attribute__((visibility("default"))) ALInterface* CreateAlgorithm() {
  try {
    return new UserAlgorithmWrapper();
  } catch (...) {
    return nullptr;
  }
}

// The interface's destructor is not virtual: delete through the concrete type.
__attribute__((visibility("default"))) void DestroyAlgorithm(ALInterface* algorithm) {
  delete static_cast<UserAlgorithmWrapper*>(algorithm);
}

__attribute__((visibility("default"))) const char* AlgorithmSelectedMode(ALInterface* algorithm) {
  return algorithm ? static_cast<UserAlgorithmWrapper*>(algorithm)->selectedMode() : "";
}

}  // extern "C"
