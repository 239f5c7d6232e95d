#pragma once
// Environment lookup with the reference's naming rule
// (blockdaglu/include/blockdaglu/env.h:15-25). A request whose name starts with
// the 8 characters "USERALGO" first tries "BLOCKDAGLU_" followed by the rest of
// the name after those 8 characters, then the name itself. The rest keeps the
// underscore that follows "USERALGO", so USERALGO_FOO is looked up as
// BLOCKDAGLU__FOO (two underscores) and then as USERALGO_FOO; BLOCKDAGLU_FOO
// is never read for it. Every other name is read as given.

#include <cstdlib>
#include <cstring>
#include <string>

inline const char* blockdaglu_getenv_compat(const char* name) {
  if (name == nullptr) return nullptr;
  constexpr std::size_t kPrefixLen = 8;  // "USERALGO", as the reference compares it
  if (std::strncmp(name, "USERALGO", kPrefixLen) == 0) {
    const std::string alias = std::string("BLOCKDAGLU_") + (name + kPrefixLen);
    if (const char* v = std::getenv(alias.c_str())) return v;
  }
  return std::getenv(name);
}
