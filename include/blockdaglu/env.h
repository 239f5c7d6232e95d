#pragma once
// Environment lookup with the reference's naming rule
// (blockdaglu/include/blockdaglu/env.h:15-25): a request for USERALGO_FOO first
// tries BLOCKDAGLU_FOO, then USERALGO_FOO. Every other name is read as given.

#include <cstdlib>
#include <cstring>
#include <string>

inline const char* blockdaglu_getenv_compat(const char* name) {
  if (name == nullptr) return nullptr;
  static const char kLegacy[] = "USERALGO_";
  const std::size_t legacy_len = sizeof(kLegacy) - 1;
  if (std::strncmp(name, kLegacy, legacy_len) == 0) {
    const std::string modern = std::string("BLOCKDAGLU_") + (name + legacy_len);
    if (const char* v = std::getenv(modern.c_str())) return v;
  }
  return std::getenv(name);
}
