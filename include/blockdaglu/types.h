#pragma once
// Public COO type of the blockdaglu API.
// Mirrors blockdaglu/include/blockdaglu/types.h:10 (same element type so callers
// of the reference move their matrices in unchanged).

#include <tuple>
#include <vector>

namespace blockdaglu {

using CooMatrix = std::vector<std::tuple<int, int, double>>;

}  // namespace blockdaglu
