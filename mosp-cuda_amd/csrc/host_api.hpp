// host_api.hpp -- host helpers shared by the engine and the C++ API.
#pragma once

#include <stdint.h>

namespace mosp {

// L = lcm(Pref) (src/combinedGraphGpu.cu:129-147): 1 for an empty Pref,
// 0 if nPref != K, a value < 1, or L > 2^20.
int64_t preference_scale(const int32_t *pref, int32_t nPref, int32_t K);

// L * (K + 1) - sum over the trees in mask of L / Pref_i
// (src/combinedGraphGpu.cu:149-159).
int64_t combined_edge_weight(uint32_t mask, const int32_t *pref, int32_t nPref,
                             int32_t K, int64_t scale);

}  // namespace mosp
