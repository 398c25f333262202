// Launcher of the packed 16-bit banded DP kernel for a runtime gap
// (2..16), shared by dp16.hip (scores and match counts) and dp16_score.hip
// (scores only), which instantiate it apart so they compile in parallel.
#pragma once
#include <algorithm>

#include "kernels.hpp"

namespace clast {

// LDS a score pass's workgroups reserve when asked (they use none): with
// 160 KiB per CU at most three fit, so three of a SIMD's four register-file
// quarters go to that score pass and the fourth stays open for the kernels
// beside it and for other searches' small launches, which otherwise wait
// for its long-lived wavefronts to finish.
constexpr size_t kScorePassLds = 41 * 1024;
static_assert(kScorePassLds * 4 > 160 * 1024 && kScorePassLds * 3 <= 160 * 1024, "three workgroups per CU");

// grid: one wavefront per chunk of 64 task pairs of all a.n tasks (the
// kernel's range is a part of them); persistent: at most grid_cap
// workgroups (about one occupancy round), which loop over the chunks.
template <int G, bool SCORE>
static void launch_dp16_gap(bool edge, bool persist, uint32_t grid_cap, bool reserve, const DPArgs& a,
                            hipStream_t st) {
  uint32_t grid = grid_for((static_cast<uint64_t>(a.n) + 1) / 2, kBlock);
  const size_t lds = SCORE && reserve ? kScorePassLds : 0;
  if (edge) {
    k_dp16<G, true, false, SCORE><<<grid, kBlock, 0, st>>>(a);
  } else if (persist) {
    if constexpr (G <= kPersistMaxGap) {
      grid = std::max<uint32_t>(1, std::min(grid, grid_cap));
      k_dp16<G, false, true, SCORE><<<grid, kBlock, lds, st>>>(a);
    } else {
      throw ClastError(1, "persistent 16-bit DP schedule over its gap limit");
    }
  } else {
    k_dp16<G, false, false, SCORE><<<grid, kBlock, lds, st>>>(a);
  }
}

template <bool SCORE>
static void launch_dp16_any(int gap, bool edge, bool persist, uint32_t grid_cap, bool reserve, const DPArgs& a,
                            hipStream_t st) {
  switch (gap) {
    case 2: launch_dp16_gap<2, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    case 3: launch_dp16_gap<3, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    case 4: launch_dp16_gap<4, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    case 5: launch_dp16_gap<5, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    case 6: launch_dp16_gap<6, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    case 7: launch_dp16_gap<7, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    case 8: launch_dp16_gap<8, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    case 9: launch_dp16_gap<9, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    case 10: launch_dp16_gap<10, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    case 11: launch_dp16_gap<11, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    case 12: launch_dp16_gap<12, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    case 13: launch_dp16_gap<13, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    case 14: launch_dp16_gap<14, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    case 15: launch_dp16_gap<15, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    case 16: launch_dp16_gap<16, SCORE>(edge, persist, grid_cap, reserve, a, st); break;
    default: throw ClastError(1, "gap out of range");
  }
}

}  // namespace clast
