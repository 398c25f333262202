// Instantiations of the packed 16-bit banded DP kernel (global alignment,
// gap 2..16), inside the sequence and with 'X' past its end.
#include <algorithm>

#include "kernels.hpp"

namespace clast {

// grid: one wavefront per chunk of 64 task pairs of all a.n tasks (the
// kernel's range is a part of them); persistent: at most grid_cap
// workgroups (about one occupancy round), which loop over the chunks.
template <int G>
static void launch(bool edge, bool persist, uint32_t grid_cap, const DPArgs& a, hipStream_t st) {
  uint32_t grid = grid_for((static_cast<uint64_t>(a.n) + 1) / 2, kBlock);
  if (edge) {
    k_dp16<G, true, false><<<grid, kBlock, 0, st>>>(a);
  } else if (persist) {
    grid = std::max<uint32_t>(1, std::min(grid, grid_cap));
    k_dp16<G, false, true><<<grid, kBlock, 0, st>>>(a);
  } else {
    k_dp16<G, false, false><<<grid, kBlock, 0, st>>>(a);
  }
}

void launch_dp16(int gap, bool edge, bool persist, uint32_t grid_cap, const DPArgs& a, hipStream_t st) {
  switch (gap) {
    case 2: launch<2>(edge, persist, grid_cap, a, st); break;
    case 3: launch<3>(edge, persist, grid_cap, a, st); break;
    case 4: launch<4>(edge, persist, grid_cap, a, st); break;
    case 5: launch<5>(edge, persist, grid_cap, a, st); break;
    case 6: launch<6>(edge, persist, grid_cap, a, st); break;
    case 7: launch<7>(edge, persist, grid_cap, a, st); break;
    case 8: launch<8>(edge, persist, grid_cap, a, st); break;
    case 9: launch<9>(edge, persist, grid_cap, a, st); break;
    case 10: launch<10>(edge, persist, grid_cap, a, st); break;
    case 11: launch<11>(edge, persist, grid_cap, a, st); break;
    case 12: launch<12>(edge, persist, grid_cap, a, st); break;
    case 13: launch<13>(edge, persist, grid_cap, a, st); break;
    case 14: launch<14>(edge, persist, grid_cap, a, st); break;
    case 15: launch<15>(edge, persist, grid_cap, a, st); break;
    case 16: launch<16>(edge, persist, grid_cap, a, st); break;
    default: throw ClastError(1, "gap out of range");
  }
}

}  // namespace clast
