// Instantiations of the packed 16-bit banded DP kernel's score pass
// (global alignment, gap 2..16, inside the sequence and with 'X' past its
// end): scores and end nodes, match counts written as 0.
#include "dp16_launch.hpp"

namespace clast {

void launch_dp16_score(int gap, bool edge, bool persist, uint32_t grid_cap, bool reserve, const DPArgs& a,
                       hipStream_t st) {
  launch_dp16_any<true>(gap, edge, persist, grid_cap, reserve, a, st);
}

}  // namespace clast
