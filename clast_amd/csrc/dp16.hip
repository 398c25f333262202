// Instantiations of the packed 16-bit banded DP kernel (global alignment,
// gap 2..16), inside the sequence and with 'X' past its end: scores, end
// nodes and match counts.
#include "dp16_launch.hpp"

namespace clast {

void launch_dp16(int gap, bool edge, bool persist, uint32_t grid_cap, bool, const DPArgs& a, hipStream_t st) {
  launch_dp16_any<false>(gap, edge, persist, grid_cap, false, a, st);
}

}  // namespace clast
