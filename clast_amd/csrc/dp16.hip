// Instantiations of the packed 16-bit banded DP kernel (global alignment,
// gap 2..16), inside the sequence and with 'X' past its end.
#include "kernels.hpp"

namespace clast {

template <int G>
static void launch(bool edge, const DPArgs& a, hipStream_t st) {
  const uint32_t grid = grid_for((static_cast<uint64_t>(a.n) + 1) / 2, kBlock);
  if (edge) k_dp16<G, true><<<grid, kBlock, 0, st>>>(a);
  else k_dp16<G, false><<<grid, kBlock, 0, st>>>(a);
}

void launch_dp16(int gap, bool edge, const DPArgs& a, hipStream_t st) {
  switch (gap) {
    case 2: launch<2>(edge, a, st); break;
    case 3: launch<3>(edge, a, st); break;
    case 4: launch<4>(edge, a, st); break;
    case 5: launch<5>(edge, a, st); break;
    case 6: launch<6>(edge, a, st); break;
    case 7: launch<7>(edge, a, st); break;
    case 8: launch<8>(edge, a, st); break;
    case 9: launch<9>(edge, a, st); break;
    case 10: launch<10>(edge, a, st); break;
    case 11: launch<11>(edge, a, st); break;
    case 12: launch<12>(edge, a, st); break;
    case 13: launch<13>(edge, a, st); break;
    case 14: launch<14>(edge, a, st); break;
    case 15: launch<15>(edge, a, st); break;
    case 16: launch<16>(edge, a, st); break;
    default: throw ClastError(1, "gap out of range");
  }
}

}  // namespace clast
