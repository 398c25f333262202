// Instantiations of the packed 16-bit banded DP kernel (global alignment,
// gap 2..16).
#include "kernels.hpp"

namespace clast {

template <int G>
static void launch(const DPArgs& a, hipStream_t st) {
  k_dp16<G><<<grid_for((static_cast<uint64_t>(a.n) + 1) / 2, kBlock), kBlock, 0, st>>>(a);
}

void launch_dp16(int gap, const DPArgs& a, hipStream_t st) {
  switch (gap) {
    case 2: launch<2>(a, st); break;
    case 3: launch<3>(a, st); break;
    case 4: launch<4>(a, st); break;
    case 5: launch<5>(a, st); break;
    case 6: launch<6>(a, st); break;
    case 7: launch<7>(a, st); break;
    case 8: launch<8>(a, st); break;
    case 9: launch<9>(a, st); break;
    case 10: launch<10>(a, st); break;
    case 11: launch<11>(a, st); break;
    case 12: launch<12>(a, st); break;
    case 13: launch<13>(a, st); break;
    case 14: launch<14>(a, st); break;
    case 15: launch<15>(a, st); break;
    case 16: launch<16>(a, st); break;
    default: throw ClastError(1, "gap out of range");
  }
}

}  // namespace clast
