// Instantiations of the lane-per-band DP kernel for long extensions (both
// modes, gap 2..16).
#include <algorithm>

#include "kernels.hpp"

namespace clast {

template <int G, bool LOCAL>
static void launch(const DPArgs& a, hipStream_t st) {
  // a.n bounds the tasks (the band task count is on the device):
  // workgroups past the last pair exit at once; at most kMaxBlocks
  // workgroups, which stride over the pairs
  constexpr uint32_t kPairsPerBlock = (kBlock / 64) * (64 / (2 * G + 1));
  constexpr uint64_t kMaxBlocks = 4096;
  const uint64_t pairs = (static_cast<uint64_t>(a.n) + 1) / 2;
  const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((pairs + kPairsPerBlock - 1) / kPairsPerBlock, kMaxBlocks));
  if (blocks) k_dp_band<G, LOCAL><<<blocks, kBlock, 0, st>>>(a);
}

template <bool LOCAL>
static void launch_gap(int gap, const DPArgs& a, hipStream_t st) {
  switch (gap) {
    case 2: launch<2, LOCAL>(a, st); break;
    case 3: launch<3, LOCAL>(a, st); break;
    case 4: launch<4, LOCAL>(a, st); break;
    case 5: launch<5, LOCAL>(a, st); break;
    case 6: launch<6, LOCAL>(a, st); break;
    case 7: launch<7, LOCAL>(a, st); break;
    case 8: launch<8, LOCAL>(a, st); break;
    case 9: launch<9, LOCAL>(a, st); break;
    case 10: launch<10, LOCAL>(a, st); break;
    case 11: launch<11, LOCAL>(a, st); break;
    case 12: launch<12, LOCAL>(a, st); break;
    case 13: launch<13, LOCAL>(a, st); break;
    case 14: launch<14, LOCAL>(a, st); break;
    case 15: launch<15, LOCAL>(a, st); break;
    case 16: launch<16, LOCAL>(a, st); break;
    default: throw ClastError(1, "gap out of range");
  }
}

void launch_dp_band(int gap, bool local, const DPArgs& a, hipStream_t st) {
  if (local) launch_gap<true>(gap, a, st);
  else launch_gap<false>(gap, a, st);
}

}  // namespace clast
