// Score threshold of one query length: the smallest integer score s with
// !(cutOff < K * D * qLen * exp(-lambda * s)) (…deleteBadHits.cuh:32-39
// with krnlCalculateEvalue.cu:21), decided by that exact expression;
// INT32_MAX when no score passes, INT32_MIN when every score does.  The
// search starts from the closed form ceil(ln(K D qLen / cutOff) / lambda)
// and steps to the boundary (keep(s) and not keep(s - 1)); a guess that is
// not finite or misses by more than a few steps falls back to bisection.
// Host code only (tests/test_host_cpu.py builds it alone and checks it
// against bisection).
#pragma once
#include <cmath>
#include <cstdint>

namespace clast {

inline int32_t score_threshold(double K, double total_db_size, double cut_off, double lambda, int qlen) {
  auto keep = [&](int64_t s) {
    const double e = K * total_db_size * qlen * std::exp(-lambda * static_cast<double>(s));
    return !(cut_off < e);
  };
  int64_t lo = -(1ll << 30), hi = (1ll << 30);
  if (!keep(hi)) return INT32_MAX;
  if (keep(lo)) return INT32_MIN;
  const double g = std::ceil(std::log(K * total_db_size * qlen / cut_off) / lambda);
  if (std::isfinite(g) && g > static_cast<double>(lo + 8) && g < static_cast<double>(hi - 8)) {
    int64_t s = static_cast<int64_t>(g);
    for (int i = 0; i < 4 && !keep(s); ++i) ++s;
    for (int i = 0; i < 4 && keep(s - 1); ++i) --s;
    if (keep(s) && !keep(s - 1)) return static_cast<int32_t>(s);
  }
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (keep(mid)) hi = mid;
    else lo = mid;
  }
  return static_cast<int32_t>(hi);
}

}  // namespace clast
