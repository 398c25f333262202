// Checks clast::score_threshold (the library's per-length E-value cut)
// against plain bisection on the same expression over a grid of parameters
// and query lengths, including cut-offs placed exactly on a score's E-value.
// Prints the number of cases and of mismatches; exit status 1 on any.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <initializer_list>

#include "../clast_amd/csrc/score_threshold.hpp"

static int32_t bisect(double K, double D, double cut, double lambda, int qlen) {
  auto keep = [&](int64_t s) { return !(cut < K * D * qlen * std::exp(-lambda * static_cast<double>(s))); };
  int64_t lo = -(1ll << 30), hi = (1ll << 30);
  if (!keep(hi)) return INT32_MAX;
  if (keep(lo)) return INT32_MIN;
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (keep(mid)) hi = mid;
    else lo = mid;
  }
  return static_cast<int32_t>(hi);
}

int main() {
  const double Ks[] = {0.711, 0.1, 1.0, 0.0};
  const double lams[] = {1.374, 0.5, 2.0, 0.01};
  const double Ds[] = {1e6, 3.0e9, 1e12, 150.0};
  const double cuts[] = {10.0, 1e-5, 1e-200, 1e3, -1.0, 0.0, 1e300};
  const int qlens[] = {0, 1, 17, 50, 150, 151, 1000, 4097, 14999, 30000, 1000000, 2147483647};
  long n = 0, bad = 0;
  auto one = [&](double K, double D, double cut, double lam, int q) {
    ++n;
    const int32_t a = clast::score_threshold(K, D, cut, lam, q), b = bisect(K, D, cut, lam, q);
    if (a != b) {
      if (++bad <= 10) std::printf("mismatch K=%g D=%g cut=%g lambda=%g qlen=%d: %d vs %d\n", K, D, cut, lam, q, a, b);
    }
  };
  for (double K : Ks)
    for (double lam : lams)
      for (double D : Ds)
        for (int q : qlens) {
          for (double cut : cuts) one(K, D, cut, lam, q);
          // cut-offs exactly on the E-value of a score near the threshold
          for (int s : {-40, -3, 0, 1, 7, 25, 60, 400}) one(K, D, K * D * q * std::exp(-lam * s), lam, q);
        }
  // every length up to 20000 under the serving defaults
  for (int q = 0; q <= 20000; ++q) one(0.711, 3.0e9, 10.0, 1.374, q);
  std::printf("cases %ld mismatches %ld\n", n, bad);
  return bad ? 1 : 0;
}
