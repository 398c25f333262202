// Autotuner of the DP stage of libclast_hip.  A DP configuration picks, for
// one search, how the banded extensions are launched:
//   sched  the packed 16-bit kernel's schedule: 0 static (one wavefront per
//          chunk of 64 task pairs, snake order over the SIMDs), 1 persistent
//          (one occupancy round of wavefronts taking chunks longest first
//          from a device counter);
//   band   which extensions of at least kBandRows rows run one lane per band
//          (k_dp_band): 0 none, 1 those the 16-bit kernel cannot take, 2
//          every one over twice a SIMD's average share of rows;
//   radix  digit bits of the task-key radix sort (4..8).
// Every configuration computes the same rows; they differ only in speed.
// The tuner keeps, per (gap, global / local, length bucket of the search's
// longest extension, size class of the search), the configuration that was
// fastest in timed trials on the first search of that kind (each the least
// of three runs), and serves it to later searches.  The size class keeps small searches, whose DP stage is
// bound by the latency of their longest extensions, apart from large ones,
// bound by throughput: the faster configuration differs between them.
// Environment: CLAST_TUNE=0 turns trials off (the defaults, sched 0, band 0,
// radix 8, are then used); CLAST_DP_SCHED=static|persist, CLAST_DP_BAND=
// off|rest|all and CLAST_RADIX_BITS=4..8 fix a knob (the tuner varies only
// the others).
//
// Two-pass DP stage (global mode with the E-value cut): the 16-bit tasks run
// first as scores only, and the tasks of the hits that pass the bad-hit
// filters run again with match counts.  It pays while few 16-bit tasks
// survive (the score pass issues 11 of the full kernel's 18 instructions per
// cell pair), so the tuner keeps, per kind, the surviving share of the
// 16-bit tasks of the searches so far, and a search runs two passes while
// that share is below kTwoPassMax (and on the first search of its kind).
// CLAST_DP_TWO_PASS=0|1 fixes the choice.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "runtime.hpp"

namespace clast {

struct DPConfig {
  int sched = 0;
  int band = 0;
  int radix = 8;
  bool operator==(const DPConfig& o) const { return sched == o.sched && band == o.band && radix == o.radix; }
};

// Length bucket of a search by its longest extension (rows).
inline int dp_bucket(int maxrows) { return maxrows < 256 ? 0 : maxrows < 2048 ? 1 : maxrows < 8192 ? 2 : 3; }

// Size class of a search by its DP task count: 0 below kSmallTasks.
constexpr uint32_t kSmallTasks = 1u << 15;
inline int dp_size_class(uint32_t ntask) { return ntask < kSmallTasks ? 0 : 1; }

// Forced knobs from the environment (-1: free).
struct DPForced {
  int sched = -1, band = -1, radix = -1, two_pass = -1;
  bool tune = true;
  static DPForced from_env() {
    DPForced f;
    if (const char* v = std::getenv("CLAST_TUNE")) f.tune = !(std::strcmp(v, "0") == 0 || std::strcmp(v, "off") == 0);
    if (const char* v = std::getenv("CLAST_DP_SCHED")) {
      if (!std::strcmp(v, "static")) f.sched = 0;
      else if (!std::strcmp(v, "persist")) f.sched = 1;
      else throw ClastError(1, "CLAST_DP_SCHED must be static or persist");
    }
    if (const char* v = std::getenv("CLAST_DP_BAND")) {
      if (!std::strcmp(v, "off")) f.band = 0;
      else if (!std::strcmp(v, "rest")) f.band = 1;
      else if (!std::strcmp(v, "all")) f.band = 2;
      else throw ClastError(1, "CLAST_DP_BAND must be off, rest or all");
    }
    if (const char* v = std::getenv("CLAST_DP_TWO_PASS")) {
      if (!std::strcmp(v, "0")) f.two_pass = 0;
      else if (!std::strcmp(v, "1")) f.two_pass = 1;
      else if (std::strcmp(v, "auto")) throw ClastError(1, "CLAST_DP_TWO_PASS must be 0, 1 or auto");
    }
    if (const char* v = std::getenv("CLAST_RADIX_BITS")) {
      const int b = std::atoi(v);
      if (b < 4 || b > 8) throw ClastError(1, "CLAST_RADIX_BITS must be in 4..8");
      f.radix = b;
    }
    return f;
  }
  DPConfig apply(DPConfig c) const {
    if (sched >= 0) c.sched = sched;
    if (band >= 0) c.band = band;
    if (radix >= 0) c.radix = radix;
    return c;
  }
};

class DPTuner {
 public:
  static DPTuner& get() {
    static DPTuner* t = new DPTuner();  // never destroyed: used from any thread until exit
    return *t;
  }
  using Key = std::tuple<int, int, int, int>;  // gap, local, bucket, size class
  bool lookup(const Key& k, DPConfig* out) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = table_.find(k);
    if (it == table_.end()) return false;
    *out = it->second.best;
    return true;
  }
  // Two passes for the next search of kind k (forced: -1 free, 0, 1)?
  bool two_pass(const Key& k, int forced) {
    if (forced >= 0) return forced == 1;
    std::lock_guard<std::mutex> lk(mu_);
    auto it = survive_.find(k);
    return it == survive_.end() || it->second < kTwoPassMax;
  }
  // A search of kind k ran `total` 16-bit tasks of which `kept` belong to
  // hits that passed the filters (running mean, recent searches weighted).
  void survived(const Key& k, uint64_t kept, uint64_t total) {
    if (total == 0) return;
    const double f = static_cast<double>(kept) / static_cast<double>(total);
    std::lock_guard<std::mutex> lk(mu_);
    auto it = survive_.find(k);
    if (it == survive_.end()) survive_[k] = f;
    else it->second = 0.75 * it->second + 0.25 * f;
  }
  static constexpr double kTwoPassMax = 0.4;

  void store(const Key& k, const DPConfig& best, const std::vector<std::pair<DPConfig, float>>& trials) {
    std::lock_guard<std::mutex> lk(mu_);
    if (table_.count(k)) return;  // another thread tuned it first
    table_[k] = Entry{best, trials};
  }
  // One line per tuned kind: "gap=G local=L bucket=B size=Z -> sched=S band=N radix=R | trials: cfg ms, ..."
  std::string report() {
    std::lock_guard<std::mutex> lk(mu_);
    std::string s;
    char buf[160];
    for (auto& kv : table_) {
      const auto& e = kv.second;
      std::snprintf(buf, sizeof buf, "gap=%d local=%d bucket=%d size=%d -> sched=%d band=%d radix=%d | trials:",
                    std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), std::get<3>(kv.first),
                    e.best.sched, e.best.band, e.best.radix);
      s += buf;
      for (auto& t : e.trials) {
        std::snprintf(buf, sizeof buf, " %d/%d/%d %.3f ms", t.first.sched, t.first.band, t.first.radix, t.second);
        s += buf;
      }
      s += "\n";
    }
    for (auto& kv : survive_) {
      std::snprintf(buf, sizeof buf, "gap=%d local=%d bucket=%d size=%d: 16-bit tasks surviving the filters %.4f\n",
                    std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), std::get<3>(kv.first),
                    kv.second);
      s += buf;
    }
    return s;
  }

 private:
  struct Entry {
    DPConfig best;
    std::vector<std::pair<DPConfig, float>> trials;
  };
  std::mutex mu_;
  std::map<Key, Entry> table_;
  std::map<Key, double> survive_;
};

}  // namespace clast
