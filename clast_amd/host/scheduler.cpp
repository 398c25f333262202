// Window scheduling over the device stages (CHostSchedular.cu:43-130).
#include <climits>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <vector>

#include "host.hpp"

namespace clast_host {

namespace {
struct IndexGuard {
  clast_index h = nullptr;
  ~IndexGuard() { if (h) clast_index_destroy(h); }
};
struct QueryGuard {
  clast_query h = nullptr;
  ~QueryGuard() { if (h) clast_query_destroy(h); }
};
struct HitsGuard {
  clast_hits h = nullptr;
  ~HitsGuard() { if (h) clast_hits_destroy(h); }
};
}  // namespace

void search(const Settings& s, const SeqList& t, const SeqList& q, ResultHolder& holder) {
  const clast_params p = s.params();
  const auto twin = target_windows(t.gateway, s.t_vram);
  const auto qwin = query_windows(q.gateway, s.q_vram);
  if (twin.empty() || qwin.empty()) return;
  // one CDeviceHashTable per target window, one CDeviceSeqList_query +
  // CDeviceHitList per query window
  for (const auto& tw : twin) {
    std::vector<int32_t> tg;
    const clast_window tw_arr = t.window(tw.first, tw.second, tg);
    IndexGuard ix;
    check(clast_index_create(&p, &tw_arr, &ix.h), "clast_index_create");
    for (const auto& qw : qwin) {
      std::vector<int32_t> qg;
      const clast_window qw_arr = q.window(2 * qw.first, 2 * qw.second, qg);
      QueryGuard qq;
      check(clast_query_create(&p, &qw_arr, &qq.h), "clast_query_create");
      HitsGuard hits;
      check(clast_search(&p, ix.h, qq.h, tw.first, 2 * qw.first, &hits.h), "clast_search");
      holder.add(hits.h, t);
    }
  }
}

clast_query query_chunk_reads(const Settings& s, const std::vector<Fasta>& reads) {
  const clast_params p = s.params();
  const int64_t joint = s.lmer > 1 ? s.lmer - 1 : 0;
  std::vector<uint8_t> bases;
  std::vector<int64_t> off{0};
  std::vector<int32_t> gateway{0};  // the query host list's gateway: '+' and '-' per read
  for (const auto& r : reads) {
    bases.insert(bases.end(), r.seq.begin(), r.seq.end());
    off.push_back(static_cast<int64_t>(bases.size()));
    const int64_t span = static_cast<int64_t>(r.seq.size()) + joint;
    for (int k = 0; k < 2; ++k) {
      if (gateway.back() + span > INT32_MAX) throw std::runtime_error("query chunk over 2^31 slots");
      gateway.push_back(static_cast<int32_t>(gateway.back() + span));
    }
  }
  const auto qwin = query_windows(gateway, s.q_vram);
  if (qwin.empty()) return nullptr;
  std::vector<int32_t> win, id_begin;
  for (const auto& w : qwin) {
    win.push_back(w.first);
    win.push_back(w.second);
    id_begin.push_back(2 * w.first);
  }
  clast_query q = nullptr;
  check(clast_query_create_reads(&p, bases.data(), off.data(), win.data(), id_begin.data(),
                                 static_cast<int32_t>(qwin.size()), &q),
        "clast_query_create_reads");
  return q;
}

void search_batched(const Settings& s, const SeqList& t, clast_query queries, ResultHolder& holder) {
  const clast_params p = s.params();
  const auto twin = target_windows(t.gateway, s.t_vram);
  if (twin.empty() || !queries) return;
  std::vector<std::vector<int32_t>> tg(twin.size());
  std::vector<clast_window> tws;
  for (size_t k = 0; k < twin.size(); ++k) tws.push_back(t.window(twin[k].first, twin[k].second, tg[k]));
  IndexGuard ix;
  check(clast_index_create_multi(&p, tws.data(), static_cast<int32_t>(tws.size()), &ix.h), "clast_index_create_multi");
  HitsGuard hits;
  check(clast_search_batch(&p, ix.h, queries, 0, &hits.h), "clast_search_batch");
  holder.add(hits.h, t);
}

}  // namespace clast_host
