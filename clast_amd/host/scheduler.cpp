// Window scheduling over the device stages (CHostSchedular.cu:43-130).
#include <memory>
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
  if (!s.batch) {
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
    return;
  }
  // batched: every window of this target chunk in one index, every query
  // window in one query set, one launch chain; rows come back in the order
  // the per-window loop would append them after fixResult's stable sort.
  std::vector<std::vector<int32_t>> tg(twin.size()), qg(qwin.size());
  std::vector<clast_window> tws, qws;
  for (size_t k = 0; k < twin.size(); ++k) tws.push_back(t.window(twin[k].first, twin[k].second, tg[k]));
  for (size_t k = 0; k < qwin.size(); ++k) qws.push_back(q.window(2 * qwin[k].first, 2 * qwin[k].second, qg[k]));
  IndexGuard ix;
  check(clast_index_create_multi(&p, tws.data(), static_cast<int32_t>(tws.size()), &ix.h), "clast_index_create_multi");
  QueryGuard qq;
  check(clast_query_create_multi(&p, qws.data(), static_cast<int32_t>(qws.size()), &qq.h), "clast_query_create_multi");
  HitsGuard hits;
  check(clast_search_batch(&p, ix.h, qq.h, 0, &hits.h), "clast_search_batch");
  holder.add(hits.h, t);
}

}  // namespace clast_host
