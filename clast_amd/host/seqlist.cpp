// Host sequence lists and windows (CHostSeqList*.cu, utilAddSequence.cu,
// utilReverseSeq.cpp, CHostSchedular.cu:12-35).
#include "host.hpp"

namespace clast_host {

static char complement(char b) {
  switch (b) {
    case 'A': return 'T';
    case 'T': return 'A';
    case 'G': return 'C';
    case 'C': return 'G';
    default: return 'N';
  }
}

// addSequence (utilAddSequence.cu:6-47): len bases, then a copy of the first
// lMer-1 (bytes past a shorter sequence's end are 'N').
void SeqList::add_sequence(const char* s, int len, bool rc) {
  const int joint = lmer - 1;
  const size_t at = base.size();
  base.resize(at + static_cast<size_t>(len) + (joint > 0 ? joint : 0));
  char* out = base.data() + at;
  if (rc)
    for (int i = 0; i < len; ++i) out[i] = complement(s[len - 1 - i]);
  else
    for (int i = 0; i < len; ++i) out[i] = s[i];
  for (int i = 0; i < joint; ++i) out[len + i] = i < len ? out[i] : 'N';
  length.push_back(len);
  if (gateway.empty()) gateway.push_back(0);
  gateway.push_back(gateway.back() + len + joint);
}

void SeqList::add_targets(const std::vector<Fasta>& f) {
  for (const auto& r : f) {
    label.push_back(r.label);
    start_idx.push_back(r.start_idx);
    add_sequence(r.seq.data(), static_cast<int>(r.seq.size()), false);
  }
}

void SeqList::add_queries(const std::vector<Fasta>& f) {
  for (const auto& r : f) {
    label.push_back(r.label);
    add_sequence(r.seq.data(), static_cast<int>(r.seq.size()), false);
    label.push_back(r.label);
    add_sequence(r.seq.data(), static_cast<int>(r.seq.size()), true);
  }
}

clast_window SeqList::window(int a, int b, std::vector<int32_t>& gw) const {
  const int32_t lo = gateway[a];
  gw.resize(b - a);
  for (int k = a; k < b; ++k) gw[k - a] = gateway[k] - lo;
  clast_window w;
  w.bases = base.data() + lo;
  w.n_slots = gateway[b] - lo;
  w.lengths = length.data() + a;
  w.gateway = gw.data();
  w.n_seq = b - a;
  w.id_begin = a;
  return w;
}

// moveTargetWindow: extend while the slots from the window start to the end
// of the next sequence stay below the VRAM budget (at least one sequence).
std::vector<std::pair<int, int>> target_windows(const std::vector<int32_t>& gateway, double vram) {
  std::vector<std::pair<int, int>> out;
  const int end_id = static_cast<int>(gateway.size()) - 1;
  for (int b = 0, e = 1; b < end_id; b = e++) {
    const int bi = gateway[b];
    while (e < end_id && gateway[e + 1] - bi < vram) ++e;
    out.emplace_back(b, e);
  }
  return out;
}

// moveQueryWindow: the same over '+'/'-' pairs.
std::vector<std::pair<int, int>> query_windows(const std::vector<int32_t>& gateway, double vram) {
  std::vector<std::pair<int, int>> out;
  const int end_id = (static_cast<int>(gateway.size()) - 1) / 2;
  for (int b = 0, e = 1; b < end_id; b = e++) {
    const int bi = gateway[2 * b];
    while (e < end_id && gateway[2 * e + 2] - bi < vram) ++e;
    out.emplace_back(b, e);
  }
  return out;
}

}  // namespace clast_host
