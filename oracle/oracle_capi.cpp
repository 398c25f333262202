// C entry points over the CPU oracle, for the Python tests (ctypes) and the
// CPU-baseline leg of bench.py — TEST INFRASTRUCTURE ONLY, never linked by
// the product.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "clast_oracle.hpp"

using namespace oracle;

extern "C" {

struct oracle_params {
  int32_t lmer, stride, cut_repeat, width, gap, local;
  double cut_off, total_db_size, K, lambda;
};

// A host sequence list in CHostSeqList layout: the concatenated slot array
// (sequence + lMer-1 joint per sequence), per-sequence lengths and the
// n_seq+1 gateway prefix sums.  IDs and intra-sequence indexes are derived.
struct oracle_list {
  const char* base;
  int64_t n_slots;
  const int32_t* length;
  const int32_t* gateway;
  int32_t n_seq;
};

enum { ORACLE_FINAL = 0, ORACLE_RAW, ORACLE_BOUNDARY, ORACLE_SORTED, ORACLE_ISOLATE, ORACLE_DUPLICATE,
       ORACLE_ALIGNED, ORACLE_FILTERED };

struct oracle_result {
  Hits hits;               // holder order (fixResult) over all units
  std::vector<int> twin;   // target window index of each final hit
  StageDump dump;          // single-unit runs only
  double index_s = 0;      // wall time building the target windows' tables
  double search_s = 0;     // wall time of the unit searches
};

static Params to_params(const oracle_params* p) {
  Params r;
  r.lmer = p->lmer; r.stride = p->stride; r.cut_repeat = p->cut_repeat; r.width = p->width;
  r.gap = p->gap; r.local = p->local; r.cut_off = p->cut_off; r.total_db_size = p->total_db_size;
  r.K = p->K; r.lambda = p->lambda;
  return r;
}

static SeqList to_list(const oracle_list* l) {
  SeqList s;
  s.base.assign(l->base, l->base + l->n_slots);
  s.length.assign(l->length, l->length + l->n_seq);
  s.gateway.assign(l->gateway, l->gateway + l->n_seq + 1);
  s.index.resize(l->n_slots);
  s.id.resize(l->n_slots);
  for (int k = 0; k < l->n_seq; ++k)
    for (int64_t i = s.gateway[k]; i < s.gateway[k + 1]; ++i) {
      s.id[i] = k;
      s.index[i] = static_cast<int>(i - s.gateway[k]);
    }
  s.label.resize(l->n_seq);
  s.start_idx.assign(l->n_seq, 0);
  return s;
}

int64_t oracle_kmer_code(const char* b, int64_t avail, int32_t lmer) { return kmer_code(b, avail, lmer); }

// Runs every (target window x query window) unit in (target window, query
// window) order, appends like CHostResultHolder::addResult and applies the
// fixResult stable sort.  twin: 2*ntw sequence-ID bounds; qwin: 2*nqw
// query-PAIR bounds.  dump != 0 keeps the stage dumps (ntw == nqw == 1 only).
void* oracle_search(const oracle_params* pp, const oracle_list* tl, const int32_t* twin, int32_t ntw,
                    const oracle_list* ql, const int32_t* qwin, int32_t nqw, int32_t dump) {
  const Params p = to_params(pp);
  const SeqList t = to_list(tl), q = to_list(ql);
  auto* r = new oracle_result();
  std::vector<int> order_tw;
  using clk = std::chrono::steady_clock;
  for (int a = 0; a < ntw; ++a) {
    const auto t0 = clk::now();
    HashTable h = build_hash_table(p, t, twin[2 * a], twin[2 * a + 1]);
    const auto t1 = clk::now();
    std::vector<Hits> res(nqw);
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nqw; ++b) {
      QueryWindow qw = make_query_window(p, q, qwin[2 * b], qwin[2 * b + 1]);
      res[b] = search(p, h, qw, twin[2 * a], qwin[2 * b] * 2,
                      (dump && ntw == 1 && nqw == 1) ? &r->dump : nullptr);
    }
    r->index_s += std::chrono::duration<double>(t1 - t0).count();
    r->search_s += std::chrono::duration<double>(clk::now() - t1).count();
    for (int b = 0; b < nqw; ++b) {
      r->hits.append(res[b]);
      order_tw.insert(order_tw.end(), res[b].size(), a);
    }
  }
  // fixResult: stable (qID/2 asc, score desc) carrying the window index along.
  std::vector<size_t> o(r->hits.size());
  for (size_t i = 0; i < o.size(); ++i) o[i] = i;
  const Hits& hh = r->hits;
  std::stable_sort(o.begin(), o.end(), [&](size_t a, size_t b) {
    const int qa = hh.qID[a] / 2, qb = hh.qID[b] / 2;
    if (qa != qb) return qa < qb;
    return hh.score[a] > hh.score[b];
  });
  Hits s;
  for (size_t i : o) {
    Hits one;
    s.tID.push_back(hh.tID[i]); s.tIdx.push_back(hh.tIdx[i]); s.qID.push_back(hh.qID[i]);
    s.qIdx.push_back(hh.qIdx[i]); s.tHL.push_back(hh.tHL[i]); s.qHL.push_back(hh.qHL[i]);
    s.match.push_back(hh.match[i]); s.score.push_back(hh.score[i]); s.evalue.push_back(hh.evalue[i]);
    r->twin.push_back(order_tw[i]);
  }
  r->hits = std::move(s);
  return r;
}

static const Seeds* stage_seeds(const oracle_result* r, int stage) {
  switch (stage) {
    case ORACLE_RAW: return &r->dump.raw;
    case ORACLE_BOUNDARY: return &r->dump.boundary;
    case ORACLE_SORTED: return &r->dump.sorted;
    case ORACLE_ISOLATE: return &r->dump.isolate;
    case ORACLE_DUPLICATE: return &r->dump.duplicate;
    default: return nullptr;
  }
}
static const Hits* stage_hits(const oracle_result* r, int stage) {
  switch (stage) {
    case ORACLE_FINAL: return &r->hits;
    case ORACLE_ALIGNED: return &r->dump.aligned;
    case ORACLE_FILTERED: return &r->dump.filtered;
    default: return nullptr;
  }
}

int64_t oracle_count(void* h, int32_t stage) {
  const auto* r = static_cast<oracle_result*>(h);
  if (const Seeds* s = stage_seeds(r, stage)) return static_cast<int64_t>(s->size());
  if (const Hits* x = stage_hits(r, stage)) return static_cast<int64_t>(x->size());
  return -1;
}

// Seed stages fill the first four arrays; hit stages fill all nine (+ twin
// for the final stage).  Null pointers are skipped.
int32_t oracle_copy(void* h, int32_t stage, int32_t* tID, int32_t* tIdx, int32_t* qID, int32_t* qIdx,
                    int32_t* tHL, int32_t* qHL, int32_t* match, int32_t* score, double* evalue,
                    int32_t* twin) {
  const auto* r = static_cast<oracle_result*>(h);
  auto cp = [](int32_t* dst, const std::vector<int>& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(int32_t));
  };
  if (const Seeds* s = stage_seeds(r, stage)) {
    cp(tID, s->tID); cp(tIdx, s->tIdx); cp(qID, s->qID); cp(qIdx, s->qIdx);
    return 0;
  }
  if (const Hits* x = stage_hits(r, stage)) {
    cp(tID, x->tID); cp(tIdx, x->tIdx); cp(qID, x->qID); cp(qIdx, x->qIdx);
    cp(tHL, x->tHL); cp(qHL, x->qHL); cp(match, x->match); cp(score, x->score);
    if (evalue && !x->evalue.empty()) std::memcpy(evalue, x->evalue.data(), x->evalue.size() * sizeof(double));
    if (twin && stage == ORACLE_FINAL) cp(twin, r->twin);
    return 0;
  }
  return -1;
}

void oracle_free(void* h) { delete static_cast<oracle_result*>(h); }

// Wall seconds of oracle_search: [0] table builds, [1] unit searches.
void oracle_times(void* h, double* out) {
  out[0] = static_cast<oracle_result*>(h)->index_s;
  out[1] = static_cast<oracle_result*>(h)->search_s;
}

// The reference index of one target window (CDeviceHashTable members).
struct oracle_index {
  HashTable h;
};
void* oracle_index_build(const oracle_params* pp, const oracle_list* tl, int32_t t_begin, int32_t t_end) {
  const Params p = to_params(pp);
  const SeqList t = to_list(tl);
  auto* r = new oracle_index();
  r->h = build_hash_table(p, t, t_begin, t_end);
  return r;
}
int64_t oracle_index_cells(void* h) { return static_cast<int64_t>(static_cast<oracle_index*>(h)->h.gateway_key.size()); }
int64_t oracle_index_size(void* h) { return static_cast<int64_t>(static_cast<oracle_index*>(h)->h.index_array.size()); }
void oracle_index_copy(void* h, int64_t* key, int32_t* cell_size, int32_t* gateway_index, int32_t* positions) {
  const HashTable& t = static_cast<oracle_index*>(h)->h;
  std::memcpy(key, t.gateway_key.data(), t.gateway_key.size() * sizeof(int64_t));
  std::memcpy(cell_size, t.cell_size.data(), t.cell_size.size() * sizeof(int32_t));
  std::memcpy(gateway_index, t.gateway_index.data(), t.gateway_index.size() * sizeof(int32_t));
  std::memcpy(positions, t.index_array.data(), t.index_array.size() * sizeof(int32_t));
}
void oracle_index_free(void* h) { delete static_cast<oracle_index*>(h); }

// Both DP directions for one seed (known-answer tests).  out = tIdx, qIdx,
// tHL, qHL, match, score.
void oracle_align_one(const oracle_params* pp, const oracle_list* tl, int32_t t_begin, int32_t t_end,
                      const oracle_list* ql, int32_t pair_begin, int32_t pair_end, int32_t tID,
                      int32_t tIdx, int32_t qID, int32_t qIdx, int32_t* out) {
  const Params p = to_params(pp);
  const SeqList t = to_list(tl), q = to_list(ql);
  const Window tw = make_window(t, t_begin, t_end);
  const Window qw = make_window(q, 2 * pair_begin, 2 * pair_end);
  const DPResult d = align_one(p, tw, qw, t_begin, 2 * pair_begin, tID, tIdx, qID, qIdx);
  out[0] = d.tIdx; out[1] = d.qIdx; out[2] = d.tHL; out[3] = d.qHL; out[4] = d.match; out[5] = d.score;
}

int64_t oracle_total_db_size(const char* const* files, int32_t n) {
  std::vector<std::string> f(files, files + n);
  return calc_total_db_size(f);
}

}  // extern "C"
