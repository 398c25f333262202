// CPU oracle for the CLAST search pipeline — TEST INFRASTRUCTURE ONLY.
// See clast_oracle.hpp for scope and the conventions used where the reference
// is undefined.  Reference citations are file:line in masayano/CLAST.
#include "clast_oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <numeric>
#include <sstream>
#include <stdexcept>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

// ---------------------------------------------------------------- host lists

// utilReverseSeq.cpp:3-19: A<->T, G<->C, anything else -> 'N', reversed.
std::string comp_seq(const std::string& seq) {
  std::string out;
  out.reserve(seq.size());
  for (auto it = seq.rbegin(); it != seq.rend(); ++it) {
    char c = *it;
    char r = 'N';
    if (c == 'A') r = 'T';
    else if (c == 'T') r = 'A';
    else if (c == 'G') r = 'C';
    else if (c == 'C') r = 'G';
    out.push_back(r);
  }
  return out;
}

// utilAddSequence.cu:6-47: append len + lMer - 1 slots (the sequence, then a
// copy of its first lMer - 1 bases); ID = previous ID + 1 (0 for the first).
void add_sequence(int lmer, const std::string& seq, SeqList& s) {
  const int len = static_cast<int>(seq.size());
  const int joint = lmer - 1;
  const int id = s.index.empty() ? 0 : s.id.back() + 1;
  for (int i = 0; i < len + joint; ++i) s.index.push_back(i);
  s.id.insert(s.id.end(), static_cast<size_t>(len + joint), id);
  s.base.insert(s.base.end(), seq.begin(), seq.end());
  for (int i = 0; i < joint; ++i) s.base.push_back(i < len ? seq[i] : 'N');
}

static void push_gateway(SeqList& s, int len, int lmer) {
  if (s.gateway.empty()) s.gateway.push_back(0);
  s.gateway.push_back(s.gateway.back() + len + lmer - 1);
}

// CHostSeqList_target.cu:8-33
void add_target(const Params& p, const std::vector<Fasta>& f, SeqList& s) {
  for (const auto& r : f) {
    const int len = static_cast<int>(r.seq.size());
    s.label.push_back(r.label);
    s.start_idx.push_back(r.start_idx);
    add_sequence(p.lmer, r.seq, s);
    s.length.push_back(len);
    push_gateway(s, len, p.lmer);
  }
}

// CHostSeqList_query.cu:11-54: '+' strand then '-' (reverse complement).
void add_query(const Params& p, const std::vector<Fasta>& f, SeqList& s) {
  for (const auto& r : f) {
    const int len = static_cast<int>(r.seq.size());
    s.label.push_back(r.label);
    add_sequence(p.lmer, r.seq, s);
    s.length.push_back(len);
    push_gateway(s, len, p.lmer);
    s.label.push_back(r.label);
    add_sequence(p.lmer, comp_seq(r.seq), s);
    s.length.push_back(len);
    push_gateway(s, len, p.lmer);
  }
}

// CHostSeqList.cu:15-70: slot arrays [gateway[a], gateway[b]); length and
// gateway [a, b), gateway rebased to the window start.
Window make_window(const SeqList& s, int a, int b) {
  Window w;
  const int lo = s.gateway[a], hi = s.gateway[b];
  w.index.assign(s.index.begin() + lo, s.index.begin() + hi);
  w.id.assign(s.id.begin() + lo, s.id.begin() + hi);
  w.base.assign(s.base.begin() + lo, s.base.begin() + hi);
  w.length.assign(s.length.begin() + a, s.length.begin() + b);
  w.gateway.assign(s.gateway.begin() + a, s.gateway.begin() + b);
  for (auto& g : w.gateway) g -= lo;
  return w;
}

// SRead2bit.cuh:6-21 + SRepelHasOddBase.cuh:6-14: MSB-first 2-bit code,
// A0 C1 G2 T3; any other byte (or a byte past `avail`) makes the k-mer -1.
int64_t kmer_code(const char* b, int64_t avail, int lmer) {
  int64_t key = 0;
  bool odd = false;
  for (int j = 0; j < lmer; ++j) {
    key <<= 2;
    if (j >= avail) { odd = true; continue; }
    switch (b[j]) {
      case 'A': key += 0; break;
      case 'C': key += 1; break;
      case 'G': key += 2; break;
      case 'T': key += 3; break;
      default: odd = true;
    }
  }
  return odd ? -1 : key;
}

// ---------------------------------------------------------------- index

// CDeviceHashTable.cu:20-174
HashTable build_hash_table(const Params& p, const SeqList& t, int t_begin, int t_end) {
  HashTable h;
  h.target = make_window(t, t_begin, t_end);
  const std::vector<char>& base = h.target.base;
  const int64_t n = static_cast<int64_t>(base.size());
  const int64_t size = n / p.stride;  // createHashIndex :34, makeIndexArray :76
  std::vector<int64_t> key(size);
  std::vector<int> pos(size);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < size; ++i) {
    const int64_t at = i * p.stride;
    key[i] = kmer_code(base.data() + at, n - at, p.lmer);
    pos[i] = static_cast<int>(at);
  }
  // sort_by_key (:159), stable: positions ascend within a key.
  std::vector<int> order(size);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  std::vector<int64_t> sk(size);
  h.index_array.resize(size);
  for (int64_t i = 0; i < size; ++i) {
    sk[i] = key[order[i]];
    h.index_array[i] = pos[order[i]];
  }
  // makeGatewayKey (:87-94), makeCellSizeArray (:96-110), makeGatewayIndex (:112-120)
  for (int64_t i = 0; i < size; ++i) {
    if (i == 0 || sk[i] != sk[i - 1]) {
      h.gateway_key.push_back(sk[i]);
      h.cell_size.push_back(0);
    }
    h.cell_size.back() += 1;
  }
  h.gateway_index.resize(h.cell_size.size());
  int run = 0;
  for (size_t k = 0; k < h.cell_size.size(); ++k) {
    h.gateway_index[k] = run;
    run += h.cell_size[k];
  }
  // removeRepeatedSequence (:122-137) after the offsets were taken
  if (p.cut_repeat != -1)
    for (auto& c : h.cell_size) c = c < p.cut_repeat ? c : 0;
  // :170-173 — the invalid cell is zeroed (guarded: an empty table has no cell 0)
  if (!h.gateway_key.empty() && h.gateway_key[0] == -1) h.cell_size[0] = 0;
  return h;
}

// CDeviceSeqList_query.cu:15-71: stride-1 codes over the whole window.
QueryWindow make_query_window(const Params& p, const SeqList& q, int pb, int pe) {
  QueryWindow qw;
  qw.w = make_window(q, 2 * pb, 2 * pe);
  const int64_t n = static_cast<int64_t>(qw.w.base.size());
  const int64_t size = n - (p.lmer - 1);
  qw.hash_index.resize(size > 0 ? size : 0);
  for (int64_t i = 0; i < size; ++i) qw.hash_index[i] = kmer_code(qw.w.base.data() + i, n - i, p.lmer);
  return qw;
}

// ---------------------------------------------------------------- seeds

// krnlBinarySearch.cu:3-33
static int binary_search(const std::vector<int64_t>& db, int64_t input) {
  int lo = 0, hi = static_cast<int>(db.size()) - 1, out = -1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int64_t v = db[mid];
    if (v < input) lo = mid + 1;
    else if (v == input) { out = mid; break; }
    else hi = mid - 1;
  }
  return out;
}

// CDeviceHitList_createRawSeedList.cu:139-232 + krnlWriteSeedList.cu:3-37
static Seeds raw_seeds(const HashTable& h, const QueryWindow& q) {
  Seeds s;
  const size_t nq = q.hash_index.size();
  for (size_t i = 0; i < nq; ++i) {
    const int cell = binary_search(h.gateway_key, q.hash_index[i]);
    const int gi = cell == -1 ? 0 : h.gateway_index[cell];
    const int cnt = cell == -1 ? 0 : h.cell_size[cell];
    for (int j = 0; j < cnt; ++j) {
      const int ref = h.index_array[gi + j];
      s.tID.push_back(h.target.id[ref]);
      s.tIdx.push_back(h.target.index[ref]);
      s.qID.push_back(q.w.id[i]);
      s.qIdx.push_back(q.w.index[i]);
    }
  }
  return s;
}

template <class Pred>
static void seed_remove_if(Seeds& s, Pred drop) {  // order-preserving remove_if
  size_t o = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    if (drop(i)) continue;
    s.tID[o] = s.tID[i]; s.tIdx[o] = s.tIdx[i]; s.qID[o] = s.qID[i]; s.qIdx[o] = s.qIdx[i];
    ++o;
  }
  s.tID.resize(o); s.tIdx.resize(o); s.qID.resize(o); s.qIdx.resize(o);
}

// CDeviceHitList_deleteSeedsOnSequenceBoundary.cu:192-269, predicates .cuh:21-49
static void boundary_filter(const Params& p, const HashTable& h, const QueryWindow& q, int t_begin,
                            int q_begin, Seeds& s) {
  const int L = p.lmer;
  seed_remove_if(s, [&](size_t i) { return L + s.qIdx[i] > q.w.length[s.qID[i] - q_begin]; });
  seed_remove_if(s, [&](size_t i) { return L + s.tIdx[i] > h.target.length[s.tID[i] - t_begin]; });
  if (!p.local) {
    seed_remove_if(s, [&](size_t i) {
      const int tLen = h.target.length[s.tID[i] - t_begin];
      const int qLen = q.w.length[s.qID[i] - q_begin];
      const int start = s.tIdx[i] - s.qIdx[i];
      if (start + p.gap < 0) return true;
      if (start + qLen - p.gap > tLen) return true;
      return false;
    });
  }
}

// CDeviceHitList_sortSeeds.cuh:14-55 (qID, tID, tIdx-qIdx, qIdx) ascending
static void sort_seeds(Seeds& s) {
  std::vector<size_t> o(s.size());
  std::iota(o.begin(), o.end(), 0);
  std::stable_sort(o.begin(), o.end(), [&](size_t a, size_t b) {
    if (s.qID[a] != s.qID[b]) return s.qID[a] < s.qID[b];
    if (s.tID[a] != s.tID[b]) return s.tID[a] < s.tID[b];
    const int da = s.tIdx[a] - s.qIdx[a], db = s.tIdx[b] - s.qIdx[b];
    if (da != db) return da < db;
    return s.qIdx[a] < s.qIdx[b];
  });
  Seeds r;
  for (size_t i : o) {
    r.tID.push_back(s.tID[i]); r.tIdx.push_back(s.tIdx[i]);
    r.qID.push_back(s.qID[i]); r.qIdx.push_back(s.qIdx[i]);
  }
  s = std::move(r);
}

// hasNearPair / hasNotNearPair core (…deleteDuplicateSeeds.cuh:29-38,
// …deleteIsolateSeeds.cuh:28-37): `post` is the predecessor element.
static bool near_pair(const Params& p, const Seeds& s, size_t cur, size_t post) {
  if (s.qID[cur] == s.qID[post] && s.tID[cur] == s.tID[post]) {
    if (p.width >= s.tIdx[cur] - s.tIdx[post] && p.width >= s.tIdx[post] - s.tIdx[cur]) {
      const int post_dia = s.tIdx[post] - s.qIdx[post];
      const int dia = s.tIdx[cur] - s.qIdx[cur];
      if (p.gap >= post_dia - dia) return true;
    }
  }
  return false;
}

// …deleteIsolateSeeds.cu:19-77 — element 0 kept; i >= 1 kept iff it has a
// near predecessor (evaluated on the pre-filter array).
static void isolate_filter(const Params& p, Seeds& s) {
  if (s.size() <= 1) return;
  std::vector<char> keep(s.size(), 0);
  keep[0] = 1;
  for (size_t i = 1; i < s.size(); ++i) keep[i] = near_pair(p, s, i, i - 1);
  seed_remove_if(s, [&](size_t i) { return !keep[i]; });
}

// …deleteDuplicateSeeds.cu:15-71 — element 0 kept; i >= 1 dropped iff near.
static void duplicate_filter(const Params& p, Seeds& s) {
  if (s.size() <= 1) return;
  std::vector<char> keep(s.size(), 0);
  keep[0] = 1;
  for (size_t i = 1; i < s.size(); ++i) keep[i] = !near_pair(p, s, i, i - 1);
  seed_remove_if(s, [&](size_t i) { return !keep[i]; });
}

// ---------------------------------------------------------------- DP

// krnlMatrix.cuh:11-67 as a class rule (verified entry-by-entry against the
// 26x26 tables): X row/column -> bad area; E F I J L O P Q Z -> -1; identical
// A/C/G/T -> match; every other pair (N vs N included) -> mismatch.
static bool is_star(char c) {
  switch (c) {
    case 'E': case 'F': case 'I': case 'J': case 'L': case 'O': case 'P': case 'Q': case 'Z': return true;
    default: return false;
  }
}
static int point_matrix(char t, char q) {
  if (t == 'X' || q == 'X') return kBadAreaPoint;
  if (is_star(t) || is_star(q)) return -1;
  if (t == q && (t == 'A' || t == 'C' || t == 'G' || t == 'T')) return kMatchPoint;
  return kMismatchPoint;
}
static int match_counter(char t, char q) {
  if (t == 'X' || q == 'X') return kBadAreaPoint;
  if (is_star(t) || is_star(q)) return -1;
  if (t == q && (t == 'A' || t == 'C' || t == 'G' || t == 'T')) return 1;
  return 0;
}

struct Band {  // one hit's column of the [2g+3][hitNum] temp-node arrays
  std::vector<int> s, v, h, m;
};

// initTempNodeArray, krnlAlignment.cu:6-38
static Band init_band(int gap) {
  const int half = gap + kMargin, tw = 1 + 2 * half;
  Band b;
  b.s.resize(tw); b.v.resize(tw); b.h.resize(tw); b.m.resize(tw);
  for (int i = 0; i < tw; ++i) {
    if (i < half) {
      b.s[i] = b.v[i] = b.h[i] = b.m[i] = -30000;
    } else if (i == half) {
      b.s[i] = 0; b.v[i] = kGapOpenPoint; b.h[i] = kGapOpenPoint; b.m[i] = 0;
    } else {
      const int k = i - half;
      const int sc = k * kGapPoint + kGapOpenPoint;
      b.s[i] = sc; b.v[i] = sc + kGapOpenPoint; b.h[i] = sc; b.m[i] = 0;
    }
  }
  return b;
}

// calcutlateCell, krnlAlignment.cu:42-77; band index `b` is the slant node,
// b+1 the vertical (old row), b-1 the horizontal (new row).
static void calc_cell(char tb, char qb, Band& n, int b) {
  const int vs = n.s[b + 1], vv = n.v[b + 1], vm = n.m[b + 1];
  const int hs = n.s[b - 1], hh = n.h[b - 1], hm = n.m[b - 1];
  n.v[b] = std::max(vs + kGapOpenPoint + kGapPoint, vv + kGapPoint);
  n.h[b] = std::max(hs + kGapOpenPoint + kGapPoint, hh + kGapPoint);
  const int point = n.s[b] + point_matrix(tb, qb);
  if (point > std::max(n.v[b], n.h[b])) {
    n.s[b] = point;
    n.m[b] += match_counter(tb, qb);
  } else if (n.v[b] > n.h[b]) {
    n.s[b] = n.v[b];
    n.m[b] = vm;
  } else {
    n.s[b] = n.h[b];
    n.m[b] = hm;
  }
}

// Target base at intra-sequence index `ti` of a sequence starting at window
// slot `start`.  `guard_hi` selects the reference's one-sided guard:
// backward kernels test ti >= tLen (krnlAlignment.cu:123,161), forward
// kernels test ti < 0 (:236,272).  Reads that leave the window array are 'X'.
static char target_base(const Window& t, int start, int ti, int tLen, bool guard_hi) {
  if (guard_hi ? ti >= tLen : ti < 0) return kBadAreaChar;
  const int64_t at = static_cast<int64_t>(start) + ti;
  if (at < 0 || at >= static_cast<int64_t>(t.base.size())) return kBadAreaChar;
  return t.base[at];
}

struct HitRef {
  int tID, qID, tIdx, qIdx, tHL, qHL, match, score;
};

// globalAlignBackward / localAlignBackward, krnlAlignment.cu:81-195, 312-423
static void align_backward(const Params& p, const Window& t, const Window& q, int t_begin,
                           int q_begin, HitRef& hit) {
  const int g = p.gap, W = 1 + 2 * g;
  const int tStart = t.gateway[hit.tID - t_begin], qStart = q.gateway[hit.qID - q_begin];
  const int tLen = t.length[hit.tID - t_begin], qLen = q.length[hit.qID - q_begin];
  std::vector<char> ring(W);
  for (int k = 0; k < W; ++k) ring[k] = target_base(t, tStart, hit.tIdx + hit.qHL - g + k, tLen, true);
  Band n = init_band(g);
  int calcIdx = -1, endPoint = p.local ? g : -1, bestS = 0, bestM = 0;
  int ri = 0;
  const int startQ = hit.qIdx + hit.qHL;
  const int rows = qLen - startQ;
  for (int r = 0; r < rows; ++r) {
    const char qb = q.base[qStart + startQ + r];
    for (int k = 0; k < W; ++k) {
      calc_cell(ring[ri++], qb, n, k + 1);
      if (ri == W) ri = 0;
      if (p.local && bestS < n.s[k + 1]) {
        bestS = n.s[k + 1]; bestM = n.m[k + 1]; calcIdx = r; endPoint = k;
      }
    }
    ring[ri] = target_base(t, tStart, hit.tIdx + hit.qHL + g + r + 1, tLen, true);
    if (++ri == W) ri = 0;
  }
  if (!p.local) {  // :170-186 select end node: strict '<' from band 0
    bestS = n.s[0]; bestM = n.m[0]; endPoint = -1;
    for (int k = 0; k < W; ++k)
      if (bestS < n.s[k + 1]) { bestS = n.s[k + 1]; bestM = n.m[k + 1]; endPoint = k; }
  }
  endPoint -= g;
  if (rows != 0) {
    const int move = p.local ? calcIdx + 1 : rows;
    hit.tHL += move + endPoint;
    hit.qHL += move;
    hit.match += bestM;
    hit.score += bestS;
  }
}

// globalAlignForward / localAlignForward, krnlAlignment.cu:197-308, 425-533
static void align_forward(const Params& p, const Window& t, const Window& q, int t_begin,
                          int q_begin, HitRef& hit) {
  const int g = p.gap, W = 1 + 2 * g;
  const int tStart = t.gateway[hit.tID - t_begin], qStart = q.gateway[hit.qID - q_begin];
  const int tLen = t.length[hit.tID - t_begin];
  std::vector<char> ring(W);
  for (int k = 0; k < W; ++k) ring[k] = target_base(t, tStart, hit.tIdx + g - k - 1, tLen, false);
  Band n = init_band(g);
  int calcIdx = -1, endPoint = p.local ? g : -1, bestS = 0, bestM = 0;
  int ri = 0;
  const int rows = hit.qIdx;
  for (int r = 0; r < rows; ++r) {
    const char qb = q.base[qStart + hit.qIdx - r - 1];
    for (int k = 0; k < W; ++k) {
      calc_cell(ring[ri++], qb, n, k + 1);
      if (ri == W) ri = 0;
      if (p.local && bestS < n.s[k + 1]) {
        bestS = n.s[k + 1]; bestM = n.m[k + 1]; calcIdx = r; endPoint = k;
      }
    }
    ring[ri] = target_base(t, tStart, hit.tIdx - g - r - 2, tLen, false);
    if (++ri == W) ri = 0;
  }
  if (!p.local) {
    bestS = n.s[0]; bestM = n.m[0]; endPoint = -1;
    for (int k = 0; k < W; ++k)
      if (bestS < n.s[k + 1]) { bestS = n.s[k + 1]; bestM = n.m[k + 1]; endPoint = k; }
  }
  endPoint -= g;
  if (rows != 0) {
    const int move = p.local ? calcIdx + 1 : rows;
    hit.tIdx -= move + endPoint;
    hit.qIdx -= move;
    hit.tHL += move + endPoint;
    hit.qHL += move;
    hit.match += bestM;
    hit.score += bestS;
  }
}

DPResult align_one(const Params& p, const Window& t, const Window& q, int t_begin, int q_begin,
                   int tID, int tIdx, int qID, int qIdx) {
  HitRef h{tID, qID, tIdx, qIdx, p.lmer, p.lmer, p.lmer, p.lmer * kMatchPoint};
  align_backward(p, t, q, t_begin, q_begin, h);
  align_forward(p, t, q, t_begin, q_begin, h);
  return DPResult{h.tIdx, h.qIdx, h.tHL, h.qHL, h.match, h.score};
}

template <class Cmp>
static void stable_sort_hits(std::vector<HitRef>& v, Cmp cmp) {
  std::stable_sort(v.begin(), v.end(), cmp);
}

void Hits::append(const Hits& o) {
  tID.insert(tID.end(), o.tID.begin(), o.tID.end());
  tIdx.insert(tIdx.end(), o.tIdx.begin(), o.tIdx.end());
  qID.insert(qID.end(), o.qID.begin(), o.qID.end());
  qIdx.insert(qIdx.end(), o.qIdx.begin(), o.qIdx.end());
  tHL.insert(tHL.end(), o.tHL.begin(), o.tHL.end());
  qHL.insert(qHL.end(), o.qHL.begin(), o.qHL.end());
  match.insert(match.end(), o.match.begin(), o.match.end());
  score.insert(score.end(), o.score.begin(), o.score.end());
  evalue.insert(evalue.end(), o.evalue.begin(), o.evalue.end());
}

static Hits to_hits(const std::vector<HitRef>& v, const std::vector<double>& e) {
  Hits h;
  for (size_t i = 0; i < v.size(); ++i) {
    h.tID.push_back(v[i].tID); h.tIdx.push_back(v[i].tIdx);
    h.qID.push_back(v[i].qID); h.qIdx.push_back(v[i].qIdx);
    h.tHL.push_back(v[i].tHL); h.qHL.push_back(v[i].qHL);
    h.match.push_back(v[i].match); h.score.push_back(v[i].score);
    h.evalue.push_back(e.empty() ? 0.0 : e[i]);
  }
  return h;
}

// utilResultSorting.cuh:6-26, stable
void result_sort(Hits& h) {
  std::vector<size_t> o(h.size());
  std::iota(o.begin(), o.end(), 0);
  std::stable_sort(o.begin(), o.end(), [&](size_t a, size_t b) {
    const int qa = h.qID[a] / 2, qb = h.qID[b] / 2;
    if (qa != qb) return qa < qb;
    return h.score[a] > h.score[b];
  });
  Hits r;
  for (size_t i : o) {
    r.tID.push_back(h.tID[i]); r.tIdx.push_back(h.tIdx[i]);
    r.qID.push_back(h.qID[i]); r.qIdx.push_back(h.qIdx[i]);
    r.tHL.push_back(h.tHL[i]); r.qHL.push_back(h.qHL[i]);
    r.match.push_back(h.match[i]); r.score.push_back(h.score[i]);
    r.evalue.push_back(h.evalue[i]);
  }
  h = std::move(r);
}

// ---------------------------------------------------------------- the chain

Hits search(const Params& p, const HashTable& h, const QueryWindow& q, int t_begin, int q_begin,
            StageDump* dump) {
  Seeds s = raw_seeds(h, q);
  if (dump) dump->raw = s;
  boundary_filter(p, h, q, t_begin, q_begin, s);
  if (dump) dump->boundary = s;
  sort_seeds(s);
  if (dump) dump->sorted = s;
  isolate_filter(p, s);
  if (dump) dump->isolate = s;
  duplicate_filter(p, s);
  if (dump) dump->duplicate = s;

  // CDeviceHitList.cu:283-286 hit initialisation
  std::vector<HitRef> hits;
  hits.reserve(s.size());
  for (size_t i = 0; i < s.size(); ++i)
    hits.push_back(HitRef{s.tID[i], s.qID[i], s.tIdx[i], s.qIdx[i], p.lmer, p.lmer, p.lmer,
                          p.lmer * kMatchPoint});

  // CDeviceHitList_alignmentHits.cu:122-348
  // sortBeforeAlignBackWard (:10-78): stable by qLen - (qIdx + qHL) descending
  stable_sort_hits(hits, [&](const HitRef& a, const HitRef& b) {
    const int sa = q.w.length[a.qID - q_begin] - (a.qIdx + a.qHL);
    const int sb = q.w.length[b.qID - q_begin] - (b.qIdx + b.qHL);
    return sa > sb;
  });
#pragma omp parallel for schedule(dynamic, 256)
  for (size_t i = 0; i < hits.size(); ++i) align_backward(p, h.target, q.w, t_begin, q_begin, hits[i]);
  // sortBeforeAlignForward (:80-118): stable by qIdx descending
  stable_sort_hits(hits, [](const HitRef& a, const HitRef& b) { return a.qIdx > b.qIdx; });
#pragma omp parallel for schedule(dynamic, 256)
  for (size_t i = 0; i < hits.size(); ++i) align_forward(p, h.target, q.w, t_begin, q_begin, hits[i]);

  // calculateEvalue, krnlCalculateEvalue.cu:5-23
  std::vector<double> ev(hits.size());
  for (size_t i = 0; i < hits.size(); ++i) {
    const int queryLength = q.w.length[hits[i].qID - q_begin];
    const int score = hits[i].score;
    ev[i] = p.K * p.total_db_size * queryLength * std::exp(-p.lambda * score);
  }
  Hits all = to_hits(hits, ev);
  if (dump) dump->aligned = all;

  // deleteHits_lowEValue / deleteHits_tooShort, …deleteBadHits.cu:7-221
  Hits kept;
  for (size_t i = 0; i < all.size(); ++i) {
    if (p.cut_off != -1 && p.cut_off < all.evalue[i]) continue;
    if (p.lmer + p.stride > all.qHL[i]) continue;
    kept.tID.push_back(all.tID[i]); kept.tIdx.push_back(all.tIdx[i]);
    kept.qID.push_back(all.qID[i]); kept.qIdx.push_back(all.qIdx[i]);
    kept.tHL.push_back(all.tHL[i]); kept.qHL.push_back(all.qHL[i]);
    kept.match.push_back(all.match[i]); kept.score.push_back(all.score[i]);
    kept.evalue.push_back(all.evalue[i]);
  }
  if (dump) dump->filtered = kept;

  // addResult (:132-225) into the empty member arrays, resultSorting (:392),
  // deleteHits_duplicateResult (:32-128) against the original predecessor.
  result_sort(kept);
  if (kept.size() <= 1) return kept;
  std::vector<char> keep(kept.size(), 1);
  for (size_t i = 1; i < kept.size(); ++i)
    keep[i] = !(kept.tID[i] == kept.tID[i - 1] && kept.tIdx[i] == kept.tIdx[i - 1] &&
                kept.qID[i] == kept.qID[i - 1] && kept.qIdx[i] == kept.qIdx[i - 1]);
  Hits out;
  for (size_t i = 0; i < kept.size(); ++i) {
    if (!keep[i]) continue;
    out.tID.push_back(kept.tID[i]); out.tIdx.push_back(kept.tIdx[i]);
    out.qID.push_back(kept.qID[i]); out.qIdx.push_back(kept.qIdx[i]);
    out.tHL.push_back(kept.tHL[i]); out.qHL.push_back(kept.qHL[i]);
    out.match.push_back(kept.match[i]); out.score.push_back(kept.score[i]);
    out.evalue.push_back(kept.evalue[i]);
  }
  return out;
}

// ---------------------------------------------------------------- driver

// CHostSetting.cu:46-58
int64_t calc_total_db_size(const std::vector<std::string>& files) {
  int64_t total = 0;
  for (const auto& f : files) {
    std::ifstream ifs(f.c_str(), std::ios::binary);
    std::string buf;
    while (ifs && std::getline(ifs, buf))
      if (buf.find(">") == std::string::npos) total += static_cast<int64_t>(buf.size());
  }
  return total;
}

// CFASTALoader.cpp:11-170
class FastaLoader {
 public:
  FastaLoader(const std::vector<std::string>& files, int avail, bool target)
      : files_(files), avail_(avail), target_(target) {}
  void load(std::vector<Fasta>& out) {
    for (; file_index_ < static_cast<int>(files_.size()); ++file_index_)
      if (!read(files_[file_index_], out)) break;
    if (file_index_ == static_cast<int>(files_.size())) file_index_ = -1;
  }
  int file_index() const { return file_index_; }

 private:
  // judgeContinueReading, CFASTALoader.cpp:43-87
  bool flush(int64_t pos, std::string& label, std::string& seq, int& start, std::vector<Fasta>& out) {
    if (!label.empty()) {
      const int len = static_cast<int>(seq.size());
      if (len > avail_) {
        std::cout << " warning : sequence " << label << " is longer than the "
                  << (target_ ? "target" : "query") << " RAM budget; skipped." << std::endl;
        label.erase(); seq.erase(); start = 0; read_pos_ = pos;
      } else if (sum_read_ + len < avail_) {
        out.push_back(Fasta{label, seq, start});
        sum_read_ += len;
        label.erase(); seq.erase(); start = 0; read_pos_ = pos;
      } else {
        sum_read_ = 0;
        return false;
      }
    }
    return true;
  }
  // CFASTALoader.cpp:89-106
  static void clean(std::string& s) {
    std::string r;
    r.reserve(s.size());
    for (char c : s) {
      if (c == ' ') continue;
      if (c > 'Z') c -= 32;
      if (c < 'A' || c > 'Z') c = 'N';
      r.push_back(c);
    }
    s.swap(r);
  }
  // CFASTALoader.cpp:112-170
  bool read(const std::string& path, std::vector<Fasta>& out) {
    std::ifstream ifs(path.c_str(), std::ios::binary);
    if (!ifs) {
      std::cout << "file \"" << path << "\" open failed." << std::endl;
      std::abort();
    }
    std::string buf, label, seq;
    int start = 0;
    ifs.seekg(read_pos_);
    int64_t temp = read_pos_;
    while (ifs && std::getline(ifs, buf)) {
      if (buf.find(">") != std::string::npos) {
        if (!flush(temp, label, seq, start, out)) return false;
        label = buf.substr(1);
      } else if (buf.find("@") != std::string::npos) {
        char* end = nullptr;
        const std::string num = buf.substr(1);
        const long v = std::strtol(num.c_str(), &end, 10);
        if (num.empty() || *end != '\0') {
          std::cout << " bad '@' line" << std::endl;
          std::abort();
        }
        start = static_cast<int>(v);
      } else if (buf.empty()) {
        if (!flush(temp, label, seq, start, out)) return false;
      } else if (!label.empty()) {
        clean(buf);
        seq += buf;
      }
      temp = ifs.tellg();
    }
    if (ifs.eof())
      if (!flush(temp, label, seq, start, out)) return false;
    read_pos_ = 0;
    return true;
  }

  std::vector<std::string> files_;
  int avail_;
  bool target_;
  int file_index_ = 0;
  int64_t read_pos_ = 0, sum_read_ = 0;
};

std::string format_row(const std::string& qlabel, int qIdx, int qHL, char strand,
                       const std::string& tlabel, int tIdx, int tHL, int match, int score,
                       double evalue) {
  std::ostringstream os;
  os << qlabel << "\t" << qIdx << "\t" << qHL << "\t" << strand << "\t" << tlabel << "\t" << tIdx
     << "\t" << tHL << "\t" << match << "(" << static_cast<double>(match * 100) / qHL << "%)"
     << "\t" << score << "\t" << evalue << "\n";
  return os.str();
}

// CHostResultHolder.{cuh,cu}
struct Holder {
  std::vector<std::string> qlabel;
  std::vector<char> qstrand;
  std::vector<std::string> tlabel;
  std::vector<int> tstart;
  Hits rows;

  explicit Holder(const std::vector<Fasta>& q) {
    for (const auto& f : q) {
      qlabel.push_back(f.label); qlabel.push_back(f.label);
      qstrand.push_back('+'); qstrand.push_back('-');
    }
  }
  void add(const Hits& h, const SeqList& t) {  // addResult + addLabel + addStartIdx
    rows.append(h);
    for (size_t i = 0; i < h.size(); ++i) {
      tlabel.push_back(t.label.at(h.tID[i]));
      tstart.push_back(t.start_idx.at(h.tID[i]));
    }
  }
  void fix() {  // fixResult :75-128 — stable (qID/2, score desc) over row indices
    std::vector<size_t> o(rows.size());
    std::iota(o.begin(), o.end(), 0);
    std::stable_sort(o.begin(), o.end(), [&](size_t a, size_t b) {
      const int qa = rows.qID[a] / 2, qb = rows.qID[b] / 2;
      if (qa != qb) return qa < qb;
      return rows.score[a] > rows.score[b];
    });
    Hits r;
    std::vector<std::string> tl;
    std::vector<int> ts;
    for (size_t i : o) {
      r.tID.push_back(rows.tID[i]); r.tIdx.push_back(rows.tIdx[i]);
      r.qID.push_back(rows.qID[i]); r.qIdx.push_back(rows.qIdx[i]);
      r.tHL.push_back(rows.tHL[i]); r.qHL.push_back(rows.qHL[i]);
      r.match.push_back(rows.match[i]); r.score.push_back(rows.score[i]);
      r.evalue.push_back(rows.evalue[i]);
      tl.push_back(tlabel[i]); ts.push_back(tstart[i]);
    }
    rows = std::move(r); tlabel.swap(tl); tstart.swap(ts);
  }
  void print(int num_out, const std::string& path) const {  // printResult :131-243
    std::ofstream ofs(path.c_str(), std::ios::app);
    // preQueryLabel starts empty (:148); the other "previous" fields are
    // uninitialised in the reference, so the first row always prints here.
    bool have_prev = false;
    std::string pql, ptl;
    int pqi = 0, pti = 0, pthl = 0, pqhl = 0, pm = 0, ps = 0;
    char pst = 0;
    double pe = 0;
    int count = 0;
    for (size_t i = 0; i < rows.size(); ++i) {
      const std::string& ql = qlabel[rows.qID[i]];
      const bool new_label = pql != ql;
      if (new_label) count = 0;
      if (num_out == -1 || new_label || count < num_out) {
        const std::string& tl = tlabel[i];
        const int qi = rows.qIdx[i];
        const char st = qstrand[rows.qID[i]];
        const int ti = rows.tIdx[i] + tstart[i];
        const int thl = rows.tHL[i], qhl = rows.qHL[i], m = rows.match[i], sc = rows.score[i];
        const double e = rows.evalue[i];
        if (!have_prev || pql != ql || pqi != qi || pst != st || ptl != tl || pti != ti ||
            pthl != thl || pqhl != qhl || pm != m || ps != sc || pe != e) {
          ofs << format_row(ql, qi, qhl, st, tl, ti, thl, m, sc, e);
          ++count;
          have_prev = true;
          pql = ql; pqi = qi; pst = st; ptl = tl; pti = ti; pthl = thl; pqhl = qhl; pm = m;
          ps = sc; pe = e;
        }
      }
    }
  }
};

// CHostSchedular.cu:12-130 with the per-query-window loop run in parallel and
// appended in window order.
static void schedule(const RunConfig& cfg, const SeqList& t, const SeqList& q, Holder& holder) {
  const Params& p = cfg.p;
  const double tv = cfg.t_vram * (1000 * 1000), qv = cfg.q_vram * (1000 * 1000);
  const int t_end_id = static_cast<int>(t.gateway.size()) - 1;
  for (int tb = 0, te = 1; tb < t_end_id; tb = te++) {
    const int tbi = t.gateway[tb];
    while (te < t_end_id && t.gateway[te + 1] - tbi < tv) ++te;
    HashTable h = build_hash_table(p, t, tb, te);
    const int q_end_id = (static_cast<int>(q.gateway.size()) - 1) / 2;
    std::vector<std::pair<int, int>> qwin;
    for (int qb = 0, qe = 1; qb < q_end_id; qb = qe++) {
      const int qbi = q.gateway[qb * 2];
      while (qe < q_end_id && q.gateway[qe * 2 + 2] - qbi < qv) ++qe;
      qwin.emplace_back(qb, qe);
    }
    std::vector<Hits> res(qwin.size());
#pragma omp parallel for schedule(dynamic, 1) if (cfg.threads != 1)
    for (size_t k = 0; k < qwin.size(); ++k) {
      QueryWindow qw = make_query_window(p, q, qwin[k].first, qwin[k].second);
      res[k] = search(p, h, qw, tb, qwin[k].first * 2);
    }
    for (auto& r : res) holder.add(r, t);
  }
}

// main.cu:16-112
int run_main(const RunConfig& cfg) {
#ifdef _OPENMP
  if (cfg.threads > 0) omp_set_num_threads(cfg.threads);
#endif
  { std::ofstream ofs(cfg.output.c_str()); }
  FastaLoader ql(cfg.queries, static_cast<int>(cfg.q_ram * (1000 * 1000) / 2), false);
  while (ql.file_index() != -1) {
    std::vector<Fasta> qf;
    ql.load(qf);
    Holder holder(qf);
    FastaLoader tl(cfg.targets, static_cast<int>(cfg.t_ram * (1000 * 1000)), true);
    while (tl.file_index() != -1) {
      std::vector<Fasta> tf;
      tl.load(tf);
      SeqList t, q;
      add_target(cfg.p, tf, t);
      add_query(cfg.p, qf, q);
      schedule(cfg, t, q, holder);
    }
    holder.fix();
    holder.print(cfg.num_out, cfg.output);
  }
  return 0;
}

}  // namespace oracle
