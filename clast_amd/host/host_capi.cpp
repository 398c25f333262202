// C entry points over the host driver code, for the Python package (list
// building, windows and chunking of in-memory sequences, flag parsing).
#include <cstring>
#include <string>
#include <vector>

#include "host.hpp"

using namespace clast_host;

extern "C" {

// Slots of a host list built from n sequences (query: '+' and '-' each).
int64_t clast_host_list_slots(const int64_t* off, int32_t n, int32_t lmer, int32_t query) {
  int64_t total = 0;
  for (int32_t i = 0; i < n; ++i) total += (off[i + 1] - off[i]) + (lmer - 1);
  return query ? 2 * total : total;
}

// CHostSeqList layout of n sequences (seqs + off, n+1 offsets); out_len and
// out_gateway have n (2n for queries) and n+1 (2n+1) entries.
int32_t clast_host_build_list(const char* seqs, const int64_t* off, int32_t n, int32_t lmer, int32_t query,
                              char* out_bases, int32_t* out_len, int32_t* out_gateway) {
  SeqList s;
  s.lmer = lmer;
  for (int32_t i = 0; i < n; ++i) {
    const int len = static_cast<int>(off[i + 1] - off[i]);
    s.add_sequence(seqs + off[i], len, false);
    if (query) s.add_sequence(seqs + off[i], len, true);
  }
  std::memcpy(out_bases, s.base.data(), s.base.size());
  std::memcpy(out_len, s.length.data(), s.length.size() * sizeof(int32_t));
  std::memcpy(out_gateway, s.gateway.data(), s.gateway.size() * sizeof(int32_t));
  return 0;
}

// Windows as [begin, end) pairs into out (2 * n_seq ints at most); returns
// the window count.  query != 0: pairs of query strands.
int32_t clast_host_windows(const int32_t* gateway, int32_t n_gate, double vram_slots, int32_t query,
                           int32_t* out) {
  std::vector<int32_t> g(gateway, gateway + n_gate);
  const auto w = query ? query_windows(g, vram_slots) : target_windows(g, vram_slots);
  for (size_t k = 0; k < w.size(); ++k) {
    out[2 * k] = w[k].first;
    out[2 * k + 1] = w[k].second;
  }
  return static_cast<int32_t>(w.size());
}

// The FASTA loader's chunking applied to in-memory records of the given
// lengths (one file): writes [begin, end) record ranges of each chunk and
// a keep flag per record (0 = skipped as longer than the budget); returns
// the chunk count.
int32_t clast_host_chunks(const int64_t* lengths, int32_t n, int64_t avail, int32_t* out_bounds, int8_t* keep) {
  int32_t chunks = 0, begin = 0;
  int64_t sum = 0;
  int32_t i = 0;
  while (i < n) {
    const int64_t len = lengths[i];
    if (len > avail || (sum == 0 && len >= avail)) {
      keep[i] = 0;
      ++i;
      continue;
    }
    if (sum + len < avail) {
      keep[i] = 1;
      sum += len;
      ++i;
      continue;
    }
    out_bounds[2 * chunks] = begin;
    out_bounds[2 * chunks + 1] = i;
    ++chunks;
    begin = i;
    sum = 0;
  }
  out_bounds[2 * chunks] = begin;
  out_bounds[2 * chunks + 1] = n;
  return chunks + 1;
}

// The FASTA loader over one file with a RAM budget of `avail` bases, load
// after load (CFASTALoader::loadFASTA, main.cu:40-67): the number of
// records of each load into counts (at most max_chunks loads); returns the
// number of loads.
int32_t clast_host_fasta_chunks(const char* path, int64_t avail, int32_t target, int32_t* counts,
                                int32_t max_chunks) {
  FastaLoader ld({std::string(path)}, static_cast<int>(avail), target != 0);
  int32_t n = 0;
  while (ld.file_index() != -1 && n < max_chunks) {
    std::vector<Fasta> out;
    ld.load(out);
    counts[n++] = static_cast<int32_t>(out.size());
  }
  return n;
}

}  // extern "C"
