// CPU oracle for the CLAST search pipeline — TEST INFRASTRUCTURE ONLY.
//
// A host C++ restatement of the reference's search path, written from the
// reference source read as text.  Each function cites the reference file:line
// it follows.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only to check results against: the
// product (libclast_hip + the clast driver) never links or calls it.
//
// Conventions for behaviour the reference leaves undefined (parity unpinned,
// fixed by this oracle's documented choice):
//   * bytes read past the end of a window's base array (last strided target
//     k-mers, CDeviceHashTable.cu:28-31) count as invalid bases -> key -1;
//   * joint bytes copied past the end of a sequence shorter than lMer-1
//     (utilAddSequence.cu:42-46) are 'N';
//   * DP target reads outside the window's base array are 'X'
//     (krnlAlignment.cu:121-127, 234-240 only guard one side each);
//   * host/device remove_if with a stencil aliasing its output compares each
//     element with its ORIGINAL predecessor (CDeviceHitList.cu:53-105);
//   * every comparator sort is stable (Thrust's sort -> stable_sort).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace oracle {

// common.hpp:12-43
constexpr int kMaxLMer = 31;
constexpr int kMaxGap = 16;
constexpr int kMargin = 1;
constexpr char kBadAreaChar = 'X';
constexpr int kBadAreaPoint = -1000;
constexpr int kMatchPoint = 1;
constexpr int kMismatchPoint = -3;
constexpr int kGapOpenPoint = -5;
constexpr int kGapPoint = -2;
constexpr double kGlobalK = 0.71;
constexpr double kGlobalLambda = 1.37;

struct Params {
  int lmer = 15, stride = 5, cut_repeat = 20, width = 100, gap = 8, local = 0;
  double cut_off = 10.0, total_db_size = 0.0, K = kGlobalK, lambda = kGlobalLambda;
};

struct Fasta {  // CHostFASTA.hpp
  std::string label, seq;
  int start_idx = 0;
};

// CHostSeqList.cuh:13-30 (+ CHostSeqList_target.cuh startIdxArray)
struct SeqList {
  std::vector<std::string> label;
  std::vector<int> index, id;
  std::vector<char> base;
  std::vector<int> length, gateway;
  std::vector<int> start_idx;  // target only
};

void add_sequence(int lmer, const std::string& seq, SeqList& s);      // utilAddSequence.cu:6-47
void add_target(const Params& p, const std::vector<Fasta>& f, SeqList& s);  // CHostSeqList_target.cu:8-33
void add_query(const Params& p, const std::vector<Fasta>& f, SeqList& s);   // CHostSeqList_query.cu:11-54
std::string comp_seq(const std::string& seq);                        // utilReverseSeq.cpp:3-19

// CDeviceSeqList.cuh:8-22 — one window of a host list
struct Window {
  std::vector<int> index, id;
  std::vector<char> base;
  std::vector<int> length, gateway;
};
Window make_window(const SeqList& s, int start_id, int end_id);  // CHostSeqList.cu:15-70

int64_t kmer_code(const char* bases, int64_t avail, int lmer);  // SRead2bit.cuh + SRepelHasOddBase.cuh

// CDeviceHashTable.cuh:10-27
struct HashTable {
  Window target;
  std::vector<int64_t> gateway_key;
  std::vector<int> cell_size, gateway_index, index_array;
};
HashTable build_hash_table(const Params& p, const SeqList& t, int t_begin, int t_end);

// CDeviceSeqList_query.cuh:11-20
struct QueryWindow {
  Window w;
  std::vector<int64_t> hash_index;
};
QueryWindow make_query_window(const Params& p, const SeqList& q, int pair_begin, int pair_end);

struct Seeds {
  std::vector<int> tID, tIdx, qID, qIdx;
  size_t size() const { return tID.size(); }
};

struct Hits {
  std::vector<int> tID, tIdx, qID, qIdx, tHL, qHL, match, score;
  std::vector<double> evalue;
  size_t size() const { return tID.size(); }
  void append(const Hits& o);
};

// Intermediate stages of one CDeviceHitList run, kept for stage-level parity.
struct StageDump {
  Seeds raw, boundary, sorted, isolate, duplicate;
  Hits aligned;   // after alignmentHits (in the reference's post-forward-sort order)
  Hits filtered;  // after E-value and length filters
};

// CDeviceHitList.cu:228-423: the whole per-(target window, query window) chain.
Hits search(const Params& p, const HashTable& h, const QueryWindow& q, int t_begin, int q_begin,
            StageDump* dump = nullptr);

// krnlAlignment.cu: one hit, both directions (exposed for known-answer tests).
// Inputs are the hit's seed coordinates; outputs are the extended hit.
struct DPResult {
  int tIdx, qIdx, tHL, qHL, match, score;
};
DPResult align_one(const Params& p, const Window& t, const Window& q, int t_begin, int q_begin,
                   int tID, int tIdx, int qID, int qIdx);

// utilResultSorting.cu:6-47 (stable by (qID/2 asc, score desc))
void result_sort(Hits& h);

// The main.cu:16-112 loop: reads FASTA files, chunks, windows, writes TSV.
struct RunConfig {
  Params p;
  double t_ram = 64, q_ram = 64, t_vram = 64, q_vram = 2;  // MBase
  int num_out = -1;
  std::vector<std::string> targets, queries;
  std::string output;
  int threads = 0;  // OpenMP threads for the per-query-window loop (0 = default)
};
int64_t calc_total_db_size(const std::vector<std::string>& files);  // CHostSetting.cu:46-58
int run_main(const RunConfig& cfg);

// Format one TSV row exactly as CHostResultHolder.cu:199-218.
std::string format_row(const std::string& qlabel, int qIdx, int qHL, char strand,
                       const std::string& tlabel, int tIdx, int tHL, int match, int score,
                       double evalue);

}  // namespace oracle
