// Host driver of the CLAST search path: settings and flags, FASTA chunk
// loading, host sequence lists, target/query windows, the result holder and
// the tab-separated writer.  Device work goes through libclast_hip's C ABI
// (include/clast_hip.h) only.
#pragma once
#include <cstdint>
#include <ostream>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "clast_hip.h"

namespace clast_host {

// CHostSetting (CHostSetting.cuh:8-49) with the defaults of common.hpp:12-25.
struct Settings {
  double t_ram = 64, q_ram = 64, t_vram = 64, q_vram = 2;  // MBase on the command line
  int lmer = 15, stride = 5, cut_repeat = 20, width = 100, gap = 8;
  int num_out = -1, local = 0, device = 0, sleep = 0;
  double cut_off = 10;
  std::vector<std::string> targets, queries;
  std::string output;
  int batch = 0;  // extension: 1 = one launch chain per (query chunk, target chunk)
  // derived (CHostSetting.cu:132-141)
  double total_db_size = 0, K = 0.71, lambda = 1.37;
  clast_params params() const;
};

// listenCommandLine (listenCommandLine.cpp:281-325) + validator
// (CHostSetting.cu:14-44) + the settings printout and unit scaling
// (:108-141).  Bad input prints the reference's message and aborts.
Settings parse_settings(int argc, const char** argv, bool quiet = false);
int64_t total_db_size(const std::vector<std::string>& files);

// CHostFASTA
struct Fasta {
  std::string label, seq;
  int start_idx = 0;
};

// CFASTALoader (CFASTALoader.cpp:11-170): reads records until the RAM
// budget (bases) would be reached, resuming where the previous chunk ended.
class FastaLoader {
 public:
  FastaLoader(std::vector<std::string> files, int avail_bases, bool target);
  void load(std::vector<Fasta>& out);
  int file_index() const { return file_index_; }

 private:
  bool read_file(const std::string& path, std::vector<Fasta>& out);
  bool flush(int64_t pos, std::string& label, std::string& seq, int& start, std::vector<Fasta>& out);
  std::vector<std::string> files_;
  int avail_;
  bool target_;
  int file_index_ = 0;
  int64_t read_pos_ = 0, sum_read_ = 0;
};

// CHostSeqList (+ _query / _target): the slot array (each sequence followed
// by a copy of its first lMer-1 bases), lengths and gateway prefix sums.
struct SeqList {
  int lmer = 15;
  std::vector<std::string> label;
  std::vector<int32_t> start_idx;  // targets: '@' offsets
  std::vector<char> base;
  std::vector<int32_t> length, gateway;
  void add_sequence(const char* s, int len, bool reverse_complement);
  void add_targets(const std::vector<Fasta>& f);
  void add_queries(const std::vector<Fasta>& f);  // '+' then '-' per query
  // CHostSeqList::get*Array(a, b) as a clast_window (gateway rebased into gw)
  clast_window window(int a, int b, std::vector<int32_t>& gw) const;
  int n_seq() const { return static_cast<int>(length.size()); }
};

// moveTargetWindow / moveQueryWindow (CHostSchedular.cu:12-35): windows as
// [begin, end) of sequence IDs (targets) or query pairs (queries).
std::vector<std::pair<int, int>> target_windows(const std::vector<int32_t>& gateway, double vram_slots);
std::vector<std::pair<int, int>> query_windows(const std::vector<int32_t>& gateway, double vram_slots);

// CHostResultHolder (CHostResultHolder.cuh:8-44).  Labels are kept as
// integer ids (equal strings, equal ids) so that fixResult's order and
// printResult's row selection run on the device (clast_result_finalize);
// only the text of the printed rows is written on the host.
class ResultHolder {
 public:
  explicit ResultHolder(const std::vector<Fasta>& queries);
  void add(clast_hits h, const SeqList& targets);  // addResult + addLabel + addStartIdx
  // the same from host arrays (n rows; cols: tID tIdx qID qIdx tHL qHL match score)
  void add_rows(const int32_t* const cols[8], const double* evalue, int64_t n, const SeqList& targets);
  // fixResult + printResult: order and selection on the device, the printed
  // rows appended to `path`; returns rows printed
  int64_t print(int num_out, const std::string& path) const;
  // the text of rows order[k] with printed[k] set, in that order
  void write_rows(std::ostream& os, const int32_t* order, const uint8_t* printed, int64_t n) const;
  int64_t size() const { return static_cast<int64_t>(col_[0].size()); }

 private:
  int32_t target_label_id(const std::string& label);
  std::vector<std::string> qlabel_;   // per query strand (qID)
  std::vector<int32_t> qlabel_id_;    // per query strand; 0: the empty label
  std::vector<char> qstrand_;
  std::vector<std::string> tname_;    // target labels by id
  std::unordered_map<std::string, int32_t> tname_id_;
  std::vector<int32_t> tlabel_;       // per row: target label id
  std::vector<int32_t> tstart_;       // per row: target start offset
  std::vector<int32_t> col_[8];       // tID tIdx qID qIdx tHL qHL match score
  std::vector<double> ev_;
};

// CHostSchedular::search (CHostSchedular.cu:43-130): every target window x
// query window, one clast_search each.
void search(const Settings& s, const SeqList& targets, const SeqList& queries, ResultHolder& holder);

// Batched path (-batch 1): the query set of a query chunk is built once, on
// the device from the reads as given (clast_query_create_reads: both strands
// and joints, windows moveQueryWindow computes from the read lengths), and
// searched against every target chunk with one launch chain each; the same
// rows as search() over the host-built query list.
clast_query query_chunk_reads(const Settings& s, const std::vector<Fasta>& reads);  // null: no windows
void search_batched(const Settings& s, const SeqList& targets, clast_query queries, ResultHolder& holder);

// Throws std::runtime_error with clast_last_error() on a failed call.
void check(int rc, const char* what);

}  // namespace clast_host
