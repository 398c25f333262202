// Result holder and TSV writer (CHostResultHolder.cu).
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "host.hpp"

namespace clast_host {

void check(int rc, const char* what) {
  if (rc != CLAST_OK) throw std::runtime_error(std::string(what) + ": " + clast_last_error());
}

// CHostResultHolder ctor (:14-22): two rows ('+', '-') per query; query
// label ids with 0 for the empty label (printResult's walk starts from it).
ResultHolder::ResultHolder(const std::vector<Fasta>& q) {
  std::unordered_map<std::string, int32_t> ids{{std::string(), 0}};
  for (const auto& f : q) {
    const auto it = ids.emplace(f.label, static_cast<int32_t>(ids.size())).first;
    for (int s = 0; s < 2; ++s) {
      qlabel_.push_back(f.label);
      qlabel_id_.push_back(it->second);
    }
    qstrand_.push_back('+');
    qstrand_.push_back('-');
  }
}

int32_t ResultHolder::target_label_id(const std::string& label) {
  const auto it = tname_id_.emplace(label, static_cast<int32_t>(tname_.size()));
  if (it.second) tname_.push_back(label);
  return it.first->second;
}

// addResult + addLabel + addStartIdx (:25-72)
void ResultHolder::add(clast_hits h, const SeqList& t) {
  const int64_t n = clast_hits_count(h);
  if (n <= 0) return;
  const size_t at = col_[0].size();
  for (auto& c : col_) c.resize(at + n);
  ev_.resize(at + n);
  check(clast_hits_copy(h, col_[0].data() + at, col_[1].data() + at, col_[2].data() + at, col_[3].data() + at,
                        col_[4].data() + at, col_[5].data() + at, col_[6].data() + at, col_[7].data() + at,
                        ev_.data() + at),
        "clast_hits_copy");
  std::vector<int32_t> ids(t.label.size(), -1);  // label id per target of this chunk, on first use
  for (size_t i = at; i < at + n; ++i) {
    const int32_t tid = col_[0][i];
    if (ids.at(tid) < 0) ids[tid] = target_label_id(t.label[tid]);
    tlabel_.push_back(ids[tid]);
    tstart_.push_back(t.start_idx.at(tid));
  }
}

void ResultHolder::add_rows(const int32_t* const cols[8], const double* evalue, int64_t n, const SeqList& t) {
  for (int64_t i = 0; i < n; ++i) {
    for (int c = 0; c < 8; ++c) col_[c].push_back(cols[c][i]);
    ev_.push_back(evalue[i]);
    tlabel_.push_back(target_label_id(t.label.at(cols[0][i])));
    tstart_.push_back(t.start_idx.at(cols[0][i]));
  }
}

// fixResult (:75-128) and printResult (:131-243): the order and the row
// selection come from the device; default ostream formatting throughout.
int64_t ResultHolder::print(int num_out, const std::string& path) const {
  const int64_t n = size();
  std::vector<int32_t> order(n), qlab(n);
  std::vector<uint8_t> printed(n);
  for (int64_t i = 0; i < n; ++i) qlab[i] = qlabel_id_.at(col_[2][i]);
  int64_t np = 0;
  if (n)
    check(clast_result_finalize(n, col_[1].data(), col_[2].data(), col_[3].data(), col_[4].data(), col_[5].data(),
                                col_[6].data(), col_[7].data(), ev_.data(), qlab.data(), tlabel_.data(),
                                tstart_.data(), num_out, order.data(), printed.data(), &np),
          "clast_result_finalize");
  std::ofstream ofs(path.c_str(), std::ios::app);
  std::ostringstream os;
  write_rows(os, order.data(), printed.data(), n);
  ofs << os.str();
  std::cout << " " << np << " hits has printed." << std::endl;
  return np;
}

void ResultHolder::write_rows(std::ostream& os, const int32_t* order, const uint8_t* printed, int64_t n) const {
  for (int64_t k = 0; k < n; ++k) {
    if (!printed[k]) continue;
    const int64_t i = order[k];
    const int32_t qid = col_[2][i];
    const int qhl = col_[5][i], m = col_[6][i];
    os << qlabel_[qid] << "\t" << col_[3][i] << "\t" << qhl << "\t" << qstrand_[qid] << "\t" << tname_[tlabel_[i]]
       << "\t" << col_[1][i] + tstart_[i] << "\t" << col_[4][i] << "\t" << m << "("
       << static_cast<double>(m * 100) / qhl << "%)"
       << "\t" << col_[7][i] << "\t" << ev_[i] << "\n";
  }
}

}  // namespace clast_host
