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

// CHostResultHolder ctor (:14-22): two rows ('+', '-') per query.
ResultHolder::ResultHolder(const std::vector<Fasta>& q) {
  for (const auto& f : q) {
    qlabel_.push_back(f.label);
    qlabel_.push_back(f.label);
    qstrand_.push_back('+');
    qstrand_.push_back('-');
  }
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
  for (size_t i = at; i < at + n; ++i) {
    tlabel_.push_back(t.label.at(col_[0][i]));
    tstart_.push_back(t.start_idx.at(col_[0][i]));
  }
}

void ResultHolder::add_rows(const int32_t* const cols[8], const double* evalue, int64_t n, const SeqList& t) {
  for (int64_t i = 0; i < n; ++i) {
    for (int c = 0; c < 8; ++c) col_[c].push_back(cols[c][i]);
    ev_.push_back(evalue[i]);
    tlabel_.push_back(t.label.at(cols[0][i]));
    tstart_.push_back(t.start_idx.at(cols[0][i]));
  }
}

// fixResult (:75-128): the device stable sort by (qID/2, score desc); as in
// the reference, the tID column carries the row index so that the target
// labels and offsets follow their rows.
void ResultHolder::fix() {
  const int64_t n = size();
  if (n <= 1) return;
  std::vector<int32_t> row(n);
  for (int64_t i = 0; i < n; ++i) row[i] = static_cast<int32_t>(i);
  check(clast_result_sort(n, row.data(), col_[1].data(), col_[2].data(), col_[3].data(), col_[4].data(),
                          col_[5].data(), col_[6].data(), col_[7].data(), ev_.data()),
        "clast_result_sort");
  std::vector<std::string> tl(n);
  std::vector<int32_t> ts(n), tid(n);
  for (int64_t i = 0; i < n; ++i) {
    tl[i] = std::move(tlabel_[row[i]]);
    ts[i] = tstart_[row[i]];
    tid[i] = col_[0][row[i]];
  }
  tlabel_.swap(tl);
  tstart_.swap(ts);
  col_[0].swap(tid);
}

// printResult (:131-243): numOut rows per query label (the first row of a
// new label always counts); a row equal in all ten printed fields to the
// previous printed row is skipped.  Default ostream formatting throughout.
int64_t ResultHolder::print(int num_out, const std::string& path) const {
  std::ofstream ofs(path.c_str(), std::ios::app);
  std::ostringstream os;
  bool have_prev = false;
  std::string pql, ptl;
  int pqi = 0, pti = 0, pthl = 0, pqhl = 0, pm = 0, ps = 0;
  char pst = 0;
  double pe = 0;
  int count = 0;
  int64_t printed = 0;
  for (int64_t i = 0; i < size(); ++i) {
    const std::string& ql = qlabel_[col_[2][i]];
    const bool new_label = pql != ql;
    if (new_label) count = 0;
    if (!(num_out == -1 || new_label || count < num_out)) continue;
    const std::string& tl = tlabel_[i];
    const int qi = col_[3][i];
    const char st = qstrand_[col_[2][i]];
    const int ti = col_[1][i] + tstart_[i];
    const int thl = col_[4][i], qhl = col_[5][i], m = col_[6][i], sc = col_[7][i];
    const double e = ev_[i];
    if (have_prev && pql == ql && pqi == qi && pst == st && ptl == tl && pti == ti && pthl == thl && pqhl == qhl &&
        pm == m && ps == sc && pe == e)
      continue;
    os << ql << "\t" << qi << "\t" << qhl << "\t" << st << "\t" << tl << "\t" << ti << "\t" << thl << "\t" << m << "("
       << static_cast<double>(m * 100) / qhl << "%)"
       << "\t" << sc << "\t" << e << "\n";
    ++count;
    ++printed;
    have_prev = true;
    pql = ql; pqi = qi; pst = st; ptl = tl; pti = ti; pthl = thl; pqhl = qhl; pm = m; ps = sc; pe = e;
  }
  ofs << os.str();
  std::cout << " " << printed << " hits has printed." << std::endl;
  return printed;
}

}  // namespace clast_host
