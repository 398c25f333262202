// Self-test of the host driver's CPU side (flags, FASTA chunk loading, host
// lists, windows, the TSV writer), for builds with the address and
// undefined-behaviour sanitizers.  No device call is made.
//   host_selftest <scratch dir>
// Prints one summary line; exits non-zero on a failed check.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "host.hpp"

using namespace clast_host;

static void write(const std::string& path, const std::string& text) {
  std::ofstream f(path.c_str(), std::ios::binary);
  f << text;
}

static int fail(const char* what) {
  std::cerr << "selftest failed: " << what << std::endl;
  return 1;
}

int main(int argc, char** argv) {
  if (argc != 2) return fail("usage: host_selftest <scratch dir>");
  const std::string dir = argv[1];
  // targets: lower case, non-letters, CRLF, an '@' offset, an empty line,
  // a record longer than the budget, one exactly the budget
  std::string t = ">chrA first\nACGTacgtNNxx\r\nGGGGCCCCAAAATTTT\n>chrB\n@120\nTTTTGGGGCCCCAAAAACGT\n\n";
  t += ">long\n" + std::string(400, 'A') + "\n>exact\n" + std::string(200, 'C') + "\n>chrC\nACGTACGTACGT\n";
  write(dir + "/t.fa", t);
  std::string q;
  for (int i = 0; i < 30; ++i) q += ">r" + std::to_string(i) + "\n" + std::string(20 + i, "ACGT"[i % 4]) + "\n";
  write(dir + "/q.fa", q);
  const std::string out = dir + "/o.tsv", tpath = dir + "/t.fa", qpath = dir + "/q.fa";
  const char* args[] = {"clast", "-t", tpath.c_str(), "-q", qpath.c_str(), "-o", out.c_str(),
                        "-lMer", "11", "-stride", "3", "-tRAM", "0.0002", "-qRAM", "0.0006", "-tVRAM", "0.00005",
                        "-qVRAM", "0.0001", "-numOut", "2"};
  Settings s = parse_settings(static_cast<int>(sizeof args / sizeof args[0]), args, true);
  if (s.lmer != 11 || s.stride != 3 || s.num_out != 2) return fail("flags");
  if (total_db_size(s.targets) <= 0) return fail("db size");
  int64_t rows = 0, loads = 0, windows = 0;
  FastaLoader ql(s.queries, static_cast<int>(s.q_ram * 1000 * 1000 / 2), false);
  while (ql.file_index() != -1) {
    std::vector<Fasta> qf;
    ql.load(qf);
    ++loads;
    ResultHolder holder(qf);
    FastaLoader tl(s.targets, static_cast<int>(s.t_ram * 1000 * 1000), true);
    while (tl.file_index() != -1) {
      std::vector<Fasta> tf;
      tl.load(tf);
      ++loads;
      SeqList ts, qs;
      ts.lmer = qs.lmer = s.lmer;
      ts.add_targets(tf);
      qs.add_queries(qf);
      const auto tw = target_windows(ts.gateway, s.t_vram);
      const auto qw = query_windows(qs.gateway, s.q_vram);
      windows += static_cast<int64_t>(tw.size() + qw.size());
      for (const auto& w : tw) {
        std::vector<int32_t> gw;
        const clast_window cw = ts.window(w.first, w.second, gw);
        if (cw.n_seq != w.second - w.first) return fail("window");
      }
      if (ts.n_seq() == 0 || qs.n_seq() == 0) continue;
      // synthetic rows: every query strand against target 0, two rows per
      // strand, the second a repeat of the first (skipped when printed)
      const int n = qs.n_seq() * 2;
      std::vector<int32_t> c[8];
      std::vector<double> ev;
      for (int k = 0; k < qs.n_seq(); ++k)
        for (int r = 0; r < 2; ++r) {
          const int32_t v[8] = {0, 1, k, 0, 10, 10, 9, 5};
          for (int j = 0; j < 8; ++j) c[j].push_back(v[j]);
          ev.push_back(1e-3);
        }
      const int32_t* cols[8];
      for (int j = 0; j < 8; ++j) cols[j] = c[j].data();
      holder.add_rows(cols, ev.data(), n, ts);
    }
    // every row in append order through the text writer (the order and
    // selection of print() come from the device, not run here)
    std::vector<int32_t> order(holder.size());
    std::vector<uint8_t> all(holder.size(), 1);
    for (int64_t i = 0; i < holder.size(); ++i) order[i] = static_cast<int32_t>(i);
    std::ofstream ofs(out.c_str(), std::ios::app);
    holder.write_rows(ofs, order.data(), all.data(), holder.size());
    rows += holder.size();
  }
  std::cout << "selftest ok: " << loads << " loads, " << windows << " windows, " << rows << " rows printed"
            << std::endl;
  return 0;
}
