// clast: the search driver (main.cu:16-112) over libclast_hip.
#include <chrono>
#include <ctime>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <vector>

#include "host.hpp"

using namespace clast_host;

int main(int argc, const char** argv) {
  std::cout << "...Loading search setting." << std::endl;
  const Settings s = parse_settings(argc, argv);
  { std::ofstream ofs(s.output.c_str()); }  // clear the result file
  try {
    check(clast_init(s.device), "clast_init");
    const auto t0 = std::chrono::steady_clock::now();
    std::cout << " << Start searching >>" << std::endl << std::endl;
    FastaLoader qloader(s.queries, static_cast<int>(s.q_ram / 2), false);
    while (qloader.file_index() != -1) {
      std::vector<Fasta> qf;
      qloader.load(qf);
      ResultHolder holder(qf);
      // batched: the query set once per query chunk, built on the device
      std::unique_ptr<clast_query_s, int (*)(clast_query)> qdev(s.batch ? query_chunk_reads(s, qf) : nullptr,
                                                                 clast_query_destroy);
      FastaLoader tloader(s.targets, static_cast<int>(s.t_ram), true);
      while (tloader.file_index() != -1) {
        std::vector<Fasta> tf;
        tloader.load(tf);
        SeqList t;
        t.lmer = s.lmer;
        t.add_targets(tf);
        if (s.batch) {
          search_batched(s, t, qdev.get(), holder);
        } else {
          SeqList q;
          q.lmer = s.lmer;
          q.add_queries(qf);
          search(s, t, q, holder);
        }
      }
      std::cout << " Fixing and printing the results." << std::endl;
      holder.print(s.num_out, s.output);  // fixResult + printResult
      std::cout << " ...finished." << std::endl;
      if (s.sleep > 0) {  // main.cu:94-103 cool-down, as a plain wait
        const std::clock_t c0 = std::clock();
        while (1000.0 * (std::clock() - c0) / CLOCKS_PER_SEC <= s.sleep * 1000.0) {
        }
      }
    }
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << std::endl << " << Searching end >>" << std::endl << "Total: " << sec << "sec." << std::endl;
  } catch (const std::exception& e) {
    std::cout << " error : " << e.what() << std::endl;
    return 2;
  }
  return 0;
}
