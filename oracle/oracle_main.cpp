// CLI over the CPU oracle (TEST INFRASTRUCTURE ONLY): the reference's flags
// (listenCommandLine.cpp:7-37, 281-325) with its defaults (common.hpp:12-25),
// plus -threads N for the OpenMP loop over query windows.
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "clast_oracle.hpp"

int main(int argc, char** argv) {
  oracle::RunConfig c;
  enum { START, VAL, TF, QF, OF } st = START;
  std::string flag;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    const bool is_flag = a == "-tRAM" || a == "-qRAM" || a == "-tVRAM" || a == "-qVRAM" || a == "-lMer" ||
                         a == "-stride" || a == "-repeat" || a == "-width" || a == "-gap" || a == "-numOut" ||
                         a == "-cutOff" || a == "-local" || a == "-device" || a == "-sleep" || a == "-threads";
    if (st == START || st == TF || st == QF) {
      if (is_flag) { flag = a; st = VAL; continue; }
      if (a == "-t" && st != TF) { st = TF; continue; }
      if (a == "-q" && st != QF) { st = QF; continue; }
      if (a == "-o") { st = OF; continue; }
      if (st == TF) { c.targets.push_back(a); continue; }
      if (st == QF) { c.queries.push_back(a); continue; }
      std::cout << "  error : Odd parameter was inputted." << std::endl;
      std::abort();
    }
    if (st == OF) { c.output = a; st = START; continue; }
    const double v = std::atof(a.c_str());
    if (flag == "-tRAM") c.t_ram = v;
    else if (flag == "-qRAM") c.q_ram = v;
    else if (flag == "-tVRAM") c.t_vram = v;
    else if (flag == "-qVRAM") c.q_vram = v;
    else if (flag == "-lMer") c.p.lmer = static_cast<int>(v);
    else if (flag == "-stride") c.p.stride = static_cast<int>(v);
    else if (flag == "-repeat") c.p.cut_repeat = static_cast<int>(v);
    else if (flag == "-width") c.p.width = static_cast<int>(v);
    else if (flag == "-gap") c.p.gap = static_cast<int>(v);
    else if (flag == "-numOut") c.num_out = static_cast<int>(v);
    else if (flag == "-cutOff") c.p.cut_off = v;
    else if (flag == "-local") c.p.local = static_cast<int>(v);
    else if (flag == "-threads") c.threads = static_cast<int>(v);
    st = START;
  }
  if (c.p.lmer > oracle::kMaxLMer || c.p.gap > oracle::kMaxGap || c.p.gap < 2 || c.targets.empty() ||
      c.queries.empty() || c.output.empty()) {
    std::cout << " error : bad settings" << std::endl;
    std::abort();
  }
  c.p.total_db_size = static_cast<double>(oracle::calc_total_db_size(c.targets));
  return oracle::run_main(c);
}
