// Flags and settings (listenCommandLine.cpp, CHostSetting.cu).
#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "host.hpp"

namespace clast_host {

namespace {

[[noreturn]] void odd(const std::string& after) {
  if (after.empty())
    std::cout << "  error : Odd parameter was inputted. Check the command line parameter." << std::endl;
  else
    std::cout << "  error : Odd parameter was inputted. Check the after of \"" << after << "\"." << std::endl;
  std::abort();
}

// boost::lexical_cast semantics: the whole argument must be the number.
bool parse_int(const std::string& a, int& out) {
  if (a.empty() || std::isspace(static_cast<unsigned char>(a[0]))) return false;
  errno = 0;
  char* end = nullptr;
  const long v = std::strtol(a.c_str(), &end, 10);
  if (*end != '\0' || errno == ERANGE || v < INT32_MIN || v > INT32_MAX) return false;
  out = static_cast<int>(v);
  return true;
}
bool parse_double(const std::string& a, double& out) {
  if (a.empty() || std::isspace(static_cast<unsigned char>(a[0]))) return false;
  char* end = nullptr;
  const double v = std::strtod(a.c_str(), &end);
  if (*end != '\0') return false;
  out = v;
  return true;
}

enum class St { Start, Value, TFile, QFile, OFile };

}  // namespace

int64_t total_db_size(const std::vector<std::string>& files) {
  // calcTotalDbSize (CHostSetting.cu:46-58): every byte of every line that
  // holds no '>' counts, '@' lines and '\r' included.
  int64_t total = 0;
  for (const auto& f : files) {
    std::ifstream ifs(f.c_str(), std::ios::binary);
    std::string line;
    while (ifs && std::getline(ifs, line))
      if (line.find('>') == std::string::npos) total += static_cast<int64_t>(line.size());
  }
  return total;
}

clast_params Settings::params() const {
  clast_params p;
  p.lmer = lmer;
  p.stride = stride;
  p.cut_repeat = cut_repeat;
  p.width = width;
  p.gap = gap;
  p.local = local;
  p.cut_off = cut_off;
  p.total_db_size = total_db_size;
  p.K = K;
  p.lambda = lambda;
  return p;
}

Settings parse_settings(int argc, const char** argv, bool quiet) {
  Settings s;
  St st = St::Start;
  std::string flag;
  auto is_value_flag = [](const std::string& a) {
    return a == "-tRAM" || a == "-qRAM" || a == "-tVRAM" || a == "-qVRAM" || a == "-lMer" || a == "-stride" ||
           a == "-repeat" || a == "-width" || a == "-gap" || a == "-numOut" || a == "-cutOff" || a == "-local" ||
           a == "-device" || a == "-sleep" || a == "-batch";
  };
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    switch (st) {
      case St::Start:
        if (is_value_flag(a)) { flag = a; st = St::Value; }
        else if (a == "-t") st = St::TFile;
        else if (a == "-q") st = St::QFile;
        else if (a == "-o") st = St::OFile;
        else odd("");
        break;
      case St::TFile:
      case St::QFile: {
        const bool tf = st == St::TFile;
        if (is_value_flag(a)) { flag = a; st = St::Value; }
        else if (a == "-t") { if (tf) odd("-t"); st = St::TFile; }
        else if (a == "-q") { if (!tf) odd("-q"); st = St::QFile; }
        else if (a == "-o") st = St::OFile;
        else (tf ? s.targets : s.queries).push_back(a);
        break;
      }
      case St::OFile:
        s.output = a;
        st = St::Start;
        break;
      case St::Value: {
        bool ok;
        if (flag == "-tRAM") ok = parse_double(a, s.t_ram);
        else if (flag == "-qRAM") ok = parse_double(a, s.q_ram);
        else if (flag == "-tVRAM") ok = parse_double(a, s.t_vram);
        else if (flag == "-qVRAM") ok = parse_double(a, s.q_vram);
        else if (flag == "-cutOff") ok = parse_double(a, s.cut_off);
        else if (flag == "-lMer") ok = parse_int(a, s.lmer);
        else if (flag == "-stride") ok = parse_int(a, s.stride);
        else if (flag == "-repeat") ok = parse_int(a, s.cut_repeat);
        else if (flag == "-width") ok = parse_int(a, s.width);
        else if (flag == "-gap") ok = parse_int(a, s.gap);
        else if (flag == "-numOut") ok = parse_int(a, s.num_out);
        else if (flag == "-local") ok = parse_int(a, s.local);
        else if (flag == "-device") ok = parse_int(a, s.device);
        else if (flag == "-sleep") ok = parse_int(a, s.sleep);
        else ok = parse_int(a, s.batch);
        if (!ok) odd(flag);
        st = St::Start;
        break;
      }
    }
  }
  // validator (CHostSetting.cu:14-44)
  auto fail = [](const char* m) {
    std::cout << " error : " << m << std::endl;
    std::abort();
  };
  if (s.lmer > 31) fail("too large lMerLength was inputted.");
  if (s.gap > 16) fail("too large allowableGap was inputted.");
  else if (s.gap < 2) fail("allowableGap must be larger than 1.");
  if (s.targets.empty()) fail("no target file was inputted.");
  if (s.queries.empty()) fail("no query file was inputted.");
  if (s.output.empty()) fail("no output file was inputted.");
  if (!quiet) {
    std::cout << " Search setting :" << std::endl
              << "  RAM usage for target (MB): " << s.t_ram << std::endl
              << "  RAM usage for query  (MB): " << s.q_ram << std::endl
              << "  VRAM usage for target(MB): " << s.t_vram << std::endl
              << "  VRAM usage for query (MB): " << s.q_vram << std::endl
              << "  l-mer length             : " << s.lmer << std::endl
              << "  stride length            : " << s.stride << std::endl
              << "  cut repeat num           : " << s.cut_repeat << std::endl
              << "  allowable seed-seed width: " << s.width << std::endl
              << "  allowable gap            : " << s.gap << std::endl
              << "  number of output         : " << s.num_out << std::endl
              << "  cut off E-Value score    : " << s.cut_off << std::endl
              << "  alignment type           : " << (s.local ? "local" : "global") << std::endl
              << "  device ID                : " << s.device << std::endl
              << "  sleeping interval time   : " << s.sleep << std::endl
              << "  launch mode              : " << (s.batch ? "batched" : "per window") << std::endl
              << "  target file list         : " << std::endl;
    for (size_t i = 0; i < s.targets.size(); ++i) std::cout << "   No." << i + 1 << " " << s.targets[i] << std::endl;
    std::cout << "  query file list          : " << std::endl;
    for (size_t i = 0; i < s.queries.size(); ++i) std::cout << "   No." << i + 1 << " " << s.queries[i] << std::endl;
    std::cout << "  output file name         : " << s.output << std::endl;
  }
  s.t_ram *= 1000 * 1000;
  s.q_ram *= 1000 * 1000;
  s.t_vram *= 1000 * 1000;
  s.q_vram *= 1000 * 1000;
  s.total_db_size = static_cast<double>(total_db_size(s.targets));
  return s;
}

}  // namespace clast_host
