// devideQuery: splits a query FASTA file into N parts of near-equal total
// length for query sharding over N GPUs (one `clast -device k` per part):
// each record goes, in file order, to the part with the smallest total so
// far (the lowest-numbered one on ties), appended to "<file>.<k>" as its
// label line, its sequence on one line and an empty line
// (accessory/devideQuery/devideQuery.cpp:1-101).
//
// Records are read as by the reference tool: a '>' line starts one, an empty
// line or the next '>' line ends it, sequence lines are joined with their
// spaces removed.  Same arguments and error behaviour: anything but two
// arguments prints the usage and aborts.

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {

int int_arg(const char* s) {
  std::size_t used = 0;
  int v = 0;
  try {
    v = std::stoi(s, &used);
  } catch (...) {
    used = 0;
  }
  if (used == 0 || s[used] != '\0' || v < 1) {
    std::cout << " the number of parts (arg1) must be a positive integer: '" << s << "'" << std::endl;
    std::abort();
  }
  return v;
}

// devideQuery.cpp:11-33
void write_record(const std::string& path, const std::string& label, const std::string& seq,
                  std::vector<long>& total) {
  std::size_t k = 0;
  for (std::size_t i = 1; i < total.size(); ++i)
    if (total[i] < total[k]) k = i;
  std::ofstream out((path + "." + std::to_string(k)).c_str(), std::ios::app);
  out << label << '\n' << seq << "\n\n";
  total[k] += static_cast<long>(seq.size());
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cout << "usage: devideQuery <number of parts> <query.fa>" << std::endl
              << "  writes <query.fa>.0 ... <query.fa>.<number of parts - 1> (appended to)" << std::endl;
    std::abort();
  }
  std::vector<long> total(static_cast<std::size_t>(int_arg(argv[1])), 0);
  const std::string path = argv[2];
  // devideQuery.cpp:48-82
  std::ifstream in(path.c_str());
  std::string line, label, seq;
  auto flush = [&] {
    if (!label.empty()) {
      write_record(path, label, seq, total);
      label.clear();
      seq.clear();
    }
  };
  while (in && std::getline(in, line)) {
    if (line.find('>') != std::string::npos) {
      flush();
      label = line;
    } else if (line.empty()) {
      flush();
    } else if (!label.empty()) {
      for (char c : line)
        if (c != ' ') seq.push_back(c);
    }
  }
  flush();
  return 0;
}
