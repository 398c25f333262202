// preprocessDB: splits every record of a target FASTA file into parts of at
// most maxLength bases that overlap by overlapLength, each written as
//   label
//   @start
//   60-column sequence lines
//   (empty line)
// so that a target sequence longer than one target chunk can be searched in
// parts; the FASTA loader reads "@start" as the part's offset and the
// result rows report target positions in the whole sequence
// (accessory/preprocessDB/preprocessDB.cpp:1-121; CFASTALoader.cpp:126-165
// reads the offset back).
//
// Same arguments, output and error behaviour as the reference tool: output
// is appended; a record is a '>' line followed by sequence lines up to the
// next '>' line or empty line (spaces removed, lines otherwise kept as
// read); parts start at 0, step maxLength - overlapLength while a whole part
// fits, and a last part holds the rest from the last start.

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

namespace {

// CFASTALoader-style integer argument: the whole string must be a number.
int int_arg(const char* s) {
  std::size_t used = 0;
  int v = 0;
  try {
    v = std::stoi(s, &used);
  } catch (...) {
    used = 0;
  }
  if (used == 0 || s[used] != '\0') {
    std::cout << " argument '" << s << "' is not an integer." << std::endl;
    std::abort();
  }
  return v;
}

// preprocessDB.cpp:8-23
void write_part(std::ofstream& out, const std::string& label, int start, const std::string& part) {
  out << label << '\n' << '@' << start << '\n';
  const int n = static_cast<int>(part.size());
  int i = 0;
  for (; i <= n - 60; i += 60) out << part.substr(static_cast<std::size_t>(i), 60) << '\n';
  out << part.substr(static_cast<std::size_t>(i)) << "\n\n";
}

// preprocessDB.cpp:25-39
void split_record(std::ofstream& out, const std::string& label, const std::string& seq, int max_len, int overlap) {
  const int n = static_cast<int>(seq.size());
  int i = 0;
  for (; i <= n - max_len; i += max_len - overlap)
    write_part(out, label, i, seq.substr(static_cast<std::size_t>(i), static_cast<std::size_t>(max_len)));
  write_part(out, label, i, seq.substr(static_cast<std::size_t>(i)));
}

// preprocessDB.cpp:54-89; the output file is opened (for appending) by the
// first part written, so an input without records creates no file
void split_file(const std::string& in_path, const std::string& out_path, int max_len, int overlap) {
  std::ifstream in(in_path.c_str());
  std::ofstream out;
  std::string line, label, seq;
  auto flush = [&] {
    if (!label.empty()) {
      if (!out.is_open()) out.open(out_path.c_str(), std::ios::app);
      split_record(out, label, seq, max_len, overlap);
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
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 5) {
    const int max_len = int_arg(argv[3]);
    const int overlap = int_arg(argv[4]);
    if (max_len <= overlap) {
      std::cout << " the part length (arg3) must be larger than the overlap (arg4)." << std::endl;
      std::abort();
    }
    split_file(argv[1], argv[2], max_len, overlap);
  } else {
    std::cout << "preprocessDB: split every sequence of a target FASTA file into overlapping parts" << std::endl
              << "of at most <part length> bases, each labelled with its offset ('@start'), for" << std::endl
              << "target sequences longer than one target chunk." << std::endl
              << std::endl
              << "usage: preprocessDB <input.fa> <output.fa> <part length> <overlap length>" << std::endl
              << "  the output file is appended to" << std::endl;
  }
  return 0;
}
