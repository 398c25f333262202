// FASTA chunk loading with the reference's budget rule (CFASTALoader.cpp).
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "host.hpp"

namespace clast_host {

FastaLoader::FastaLoader(std::vector<std::string> files, int avail_bases, bool target)
    : files_(std::move(files)), avail_(avail_bases), target_(target) {}

// loadFASTA (CFASTALoader.cpp:27-36): continue through the files until one
// stops on a full chunk; file_index() becomes -1 after the last file.
void FastaLoader::load(std::vector<Fasta>& out) {
  for (; file_index_ < static_cast<int>(files_.size()); ++file_index_)
    if (!read_file(files_[file_index_], out)) break;
  if (file_index_ == static_cast<int>(files_.size())) file_index_ = -1;
}

// judgeContinueReading (CFASTALoader.cpp:43-87).  A finished record joins
// the chunk if sum + len < budget; one longer than the whole budget is
// skipped with a warning; otherwise the chunk ends and the next load
// restarts at this record's header.
bool FastaLoader::flush(int64_t pos, std::string& label, std::string& seq, int& start, std::vector<Fasta>& out) {
  if (label.empty()) return true;
  const int64_t len = static_cast<int64_t>(seq.size());
  if (len > avail_) {
    std::cout << " warning : " << (target_ ? "Reference" : "Query") << " sequence " << label
              << " is longer than \"CPU-RAM usage for " << (target_ ? "target" : "query") << "\"." << std::endl
              << (target_ ? " No query will be mapped to this sequence." : " This query will not be mapped to any reference sequence.")
              << std::endl;
  } else if (sum_read_ + len < avail_) {
    out.push_back(Fasta{label, seq, start});
    sum_read_ += len;
  } else if (sum_read_ == 0) {
    // A record of exactly the budget never fits (the reference would reload
    // it forever); it is skipped like an over-long one.
    std::cout << " warning : sequence " << label << " fills the whole RAM budget; skipped." << std::endl;
  } else {
    sum_read_ = 0;
    return false;
  }
  label.clear();
  seq.clear();
  start = 0;
  read_pos_ = pos;
  return true;
}

// read (CFASTALoader.cpp:112-170): '>' anywhere opens a record (label = the
// line after its first byte), '@' anywhere sets the record's start offset,
// an empty line closes a record; sequence lines drop spaces, fold lower case
// and map anything outside A..Z to 'N' (:89-106).
bool FastaLoader::read_file(const std::string& path, std::vector<Fasta>& out) {
  std::ifstream ifs(path.c_str(), std::ios::binary);
  if (!ifs) {
    std::cout << "file \"" << path << "\" open failed." << std::endl;
    std::abort();
  }
  std::string line, label, seq;
  int start = 0;
  ifs.seekg(read_pos_);
  int64_t pos = read_pos_;
  while (ifs && std::getline(ifs, line)) {
    if (line.find('>') != std::string::npos) {
      if (!flush(pos, label, seq, start, out)) return false;
      label = line.substr(1);
    } else if (line.find('@') != std::string::npos) {
      const std::string num = line.substr(1);
      char* end = nullptr;
      const long v = std::strtol(num.c_str(), &end, 10);
      if (num.empty() || std::isspace(static_cast<unsigned char>(num[0])) || *end != '\0') {
        std::cout << " at CFASTALoader::CFASTAFilteLoader::read() : bad file type." << std::endl
                  << " Something was strange at the line which contains '@'." << std::endl;
        std::abort();
      }
      start = static_cast<int>(v);
    } else if (line.empty()) {
      if (!flush(pos, label, seq, start, out)) return false;
    } else if (!label.empty()) {
      seq.reserve(seq.size() + line.size());
      for (char c : line) {
        if (c == ' ') continue;
        if (c > 'Z') c = static_cast<char>(c - 32);
        if (c < 'A' || c > 'Z') c = 'N';
        seq.push_back(c);
      }
    }
    pos = ifs.tellg();
  }
  if (ifs.eof())
    if (!flush(pos, label, seq, start, out)) return false;
  read_pos_ = 0;
  return true;
}

}  // namespace clast_host
