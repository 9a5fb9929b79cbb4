// Fortran list-directed input over the records of a text file, as the data files of the CRP library and the SCM
// correction are read (src/fortran_text/list_directed_reader.cpp): every READ starts at the next record, takes
// values separated by blanks or commas from as many records as it needs (r*c repeats, a slash ends it) and skips
// the rest of its last record. Errors are the CRP library's DataError, naming the file.
#pragma once

#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <optional>
#include <string>
#include <vector>

#include "crp.hpp"

namespace spo::crp {

class ListReader {
 public:
  explicit ListReader(const std::filesystem::path& file) : file_(file.string()) {
    std::ifstream in(file);
    if (!in) throw DataError(file_ + ": cannot open the file");
    std::string line;
    while (std::getline(in, line)) records_.push_back(line);
  }

  // n values, or nullopt when the file ends first.
  std::optional<std::vector<std::string>> try_read(int n) {
    std::vector<std::string> values;
    if (next_ >= records_.size()) return std::nullopt;
    bool ended = false;
    while (static_cast<int>(values.size()) < n && !ended) {
      if (next_ >= records_.size()) return std::nullopt;
      const std::string& rec = records_[next_++];
      std::size_t i = 0;
      while (i < rec.size() && static_cast<int>(values.size()) < n) {
        while (i < rec.size() && (rec[i] == ' ' || rec[i] == '\t' || rec[i] == ',' || rec[i] == '\r')) ++i;
        if (i >= rec.size()) break;
        if (rec[i] == '/') {
          ended = true;
          break;
        }
        std::string token;
        if (rec[i] == '\'' || rec[i] == '"') {
          const char q = rec[i++];
          while (i < rec.size()) {
            if (rec[i] == q) {
              if (i + 1 < rec.size() && rec[i + 1] == q) {
                token += q;
                i += 2;
                continue;
              }
              ++i;
              break;
            }
            token += rec[i++];
          }
        } else {
          while (i < rec.size() && rec[i] != ' ' && rec[i] != '\t' && rec[i] != ',' && rec[i] != '/' &&
                 rec[i] != '\r')
            token += rec[i++];
        }
        const std::size_t star = token.find('*');
        if (star != std::string::npos && star > 0 &&
            token.find_first_not_of("0123456789") == star) {  // r*c: r copies of c
          const int repeat = std::atoi(token.substr(0, star).c_str());
          for (int k = 0; k < repeat && static_cast<int>(values.size()) < n; ++k)
            values.push_back(token.substr(star + 1));
        } else {
          values.push_back(token);
        }
      }
    }
    return values;
  }

  std::vector<std::string> read(int n) {
    if (n == 0) {  // an empty READ still skips a record
      if (next_ < records_.size()) ++next_;
      return {};
    }
    auto v = try_read(n);
    if (!v) throw DataError(file_ + ": unexpected end of file");
    if (static_cast<int>(v->size()) < n) throw DataError(file_ + ": a slash ended a READ before all values");
    return *v;
  }

  double real(const std::string& token) const {
    std::string t = token;
    for (char& c : t)
      if (c == 'd' || c == 'D') c = 'e';
    char* end = nullptr;
    const double x = std::strtod(t.c_str(), &end);
    if (t.empty() || *end != '\0') throw DataError(file_ + ": '" + token + "' is not a real number");
    return x;
  }
  int integer(const std::string& token) const {
    char* end = nullptr;
    const long x = std::strtol(token.c_str(), &end, 10);
    if (token.empty() || *end != '\0') throw DataError(file_ + ": '" + token + "' is not an integer");
    return static_cast<int>(x);
  }
  bool logical(const std::string& token) const {
    std::size_t i = token.empty() || token[0] != '.' ? 0 : 1;
    if (i < token.size() && (token[i] == 'T' || token[i] == 't')) return true;
    if (i < token.size() && (token[i] == 'F' || token[i] == 'f')) return false;
    throw DataError(file_ + ": '" + token + "' is not a logical value");
  }
  const std::string& name() const { return file_; }
  // The next record as text (an A-format READ of a whole line), or nullopt at the end of the file.
  std::optional<std::string> read_line() {
    if (next_ >= records_.size()) return std::nullopt;
    return records_[next_++];
  }

 private:
  std::string file_;
  std::vector<std::string> records_;
  std::size_t next_ = 0;
};

}  // namespace spo::crp
