#include "qsc.hpp"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace qsh {

bool ParamFile::open(const std::string& path, char comment) {
  std::ifstream f(path);
  if (!f) return false;
  lines_.clear();
  std::string line;
  while (std::getline(f, line)) {
    const size_t c = line.find(comment);
    if (c != std::string::npos) line.resize(c);
    lines_.push_back(line);
  }
  pos_ = 0;
  path_ = path;
  return true;
}

bool ParamFile::read(const std::string& title, std::string* out, bool wrap) {
  auto scan = [&](size_t from) -> bool {
    for (size_t i = from; i < lines_.size(); i++) {
      const size_t k = lines_[i].find(title);
      if (k != std::string::npos) {
        if (out) *out = lines_[i].substr(k + title.size());
        pos_ = i + 1;
        return true;
      }
    }
    pos_ = lines_.size();
    return false;
  };
  if (scan(pos_)) return true;
  if (wrap) return scan(0);
  return false;
}

bool first_word(const std::string& s, std::string* w) {
  std::istringstream is(s);
  return static_cast<bool>(is >> *w);
}

bool first_double(const std::string& s, double* v) {
  const char* p = s.c_str();
  char* end = nullptr;
  const double d = strtod(p, &end);
  if (end == p) return false;
  *v = d;
  return true;
}

bool first_int(const std::string& s, int* v) {
  const char* p = s.c_str();
  char* end = nullptr;
  const long d = strtol(p, &end, 10);
  if (end == p) return false;
  *v = (int)d;
  return true;
}

bool yes(const std::string& s) {
  std::string w;
  if (!first_word(s, &w) || w.empty()) return false;
  return std::tolower((unsigned char)w[0]) == 'y';
}

}  // namespace qsh
