// qsc.hpp -- ".qsc" parameter files with QSTEM's readparam semantics
// (libs/readparams.cpp:173-245): each call searches forward from the current
// line for the first line containing `title` (after cutting the comment at
// the comment character), optionally wrapping to the start of the file, and
// returns the text after the title.
#pragma once
#include <string>
#include <vector>

namespace qsh {

class ParamFile {
 public:
  bool open(const std::string& path, char comment = '%');
  // true and *out = text after the title when found
  bool read(const std::string& title, std::string* out, bool wrap = true);
  void reset() { pos_ = 0; }
  const std::string& path() const { return path_; }

 private:
  std::vector<std::string> lines_;
  size_t pos_ = 0;
  std::string path_;
};

// sscanf-like helpers on the returned text
bool first_word(const std::string& s, std::string* w);
bool first_double(const std::string& s, double* v);
bool first_int(const std::string& s, int* v);
bool yes(const std::string& s);  // tolower(first char of first word) == 'y'

}  // namespace qsh
