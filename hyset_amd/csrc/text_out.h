// Buffered "r s\n" / "r s overlap\n" text output: the line format of the
// reference's pairs sink (include/hyset/output.hpp:39-76), written through a
// 1 MiB buffer with its own decimal formatting instead of one fprintf per
// line.
#pragma once

#include <cstdint>
#include <cstdio>
#include <vector>

namespace hs {

class PairTextWriter {
 public:
  explicit PairTextWriter(const char* path) : f_(path ? std::fopen(path, "wb") : nullptr), buf_(1u << 20) {}
  PairTextWriter(const PairTextWriter&) = delete;
  PairTextWriter& operator=(const PairTextWriter&) = delete;
  ~PairTextWriter() {
    if (f_) std::fclose(f_);
  }
  bool opened() const { return f_ != nullptr; }

  void add(const uint32_t* r, const uint32_t* s, const uint32_t* ov, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
      if (len_ + kMaxLine > buf_.size()) flush();
      char* p = buf_.data() + len_;
      p = put(p, r[i]);
      *p++ = ' ';
      p = put(p, s[i]);
      if (ov) {
        *p++ = ' ';
        p = put(p, ov[i]);
      }
      *p++ = '\n';
      len_ = (size_t)(p - buf_.data());
    }
  }

  // Flushes and closes; false if any write or the close failed.
  bool close() {
    flush();
    const bool ok = good_ && f_ && std::fclose(f_) == 0;
    f_ = nullptr;
    return ok;
  }

 private:
  static constexpr size_t kMaxLine = 3 * 11;  // three 10-digit numbers and their separators
  static char* put(char* p, uint32_t v) {
    char t[10];
    int k = 0;
    do {
      t[k++] = (char)('0' + v % 10u);
      v /= 10u;
    } while (v);
    while (k) *p++ = t[--k];
    return p;
  }
  void flush() {
    if (len_ && f_) good_ = good_ && std::fwrite(buf_.data(), 1, len_, f_) == len_;
    len_ = 0;
  }
  FILE* f_;
  std::vector<char> buf_;
  size_t len_ = 0;
  bool good_ = true;
};

}  // namespace hs
