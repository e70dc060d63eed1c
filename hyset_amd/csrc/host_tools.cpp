// Text input and output in the reference's formats.
//
// Input (include/hyset/collection.hpp:92-177): one set per line, tokens as
// space-separated unsigned integers, empty lines skipped, set ids are the
// order of the non-empty lines.  The reference trusts the file to be
// preprocessed (tokens ascending, lines sorted by size); here that is
// checked when the collection is uploaded.
// Output (include/hyset/output.hpp:39-76): one pair per line, "r s\n", with
// an optional third column holding the overlap.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hyset_amd.h"
#include "text_out.h"

extern "C" {

int hs_read_collection_text(const char* path, hs_host_collection* out) {
  if (!path || !out) return HS_E_INVALID;
  std::memset(out, 0, sizeof(*out));
  FILE* f = std::fopen(path, "rb");
  if (!f) return HS_E_INVALID;
  std::vector<uint32_t> tokens, starts, sizes;
  std::vector<char> line(1 << 16);
  uint32_t maxtok = 0;
  int rc = HS_OK;
  size_t len = 0;
  bool eof = false;
  while (!eof) {
    // read one line of arbitrary length
    len = 0;
    int c;
    while ((c = std::fgetc(f)) != EOF && c != '\n') {
      if (len + 1 >= line.size()) line.resize(line.size() * 2);
      line[len++] = (char)c;
    }
    if (c == EOF) eof = true;
    line[len] = '\0';
    const char* p = line.data();
    const size_t before = tokens.size();
    while (*p) {
      while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
      if (!*p) break;
      char* end = nullptr;
      errno = 0;
      unsigned long v = std::strtoul(p, &end, 10);
      if (end == p || errno != 0 || v > 0xFFFFFFFEul) {
        rc = HS_E_INVALID;
        break;
      }
      tokens.push_back((uint32_t)v);
      if (v > maxtok) maxtok = (uint32_t)v;
      p = end;
    }
    if (rc != HS_OK) break;
    if (tokens.size() > before) {
      starts.push_back((uint32_t)before);
      sizes.push_back((uint32_t)(tokens.size() - before));
    }
  }
  std::fclose(f);
  if (rc != HS_OK) return rc;
  const uint32_t n = (uint32_t)sizes.size();
  out->n = n;
  out->n_tokens = tokens.size();
  out->universe = n ? maxtok + 1 : 1;
  out->tokens = (uint32_t*)std::malloc(sizeof(uint32_t) * (tokens.size() ? tokens.size() : 1));
  out->starts = (uint32_t*)std::malloc(sizeof(uint32_t) * (n ? n : 1));
  out->sizes = (uint32_t*)std::malloc(sizeof(uint32_t) * (n ? n : 1));
  if (!out->tokens || !out->starts || !out->sizes) {
    hs_host_collection_free(out);
    return HS_E_OOM;
  }
  if (!tokens.empty()) std::memcpy(out->tokens, tokens.data(), tokens.size() * 4);
  if (n) {
    std::memcpy(out->starts, starts.data(), (size_t)n * 4);
    std::memcpy(out->sizes, sizes.data(), (size_t)n * 4);
  }
  return HS_OK;
}

int hs_write_collection_text(const char* path, const uint32_t* tokens, const uint32_t* starts,
                             const uint32_t* sizes, uint32_t n) {
  if (!path || (n && (!tokens || !starts || !sizes))) return HS_E_INVALID;
  FILE* f = std::fopen(path, "wb");
  if (!f) return HS_E_INVALID;
  for (uint32_t i = 0; i < n; ++i) {
    for (uint32_t k = 0; k < sizes[i]; ++k) std::fprintf(f, k ? " %u" : "%u", tokens[starts[i] + k]);
    std::fputc('\n', f);
  }
  return std::fclose(f) == 0 ? HS_OK : HS_E_INVALID;
}

int hs_write_pairs_text(const char* path, const uint32_t* r, const uint32_t* s, const uint32_t* overlap,
                        uint64_t count) {
  if (!path || (count && (!r || !s))) return HS_E_INVALID;
  hs::PairTextWriter w(path);
  if (!w.opened()) return HS_E_INVALID;
  w.add(r, s, overlap, count);
  return w.close() ? HS_OK : HS_E_INVALID;
}

}  // extern "C"
