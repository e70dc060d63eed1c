// Dataset preparation: raw records to the collection format the join reads,
// and a binary cache of that format.
//
// The reference reads preprocessed files only (include/hyset/collection.hpp:
// 117-135: tokens ascending inside a line, lines sorted by size, token ids
// renumbered by ascending global frequency) and leaves the preparation to
// the external scripts of the set-similarity-join benchmark it points to
// (README.md:14).  This file does that preparation:
//   1. one record per line; its tokens are the whitespace-separated words of
//      the line, its character q-grams (the line padded with q-1 marks on
//      each side), or integer ids;
//   2. a token that occurs k times in a record becomes k distinct tokens
//      (its 1st, 2nd, ... occurrence), so every record is a set and the
//      overlap of two records counts repeated tokens as often as both have
//      them;
//   3. tokens are ranked by ascending number of records holding them, ties
//      by first appearance, with ids from 1 (id 0 stays unused, as in the
//      generator);
//   4. tokens ascending inside each set, sets sorted by (size, tokens);
//      records without tokens are dropped.  record_of[i] is the line number
//      (0-based) of set i.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "hyset_amd.h"

namespace {

constexpr char kMagic[8] = {'H', 'S', 'C', 'S', 'R', '1', 0, 0};

// Dictionary of token keys (the token text, or the integer id, with its
// occurrence number inside the record) to dense first-appearance ids.
struct Dict {
  std::unordered_map<std::string, uint32_t> ids;
  std::vector<uint64_t> freq;  // records holding each token
  uint32_t get(const std::string& key) {
    auto it = ids.find(key);
    if (it != ids.end()) {
      freq[it->second]++;
      return it->second;
    }
    const uint32_t id = (uint32_t)freq.size();
    ids.emplace(key, id);
    freq.push_back(1);
    return id;
  }
};

// Tokens of one record as keys "<token>\x1f<occurrence>" (repeats numbered).
void record_keys(const char* b, const char* e, const hs_prep_params& p, std::vector<std::string>& keys) {
  keys.clear();
  std::unordered_map<std::string, uint32_t> seen;
  auto add = [&](std::string tok) {
    const uint32_t k = ++seen[tok];
    tok.push_back('\x1f');
    tok += std::to_string(k);
    keys.push_back(std::move(tok));
  };
  std::string line(b, e);
  if (!line.empty() && line.back() == '\r') line.pop_back();
  if (p.lowercase)
    for (char& c : line) c = (char)std::tolower((unsigned char)c);
  if (p.mode == HS_TOKENS_QGRAMS) {
    if (line.empty()) return;
    const uint32_t q = p.q ? p.q : 3;
    const std::string padded = std::string(q - 1, '\x02') + line + std::string(q - 1, '\x03');
    for (size_t i = 0; i + q <= padded.size(); ++i) add(padded.substr(i, q));
    return;
  }
  size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
    size_t j = i;
    while (j < line.size() && !std::isspace((unsigned char)line[j])) ++j;
    if (j > i) {
      std::string tok = line.substr(i, j - i);
      if (p.mode == HS_TOKENS_INTEGERS) {
        char* end = nullptr;
        const unsigned long long v = std::strtoull(tok.c_str(), &end, 10);
        if (end == tok.c_str() || *end != '\0') throw HS_E_INVALID;
        tok = std::to_string(v);  // "007" and "7" are the same id
      }
      add(std::move(tok));
    }
    i = j;
  }
}

int preprocess(const char* text, uint64_t len, const hs_prep_params& p, hs_host_collection* out,
               uint32_t** record_of) {
  if (p.mode != HS_TOKENS_WORDS && p.mode != HS_TOKENS_QGRAMS && p.mode != HS_TOKENS_INTEGERS)
    return HS_E_INVALID;
  if (p.mode == HS_TOKENS_QGRAMS && p.q > 64) return HS_E_INVALID;
  Dict dict;
  std::vector<uint32_t> raw, rstart, rsize, rline;
  std::vector<std::string> keys;
  const char* cur = text;
  const char* end = text + len;
  uint32_t line_no = 0;
  while (cur < end) {
    const char* nl = static_cast<const char*>(std::memchr(cur, '\n', (size_t)(end - cur)));
    const char* le = nl ? nl : end;
    record_keys(cur, le, p, keys);
    if (!keys.empty()) {
      rstart.push_back((uint32_t)raw.size());
      rsize.push_back((uint32_t)keys.size());
      rline.push_back(line_no);
      for (const auto& k : keys) raw.push_back(dict.get(k));
      if (raw.size() >= 0xFFFFFFFFull) return HS_E_UNSUPPORTED;
    }
    ++line_no;
    cur = nl ? nl + 1 : end;
  }
  // ascending frequency, ties by first appearance; ids from 1
  const uint32_t U = (uint32_t)dict.freq.size();
  std::vector<uint32_t> order(U);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return dict.freq[a] < dict.freq[b]; });
  std::vector<uint32_t> rank(U);
  for (uint32_t r = 0; r < U; ++r) rank[order[r]] = r + 1;
  for (uint32_t& t : raw) t = rank[t];
  const uint32_t n = (uint32_t)rsize.size();
  for (uint32_t i = 0; i < n; ++i) std::sort(raw.begin() + rstart[i], raw.begin() + rstart[i] + rsize[i]);
  std::vector<uint32_t> sets(n);
  std::iota(sets.begin(), sets.end(), 0u);
  std::stable_sort(sets.begin(), sets.end(), [&](uint32_t a, uint32_t b) {
    if (rsize[a] != rsize[b]) return rsize[a] < rsize[b];
    return std::lexicographical_compare(raw.begin() + rstart[a], raw.begin() + rstart[a] + rsize[a],
                                        raw.begin() + rstart[b], raw.begin() + rstart[b] + rsize[b]);
  });
  std::memset(out, 0, sizeof(*out));
  out->n = n;
  out->n_tokens = raw.size();
  out->universe = U + 1;
  out->tokens = (uint32_t*)std::malloc(sizeof(uint32_t) * std::max<size_t>(raw.size(), 1));
  out->starts = (uint32_t*)std::malloc(sizeof(uint32_t) * std::max<uint32_t>(n, 1));
  out->sizes = (uint32_t*)std::malloc(sizeof(uint32_t) * std::max<uint32_t>(n, 1));
  uint32_t* rec = record_of ? (uint32_t*)std::malloc(sizeof(uint32_t) * std::max<uint32_t>(n, 1)) : nullptr;
  if (!out->tokens || !out->starts || !out->sizes || (record_of && !rec)) {
    hs_host_collection_free(out);
    std::free(rec);
    return HS_E_OOM;
  }
  uint64_t pos = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t src = sets[i];
    out->starts[i] = (uint32_t)pos;
    out->sizes[i] = rsize[src];
    std::memcpy(out->tokens + pos, raw.data() + rstart[src], sizeof(uint32_t) * rsize[src]);
    pos += rsize[src];
    if (rec) rec[i] = rline[src];
  }
  if (record_of) *record_of = rec;
  return HS_OK;
}

}  // namespace

extern "C" {

int hs_preprocess_text(const char* text, uint64_t len, const hs_prep_params* params, hs_host_collection* out,
                       uint32_t** record_of) {
  if ((!text && len) || !out) return HS_E_INVALID;
  hs_prep_params p{};
  if (params) p = *params;
  try {
    return preprocess(text ? text : "", len, p, out, record_of);
  } catch (int code) {
    return code;
  } catch (const std::bad_alloc&) {
    return HS_E_OOM;
  } catch (...) {
    return HS_E_INVALID;
  }
}

int hs_preprocess_file(const char* path, const hs_prep_params* params, hs_host_collection* out,
                       uint32_t** record_of) {
  if (!path || !out) return HS_E_INVALID;
  FILE* f = std::fopen(path, "rb");
  if (!f) return HS_E_INVALID;
  std::vector<char> buf;
  char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
  const bool bad = std::ferror(f) != 0;
  std::fclose(f);
  if (bad) return HS_E_INVALID;
  return hs_preprocess_text(buf.data(), buf.size(), params, out, record_of);
}

void hs_free_array(void* p) { std::free(p); }

int hs_write_collection_binary(const char* path, const uint32_t* tokens, const uint32_t* starts,
                               const uint32_t* sizes, uint32_t n) {
  if (!path || (n && !sizes)) return HS_E_INVALID;
  uint64_t nt = 0;
  uint32_t maxtok = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint64_t s = starts ? starts[i] : nt;
    for (uint32_t k = 0; k < sizes[i]; ++k) maxtok = std::max(maxtok, tokens[s + k]);
    nt += sizes[i];
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return HS_E_INVALID;
  const uint32_t universe = n ? maxtok + 1 : 1;
  bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(&n, 4, 1, f) == 1 &&
            std::fwrite(&universe, 4, 1, f) == 1 && std::fwrite(&nt, 8, 1, f) == 1;
  if (ok && n) ok = std::fwrite(sizes, 4, n, f) == n;
  if (starts) {
    for (uint32_t i = 0; ok && i < n; ++i) ok = std::fwrite(tokens + starts[i], 4, sizes[i], f) == sizes[i];
  } else if (ok && nt) {
    ok = std::fwrite(tokens, 4, nt, f) == nt;  // dense CSR
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? HS_OK : HS_E_INVALID;
}

int hs_read_collection_binary(const char* path, hs_host_collection* out) {
  if (!path || !out) return HS_E_INVALID;
  std::memset(out, 0, sizeof(*out));
  FILE* f = std::fopen(path, "rb");
  if (!f) return HS_E_INVALID;
  char magic[8];
  uint32_t n = 0, universe = 0;
  uint64_t nt = 0;
  int rc = HS_OK;
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kMagic, 8) != 0 || std::fread(&n, 4, 1, f) != 1 ||
      std::fread(&universe, 4, 1, f) != 1 || std::fread(&nt, 8, 1, f) != 1 || nt >= 0xFFFFFFFFull) {
    rc = HS_E_INVALID;
  }
  std::vector<uint32_t> sizes, tokens;
  if (rc == HS_OK) {
    sizes.resize(n);
    if (n && std::fread(sizes.data(), 4, n, f) != n) rc = HS_E_INVALID;
    uint64_t sum = 0;
    for (uint32_t s : sizes) sum += s;
    if (sum != nt) rc = HS_E_INVALID;
  }
  if (rc == HS_OK) {
    tokens.resize(nt);
    if (nt && std::fread(tokens.data(), 4, nt, f) != nt) rc = HS_E_INVALID;
    char extra;
    if (std::fread(&extra, 1, 1, f) != 0) rc = HS_E_INVALID;  // trailing bytes
    for (uint32_t t : tokens)
      if (t >= universe) rc = HS_E_INVALID;
  }
  std::fclose(f);
  if (rc != HS_OK) return rc;
  out->n = n;
  out->n_tokens = nt;
  out->universe = universe;
  out->tokens = (uint32_t*)std::malloc(sizeof(uint32_t) * std::max<uint64_t>(nt, 1));
  out->starts = (uint32_t*)std::malloc(sizeof(uint32_t) * std::max<uint32_t>(n, 1));
  out->sizes = (uint32_t*)std::malloc(sizeof(uint32_t) * std::max<uint32_t>(n, 1));
  if (!out->tokens || !out->starts || !out->sizes) {
    hs_host_collection_free(out);
    return HS_E_OOM;
  }
  if (nt) std::memcpy(out->tokens, tokens.data(), nt * 4);
  uint64_t pos = 0;
  for (uint32_t i = 0; i < n; ++i) {
    out->starts[i] = (uint32_t)pos;
    out->sizes[i] = sizes[i];
    pos += sizes[i];
  }
  return HS_OK;
}

}  // extern "C"
