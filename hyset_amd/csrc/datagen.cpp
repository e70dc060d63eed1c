// Synthetic collection generator for the set-similarity join.
//
// Produces collections in the input format the join expects
// (reference: include/hyset/collection.hpp:92-177 reads them from text):
//   * one set per record, tokens strictly ascending inside a set;
//   * token ids renumbered by ascending global frequency, starting at 1
//     (the reference's preprocessing contract, README.md:14);
//   * sets ordered by size ascending, ties broken lexicographically.
//
// The random stream is xoshiro256** seeded through splitmix64, so the same
// parameters give the same collection on every machine (std:: distributions
// are implementation-defined and are not used).  Zipf draws use Vose's alias
// method over ranks 1..U with weight k^-s.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "hyset_amd.h"

namespace {

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) s[i] = splitmix(x);
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  // uniform double in [0,1) with 53 random bits
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // uniform integer in [0, m)
  uint32_t below(uint32_t m) { return (uint32_t)(((next() >> 32) * (uint64_t)m) >> 32); }
  double normal() {
    // Box-Muller; the second value is discarded to keep the stream simple.
    double u1 = uniform();
    double u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

struct ZipfAlias {
  uint32_t n = 0;
  std::vector<double> prob;
  std::vector<uint32_t> alias;
  ZipfAlias(uint32_t universe, double s) : n(universe), prob(universe), alias(universe) {
    std::vector<double> w(universe);
    double total = 0.0;
    for (uint32_t k = 0; k < universe; ++k) {
      w[k] = std::pow((double)(k + 1), -s);
      total += w[k];
    }
    std::vector<uint32_t> small, large;
    small.reserve(universe);
    large.reserve(universe);
    for (uint32_t k = 0; k < universe; ++k) {
      w[k] = w[k] * universe / total;
      (w[k] < 1.0 ? small : large).push_back(k);
    }
    while (!small.empty() && !large.empty()) {
      uint32_t l = small.back();
      small.pop_back();
      uint32_t g = large.back();
      prob[l] = w[l];
      alias[l] = g;
      w[g] = (w[g] + w[l]) - 1.0;
      if (w[g] < 1.0) {
        large.pop_back();
        small.push_back(g);
      }
    }
    for (uint32_t g : large) { prob[g] = 1.0; alias[g] = g; }
    for (uint32_t l : small) { prob[l] = 1.0; alias[l] = l; }
  }
  // returns a rank in 1..n
  uint32_t draw(Rng& rng) const {
    uint64_t r = rng.next();
    uint32_t col = (uint32_t)(((r >> 32) * (uint64_t)n) >> 32);
    double coin = (double)(r & 0xFFFFFFFFull) * (1.0 / 4294967296.0);
    return (coin < prob[col] ? col : alias[col]) + 1;
  }
};

// Raw collection: each set is a list of distinct Zipf ranks (unsorted).
struct Raw {
  std::vector<uint32_t> tokens;
  std::vector<uint64_t> starts;
  std::vector<uint32_t> sizes;
};

uint32_t draw_size(const hs_gen_params& p, Rng& rng) {
  long v;
  if (p.size_dist == HS_SIZE_UNIFORM) {
    v = (long)p.size_lo + (long)rng.below(p.size_hi - p.size_lo + 1);
  } else {
    double x = std::exp(std::log(p.size_median) + p.size_sigma * rng.normal());
    v = std::lround(x);
  }
  if (v < (long)p.size_lo) v = p.size_lo;
  if (v > (long)p.size_hi) v = p.size_hi;
  return (uint32_t)v;
}

// Marks used to reject duplicate tokens inside one set.
struct Marks {
  std::vector<uint32_t> stamp;
  uint32_t epoch = 0;
  explicit Marks(uint32_t universe) : stamp(universe + 2, 0) {}
  void next_set() {
    if (++epoch == 0) {
      std::fill(stamp.begin(), stamp.end(), 0);
      epoch = 1;
    }
  }
  bool insert(uint32_t t) {
    if (stamp[t] == epoch) return false;
    stamp[t] = epoch;
    return true;
  }
};

void draw_set(const ZipfAlias& z, Rng& rng, Marks& marks, uint32_t size, std::vector<uint32_t>& out) {
  marks.next_set();
  size_t base = out.size();
  while (out.size() - base < size) {
    uint32_t t = z.draw(rng);
    if (marks.insert(t)) out.push_back(t);
  }
}

Raw generate_raw(const hs_gen_params& p, Rng& rng, const ZipfAlias& z, Marks& marks) {
  Raw raw;
  raw.sizes.resize(p.n);
  raw.starts.resize(p.n);
  for (uint32_t i = 0; i < p.n; ++i) {
    uint32_t sz = draw_size(p, rng);
    if (sz > p.universe) sz = p.universe;
    raw.sizes[i] = sz;
  }
  uint64_t total = 0;
  for (uint32_t i = 0; i < p.n; ++i) { raw.starts[i] = total; total += raw.sizes[i]; }
  raw.tokens.reserve(total);
  for (uint32_t i = 0; i < p.n; ++i) draw_set(z, rng, marks, raw.sizes[i], raw.tokens);
  return raw;
}

// Ranking by ascending frequency over `ranked`, ties by original rank.  Tokens
// that never occur in `ranked` but occur in `others` get ids after all ranked
// tokens, in original-rank order.  Returns map from original rank to new id.
std::vector<uint32_t> build_rerank(uint32_t universe, const std::vector<const Raw*>& ranked,
                                   const std::vector<const Raw*>& others, uint32_t* n_ranked) {
  std::vector<uint64_t> freq(universe + 1, 0);
  for (const Raw* r : ranked)
    for (uint32_t t : r->tokens) freq[t]++;
  std::vector<uint32_t> present;
  for (uint32_t t = 1; t <= universe; ++t)
    if (freq[t]) present.push_back(t);
  std::stable_sort(present.begin(), present.end(),
                   [&](uint32_t a, uint32_t b) { return freq[a] < freq[b]; });
  std::vector<uint32_t> map(universe + 1, 0);
  uint32_t next = 1;
  for (uint32_t t : present) map[t] = next++;
  *n_ranked = next - 1;
  for (const Raw* r : others)
    for (uint32_t t : r->tokens)
      if (map[t] == 0) map[t] = UINT32_MAX;  // mark
  for (uint32_t t = 1; t <= universe; ++t)
    if (map[t] == UINT32_MAX) map[t] = next++;
  return map;
}

int finalize(const Raw& raw, const std::vector<uint32_t>& map, hs_host_collection* out) {
  const uint32_t n = (uint32_t)raw.sizes.size();
  std::vector<uint32_t> toks(raw.tokens.size());
  uint32_t maxtok = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t b = raw.starts[i], e = b + raw.sizes[i];
    for (uint64_t k = b; k < e; ++k) {
      toks[k] = map[raw.tokens[k]];
      maxtok = std::max(maxtok, toks[k]);
    }
    std::sort(toks.begin() + b, toks.begin() + e);
  }
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    if (raw.sizes[a] != raw.sizes[b]) return raw.sizes[a] < raw.sizes[b];
    const uint32_t* pa = toks.data() + raw.starts[a];
    const uint32_t* pb = toks.data() + raw.starts[b];
    for (uint32_t k = 0; k < raw.sizes[a]; ++k)
      if (pa[k] != pb[k]) return pa[k] < pb[k];
    return a < b;
  });
  out->n = n;
  out->n_tokens = toks.size();
  out->universe = maxtok + 1;
  out->tokens = (uint32_t*)std::malloc(sizeof(uint32_t) * (toks.size() ? toks.size() : 1));
  out->starts = (uint32_t*)std::malloc(sizeof(uint32_t) * (n ? n : 1));
  out->sizes = (uint32_t*)std::malloc(sizeof(uint32_t) * (n ? n : 1));
  if (!out->tokens || !out->starts || !out->sizes) return HS_E_OOM;
  uint64_t pos = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t src = order[i];
    out->starts[i] = (uint32_t)pos;
    out->sizes[i] = raw.sizes[src];
    std::memcpy(out->tokens + pos, toks.data() + raw.starts[src], sizeof(uint32_t) * raw.sizes[src]);
    pos += raw.sizes[src];
  }
  return HS_OK;
}

bool valid(const hs_gen_params* p) {
  if (!p || p->universe == 0 || p->size_lo == 0 || p->size_hi < p->size_lo) return false;
  if (p->size_dist != HS_SIZE_UNIFORM && p->size_dist != HS_SIZE_LOGNORMAL) return false;
  if (p->size_dist == HS_SIZE_LOGNORMAL && !(p->size_median > 0.0)) return false;
  if ((uint64_t)p->n * p->size_hi > 0xFFFFFFFFull * 4) return false;
  return true;
}

}  // namespace

extern "C" {

int hs_generate_self(const hs_gen_params* p, hs_host_collection* out) {
  if (!valid(p) || !out) return HS_E_INVALID;
  std::memset(out, 0, sizeof(*out));
  try {
    Rng rng(p->seed);
    ZipfAlias z(p->universe, p->zipf_s);
    Marks marks(p->universe);
    Raw raw = generate_raw(*p, rng, z, marks);
    if (raw.tokens.size() > 0xFFFFFFFFull) return HS_E_INVALID;
    uint32_t nr = 0;
    std::vector<uint32_t> map = build_rerank(p->universe, {&raw}, {}, &nr);
    return finalize(raw, map, out);
  } catch (...) {
    return HS_E_OOM;
  }
}

int hs_generate_rs(const hs_gen_params* pr, const hs_gen_params* ps, double dup_frac, double edit_lo,
                   double edit_hi, hs_host_collection* out_r, hs_host_collection* out_s) {
  if (!valid(pr) || !valid(ps) || !out_r || !out_s || pr->universe != ps->universe) return HS_E_INVALID;
  if (dup_frac < 0.0 || dup_frac > 1.0 || edit_lo < 0.0 || edit_hi < edit_lo) return HS_E_INVALID;
  std::memset(out_r, 0, sizeof(*out_r));
  std::memset(out_s, 0, sizeof(*out_s));
  try {
    ZipfAlias z(pr->universe, pr->zipf_s);
    Marks marks(pr->universe);
    Rng rr(pr->seed);
    Raw R = generate_raw(*pr, rr, z, marks);
    Rng rs(ps->seed);
    Raw S = generate_raw(*ps, rs, z, marks);
    // Plant near-duplicates: the first round(dup_frac * |S|) raw S sets are
    // replaced by edited copies of uniformly chosen R sets.
    uint32_t ndup = (uint32_t)std::lround(dup_frac * ps->n);
    if (ndup > ps->n) ndup = ps->n;
    Raw S2;
    S2.sizes.resize(ps->n);
    S2.starts.resize(ps->n);
    S2.tokens.reserve(S.tokens.size());
    std::vector<uint32_t> cur;
    for (uint32_t i = 0; i < ps->n; ++i) {
      S2.starts[i] = S2.tokens.size();
      if (i < ndup && pr->n > 0) {
        uint32_t src = rs.below(pr->n);
        cur.assign(R.tokens.begin() + R.starts[src], R.tokens.begin() + R.starts[src] + R.sizes[src]);
        double f = edit_lo + (edit_hi - edit_lo) * rs.uniform();
        long k = std::lround(f * (double)cur.size());
        marks.next_set();
        for (uint32_t t : cur) marks.insert(t);
        for (long e = 0; e < k; ++e) {
          uint32_t kind = rs.below(3);
          if (kind == 1 && cur.size() > 1) {  // deletion (tokens stay marked: never re-inserted)
            uint32_t at = rs.below((uint32_t)cur.size());
            cur[at] = cur.back();
            cur.pop_back();
          } else {
            uint32_t fresh;
            int guard = 0;
            do { fresh = z.draw(rs); } while (!marks.insert(fresh) && ++guard < 1000000);
            if (guard >= 1000000) continue;
            if (kind == 0 && !cur.empty()) cur[rs.below((uint32_t)cur.size())] = fresh;  // substitution
            else cur.push_back(fresh);                                                    // insertion
          }
        }
        S2.tokens.insert(S2.tokens.end(), cur.begin(), cur.end());
        S2.sizes[i] = (uint32_t)cur.size();
      } else {
        S2.tokens.insert(S2.tokens.end(), S.tokens.begin() + S.starts[i], S.tokens.begin() + S.starts[i] + S.sizes[i]);
        S2.sizes[i] = S.sizes[i];
      }
    }
    uint32_t nr = 0;
    std::vector<uint32_t> map = build_rerank(pr->universe, {&R, &S2}, {}, &nr);
    int rc = finalize(R, map, out_r);
    if (rc != HS_OK) return rc;
    return finalize(S2, map, out_s);
  } catch (...) {
    return HS_E_OOM;
  }
}

int hs_generate_queries(const hs_gen_params* pi, const hs_gen_params* pq, hs_host_collection* out_index,
                        hs_host_collection* out_queries) {
  if (!valid(pi) || !valid(pq) || !out_index || !out_queries || pi->universe != pq->universe) return HS_E_INVALID;
  std::memset(out_index, 0, sizeof(*out_index));
  std::memset(out_queries, 0, sizeof(*out_queries));
  try {
    ZipfAlias z(pi->universe, pi->zipf_s);
    Marks marks(pi->universe);
    Rng ri(pi->seed);
    Raw I = generate_raw(*pi, ri, z, marks);
    ZipfAlias zq(pq->universe, pq->zipf_s);
    Rng rq(pq->seed);
    Raw Q = generate_raw(*pq, rq, zq, marks);
    uint32_t nr = 0;
    std::vector<uint32_t> map = build_rerank(pi->universe, {&I}, {&Q}, &nr);
    int rc = finalize(I, map, out_index);
    if (rc != HS_OK) return rc;
    return finalize(Q, map, out_queries);
  } catch (...) {
    return HS_E_OOM;
  }
}

void hs_host_collection_free(hs_host_collection* c) {
  if (!c) return;
  std::free(c->tokens);
  std::free(c->starts);
  std::free(c->sizes);
  std::memset(c, 0, sizeof(*c));
}

}  // extern "C"
