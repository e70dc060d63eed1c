// The CPU partition of the hybrid join: prefix-filter join on host cores.
//
// Reference: host::allpairs and host::ppjoin (include/hyset/
// host_algorithms.hpp:16-250) and the cpu_multi driver's thread-per-
// partition layout (src/cpu_multi.cu:118-207).  Used by the CPU/GPU split to
// run the units it gives the CPU and by co-processing's host threads;
// results are the same (r, s, overlap) triples the device techniques emit.
// Per probe: candidates from the id-sorted posting lists of its prefix
// tokens inside the length window, deduplicated with a per-thread stamp
// array (PPJoin: a candidate is taken at its first common prefix token only
// if the positional filter passes there, host_algorithms.hpp:183-192); each
// candidate is verified by a merge with the verify_pair early-termination
// bound (verification.hpp:9-31), run to the end for passing pairs so the
// overlap is exact.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "ctx.h"

using namespace hs;

namespace {

struct HostIndex {
  std::vector<uint32_t> offsets, ids, metas;
};

// Inverted index of the prefix entries of sets [first, last] (counting sort
// in set-id order, so every list is ascending).
void build_host_index(const hs_coll& S, uint32_t first, uint32_t last, const std::vector<uint32_t>& plen,
                      HostIndex& X) {
  X.offsets.assign((size_t)S.universe + 1, 0);
  for (uint32_t i = first; i <= last; ++i) {
    const uint32_t* t = S.h_tokens.data() + S.h_starts[i];
    for (uint32_t k = 0; k < plen[S.h_sizes[i]]; ++k) X.offsets[t[k] + 1]++;
  }
  for (uint32_t u = 0; u < S.universe; ++u) X.offsets[u + 1] += X.offsets[u];
  X.ids.resize(X.offsets[S.universe]);
  X.metas.resize(X.offsets[S.universe]);
  std::vector<uint32_t> cur(X.offsets.begin(), X.offsets.end() - 1);
  for (uint32_t i = first; i <= last; ++i) {
    const uint32_t* t = S.h_tokens.data() + S.h_starts[i];
    const uint32_t l = S.h_sizes[i];
    for (uint32_t k = 0; k < plen[l]; ++k) {
      const uint32_t at = cur[t[k]]++;
      X.ids[at] = i;
      X.metas[at] = (std::min(l, 0xFFFFu) << 16) | std::min(k, 0xFFFFu);
    }
  }
}

void join_probes(const hs_coll& P, const hs_coll& S, const HostIndex& X, const BoundTables& T, bool self,
                 const hs_range& ir, uint32_t first, uint32_t last, bool ppjoin, Triples& out,
                 hs::HostCounts& counts) {
  std::vector<uint32_t> stamp(S.n, 0xFFFFFFFFu);
  hs::host_probe_range(P, S, X.offsets.data(), X.ids.data(), ppjoin ? X.metas.data() : nullptr, T, self, ir,
                       first, last, out, stamp, &counts);
}

}  // namespace

void hs::host_probe_range(const hs_coll& P, const hs_coll& S, const uint32_t* offsets, const uint32_t* ids,
                          const uint32_t* metas, const BoundTables& T, bool self, const hs_range& ir,
                          uint32_t first, uint32_t last, Triples& out, std::vector<uint32_t>& stamp,
                          HostCounts* counts) {
  uint64_t n_visit = 0, n_cand = 0;
  std::vector<uint32_t> cand;
  if (stamp.size() < S.n) stamp.assign(S.n, 0xFFFFFFFFu);
  const uint32_t last_len = (uint32_t)S.h_size_first.size() - 1;
  for (uint32_t p = first; p <= last && first <= last; ++p) {
    const uint32_t l = P.h_sizes[p];
    const uint32_t* rt = P.h_tokens.data() + P.h_starts[p];
    const uint32_t mins = std::min(T.minsize[l], last_len);
    const uint32_t maxs1 = std::min(T.maxsize[l], last_len - 1) + 1;
    const uint32_t lo = std::max(S.h_size_first[mins], ir.first);
    uint32_t hi = std::min(S.h_size_first[maxs1], ir.last + 1);
    if (self) hi = std::min(hi, p);
    if (lo >= hi) continue;
    cand.clear();
    for (uint32_t j = 0; j < T.maxprefix[l]; ++j) {
      const uint32_t tok = rt[j];
      if (tok >= S.universe) continue;
      const uint32_t* a = ids + offsets[tok];
      const uint32_t* e = ids + offsets[tok + 1];
      for (const uint32_t* it = std::lower_bound(a, e, lo); it != e && *it < hi; ++it) {
        ++n_visit;
        if (stamp[*it] == p) continue;
        if (metas) {
          // PPJoin (host_algorithms.hpp:183-192): a candidate is taken at a
          // common prefix token only if the positional bound passes there
          const uint32_t ip = metas[it - ids] & 0xFFFFu;
          if (ip != 0xFFFFu) {
            const uint32_t ls = S.h_sizes[*it];
            if (T.minoverlap(l, ls) > std::min(l - j, ls - ip)) continue;
          }
        }
        stamp[*it] = p;
        cand.push_back(*it);
      }
    }
    n_cand += cand.size();
    for (uint32_t sid : cand) {
      const uint32_t ls = S.h_sizes[sid];
      const uint32_t mo = T.minoverlap(l, ls);
      const uint32_t* st = S.h_tokens.data() + S.h_starts[sid];
      uint32_t i = 0, k = 0, found = 0;
      bool fail = false;
      while (i < l && k < ls) {
        if (found + std::min(l - i, ls - k) < mo) {
          fail = true;
          break;
        }
        if (rt[i] == st[k]) {
          ++found;
          ++i;
          ++k;
        } else if (rt[i] < st[k]) {
          ++i;
        } else {
          ++k;
        }
      }
      if (!fail && found >= mo) {
        out.r.push_back(p);
        out.s.push_back(sid);
        out.ov.push_back(found);
      }
    }
  }
  if (counts) {
    counts->visits += n_visit;
    counts->candidates += n_cand;
  }
}

namespace hs {

uint64_t host_join(const hs_coll& indexed, const hs_coll& probe, bool self, hs_range ir, hs_range pr,
                   const BoundTables& T, uint32_t threads, std::vector<Triples>* parts_out, bool ppjoin,
                   HostCounts* counts) {
  HostIndex X;
  if (!range_empty(ir)) build_host_index(indexed, ir.first, ir.last, self ? T.midprefix : T.maxprefix, X);
  else X.offsets.assign((size_t)indexed.universe + 1, 0);
  const uint32_t nt = std::max<uint32_t>(1, threads);
  std::vector<Triples> parts(nt);
  std::vector<HostCounts> pc(nt);
  if (!range_empty(pr) && !range_empty(ir)) {
    // probe partitions: sqrt-spaced for a self-join (src/cpu_multi.cu:118-139)
    std::vector<uint32_t> pts{pr.first};
    const double n = (double)pr.last - pr.first + 1;
    for (uint32_t i = 1; i < nt; ++i)
      pts.push_back(pr.first + (uint32_t)(n * (self ? std::sqrt((double)i / nt) : (double)i / nt)));
    pts.push_back(pr.last + 1);
    if (nt == 1) {
      join_probes(probe, indexed, X, T, self, ir, pts[0], pts[1] - 1, ppjoin, parts[0], pc[0]);
    } else {
      std::vector<std::thread> pool;
      for (uint32_t i = 0; i < nt; ++i)
        if (pts[i + 1] > pts[i])
          pool.emplace_back(join_probes, std::cref(probe), std::cref(indexed), std::cref(X), std::cref(T), self,
                            std::cref(ir), pts[i], pts[i + 1] - 1, ppjoin, std::ref(parts[i]), std::ref(pc[i]));
      for (auto& th : pool) th.join();
    }
  }
  uint64_t count = 0;
  for (auto& p : parts) count += p.r.size();
  if (counts)
    for (const auto& c : pc) {
      counts->visits += c.visits;
      counts->candidates += c.candidates;
    }
  if (parts_out) *parts_out = std::move(parts);
  return count;
}

}  // namespace hs

extern "C" {

int hs_join_host(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, const hs_range* indexed_range,
                 const hs_range* probe_range, uint32_t threads, hs_result* result) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    HS_REQUIRE(indexed && probe && result, HS_E_INVALID, "NULL argument");
    HS_REQUIRE(indexed->n_joins == 0 && probe->n_joins == 0, HS_E_INVALID, "batched collections are not supported");
    HS_REQUIRE(ctx->have_threshold, HS_E_INVALID, "hs_set_threshold has not been called");
    const bool self = indexed == probe;
    const uint32_t L = std::max(indexed->maxlen, probe->maxlen);
    BoundTables T;
    if (ctx->tables_valid && ctx->tables.L >= L) T = ctx->tables;
    else T.build(ctx->kind, ctx->mode, ctx->t, std::max<uint32_t>(L, 64));
    auto t0 = std::chrono::steady_clock::now();
    std::vector<Triples> parts;
    HostCounts hc;
    const uint64_t count = host_join(*indexed, *probe, self, clamp_range(indexed_range, indexed->n),
                                     clamp_range(probe_range, probe->n), T, threads, &parts,
                                     ctx->host_algo == HS_HOST_PPJOIN, &hc);
    result->count = count;
    ctx->stats = hs_stats{};
    ctx->stats.pairs = count;
    ctx->stats.posting_visits = hc.visits;
    ctx->stats.candidates = hc.candidates;
    ctx->stats.host_filter_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->stats.total_ms = ctx->stats.host_filter_ms;
    if (!result->count_only && count) {
      HS_REQUIRE(count <= result->capacity && result->r && result->s, HS_E_OVERFLOW, "result capacity too small");
      uint64_t o = 0;
      for (auto& p : parts) {
        std::copy(p.r.begin(), p.r.end(), result->r + o);
        std::copy(p.s.begin(), p.s.end(), result->s + o);
        if (result->overlap) std::copy(p.ov.begin(), p.ov.end(), result->overlap + o);
        o += p.r.size();
      }
    }
  });
}

}  // extern "C"
