// Serving: query sets joined against a device-resident index with
// continuous batching.
//
// The reference has no serving path (its drivers run one offline join,
// src/*.cu); this runtime is built on the same device path.  The indexed
// collection stays in HBM with an inverted index over prefixes of the
// loosest supported threshold (t_min), which is a valid index prefix for
// every threshold >= t_min.  Requests (a set of query sets and a threshold)
// enter a FIFO; `slots` worker threads, each with its own HIP stream, take
// as many queued requests as fit the per-batch query limit and the device
// byte budget, run them as one prefix-filter launch (every probe carries its
// own threshold table), and hand each request its pairs.  A request that
// arrives while batches are on the device is taken by the next worker that
// frees up, so batches of different requests overlap on the GPU.  A request
// that does not fit next to the batches in flight waits (deferred); one that
// could never fit the budget is refused with HS_E_BUSY.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ctx.h"

using namespace hs;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
  uint64_t id = 0;
  uint32_t n = 0;
  uint8_t tidx = 0;
  std::vector<uint32_t> tokens, sizes;
  uint64_t bytes = 0;
  Clock::time_point t_submit, t_done;
  bool done = false, deferred = false;
  int status = HS_OK;
  std::vector<uint32_t> r, s, ov;
};

struct Slot {
  hipStream_t stream = nullptr;
  // a batch's probe arrays in one staging region (one copy per batch), laid
  // out by StageLayout; h_stage is its pinned host side
  DevBuf stage, sigs, out, counters;
  PinnedBuf h_stage, h_counters, h_out;
  // out holds the result triples as three rows of out_cap ids (r, s,
  // overlap), h_out as three rows of h_cap: one 2D copy moves a prefix of
  // all three
  uint64_t out_cap = 0, h_cap = 0;
};

// Grows a slot buffer to at least `need` bytes, at least doubling it: the
// free inside a reallocation synchronizes the whole device (every slot's
// work in flight), so it has to stay rare.
template <class B>
void grow(B& b, size_t need) {
  if (need > b.bytes || !b.ptr) b.ensure(std::max(need, 2 * b.bytes));
}

// Grows the device (host) result rows of a slot to at least `cap` triples.
void grow_out(Slot& sl, uint64_t cap) {
  grow(sl.out, cap * 12);
  sl.out_cap = sl.out.bytes / 12;
}
void grow_h_out(Slot& sl, uint64_t cap) {
  grow(sl.h_out, cap * 12);
  sl.h_cap = sl.h_out.bytes / 12;
}
// Copies triples [from, to) of the device rows into the host rows.
void copy_triples(Slot& sl, uint64_t from, uint64_t to, hipStream_t s) {
  if (to <= from) return;
  HS_CHECK_HIP(hipMemcpy2DAsync(sl.h_out.as<uint32_t>() + from, sl.h_cap * 4, sl.out.as<uint32_t>() + from,
                                sl.out_cap * 4, (to - from) * 4, 3, hipMemcpyDeviceToHost, s));
}

// Byte offsets of a batch of nq queries and ntok tokens in the staging
// region: starts, sizes, tile starts (at most nq + 1), threshold indices,
// then the tokens (4-byte aligned).
struct StageLayout {
  size_t starts, sizes, tiles, tidx, tokens, bytes;
  StageLayout(uint64_t nq, uint64_t ntok) {
    starts = 0;
    sizes = 4 * nq;
    tiles = 8 * nq;
    tidx = tiles + 4 * (nq + 1);
    tokens = (tidx + nq + 3) & ~size_t(3);
    bytes = tokens + 4 * ntok;
  }
};

}  // namespace

struct hs_server {
  int device = 0;
  hs_coll* index = nullptr;
  int kind = HS_JACCARD, mode = HS_BOUNDS_REFERENCE;
  hs_server_config cfg{};
  uint32_t L = 0;  // largest set size the tables cover
  uint32_t num_cus = 256;
  // threshold tables: entry 0 is t_min, others are added on first use
  std::mutex tmu;
  std::vector<double> thresholds;
  std::vector<std::unique_ptr<BoundTables>> tables;
  std::vector<std::unique_ptr<DevBuf>> dtables;
  DevBounds hbounds[MAX_THRESHOLDS] = {};
  DevBuf d_bounds;
  IndexBuild ib;
  // parity signatures of the indexed sets for the fgss signature test (the
  // server's own copy: joins on a context may rewrite the collection's)
  uint32_t sig_words = 0;
  DevBuf isigs;
  // requests
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> pending;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
  uint64_t next_id = 1;
  uint64_t bytes_in_flight = 0;
  uint32_t in_flight = 0;
  bool stop = false;
  hs_server_stats stats{};
  std::string err;
  std::vector<std::thread> workers;
};

namespace {

// Device bound tables for threshold t; returns its table index.
int table_for(hs_server* S, double t) {
  std::lock_guard<std::mutex> g(S->tmu);
  HS_CHECK_HIP(hipSetDevice(S->device));
  for (size_t i = 0; i < S->thresholds.size(); ++i)
    if (S->thresholds[i] == t) return (int)i;
  HS_REQUIRE(S->thresholds.size() < MAX_THRESHOLDS, HS_E_UNSUPPORTED, "at most 8 distinct thresholds per server");
  auto T = std::make_unique<BoundTables>();
  T->build(S->kind, S->mode, t, S->L);
  const size_t nl = T->L + 2, total = 4 * nl + T->mo.size();
  std::vector<uint32_t> flat(total);
  std::copy(T->minsize.begin(), T->minsize.end(), flat.begin());
  std::copy(T->maxsize.begin(), T->maxsize.end(), flat.begin() + nl);
  std::copy(T->maxprefix.begin(), T->maxprefix.end(), flat.begin() + 2 * nl);
  std::copy(T->midprefix.begin(), T->midprefix.end(), flat.begin() + 3 * nl);
  std::copy(T->mo.begin(), T->mo.end(), flat.begin() + 4 * nl);
  auto D = std::make_unique<DevBuf>();
  D->ensure(total * 4);
  HS_CHECK_HIP(hipMemcpy(D->ptr, flat.data(), total * 4, hipMemcpyHostToDevice));
  const uint32_t* b = D->as<uint32_t>();
  const size_t k = S->thresholds.size();
  S->hbounds[k] = DevBounds{b, b + nl, b + 2 * nl, b + 3 * nl, b + 4 * nl, T->L, T->kind};
  // entry k is new: no launch in flight reads it
  HS_CHECK_HIP(hipMemcpy(S->d_bounds.as<DevBounds>() + k, &S->hbounds[k], sizeof(DevBounds), hipMemcpyHostToDevice));
  S->thresholds.push_back(t);
  S->tables.push_back(std::move(T));
  S->dtables.push_back(std::move(D));
  return (int)S->thresholds.size() - 1;
}

// Sizes a slot for a batch of max_batch_queries queries of 16 tokens on
// average with 4 result triples per query before it serves anything.
void reserve_slot(const hs_server* S, Slot& sl) {
  const size_t q = std::max<uint32_t>(S->cfg.max_batch_queries, 1), tok = q * 16, out = 4 * q;
  const StageLayout lay(q, tok);
  grow(sl.h_stage, lay.bytes);
  grow(sl.stage, lay.bytes);
  grow(sl.counters, CNT_WORDS * 8);
  grow(sl.h_counters, CNT_WORDS * 8);
  grow_out(sl, out);
  grow_h_out(sl, out);
  if (S->sig_words) grow(sl.sigs, q * S->sig_words * 8);
}

void run_batch(hs_server* S, Slot& sl, std::vector<Request*>& batch) {
  // concatenated probe collection of the batch
  uint64_t ntok = 0;
  uint32_t nq = 0;
  for (Request* r : batch) {
    ntok += r->tokens.size();
    nq += r->n;
  }
  const StageLayout lay(std::max<uint32_t>(nq, 1), std::max<uint64_t>(ntok, 1));
  grow(sl.h_stage, lay.bytes);
  std::vector<uint32_t> first(batch.size() + 1, 0);
  uint8_t* hb = sl.h_stage.as<uint8_t>();
  uint32_t* tk = reinterpret_cast<uint32_t*>(hb + lay.tokens);
  uint32_t* st = reinterpret_cast<uint32_t*>(hb + lay.starts);
  uint32_t* sz = reinterpret_cast<uint32_t*>(hb + lay.sizes);
  uint8_t* ti = hb + lay.tidx;
  uint64_t pos = 0;
  uint32_t q = 0;
  for (size_t k = 0; k < batch.size(); ++k) {
    Request* r = batch[k];
    first[k] = q;
    std::copy(r->tokens.begin(), r->tokens.end(), tk + pos);
    for (uint32_t i = 0; i < r->n; ++i) {
      st[q] = (uint32_t)pos;
      sz[q] = r->sizes[i];
      ti[q] = r->tidx;
      pos += r->sizes[i];
      ++q;
    }
  }
  first[batch.size()] = q;
  // tile plan of the batch: <= TILE_TOKENS tokens and at most enough probes
  // per tile that the batch spreads over about 8 tiles per CU (a lone
  // request of a few thousand queries in 256-probe tiles would occupy a
  // handful of CUs and set the latency tail)
  const uint32_t ppt = std::clamp<uint32_t>((nq + 8 * S->num_cus - 1) / (8 * S->num_cus), 1, TILE_PROBES);
  std::vector<uint32_t> tiles;
  for (uint32_t i = 0; i < nq;) {
    tiles.push_back(i);
    uint64_t t = 0;
    uint32_t c = 0;
    while (i < nq && c < ppt && (c == 0 || t + sz[i] <= TILE_TOKENS)) {
      t += sz[i];
      ++c;
      ++i;
    }
  }
  tiles.push_back(nq);
  // a tile holds at least one probe: at most nq + 1 starts, as laid out
  std::copy(tiles.begin(), tiles.end(), reinterpret_cast<uint32_t*>(hb + lay.tiles));
  grow(sl.stage, lay.bytes);
  grow(sl.counters, CNT_WORDS * 8);
  grow(sl.h_counters, CNT_WORDS * 8);
  if (sl.out_cap < 4ull * nq) grow_out(sl, std::max<uint64_t>(1u << 16, 4ull * nq));
  // result triples fetched with the counters, before their count is known
  // (about one pair per query in the measured mixes): most batches need one
  // round trip to the host, not two
  const uint64_t guess = std::max<uint64_t>(nq, 1024);
  if (sl.h_cap < guess) grow_h_out(sl, guess);
  if (S->sig_words) grow(sl.sigs, (size_t)std::max<uint32_t>(nq, 1) * S->sig_words * 8);
  hipStream_t s = sl.stream;
  // one copy for every probe array of the batch
  HS_CHECK_HIP(hipMemcpyAsync(sl.stage.ptr, sl.h_stage.ptr, lay.tokens + ntok * 4, hipMemcpyHostToDevice, s));
  uint8_t* db = sl.stage.as<uint8_t>();
  hs_coll* I = S->index;
  unsigned long long cnt[CNT_WORDS];
  for (int attempt = 0; attempt < 2; ++attempt) {
    launch_fill_u64(s, sl.counters.as<unsigned long long>(), CNT_WORDS, 0);
    FgssArgs a{};
    a.probe = DevColl{reinterpret_cast<uint32_t*>(db + lay.tokens), reinterpret_cast<uint32_t*>(db + lay.starts),
                      reinterpret_cast<uint32_t*>(db + lay.sizes), nq};
    a.indexed = I->dev();
    a.index = dev_index(S->ib);
    a.bounds = S->hbounds[0];
    a.size_first = I->size_first.as<uint32_t>();
    a.size_first_len = (uint32_t)I->h_size_first.size();
    a.probe_first = 0;
    a.probe_last = nq - 1;
    a.idx_first = 0;
    a.idx_last = I->n - 1;
    a.self = 0;
    a.tile_starts = reinterpret_cast<const uint32_t*>(db + lay.tiles);
    a.first_tile = 0;
    a.n_plan_tiles = (uint32_t)tiles.size() - 1;
    a.probe_bounds = S->d_bounds.as<DevBounds>();
    a.probe_tidx = db + lay.tidx;
    if (S->sig_words) {
      if (attempt == 0) launch_signatures(s, a.probe, S->sig_words, sl.sigs.as<uint64_t>());
      a.sig_words = S->sig_words;
      a.probe_sigs = sl.sigs.as<uint64_t>();
      a.indexed_sigs = S->isigs.as<uint64_t>();
    }
    uint32_t* dout = sl.out.as<uint32_t>();
    a.out = DevOut{dout, dout + sl.out_cap, dout + 2 * sl.out_cap, sl.counters.as<unsigned long long>(), sl.out_cap};
    launch_fgss(s, a, 0);
    HS_CHECK_HIP(hipMemcpyAsync(sl.h_counters.ptr, sl.counters.ptr, CNT_WORDS * 8, hipMemcpyDeviceToHost, s));
    copy_triples(sl, 0, std::min(guess, sl.out_cap), s);
    HS_CHECK_HIP(hipStreamSynchronize(s));
    std::memcpy(cnt, sl.h_counters.ptr, sizeof(cnt));
    fold_counters(cnt);
    if (cnt[CNT_PAIRS] <= sl.out_cap) break;
    HS_REQUIRE(attempt == 0, HS_E_DEVICE, "result count changed between runs");
    grow_out(sl, cnt[CNT_PAIRS] + cnt[CNT_PAIRS] / 4 + 1024);
  }
  const uint64_t np = cnt[CNT_PAIRS];
  const uint64_t fetched = std::min(guess, sl.out_cap);
  if (np > fetched) {
    // the rest; a larger host buffer starts empty, so then every triple
    uint64_t from = fetched;
    if (sl.h_cap < np) {
      grow_h_out(sl, np);
      from = 0;
    }
    copy_triples(sl, from, np, s);
    HS_CHECK_HIP(hipStreamSynchronize(s));
  }
  const uint32_t* R = sl.h_out.as<uint32_t>();
  const uint32_t* Sd = R + sl.h_cap;
  const uint32_t* O = R + 2 * sl.h_cap;
  for (uint64_t i = 0; i < np; ++i) {
    const size_t k = std::upper_bound(first.begin(), first.end(), R[i]) - first.begin() - 1;
    Request* r = batch[k];
    r->r.push_back(R[i] - first[k]);
    r->s.push_back(Sd[i]);
    r->ov.push_back(O[i]);
  }
}

void worker(hs_server* S) {
  Slot sl;
  std::string fail;
  if (hipSetDevice(S->device) != hipSuccess || hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess)
    fail = "cannot create a stream";
  if (fail.empty()) {
    try {
      reserve_slot(S, sl);
    } catch (const Error& e) {
      fail = e.msg;
    }
  }
  for (;;) {
    std::vector<Request*> batch;
    uint64_t bytes = 0;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      for (;;) {
        if (S->stop && S->pending.empty()) break;
        if (!S->pending.empty()) {
          Request* head = S->pending.front();
          if (head->bytes > S->cfg.hbm_budget_bytes) {  // can never fit: refuse
            S->pending.pop_front();
            head->status = HS_E_BUSY;
            head->done = true;
            head->t_done = Clock::now();
            S->stats.rejected++;
            S->cv_done.notify_all();
            continue;
          }
          uint32_t nq = 0;
          while (!S->pending.empty()) {
            Request* r = S->pending.front();
            if (!batch.empty() && nq + r->n > S->cfg.max_batch_queries) break;
            if (S->bytes_in_flight + bytes + r->bytes > S->cfg.hbm_budget_bytes) break;
            S->pending.pop_front();
            batch.push_back(r);
            bytes += r->bytes;
            nq += r->n;
          }
          if (!batch.empty()) break;
          if (!head->deferred) {  // waits for memory held by batches in flight
            head->deferred = true;
            S->stats.deferred++;
          }
        }
        S->cv_work.wait(lk);
      }
      if (batch.empty()) break;
      S->bytes_in_flight += bytes;
      S->in_flight++;
      S->stats.batches++;
      S->stats.max_in_flight = std::max<uint64_t>(S->stats.max_in_flight, S->in_flight);
      S->stats.max_bytes_in_flight = std::max<uint64_t>(S->stats.max_bytes_in_flight, S->bytes_in_flight);
    }
    int status = HS_OK;
    if (!fail.empty()) {
      status = HS_E_DEVICE;
    } else {
      try {
        run_batch(S, sl, batch);
      } catch (const Error& e) {
        status = e.code;
        std::lock_guard<std::mutex> g(S->mu);
        S->err = e.msg;
      } catch (...) {
        status = HS_E_DEVICE;
      }
    }
    {
      std::lock_guard<std::mutex> g(S->mu);
      const auto now = Clock::now();
      for (Request* r : batch) {
        r->status = status;
        r->done = true;
        r->t_done = now;
      }
      S->bytes_in_flight -= bytes;
      S->in_flight--;
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();
  }
  if (sl.stream) {
    (void)hipStreamSynchronize(sl.stream);
    (void)hipStreamDestroy(sl.stream);
  }
}

// Device bytes a request holds while in flight: its query CSR, threshold
// indices, tile plan and signatures, plus room for 8 result triples per query.
uint64_t request_bytes(uint32_t n, uint64_t ntok, uint32_t sig_words) {
  return 4 * ntok + (13ull + 8ull * sig_words) * n + 96ull * n + 4096;
}

}  // namespace

extern "C" {

int hs_server_create(int device, hs_coll* index, int32_t sim, int32_t bounds_mode, const hs_server_config* cfg,
                     hs_server** out) {
  if (!index || !out) return HS_E_INVALID;
  *out = nullptr;
  auto S = std::make_unique<hs_server>();
  int rc = guarded(nullptr, [&] {
    HS_REQUIRE(index->n_joins == 0 && index->n > 0, HS_E_INVALID, "the index must be a non-empty collection");
    HS_REQUIRE(index->device == device, HS_E_INVALID, "index collection lives on another device");
    HS_CHECK_HIP(hipSetDevice(device));
    S->device = device;
    S->index = index;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
      S->num_cus = (uint32_t)cus;
    S->kind = sim;
    S->mode = bounds_mode;
    hs_server_config c{0.6, 4, 1ull << 30, 65536, 0};
    if (cfg) {
      if (cfg->min_threshold > 0) c.min_threshold = cfg->min_threshold;
      if (cfg->slots) c.slots = cfg->slots;
      if (cfg->hbm_budget_bytes) c.hbm_budget_bytes = cfg->hbm_budget_bytes;
      if (cfg->max_batch_queries) c.max_batch_queries = cfg->max_batch_queries;
      c.max_query_size = cfg->max_query_size;
    }
    HS_REQUIRE(c.slots <= 32, HS_E_INVALID, "at most 32 slots");
    S->cfg = c;
    const uint32_t qmax = c.max_query_size ? c.max_query_size : (sim == HS_COSINE ? 1024u : 4096u);
    S->L = std::max(index->maxlen, qmax);
    S->d_bounds.ensure(MAX_THRESHOLDS * sizeof(DevBounds));
    HS_CHECK_HIP(hipMemset(S->d_bounds.ptr, 0, MAX_THRESHOLDS * sizeof(DevBounds)));
    table_for(S.get(), c.min_threshold);
    // index over maxprefix(t_min): valid for every threshold >= t_min
    const auto& ptab = S->tables[0]->maxprefix;
    const auto& sf = index->h_size_first;
    std::vector<uint32_t> cum(sf.size());
    uint64_t acc = 0;
    for (size_t l = 0; l + 1 < sf.size(); ++l) {
      cum[l] = (uint32_t)acc;
      acc += (uint64_t)(sf[l + 1] - sf[l]) * ptab[l];
      HS_REQUIRE(acc < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "index limited to 2^32-1 entries");
    }
    cum.back() = (uint32_t)acc;
    DevBuf d_cum;
    d_cum.ensure(cum.size() * 4);
    HS_CHECK_HIP(hipMemcpy(d_cum.ptr, cum.data(), cum.size() * 4, hipMemcpyHostToDevice));
    hipStream_t st;
    HS_CHECK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    S->sig_words = auto_signature_words(index->n ? (double)index->n_tokens / index->n : 0.0);
    S->stats.signature_bits = S->sig_words * 64;
    S->isigs.ensure((size_t)std::max<uint32_t>(index->n, 1) * S->sig_words * 8);
    launch_signatures(st, index->dev(), S->sig_words, S->isigs.as<uint64_t>());
    build_index(st, index->dev(), 0, index->n - 1, S->hbounds[0].maxprefix, index->size_first.as<uint32_t>(),
                d_cum.as<uint32_t>(), index->universe, acc, S->ib,
                S->sig_words == 1 ? S->isigs.as<uint64_t>() : nullptr);
    HS_CHECK_HIP(hipStreamSynchronize(st));
    HS_CHECK_HIP(hipStreamDestroy(st));
    for (uint32_t k = 0; k < c.slots; ++k) S->workers.emplace_back(worker, S.get());
  });
  if (rc != HS_OK) return rc;
  *out = S.release();
  return HS_OK;
}

int hs_server_submit(hs_server* S, const uint32_t* tokens, const uint32_t* sizes, uint32_t n_queries,
                     uint64_t n_tokens, double threshold, uint64_t* request_id) {
  if (!S || !request_id) return HS_E_INVALID;
  return guarded(nullptr, [&] {
    HS_REQUIRE(n_queries > 0 && tokens && sizes, HS_E_INVALID, "empty request");
    HS_REQUIRE(threshold >= S->cfg.min_threshold && threshold <= 1.0, HS_E_INVALID,
               "threshold below the server's minimum");
    auto r = std::make_unique<Request>();
    r->n = n_queries;
    r->sizes.assign(sizes, sizes + n_queries);
    uint64_t total = 0;
    for (uint32_t i = 0; i < n_queries; ++i) {
      const uint32_t l = sizes[i];
      HS_REQUIRE(l >= 1 && l <= S->L, HS_E_UNSUPPORTED, "query set size outside the server's tables");
      const uint32_t* t = tokens + total;
      for (uint32_t k = 1; k < l; ++k) HS_REQUIRE(t[k - 1] < t[k], HS_E_INVALID, "query tokens not ascending");
      total += l;
    }
    HS_REQUIRE(total == n_tokens, HS_E_INVALID, "n_tokens does not match the sizes");
    r->tokens.assign(tokens, tokens + n_tokens);
    r->tidx = (uint8_t)table_for(S, threshold);
    r->bytes = request_bytes(n_queries, n_tokens, S->sig_words);
    std::lock_guard<std::mutex> g(S->mu);
    HS_REQUIRE(!S->stop, HS_E_INVALID, "server is shutting down");
    r->id = S->next_id++;
    r->t_submit = Clock::now();
    *request_id = r->id;
    S->pending.push_back(r.get());
    S->requests[r->id] = std::move(r);
    S->stats.submitted++;
    S->cv_work.notify_one();
  });
}

int hs_server_wait(hs_server* S, uint64_t request_id, hs_result* result, double* latency_ms) {
  if (!S || !result) return HS_E_INVALID;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->requests.find(request_id);
  if (it == S->requests.end()) return HS_E_INVALID;
  Request* r = it->second.get();
  S->cv_done.wait(lk, [&] { return r->done; });
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  const int status = r->status;
  const uint64_t count = r->r.size();
  result->count = count;
  if (status == HS_OK && !result->count_only && count) {
    // too small: keep the answer so the caller can ask again with room for it
    if (count > result->capacity || !result->r || !result->s) return HS_E_OVERFLOW;
    std::copy(r->r.begin(), r->r.end(), result->r);
    std::copy(r->s.begin(), r->s.end(), result->s);
    if (result->overlap) std::copy(r->ov.begin(), r->ov.end(), result->overlap);
  }
  S->requests.erase(it);
  return status;
}

int hs_server_get_stats(hs_server* S, hs_server_stats* out) {
  if (!S || !out) return HS_E_INVALID;
  std::lock_guard<std::mutex> g(S->mu);
  *out = S->stats;
  return HS_OK;
}

void hs_server_destroy(hs_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  for (auto& t : S->workers) t.join();
  delete S;
}

}  // extern "C"
