// C ABI of libhyset_amd.so: contexts, device-resident collections, bound
// tables and the join drivers for the three techniques.
//
// Reference counterparts (paths relative to the reference repository):
//   hs_create / hs_destroy     bitmap::handler, fgssjoin::handler and
//                              hybrid::device_handler constructors/destructors
//                              (include/hyset/device_algorithms.hpp:138-171,
//                              :322-348; include/hyset/hybrid_algorithms.hpp:41-61)
//   hs_upload_collection       device_collection (include/hyset/collection.hpp:54-71)
//   hs_join                    handler::join(block*, block*) and
//                              join(partition*, partition*) (device_algorithms.hpp:177-247,
//                              :363-432); the technique switch of src/queue.cu:181-214
//   hs_verify_candidates       device_handler::addCandidates / join / invoke / flush
//                              (hybrid_algorithms.hpp:71-200)
// The reference throws on CUDA errors and downcasts its output sink; here
// every call returns a code and the sink is a plain struct.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "ctx.h"
#include "text_out.h"

using namespace hs;

// ---------------------------------------------------------------- DevBuf

namespace {
std::atomic<uint64_t> g_buffer_epoch{0};
}

uint64_t hs::buffer_epoch() { return g_buffer_epoch.load(std::memory_order_acquire); }
void hs::bump_buffer_epoch() { g_buffer_epoch.fetch_add(1, std::memory_order_acq_rel); }

void hs::DevBuf::ensure(size_t need) {
  if (need <= bytes && ptr) return;
  release();
  bump_buffer_epoch();
  size_t want = need < 256 ? 256 : need + need / 8;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error{HS_E_OOM, std::string("hipMalloc of ") + std::to_string(want) + " bytes failed"};
  }
  ptr = p;
  bytes = want;
}

void hs::DevBuf::release() {
  if (ptr) {
    (void)hipFree(ptr);
    bump_buffer_epoch();
  }
  ptr = nullptr;
  bytes = 0;
}

// ------------------------------------------------------- result stream

void hs::stream_pairs(hs_ctx* ctx, uint64_t count, bool with_overlap, uint64_t chunk,
                      const std::function<void(const uint32_t*, const uint32_t*, const uint32_t*, uint64_t,
                                               uint64_t)>& consume) {
  if (count == 0) return;
  hipStream_t st = ctx->stream;
  chunk = std::min(chunk ? chunk : kStreamChunk, count);
  const size_t arrays = with_overlap ? 3 : 2;
  for (int k = 0; k < 2; ++k) {
    if (!ctx->stage_ev[k]) HS_CHECK_HIP(hipEventCreateWithFlags(&ctx->stage_ev[k], hipEventDisableTiming));
  }
  const uint64_t nch = (count + chunk - 1) / chunk;
  // a slot holds [r | s | overlap] of one chunk; the second slot is needed
  // only when there is a second chunk
  for (int k = 0; k < (nch > 1 ? 2 : 1); ++k) {
    if (ctx->stage[k].bytes < chunk * 4 * arrays) {
      HS_CHECK_HIP(hipStreamSynchronize(st));  // nothing may still be copying into the old buffer
      ctx->stage[k].ensure(chunk * 4 * arrays);
    }
  }
  auto issue = [&](uint64_t c) {
    const uint64_t off = c * chunk, n = std::min(chunk, count - off);
    // rows r, s (and overlap) of the chunk in one 2D copy
    HS_CHECK_HIP(hipMemcpy2DAsync(ctx->stage[c & 1].ptr, chunk * 4, ctx->out.as<uint32_t>() + off, ctx->out_cap * 4,
                                  n * 4, arrays, hipMemcpyDeviceToHost, st));
    HS_CHECK_HIP(hipEventRecord(ctx->stage_ev[c & 1], st));
  };
  try {
    issue(0);
    for (uint64_t c = 0; c < nch; ++c) {
      // slot (c+1)&1 held chunk c-1, which consume has finished with
      if (c + 1 < nch) issue(c + 1);
      HS_CHECK_HIP(hipEventSynchronize(ctx->stage_ev[c & 1]));
      const uint32_t* b = ctx->stage[c & 1].as<uint32_t>();
      consume(b, b + chunk, with_overlap ? b + 2 * chunk : nullptr, c * chunk, std::min(chunk, count - c * chunk));
    }
  } catch (...) {
    (void)hipStreamSynchronize(st);  // no copy may land in a slot after the caller has moved on
    throw;
  }
}

void hs::stream_result(hs_ctx* ctx, bool with_overlap, uint64_t chunk,
                       const std::function<void(const uint32_t*, const uint32_t*, const uint32_t*, uint64_t,
                                                uint64_t)>& consume) {
  uint64_t tail = 0;
  for (const auto& t : ctx->host_tail) tail += t.r.size();
  HS_REQUIRE(tail <= ctx->last_count, HS_E_DEVICE, "inconsistent result");
  const uint64_t dev = ctx->last_count - tail;
  stream_pairs(ctx, dev, with_overlap, chunk, consume);
  chunk = chunk ? chunk : kStreamChunk;
  uint64_t at = dev;
  for (const auto& t : ctx->host_tail)
    for (uint64_t o = 0; o < t.r.size(); o += chunk) {
      const uint64_t n = std::min<uint64_t>(chunk, t.r.size() - o);
      consume(t.r.data() + o, t.s.data() + o, with_overlap ? t.ov.data() + o : nullptr, at, n);
      at += n;
    }
}

void hs::copy_pairs(hs_ctx* ctx, uint64_t count, hs_result* res) {
  if (count == 0) return;
  HS_REQUIRE(res->r && res->s, HS_E_INVALID, "result arrays are NULL");
  HS_REQUIRE(count == ctx->last_count, HS_E_DEVICE, "copy of a result that is not the last one");
  stream_result(ctx, res->overlap != nullptr, 0,
               [&](const uint32_t* r, const uint32_t* s, const uint32_t* ov, uint64_t first, uint64_t n) {
                 std::memcpy(res->r + first, r, n * 4);
                 std::memcpy(res->s + first, s, n * 4);
                 if (ov) std::memcpy(res->overlap + first, ov, n * 4);
               });
}

namespace {

// ------------------------------------------------------- hybrid pipeline

struct HybridShared {
  hs_ctx* ctx;
  hs_coll* indexed;
  hs_coll* probe;
  bool self;
  hs_range ir, pr;
  int scenario;
  uint64_t batch;
  // the host copy of the index: offsets over the universe, id-ascending
  // lists, and (PPJoin filtering, else NULL) the entries' metas
  const uint32_t* off = nullptr;
  const uint32_t* ids = nullptr;
  const uint32_t* metas = nullptr;
  std::atomic<uint64_t> candidates{0}, visits{0};
};

// Copies one candidate batch into a pinned slot and queues its
// verification on the context stream; the host returns to filtering while
// the device verifies (the reference double-buffers with a thread per
// batch, hybrid_algorithms.hpp:107-120).
void submit_batch(HybridShared& H, const std::vector<uint32_t>& probes, const std::vector<uint32_t>& offsets,
                  const std::vector<uint32_t>& cands, uint32_t& next_slot) {
  if (probes.empty()) return;
  hs_ctx* ctx = H.ctx;
  std::lock_guard<std::mutex> lock(ctx->submit_mu);
  HybridSlot& sl = ctx->slots[next_slot];
  next_slot ^= 1u;
  if (sl.used) HS_CHECK_HIP(hipEventSynchronize(sl.done));
  const size_t np = probes.size(), nc = cands.size();
  sl.h_probes.ensure(np * 4);
  sl.h_offsets.ensure(np * 4);
  sl.h_cands.ensure(nc * 4);
  sl.d_probes.ensure(np * 4);
  sl.d_offsets.ensure(np * 4);
  sl.d_cands.ensure(nc * 4);
  std::memcpy(sl.h_probes.ptr, probes.data(), np * 4);
  std::memcpy(sl.h_offsets.ptr, offsets.data(), np * 4);
  std::memcpy(sl.h_cands.ptr, cands.data(), nc * 4);
  HS_CHECK_HIP(hipMemcpyAsync(sl.d_probes.ptr, sl.h_probes.ptr, np * 4, hipMemcpyHostToDevice, ctx->stream));
  HS_CHECK_HIP(hipMemcpyAsync(sl.d_offsets.ptr, sl.h_offsets.ptr, np * 4, hipMemcpyHostToDevice, ctx->stream));
  HS_CHECK_HIP(hipMemcpyAsync(sl.d_cands.ptr, sl.h_cands.ptr, nc * 4, hipMemcpyHostToDevice, ctx->stream));
  hs::VerifyArgs va{H.probe->dev(), H.indexed->dev(), ctx->dev_bounds, sl.d_probes.as<uint32_t>(),
                    sl.d_offsets.as<uint32_t>(), sl.d_cands.as<uint32_t>(), (uint32_t)np, (uint64_t)nc,
                    dev_out(ctx)};
  hs::launch_verify(ctx->stream, va, H.scenario);
  HS_CHECK_HIP(hipEventRecord(sl.done, ctx->stream));
  sl.used = true;
}

// CPU candidate generation for probes [first, last] (hybrid::allpairs and
// hybrid::ppjoin, include/hyset/hybrid_algorithms.hpp:209-387): every
// posting of a probe prefix token inside the length window, deduplicated
// per probe, no verification; PPJoin takes a candidate at a common prefix
// token only if the positional filter passes there (:333-348).  Lists are
// id-sorted, so the window start is found by binary search instead of the
// reference's linear skip.
void host_filter(HybridShared& H, uint32_t first, uint32_t last, uint32_t& next_slot) {
  hs_ctx* ctx = H.ctx;
  const hs_coll& P = *H.probe;
  const hs_coll& S = *H.indexed;
  const auto& T = ctx->tables;
  std::vector<uint32_t> stamp(S.n, 0xFFFFFFFFu);
  std::vector<uint32_t> probes, offsets, cands;
  probes.reserve(1 << 16);
  offsets.reserve(1 << 16);
  cands.reserve(std::min<uint64_t>(H.batch, 1ull << 26));
  uint64_t n_cand = 0, n_visit = 0;
  const uint32_t last_len = (uint32_t)S.h_size_first.size() - 1;
  const uint32_t* metas = H.metas;
  for (uint32_t p = first; p <= last && first <= last; ++p) {
    const uint32_t l = P.h_sizes[p];
    const uint32_t* rt = P.h_tokens.data() + P.h_starts[p];
    const uint32_t mins = std::min(T.minsize[l], last_len);
    const uint32_t maxs1 = std::min(T.maxsize[l], last_len - 1) + 1;
    uint32_t lo = std::max(S.h_size_first[mins], H.ir.first);
    uint32_t hi = std::min(S.h_size_first[maxs1], H.ir.last + 1);
    if (H.self) hi = std::min(hi, p);
    if (lo >= hi) continue;
    const uint32_t pre = T.maxprefix[l];
    const size_t row_begin = cands.size();
    for (uint32_t j = 0; j < pre; ++j) {
      const uint32_t tok = rt[j];
      if (tok >= S.universe) continue;
      const uint32_t* a = H.ids + H.off[tok];
      const uint32_t* e = H.ids + H.off[tok + 1];
      for (const uint32_t* it = std::lower_bound(a, e, lo); it != e && *it < hi; ++it) {
        ++n_visit;
        if (stamp[*it] == p) continue;
        if (metas) {
          const uint32_t ip = metas[it - H.ids] & 0xFFFFu;  // saturated: kept
          if (ip != 0xFFFFu) {
            const uint32_t ls = S.h_sizes[*it];
            if (T.minoverlap(l, ls) > std::min(l - j, ls - ip)) continue;
          }
        }
        stamp[*it] = p;
        cands.push_back(*it);
      }
    }
    if (cands.size() > row_begin) {
      probes.push_back(p);
      offsets.push_back((uint32_t)cands.size());
      n_cand += cands.size() - row_begin;
      if (cands.size() >= H.batch) {
        submit_batch(H, probes, offsets, cands, next_slot);
        probes.clear();
        offsets.clear();
        cands.clear();
      }
    }
  }
  submit_batch(H, probes, offsets, cands, next_slot);
  H.candidates += n_cand;
  H.visits += n_visit;
}

// Probe partition boundaries: sqrt-spaced for a self-join, where the work
// of probe r grows with r (src/cpu_multi.cu:118-139), even otherwise.
std::vector<uint32_t> probe_splits(const hs_range& pr, uint32_t parts, bool self) {
  std::vector<uint32_t> pts;
  const uint64_t n = (uint64_t)pr.last - pr.first + 1;
  pts.push_back(pr.first);
  for (uint32_t i = 1; i < parts; ++i) {
    double f = self ? std::sqrt((double)i / parts) : (double)i / parts;
    pts.push_back(pr.first + (uint32_t)(n * f));
  }
  pts.push_back(pr.last + 1);
  return pts;
}

}  // namespace

void hs::hybrid_filter_verify(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, bool self, const hs_range& ir,
                              const hs_range& pr, const hs_tune& tune, const uint32_t* off, const uint32_t* ids,
                              const uint32_t* metas, HostCounts* counts) {
  HybridShared H;
  H.ctx = ctx;
  H.indexed = indexed;
  H.probe = probe;
  H.self = self;
  H.ir = ir;
  H.pr = pr;
  H.off = off;
  H.ids = ids;
  H.metas = metas;
  const double avg = probe->n ? (double)probe->n_tokens / probe->n : 0.0;
  H.scenario = tune.verify_scenario ? (int)tune.verify_scenario : (avg > 64.0 ? 2 : 1);
  H.batch = tune.batch_candidates ? tune.batch_candidates : (1ull << 22);
  for (auto& sl : ctx->slots)
    if (!sl.done) HS_CHECK_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
  const uint32_t threads = std::max<uint32_t>(1, tune.host_threads);
  if (!range_empty(pr)) {
    if (threads == 1) {
      uint32_t slot = 0;
      host_filter(H, pr.first, pr.last, slot);
    } else {
      std::vector<uint32_t> pts = probe_splits(pr, threads, self);
      std::vector<std::thread> pool;
      std::vector<Error> errs;
      std::mutex emu;
      for (uint32_t i = 0; i < threads; ++i) {
        if (pts[i + 1] <= pts[i]) continue;
        pool.emplace_back([&, i] {
          try {
            uint32_t slot = i & 1u;
            host_filter(H, pts[i], pts[i + 1] - 1, slot);
          } catch (const Error& e) {
            std::lock_guard<std::mutex> g(emu);
            errs.push_back(e);
          } catch (...) {
            std::lock_guard<std::mutex> g(emu);
            errs.push_back(Error{HS_E_INVALID, "host filter failed"});
          }
        });
      }
      for (auto& th : pool) th.join();
      if (!errs.empty()) throw errs.front();
    }
  }
  if (counts) {
    counts->candidates += H.candidates.load();
    counts->visits += H.visits.load();
  }
}

namespace {

void run_hybrid(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, bool self, const hs_range& ir, const hs_range& pr,
                const hs_tune& tune) {
  // host copy of the device-built index (the reference builds its CPU index
  // on the device and copies it back, index.hpp:250-282)
  const uint32_t U = indexed->universe;
  ctx->h_offsets.resize((size_t)U + 1);
  HS_CHECK_HIP(hipMemcpyAsync(ctx->h_offsets.data(), ctx->ib.offsets.ptr, ((size_t)U + 1) * 4, hipMemcpyDeviceToHost,
                              ctx->stream));
  HS_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  const uint32_t entries = ctx->h_offsets[U];
  ctx->h_set_ids.resize(std::max<uint32_t>(entries, 1));
  if (entries)
    HS_CHECK_HIP(hipMemcpyAsync(ctx->h_set_ids.data(), ctx->ib.set_ids.ptr, (size_t)entries * 4,
                                hipMemcpyDeviceToHost, ctx->stream));
  const bool ppjoin = ctx->host_algo == HS_HOST_PPJOIN;
  if (ppjoin) {  // positions for the positional filter
    ctx->h_metas.resize(std::max<uint32_t>(entries, 1));
    if (entries)
      HS_CHECK_HIP(hipMemcpyAsync(ctx->h_metas.data(), ctx->ib.metas.ptr, (size_t)entries * 4, hipMemcpyDeviceToHost,
                                  ctx->stream));
  }
  HS_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  ctx->stats.index_entries = entries;

  HostCounts hc;
  auto t0 = std::chrono::steady_clock::now();
  hs::hybrid_filter_verify(ctx, indexed, probe, self, ir, pr, tune, ctx->h_offsets.data(), ctx->h_set_ids.data(),
                           ppjoin ? ctx->h_metas.data() : nullptr, &hc);
  HS_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  auto t1 = std::chrono::steady_clock::now();
  ctx->stats.host_filter_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  ctx->stats.candidates = hc.candidates;
  ctx->stats.posting_visits = hc.visits;
}

void check_pair(hs_ctx* ctx, hs_coll* a, hs_coll* b) {
  HS_REQUIRE(ctx && a && b, HS_E_INVALID, "NULL context or collection");
  HS_REQUIRE(ctx->device >= 0, HS_E_INVALID, "a host-only context runs hs_join_host only");
  HS_REQUIRE(a->n_joins == 0 && b->n_joins == 0, HS_E_INVALID, "batched collections join with hs_join_batch");
  HS_REQUIRE(a->device == ctx->device && b->device == ctx->device, HS_E_INVALID,
             "collection lives on another device");
}

}  // namespace

// ================================================================ C ABI

extern "C" {

int hs_create(int device, hs_ctx** out) {
  if (!out) return HS_E_INVALID;
  *out = nullptr;
  hs_ctx* ctx = nullptr;
  try {
    ctx = new hs_ctx();
  } catch (...) {
    return HS_E_OOM;
  }
  ctx->device = device;
  if (device < 0) {  // host-only context: CPU joins, bounds and cost models
    *out = ctx;
    return HS_OK;
  }
  int rc = guarded(ctx, [&] {
    int count = 0;
    HS_CHECK_HIP(hipGetDeviceCount(&count));
    HS_REQUIRE(device >= 0 && device < count, HS_E_INVALID, "no such device");
    HS_CHECK_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    for (auto& e : ctx->ev) HS_CHECK_HIP(hipEventCreate(&e));
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
      ctx->num_cus = cus;
    ctx->counters.ensure(hs::CNT_WORDS * sizeof(unsigned long long));
    ctx->h_counters.ensure(hs::CNT_WORDS * sizeof(unsigned long long));
    ensure_output(ctx, 1u << 20);
    // the allocation-free join's scratch address on this device (once per
    // process; if the runtime refuses, hs_join_noalloc reports it)
    (void)hs::noalloc_init(device);
  });
  if (rc != HS_OK) {
    hs_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return HS_OK;
}

void hs_destroy(hs_ctx* ctx) {
  if (!ctx) return;
  if (ctx->device < 0) {
    delete ctx;
    return;
  }
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  if (ctx->jgraph_exec) (void)hipGraphExecDestroy(ctx->jgraph_exec);
  if (ctx->jgraph) (void)hipGraphDestroy(ctx->jgraph);
  for (auto& sl : ctx->slots)
    if (sl.done) (void)hipEventDestroy(sl.done);
  for (auto& e : ctx->stage_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->claim_ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->icum_ev) (void)hipEventDestroy(ctx->icum_ev);
  if (ctx->index_built) (void)hipEventDestroy(ctx->index_built);
  if (ctx->index_copied) (void)hipEventDestroy(ctx->index_copied);
  if (ctx->copy_stream) {
    (void)hipStreamSynchronize(ctx->copy_stream);
    (void)hipStreamDestroy(ctx->copy_stream);
  }
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* hs_last_error(const hs_ctx* ctx) { return ctx ? ctx->err.c_str() : "NULL context"; }

int hs_upload_collection(hs_ctx* ctx, const uint32_t* tokens, const uint32_t* starts, const uint32_t* sizes,
                         uint32_t n, uint64_t n_tokens, hs_coll** out) {
  if (!ctx || !out) return HS_E_INVALID;
  *out = nullptr;
  std::unique_ptr<hs_coll> c;
  int rc = guarded(ctx, [&] {
    HS_REQUIRE(n == 0 || (sizes && tokens), HS_E_INVALID, "NULL arrays");
    HS_REQUIRE(n_tokens < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "collections are limited to 2^32-1 tokens");
    c.reset(new hs_coll());
    c->device = ctx->device;
    c->n = n;
    c->h_sizes.assign(sizes, sizes + n);
    c->h_starts.resize(n);
    uint64_t pos = 0;
    uint32_t maxtok = 0, maxlen = 0, prev = 0;
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t l = sizes[i];
      HS_REQUIRE(l >= 1, HS_E_INVALID, "empty set at id " + std::to_string(i));
      HS_REQUIRE(l >= prev, HS_E_INVALID, "sets are not sorted by size at id " + std::to_string(i));
      prev = l;
      const uint64_t src = starts ? starts[i] : pos;
      HS_REQUIRE(src + l <= n_tokens, HS_E_INVALID, "set " + std::to_string(i) + " exceeds the token array");
      const uint32_t* t = tokens + src;
      for (uint32_t k = 1; k < l; ++k)
        HS_REQUIRE(t[k - 1] < t[k], HS_E_INVALID, "tokens not strictly ascending in set " + std::to_string(i));
      maxtok = std::max(maxtok, t[l - 1]);
      maxlen = std::max(maxlen, l);
      c->h_starts[i] = (uint32_t)pos;
      pos += l;
    }
    HS_REQUIRE(maxtok < 0xFFFFFFFFu, HS_E_UNSUPPORTED, "token id 2^32-1 is reserved");
    c->n_tokens = pos;
    c->maxlen = maxlen;
    c->universe = n ? maxtok + 1 : 1;
    c->h_tokens.resize(pos);
    for (uint32_t i = 0; i < n; ++i) {
      const uint64_t src = starts ? starts[i] : c->h_starts[i];
      std::memcpy(c->h_tokens.data() + c->h_starts[i], tokens + src, sizeof(uint32_t) * sizes[i]);
    }
    c->h_size_first.assign((size_t)maxlen + 2, n);
    for (uint32_t l = 0, i = 0; l <= maxlen + 1; ++l) {
      while (i < n && c->h_sizes[i] < l) ++i;
      c->h_size_first[l] = i;
    }
    // tile plan for the prefix-filter kernel: greedy runs of at most
    // TILE_PROBES probes and TILE_TOKENS tokens (a longer set gets a tile
    // of its own)
    c->h_tiles.clear();
    for (uint32_t i = 0; i < n;) {
      c->h_tiles.push_back(i);
      uint64_t tok = 0;
      uint32_t cnt = 0;
      while (i < n && cnt < hs::TILE_PROBES && (cnt == 0 || tok + c->h_sizes[i] <= hs::TILE_TOKENS)) {
        tok += c->h_sizes[i];
        ++cnt;
        ++i;
      }
    }
    c->h_tiles.push_back(n);
    if (ctx->device < 0) return;  // host-only collection
    c->tiles.ensure(c->h_tiles.size() * 4);
    c->tokens.ensure(std::max<uint64_t>(pos, 1) * 4);
    c->starts.ensure(std::max<uint32_t>(n, 1) * 4);
    c->sizes.ensure(std::max<uint32_t>(n, 1) * 4);
    c->size_first.ensure(c->h_size_first.size() * 4);
    if (pos) HS_CHECK_HIP(hipMemcpy(c->tokens.ptr, c->h_tokens.data(), pos * 4, hipMemcpyHostToDevice));
    if (n) {
      HS_CHECK_HIP(hipMemcpy(c->starts.ptr, c->h_starts.data(), (size_t)n * 4, hipMemcpyHostToDevice));
      HS_CHECK_HIP(hipMemcpy(c->sizes.ptr, c->h_sizes.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    }
    HS_CHECK_HIP(hipMemcpy(c->size_first.ptr, c->h_size_first.data(), c->h_size_first.size() * 4,
                           hipMemcpyHostToDevice));
    HS_CHECK_HIP(hipMemcpy(c->tiles.ptr, c->h_tiles.data(), c->h_tiles.size() * 4, hipMemcpyHostToDevice));
  });
  if (rc == HS_OK) *out = c.release();
  return rc;
}

void hs_free_collection(hs_coll* coll) {
  if (!coll) return;
  if (coll->device >= 0) {
    (void)hipSetDevice(coll->device);
    (void)hipDeviceSynchronize();
  }
  delete coll;
}

int hs_set_host_algorithm(hs_ctx* ctx, int32_t algorithm) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    HS_REQUIRE(algorithm == HS_HOST_ALLPAIRS || algorithm == HS_HOST_PPJOIN, HS_E_INVALID,
               "host algorithm must be HS_HOST_ALLPAIRS or HS_HOST_PPJOIN");
    ctx->host_algo = algorithm;
  });
}

int hs_set_threshold(hs_ctx* ctx, int32_t sim, double threshold, int32_t bounds_mode) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    hs::BoundTables probe_tables;
    probe_tables.build(sim, bounds_mode, threshold, 4);  // validates the combination
    ctx->kind = sim;
    ctx->mode = bounds_mode;
    ctx->t = threshold;
    ctx->have_threshold = true;
    ctx->tables_valid = false;
  });
}

}  // extern "C"

// ------------------------------------------- device join, in two steps

void hs::prepare_device_join(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, hs_range ir, hs_range pr, int32_t algo,
                             const hs_tune& tune, bool with_index, DevicePlan& P) {
  hipStream_t st = ctx->stream;
  // index_ms: ev[1] takes the end of the index build (stage timing only)
  const bool timed = ctx->stage_times && !ctx->capturing_join;
  P = DevicePlan{};
  const bool self = indexed == probe;
  P.self = self;
  P.ir = ir;
  // the token-major plan: a fgss self-join of one range with itself, its
  // index needed by nobody else (co-processing reads the index prefixes on
  // the host), sizes that fit the 16-bit fields of the filter's metas, and
  // entries whose fields fit one 64-bit word
  P.token_major = algo == HS_ALGO_FGSS && self && !with_index && !range_empty(ir) && pr.first == ir.first &&
                  pr.last == ir.last && tune.probes_per_tile == 0 &&
                  !(tune.flags & (HS_TUNE_STATIC_TILES | HS_TUNE_TILE_QUEUE | HS_TUNE_BALANCED_TILES |
                                  HS_TUNE_ITEM_MAJOR)) &&
                  indexed->maxlen < 0xFFFFu;
  if (P.token_major) {
    const auto& sf = indexed->h_size_first;
    uint64_t probe_entries = 0;
    for (uint32_t l = 0; l + 1 < (uint32_t)sf.size(); ++l) {
      const uint32_t lo = std::max(ir.first, sf[l]);
      const uint32_t hi = std::min(ir.last + 1, sf[l + 1]);
      if (hi > lo) probe_entries += (uint64_t)(hi - lo) * ctx->tables.maxprefix[l];
    }
    P.token_major = hs::tok_pack_plan(indexed->n, indexed->maxlen, indexed->universe, probe_entries, &P.pk);
  }
  // ---- parity signatures for the fgss signature test (timed with the index)
  bool sig_in_emit = false;  // the index's emit pass computes them (self-join over the whole collection)
  if (algo == HS_ALGO_FGSS && !(tune.flags & HS_TUNE_NO_SIGNATURE) && !range_empty(ir) && !range_empty(pr)) {
    P.sig_words = hs::fgss_signature_words(indexed, tune.signature_bits);
    sig_in_emit = P.sig_words == 1 && self && ir.first == 0 && ir.last + 1 == indexed->n;
    if (sig_in_emit) {
      indexed->sigs.ensure((size_t)std::max<uint32_t>(indexed->n, 1) * 8);
      indexed->sig_words = 1;
    } else {
      hs::compute_signatures(st, indexed, P.sig_words);
      if (!self) hs::compute_signatures(st, probe, P.sig_words);
    }
  }

  // ---- index over the indexed range (not needed by the bitmap technique)
  // the index prefix of a self-join is midprefix; the token-major plan
  // indexes the probe prefixes (maxprefix), of which those are a part
  const bool mid_index = self && !P.token_major;
  const uint32_t* prefix_tab = mid_index ? ctx->dev_bounds.midprefix : ctx->dev_bounds.maxprefix;
  if ((algo != HS_ALGO_BITMAP || with_index) && !range_empty(ir)) {
    // entries per size class of the range (sets are sorted by size)
    const auto& ptab = mid_index ? ctx->tables.midprefix : ctx->tables.maxprefix;
    const auto& sf = indexed->h_size_first;
    const uint32_t nl = (uint32_t)sf.size();  // maxlen + 2
    std::vector<uint32_t>& cum = ctx->jcum_tmp;
    cum.resize(nl);
    uint64_t acc = 0, side = 0;
    for (uint32_t l = 0; l + 1 < nl; ++l) {
      cum[l] = (uint32_t)acc;
      const uint32_t lo = std::max(ir.first, sf[l]);
      const uint32_t hi = std::min(ir.last + 1, sf[l + 1]);
      if (hi > lo) {
        acc += (uint64_t)(hi - lo) * ptab[l];
        side += (uint64_t)(hi - lo) * (self ? ctx->tables.midprefix[l] : ctx->tables.maxprefix[l]);
      }
      HS_REQUIRE(acc < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "index limited to 2^32-1 entries");
    }
    cum[nl - 1] = (uint32_t)acc;
    P.index_entries = acc;
    P.index_side_entries = side;
    ctx->d_jcum.ensure((size_t)nl * 4);
    // the same table as the one the device holds: no copy (repeated joins)
    const bool held = !ctx->capturing_join && ctx->jcum_valid && ctx->jcum_epoch == hs::buffer_epoch() &&
                      ctx->jcum_host == cum;
    if (!held) {
      PinnedBuf& hcum = ctx->capturing_join ? ctx->jgraph_cum : ctx->h_cum;
      hcum.ensure((size_t)nl * 4);
      std::memcpy(hcum.ptr, cum.data(), (size_t)nl * 4);
      HS_CHECK_HIP(hipMemcpyAsync(ctx->d_jcum.ptr, hcum.ptr, (size_t)nl * 4, hipMemcpyHostToDevice, st));
      ctx->jcum_valid = !ctx->capturing_join;
      ctx->jcum_host = cum;
      ctx->jcum_epoch = hs::buffer_epoch();
    }
    if (P.token_major) {
      hs::build_token_index(st, indexed->dev(), ir.first, ir.last, prefix_tab, indexed->size_first.as<uint32_t>(),
                            ctx->d_jcum.as<uint32_t>(), acc, P.pk, ctx->ib, ctx->dev_bounds.midprefix,
                            sig_in_emit ? indexed->sigs.as<uint64_t>() : nullptr,
                            ctx->counters.as<unsigned long long>(), (uint32_t)ctx->num_cus,
                            timed ? ctx->ev[1] : nullptr);
      P.counters_ready = true;
    } else {
      hs::build_index(st, indexed->dev(), ir.first, ir.last, prefix_tab, indexed->size_first.as<uint32_t>(),
                      ctx->d_jcum.as<uint32_t>(), indexed->universe, acc, ctx->ib,
                      P.sig_words == 1 ? indexed->sigs.as<uint64_t>() : nullptr,
                      sig_in_emit ? indexed->sigs.as<uint64_t>() : nullptr,
                      algo == HS_ALGO_FGSS ? ctx->counters.as<unsigned long long>() : nullptr,
                      timed ? ctx->ev[1] : nullptr);
      P.counters_ready = algo == HS_ALGO_FGSS && acc > 0 && indexed->n > 0;
    }
  } else {
    ctx->ib.universe = 0;
    if (timed) HS_CHECK_HIP(hipEventRecord(ctx->ev[1], st));
  }
}

uint32_t hs::launch_device_join(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, DevicePlan& P, hs_range pr,
                                int32_t algo, const hs_tune& tune) {
  hipStream_t st = ctx->stream;
  const bool self = P.self;
  const hs_range ir = P.ir;
  uint32_t tile_plan = 0;
  const bool timed = ctx->stage_times && !ctx->capturing_join;
  if (range_empty(pr) || range_empty(ir)) {
    if (timed) {
      HS_CHECK_HIP(hipEventRecord(ctx->ev[5], st));
      HS_CHECK_HIP(hipEventRecord(ctx->ev[2], st));
    }
  } else if (algo == HS_ALGO_FGSS && P.token_major) {
    hs::TokArgs a{};
    a.c = probe->dev();
    a.ent = ctx->ib.packed_sorted;
    a.pk = P.pk;
    a.n_ent = (uint32_t)P.index_entries;
    a.bounds = ctx->dev_bounds;
    a.size_first = indexed->size_first.as<uint32_t>();
    a.size_first_len = (uint32_t)indexed->h_size_first.size();
    a.first = pr.first;
    a.last = pr.last;
    a.sig_words = P.sig_words;
    a.sigs = P.sig_words ? indexed->sigs.as<uint64_t>() : nullptr;
    // survivors of both filters: a first guess of half the entries, grown
    // after a join that found more (the join then runs again)
    if (ctx->tok_queue_cap == 0) {
      const uint64_t cap = std::max<uint64_t>(1u << 18, P.index_entries / 2);
      ctx->tok_queue.ensure(cap * sizeof(uint4));
      ctx->tok_queue_cap = (uint32_t)std::min<uint64_t>(ctx->tok_queue.bytes / sizeof(uint4), 0xFFFFFFFFull);
    }
    a.queue = ctx->tok_queue.as<uint4>();
    a.queue_cap = ctx->tok_queue_cap;
    // chunks of heavy blocks: a first guess of one per block, grown like the queue
    if (ctx->tok_heavy_cap == 0) {
      const uint64_t cap = std::max<uint64_t>(4096, (P.index_entries + hs::TOK_BLOCK - 1) / hs::TOK_BLOCK);
      ctx->tok_heavy.ensure(cap * sizeof(uint4));
      ctx->tok_heavy_cap = (uint32_t)std::min<uint64_t>(ctx->tok_heavy.bytes / sizeof(uint4), 0xFFFFFFFFull);
    }
    a.heavy = ctx->tok_heavy.as<uint4>();
    a.heavy_cap = ctx->tok_heavy_cap;
    ctx->tok_aoff.ensure(std::max<uint64_t>(P.index_entries, 1) * sizeof(uint2));
    a.aoff = ctx->tok_aoff.as<uint2>();
    ctx->tok_hdrs.ensure(std::max<uint64_t>((P.index_entries + hs::TOK_BLOCK - 1) / hs::TOK_BLOCK, 1) *
                         sizeof(uint2));
    a.hdrs = ctx->tok_hdrs.as<uint2>();
    if (ctx->stage_times && !ctx->capturing_join) {
      // stage-timed joins also count |r| over the distinct probes with candidates
      if (ctx->tok_stamp_n < probe->n || ctx->tok_epoch == 0xFFFFFFFFu) {
        ctx->tok_stamp.ensure((size_t)std::max<uint32_t>(probe->n, 1) * 4);
        ctx->tok_stamp_n = ctx->tok_stamp.bytes / 4;
        hs::launch_fill_u32(st, ctx->tok_stamp.as<uint32_t>(), ctx->tok_stamp_n, 0);
        ctx->tok_epoch = 0;
      }
      a.stamp = ctx->tok_stamp.as<uint32_t>();
      a.epoch = ++ctx->tok_epoch;
    }
    a.out = dev_out(ctx);
    tile_plan = HS_PLAN_TOKEN;
    hs::launch_tok_join(st, a, (uint32_t)ctx->num_cus, timed ? ctx->ev[5] : nullptr, timed ? ctx->ev[2] : nullptr,
                        timed ? &ctx->ev[6] : nullptr);
  } else if (algo == HS_ALGO_FGSS) {
    hs::FgssArgs a{};
    a.probe = probe->dev();
    a.indexed = indexed->dev();
    a.index = hs::dev_index(ctx->ib);
    a.bounds = ctx->dev_bounds;
    a.size_first = indexed->size_first.as<uint32_t>();
    a.size_first_len = (uint32_t)indexed->h_size_first.size();
    a.probe_first = pr.first;
    a.probe_last = pr.last;
    a.idx_first = ir.first;
    a.idx_last = ir.last;
    a.self = self;
    a.out = dev_out(ctx);
    a.sig_words = P.sig_words;
    a.probe_sigs = P.sig_words ? probe->sigs.as<uint64_t>() : nullptr;
    a.indexed_sigs = P.sig_words ? indexed->sigs.as<uint64_t>() : nullptr;
    const uint32_t ppt = tune.probes_per_tile;
    if (ppt == 0) {
      // the collection's tile plan, clipped to the probe range
      const auto& T = probe->h_tiles;
      const uint32_t ft = (uint32_t)(std::upper_bound(T.begin(), T.end(), pr.first) - T.begin()) - 1;
      const uint32_t lt = (uint32_t)(std::upper_bound(T.begin(), T.end(), pr.last) - T.begin()) - 1;
      a.tile_starts = probe->tiles.as<uint32_t>();
      a.first_tile = ft;
      a.n_plan_tiles = lt - ft + 1;
    }
    uint32_t wgs = 0;
    // Tile plan.  The work-balanced plan costs a pass over every probe-prefix
    // item; it pays when tiles differ much in work, which follows the mean
    // posting-list length (index entries per token of the universe): the
    // visits of a tile grow with it.  Measured on 1M sets (sizes median 10,
    // Zipf 1.1, 647,502 distinct tokens): upload-time tiles win at mean length
    // 1.8 and 2.9 (Jaccard 0.9, 0.8), balanced tiles at 4.1 and 5.6 (0.7, 0.6)
    // and on 200K x 200K sets of 100-700 tokens (length ~600); small joins keep
    // the upload-time tiles.
    bool balanced = false;
    const double mean_list = (double)P.index_entries / std::max<uint32_t>(indexed->universe, 1);
    if (ppt == 0 && (uint64_t)(pr.last - pr.first + 1) * 1024 < 0xFFFFFFFFull && probe->maxlen < 0xFFFFu) {
      if (tune.flags & HS_TUNE_BALANCED_TILES) {
        balanced = true;
      } else if (!(tune.flags & (HS_TUNE_STATIC_TILES | HS_TUNE_TILE_QUEUE))) {
        const auto& psf = probe->h_size_first;
        uint64_t items = 0;
        for (uint32_t l = 0; l + 1 < (uint32_t)psf.size(); ++l) {
          const uint32_t lo = std::max(pr.first, psf[l]), hi = std::min(pr.last + 1, psf[l + 1]);
          if (hi > lo) items += (uint64_t)(hi - lo) * ctx->tables.maxprefix[std::min(l, ctx->tables.L)];
        }
        balanced = items >= (1u << 21) && mean_list >= 3.5;
      }
    }
    // Very long posting lists (tens of entries per token and more) put one
    // heavy probe in each balanced tile; such tiles run in 256-lane
    // workgroups, where a round of visits waits on 4 waves' random
    // signature gathers instead of 8 (measured on 200K x 200K sets of
    // 100-700 tokens, ~580 entries per token).
    a.narrow_tiles = balanced && mean_list >= 32.0;
    uint32_t nl_plan = 0;
    uint64_t items_plan = 0;
    if (balanced) {
      // item slices first, then tiles of equal work claimed by a persistent grid
      // per size class of the probe range: first item slot, first probe,
      // window of indexed ids, maxprefix
      const auto& T = ctx->tables;
      const auto& psf = probe->h_size_first;
      const auto& isf = indexed->h_size_first;
      const uint32_t nl = (uint32_t)psf.size();
      const uint32_t ilast = (uint32_t)isf.size() - 1;
      std::vector<uint32_t>& tab = ctx->icum_tmp;
      tab.resize((size_t)nl * 5);
      uint64_t items = 0;
      for (uint32_t l = 0; l < nl; ++l) {
        tab[l] = (uint32_t)items;
        tab[nl + l] = std::max(pr.first, psf[l]);
        const uint32_t mins = std::min(T.minsize[std::min(l, T.L)], ilast);
        const uint32_t maxs1 = std::min(T.maxsize[std::min(l, T.L)], ilast - 1) + 1;
        tab[2 * nl + l] = std::max(isf[mins], ir.first);
        tab[3 * nl + l] = std::min(isf[maxs1], ir.last + 1);
        tab[4 * nl + l] = T.maxprefix[std::min(l, T.L)];
        if (l + 1 < nl) {
          const uint32_t hi = std::min(pr.last + 1, psf[l + 1]);
          if (hi > tab[nl + l]) items += (uint64_t)(hi - tab[nl + l]) * tab[4 * nl + l];
          HS_REQUIRE(items < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "probe range limited to 2^32-1 prefix items");
        }
      }
      ctx->d_icum.ensure((size_t)nl * 20);
      // the table the device already holds (repeated joins): no upload, and
      // no marker event between the index build and the planning kernels
      const bool held = !ctx->capturing_join && ctx->icum_valid && ctx->icum_epoch == hs::buffer_epoch() &&
                        ctx->icum_host == tab;
      if (!held) {
        // a previous launch's copy out of the staging table must have run
        // before the table is rewritten (launches may be queued back to back)
        if (ctx->icum_pending) HS_CHECK_HIP(hipEventSynchronize(ctx->icum_ev));
        ctx->icum_pending = false;
        ctx->h_icum.ensure((size_t)nl * 20);
        std::memcpy(ctx->h_icum.ptr, tab.data(), (size_t)nl * 20);
        HS_CHECK_HIP(hipMemcpyAsync(ctx->d_icum.ptr, ctx->h_icum.ptr, (size_t)nl * 20, hipMemcpyHostToDevice, st));
        if (!ctx->icum_ev) HS_CHECK_HIP(hipEventCreateWithFlags(&ctx->icum_ev, hipEventDisableTiming));
        HS_CHECK_HIP(hipEventRecord(ctx->icum_ev, st));
        ctx->icum_pending = true;
        ctx->icum_valid = !ctx->capturing_join;
        ctx->icum_host = tab;
        ctx->icum_epoch = hs::buffer_epoch();
      }
      nl_plan = nl;
      items_plan = items;
    }
    if (ppt == 0 && (balanced || (tune.flags & HS_TUNE_TILE_QUEUE))) {
      // a persistent grid claims the tiles from per-XCD queues
      ctx->batch_queue.ensure(8 * sizeof(uint32_t));
      a.tile_queue = ctx->batch_queue.as<uint32_t>();
      // as many workgroups per CU as fit: two of 512 lanes, four of 256
      wgs = (uint32_t)ctx->num_cus * (tune.workgroups_per_cu ? tune.workgroups_per_cu : a.narrow_tiles ? 4 : 2);
    }
    if (balanced) {
      hs::BalancedPlan bp{ctx->d_icum.as<uint32_t>(), nl_plan, items_plan,
                          &ctx->bal_slice_a, &ctx->bal_slice_n, &ctx->bal_cost, &ctx->bal_tiles,
                          &ctx->bal_ntiles, &ctx->ib.scan_scratch, &ctx->bal_rec, &ctx->bal_slot};
      hs::plan_balanced_tiles(st, a, bp);  // also zeroes the tile queue
    }
    tile_plan = ppt ? HS_PLAN_FIXED : balanced ? HS_PLAN_BALANCED
                                    : (tune.flags & HS_TUNE_TILE_QUEUE) ? HS_PLAN_QUEUE : HS_PLAN_UPLOAD;
    // kernel_ms: the k_fgss dispatch alone, timed by the dispatch itself
    hs::launch_fgss(st, a, ppt, wgs, timed ? ctx->ev[5] : nullptr, timed ? ctx->ev[2] : nullptr);
  } else if (algo == HS_ALGO_BITMAP) {
    // the width the fgss signature test would take: 64 bits for sets of a
    // few dozen tokens, up to 1024 for sets of hundreds
    const uint32_t bits = tune.bitmap_bits ? tune.bitmap_bits : 64 * hs::fgss_signature_words(indexed, 0);
    HS_REQUIRE(bits == 64 || bits == 128 || bits == 256 || bits == 512 || bits == 1024, HS_E_INVALID,
               "bitmap width must be 64, 128, 256, 512 or 1024");
    const uint32_t words = bits / 64;
    if (P.bitmap_words != words) {  // once per join, whatever the number of launches
      hs::compute_signatures(st, indexed, words);
      if (!self) hs::compute_signatures(st, probe, words);
      P.bitmap_words = words;
    }
    hs::BitmapArgs b{};
    b.probe = probe->dev();
    b.indexed = indexed->dev();
    b.bounds = ctx->dev_bounds;
    b.probe_sigs = probe->sigs.as<uint64_t>();
    b.indexed_sigs = indexed->sigs.as<uint64_t>();
    b.size_first = indexed->size_first.as<uint32_t>();
    b.size_first_len = (uint32_t)indexed->h_size_first.size();
    b.words = words;
    b.probe_first = pr.first;
    b.probe_last = pr.last;
    b.idx_first = ir.first;
    b.idx_last = ir.last;
    b.self = self;
    b.out = dev_out(ctx);
    // about 16 waves per SIMD: fewer probe tiles share their windows out
    const uint64_t waves = ((uint64_t)pr.last - pr.first + 1 + 63) / 64;
    b.parts = (uint32_t)std::clamp<uint64_t>((uint64_t)ctx->num_cus * 64 / std::max<uint64_t>(waves, 1), 1, 64);
    if (timed) HS_CHECK_HIP(hipEventRecord(ctx->ev[5], st));
    hs::launch_bitmap(st, b);
    if (timed) HS_CHECK_HIP(hipEventRecord(ctx->ev[2], st));
  } else {
    HS_CHECK_HIP(hipEventRecord(ctx->ev[5], st));
    run_hybrid(ctx, indexed, probe, self, ir, pr, tune);
    HS_CHECK_HIP(hipEventRecord(ctx->ev[2], st));
  }
  return tile_plan;
}

namespace {

// Identity of a join for graph replay: collections, ranges, technique,
// knobs, bound tables and result capacity.
std::string join_graph_key(const hs_ctx* ctx, const hs_coll* indexed, const hs_coll* probe, hs_range ir,
                           hs_range pr, int32_t algo, const hs_tune& t) {
  char buf[512];
  std::snprintf(buf, sizeof buf, "%p/%p/%u-%u/%u-%u/%d/%u/%u/%u/%u/%u/%u/%d/%d/%.17g/%u/%llu", (const void*)indexed,
                (const void*)probe, ir.first, ir.last, pr.first, pr.last, algo, t.bitmap_bits, t.probes_per_tile,
                t.tile_tokens, t.workgroups_per_cu, t.signature_bits, t.flags, ctx->kind, ctx->mode, ctx->t,
                ctx->tables.L, (unsigned long long)ctx->out_cap);
  return buf;
}

void drop_join_graph(hs_ctx* ctx) {
  if (ctx->jgraph_exec) (void)hipGraphExecDestroy(ctx->jgraph_exec);
  if (ctx->jgraph) (void)hipGraphDestroy(ctx->jgraph);
  ctx->jgraph_exec = nullptr;
  ctx->jgraph = nullptr;
  ctx->jgraph_key.clear();
}

// Records the join's device sequence (events, signatures, index, counters,
// launch) into ctx->jgraph.  The bound-entry table comes from a pinned
// buffer of the graph's own: a graph reads host memory when it replays.  A
// capture that fails leaves no graph and no error: the join's result is
// already there.
template <typename Reset>
bool capture_join_graph(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, hs_range ir, hs_range pr, int32_t algo,
                        const hs_tune& tune, const std::string& key, Reset&& reset_for_launch) {
  drop_join_graph(ctx);
  hipStream_t st = ctx->stream;
  ctx->jgraph_cum.ensure(indexed->h_size_first.size() * 4);
  HS_CHECK_HIP(hipStreamSynchronize(st));
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  hs::DevicePlan Q;
  uint32_t plan = 0;
  bool ok = true;
  ctx->capturing_join = true;
  try {
    // no event records in the graph: the replay is timed from outside it
    hs::prepare_device_join(ctx, indexed, probe, ir, pr, algo, tune, false, Q);
    if (!Q.counters_ready) reset_for_launch(Q);
    plan = hs::launch_device_join(ctx, indexed, probe, Q, pr, algo, tune);
  } catch (...) {
    ok = false;
  }
  ctx->capturing_join = false;
  hipGraph_t g = nullptr;
  if (hipStreamEndCapture(st, &g) != hipSuccess || !ok) {
    (void)hipGetLastError();
    if (g) (void)hipGraphDestroy(g);
    return false;
  }
  hipGraphExec_t ge = nullptr;
  if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipGraphDestroy(g);
    return false;
  }
  ctx->jgraph = g;
  ctx->jgraph_exec = ge;
  ctx->jgraph_key = key;
  ctx->jgraph_plan = Q;
  ctx->jgraph_tile_plan = plan;
  ctx->jgraph_epoch = buffer_epoch();
  return true;
}

}  // namespace

extern "C" {

int hs_join(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, const hs_range* indexed_range,
            const hs_range* probe_range, int32_t algo, const hs_tune* tune_in, hs_result* result) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    check_pair(ctx, indexed, probe);
    HS_REQUIRE(result, HS_E_INVALID, "NULL result");
    HS_REQUIRE(algo == HS_ALGO_FGSS || algo == HS_ALGO_BITMAP || algo == HS_ALGO_HYBRID, HS_E_INVALID,
               "unknown algorithm");
    hs_tune tune{};
    if (tune_in) tune = *tune_in;
    const bool self = indexed == probe;
    const hs_range ir = clamp_range(indexed_range, indexed->n);
    const hs_range pr = clamp_range(probe_range, probe->n);
    ensure_bounds(ctx, std::max(indexed->maxlen, probe->maxlen));
    ctx->stats = hs_stats{};
    hipStream_t st = ctx->stream;
    unsigned long long cnt[hs::CNT_WORDS];
    uint32_t tile_plan = 0;
    hs::DevicePlan P;
    // counters cleared and the index size placed where the kernels read it
    auto reset_for_launch = [&](const hs::DevicePlan& Q) {
      reset_counters(ctx);
      if (algo == HS_ALGO_FGSS && !Q.token_major && !range_empty(Q.ir))  // index size, read back with the counters
        HS_CHECK_HIP(hipMemcpyAsync(ctx->counters.as<unsigned long long>() + hs::CNT_ENTRIES,
                                    ctx->ib.offsets.as<uint32_t>() + ctx->ib.universe, 4, hipMemcpyDeviceToDevice,
                                    st));
    };
    // the result rows and (token-major) the survivor queue hold the counts
    auto fits = [&](const hs::DevicePlan& Q, const unsigned long long* c) {
      return c[hs::CNT_PAIRS] <= ctx->out_cap &&
             (!Q.token_major || (c[hs::CNT_QUEUE] <= ctx->tok_queue_cap && c[hs::CNT_HEAVY] <= ctx->tok_heavy_cap));
    };
    auto grow_for = [&](const hs::DevicePlan& Q, const unsigned long long* c) {
      if (c[hs::CNT_PAIRS] > ctx->out_cap) ensure_output(ctx, c[hs::CNT_PAIRS] + c[hs::CNT_PAIRS] / 4 + 1024);
      if (Q.token_major && c[hs::CNT_QUEUE] > ctx->tok_queue_cap) {
        const uint64_t cap = c[hs::CNT_QUEUE] + c[hs::CNT_QUEUE] / 4 + 1024;
        HS_REQUIRE(cap < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "more than 2^32 candidates");
        ctx->tok_queue.ensure(cap * sizeof(uint4));
        ctx->tok_queue_cap = (uint32_t)std::min<uint64_t>(ctx->tok_queue.bytes / sizeof(uint4), 0xFFFFFFFFull);
      }
      if (Q.token_major && c[hs::CNT_HEAVY] > ctx->tok_heavy_cap) {
        const uint64_t cap = c[hs::CNT_HEAVY] + c[hs::CNT_HEAVY] / 4 + 1024;
        HS_REQUIRE(cap < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "more than 2^32 chunks");
        ctx->tok_heavy.ensure(cap * sizeof(uint4));
        ctx->tok_heavy_cap = (uint32_t)std::min<uint64_t>(ctx->tok_heavy.bytes / sizeof(uint4), 0xFFFFFFFFull);
      }
    };
    // a launch sequence captured by an earlier call with the same join,
    // replayed while no buffer it recorded has moved (HS_TUNE_GRAPH)
    const bool want_graph = (tune.flags & HS_TUNE_GRAPH) && algo != HS_ALGO_HYBRID;
    const std::string key = want_graph ? join_graph_key(ctx, indexed, probe, ir, pr, algo, tune) : std::string();
    const bool replay = want_graph && ctx->jgraph_exec && ctx->jgraph_key == key && ctx->jgraph_epoch == buffer_epoch();
    uint32_t graph_state = 0;
    // Marker events between kernels cost the device a few microseconds of
    // idle time each, so by default the join is timed as a whole: ev[0]
    // before the first launch, ev[3] after the counters' copy back.  With
    // HS_TUNE_STAGE_TIMES ev[1] takes the end of the index build and
    // ev[5] / ev[2] the start and end of the main kernel (fgss: from the
    // dispatches themselves, bitmap and hybrid: markers).  The graph replay
    // is timed as a whole.
    const bool stages = (tune.flags & HS_TUNE_STAGE_TIMES) && algo != HS_ALGO_HYBRID;
    ctx->stage_times = stages || algo == HS_ALGO_HYBRID;
    // Pairs wanted on the host: a prefix of the result rows (as many as the
    // last such join found, and a quarter more) is stored into mapped host
    // memory by a kernel queued right after the launches, so it arrives with
    // the counters; only a larger result needs the second round trip of
    // copy_pairs.
    const bool want_pairs =
        !result->count_only && result->capacity > 0 && result->r && result->s && algo != HS_ALGO_HYBRID;
    uint64_t nspec = 0;
    auto speculate = [&]() {
      if (!want_pairs) return;
      nspec = std::min<uint64_t>(std::min<uint64_t>(result->capacity, ctx->out_cap),
                                 std::max<uint64_t>(ctx->spec_guess + ctx->spec_guess / 4, 4096));
      if (ctx->spec_cap < nspec) {
        HS_CHECK_HIP(hipStreamSynchronize(st));  // no copy may still land in the old buffer
        ctx->spec.ensure(nspec * 12);
        ctx->spec_cap = ctx->spec.bytes / 12;
      }
      void* dst = nullptr;
      HS_CHECK_HIP(hipHostGetDevicePointer(&dst, ctx->spec.ptr, 0));
      hs::launch_rows_to_host(st, ctx->out.as<uint32_t>(), ctx->out_cap, ctx->counters.as<unsigned long long>(), nspec,
                              static_cast<uint32_t*>(dst), ctx->spec_cap);
    };
    if (replay) {
      graph_state = 2;
      HS_CHECK_HIP(hipEventRecord(ctx->ev[0], st));
      HS_CHECK_HIP(hipGraphLaunch(ctx->jgraph_exec, st));
      ctx->jcum_valid = false;  // the graph rewrote the entry table
      P = ctx->jgraph_plan;
      tile_plan = ctx->jgraph_tile_plan;
      speculate();
      read_counters(ctx, cnt, ctx->ev[3]);
      HS_REQUIRE(cnt[hs::CNT_INDEX_ERR] == 0, HS_E_DEVICE, "token-major index sort did not complete");
      // more pairs (or survivors, or chunks) than the buffers hold: grow,
      // run again eagerly (a full chunk or survivor queue leaves the later
      // counts short, so up to three growths may follow one another)
      for (int again = 0; !fits(P, cnt); ++again) {
        HS_REQUIRE(again < 3, HS_E_DEVICE, "result count changed between runs");
        drop_join_graph(ctx);
        grow_for(P, cnt);
        reset_for_launch(P);
        tile_plan = hs::launch_device_join(ctx, indexed, probe, P, pr, algo, tune);
        speculate();
        read_counters(ctx, cnt, ctx->ev[3]);
      }
    } else {
      HS_CHECK_HIP(hipEventRecord(ctx->ev[0], st));
      hs::prepare_device_join(ctx, indexed, probe, ir, pr, algo, tune, false, P);

      // ---- candidate generation + verification
      // at most four runs: a full chunk queue leaves the survivor count
      // short and a full survivor queue the pair count, so the survivor
      // queue and the result rows may have to grow after it in turn
      for (int attempt = 0; attempt < 4; ++attempt) {
        if (attempt > 0 || !P.counters_ready) reset_for_launch(P);
        tile_plan = hs::launch_device_join(ctx, indexed, probe, P, pr, algo, tune);
        speculate();
        read_counters(ctx, cnt, ctx->ev[3]);
        HS_REQUIRE(attempt > 0 || cnt[hs::CNT_INDEX_ERR] == 0, HS_E_DEVICE,
                   "token-major index sort did not complete");
        if (fits(P, cnt)) break;
        HS_REQUIRE(attempt < 3, HS_E_DEVICE, "result count changed between runs");
        grow_for(P, cnt);
      }
      // Capture the sequence just run for the next calls.  The eager run
      // allocated every buffer, so the capture allocates nothing.  Plans that
      // stage per-launch tables through host buffers (balanced, tile queue)
      // stay eager.
      if (want_graph && (algo == HS_ALGO_BITMAP || tile_plan == HS_PLAN_UPLOAD || tile_plan == HS_PLAN_FIXED ||
                         tile_plan == HS_PLAN_TOKEN))
        if (capture_join_graph(ctx, indexed, probe, ir, pr, algo, tune, key, reset_for_launch)) graph_state = 1;
    }
    const uint64_t count = cnt[hs::CNT_PAIRS];
    set_device_result(ctx, count);
    result->count = count;
    // the stage times, work counters and plan, also when the pairs then
    // overflow the caller's arrays
    ctx->stats.total_ms = elapsed(ctx->ev[0], ctx->ev[3]);
    if (graph_state == 2) {
      ctx->stats.filter_ms = elapsed(ctx->ev[0], ctx->ev[3]);
    } else if (ctx->stage_times) {
      ctx->stats.index_ms = elapsed(ctx->ev[0], ctx->ev[1]);
      ctx->stats.filter_ms = elapsed(ctx->ev[1], ctx->ev[2]);
      // the token-major plan's three kernels: the sum of their dispatches'
      // own durations (the dispatches with events wait ~12 us between
      // them, which plain joins do not)
      ctx->stats.kernel_ms = tile_plan == HS_PLAN_TOKEN ? elapsed(ctx->ev[5], ctx->ev[6]) +
                                                              elapsed(ctx->ev[7], ctx->ev[8]) +
                                                              elapsed(ctx->ev[9], ctx->ev[2])
                                                        : elapsed(ctx->ev[5], ctx->ev[2]);
    }
    ctx->stage_times = false;
    ctx->stats.pairs = count;
    ctx->stats.graph = graph_state;
    if (algo != HS_ALGO_HYBRID) {
      ctx->stats.posting_visits = cnt[hs::CNT_VISITS];
      ctx->stats.candidates = cnt[hs::CNT_CANDS];
      ctx->stats.candidate_tokens = cnt[hs::CNT_CAND_TOKENS];
      ctx->stats.probe_prefix_tokens = cnt[hs::CNT_ITEMS];
    }
    if (algo == HS_ALGO_FGSS) {
      ctx->stats.tile_plan = tile_plan;
      // the index side's entries (the reference's index size); the
      // token-major index also holds the probe-only prefix entries
      ctx->stats.index_entries = P.token_major ? P.index_side_entries : cnt[hs::CNT_ENTRIES];
      ctx->stats.signature_tests = cnt[hs::CNT_SIG_TESTS];
      ctx->stats.signature_bits = P.sig_words * 64;
      ctx->stats.candidate_probe_tokens = cnt[hs::CNT_CAND_PROBES];
    }
    if (want_pairs) ctx->spec_guess = count;
    if (!result->count_only && count > 0) {
      HS_REQUIRE(count <= result->capacity, HS_E_OVERFLOW, "result capacity too small");
      if (count <= nspec) {  // the whole result came with the counters
        const uint32_t* rows = ctx->spec.as<uint32_t>();
        std::memcpy(result->r, rows, count * 4);
        std::memcpy(result->s, rows + ctx->spec_cap, count * 4);
        if (result->overlap) std::memcpy(result->overlap, rows + 2 * ctx->spec_cap, count * 4);
      } else {
        copy_pairs(ctx, count, result);
      }
      HS_CHECK_HIP(hipEventRecord(ctx->ev[4], st));
      HS_CHECK_HIP(hipEventSynchronize(ctx->ev[4]));
      ctx->stats.transfer_ms = elapsed(ctx->ev[3], ctx->ev[4]);
      ctx->stats.total_ms = elapsed(ctx->ev[0], ctx->ev[4]);
    }
  });
}

int hs_verify_candidates(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, const uint32_t* probes,
                         const uint32_t* offsets, const uint32_t* candidates, uint32_t n_probes, int32_t scenario,
                         hs_result* result) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    check_pair(ctx, indexed, probe);
    HS_REQUIRE(result, HS_E_INVALID, "NULL result");
    HS_REQUIRE(n_probes == 0 || (probes && offsets), HS_E_INVALID, "NULL batch arrays");
    const uint64_t nc = n_probes ? offsets[n_probes - 1] : 0;
    HS_REQUIRE(nc == 0 || candidates, HS_E_INVALID, "NULL candidates");
    uint32_t prev = 0;
    for (uint32_t i = 0; i < n_probes; ++i) {
      HS_REQUIRE(offsets[i] >= prev, HS_E_INVALID, "offsets must be non-decreasing");
      HS_REQUIRE(probes[i] < probe->n, HS_E_INVALID, "probe id out of range");
      prev = offsets[i];
    }
    for (uint64_t c = 0; c < nc; ++c) HS_REQUIRE(candidates[c] < indexed->n, HS_E_INVALID, "candidate out of range");
    ensure_bounds(ctx, std::max(indexed->maxlen, probe->maxlen));
    ctx->stats = hs_stats{};
    hipStream_t st = ctx->stream;
    HS_CHECK_HIP(hipEventRecord(ctx->ev[0], st));
    HS_CHECK_HIP(hipEventRecord(ctx->ev[1], st));
    HybridSlot& sl = ctx->slots[0];
    if (!sl.done) HS_CHECK_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    if (sl.used) HS_CHECK_HIP(hipEventSynchronize(sl.done));
    sl.d_probes.ensure((size_t)std::max<uint32_t>(n_probes, 1) * 4);
    sl.d_offsets.ensure((size_t)std::max<uint32_t>(n_probes, 1) * 4);
    sl.d_cands.ensure((size_t)std::max<uint64_t>(nc, 1) * 4);
    if (n_probes) {
      HS_CHECK_HIP(hipMemcpyAsync(sl.d_probes.ptr, probes, (size_t)n_probes * 4, hipMemcpyHostToDevice, st));
      HS_CHECK_HIP(hipMemcpyAsync(sl.d_offsets.ptr, offsets, (size_t)n_probes * 4, hipMemcpyHostToDevice, st));
    }
    if (nc) HS_CHECK_HIP(hipMemcpyAsync(sl.d_cands.ptr, candidates, nc * 4, hipMemcpyHostToDevice, st));
    const double avg = probe->n ? (double)probe->n_tokens / probe->n : 0.0;
    const int scen = scenario ? scenario : (avg > 64.0 ? 2 : 1);
    unsigned long long cnt[hs::CNT_WORDS];
    for (int attempt = 0; attempt < 2; ++attempt) {
      reset_counters(ctx);
      HS_CHECK_HIP(hipEventRecord(ctx->ev[5], st));
      hs::VerifyArgs va{probe->dev(), indexed->dev(), ctx->dev_bounds, sl.d_probes.as<uint32_t>(),
                        sl.d_offsets.as<uint32_t>(), sl.d_cands.as<uint32_t>(), n_probes, nc, dev_out(ctx)};
      hs::launch_verify(st, va, scen);
      HS_CHECK_HIP(hipEventRecord(ctx->ev[2], st));
      read_counters(ctx, cnt);
      if (cnt[hs::CNT_PAIRS] <= ctx->out_cap) break;
      HS_REQUIRE(attempt == 0, HS_E_DEVICE, "result count changed between runs");
      ensure_output(ctx, cnt[hs::CNT_PAIRS] + cnt[hs::CNT_PAIRS] / 4 + 1024);
    }
    sl.used = false;
    const uint64_t count = cnt[hs::CNT_PAIRS];
    set_device_result(ctx, count);
    result->count = count;
    HS_CHECK_HIP(hipEventRecord(ctx->ev[3], st));
    if (!result->count_only && count > 0) {
      HS_REQUIRE(count <= result->capacity, HS_E_OVERFLOW, "result capacity too small");
      copy_pairs(ctx, count, result);
    }
    HS_CHECK_HIP(hipEventRecord(ctx->ev[4], st));
    HS_CHECK_HIP(hipEventSynchronize(ctx->ev[4]));
    ctx->stats.filter_ms = elapsed(ctx->ev[1], ctx->ev[2]);
    ctx->stats.kernel_ms = elapsed(ctx->ev[5], ctx->ev[2]);
    ctx->stats.transfer_ms = elapsed(ctx->ev[3], ctx->ev[4]);
    ctx->stats.total_ms = elapsed(ctx->ev[0], ctx->ev[4]);
    ctx->stats.pairs = count;
    ctx->stats.candidates = nc;
  });
}

int hs_copy_last_result(hs_ctx* ctx, hs_result* result) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    HS_REQUIRE(result, HS_E_INVALID, "NULL result");
    result->count = ctx->last_count;
    HS_REQUIRE(ctx->last_count <= result->capacity, HS_E_OVERFLOW, "result capacity too small");
    copy_pairs(ctx, ctx->last_count, result);
  });
}

int hs_stream_last_result(hs_ctx* ctx, uint64_t chunk_pairs, int32_t with_overlap, hs_pair_sink sink, void* user) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    HS_REQUIRE(sink, HS_E_INVALID, "NULL sink");
    HS_REQUIRE(ctx->device >= 0, HS_E_INVALID, "a host-only context holds no device result");
    stream_result(ctx, with_overlap != 0, chunk_pairs,
                 [&](const uint32_t* r, const uint32_t* s, const uint32_t* ov, uint64_t first, uint64_t n) {
                   const int rc = sink(user, r, s, ov, n);
                   HS_REQUIRE(rc == 0, HS_E_INVALID,
                              "the sink stopped the stream at pair " + std::to_string(first) + " (it returned " +
                                  std::to_string(rc) + ")");
                 });
  });
}

int hs_write_last_result_text(hs_ctx* ctx, const char* path, int32_t with_overlap) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    HS_REQUIRE(path, HS_E_INVALID, "NULL path");
    HS_REQUIRE(ctx->device >= 0, HS_E_INVALID, "a host-only context holds no device result");
    hs::PairTextWriter w(path);
    HS_REQUIRE(w.opened(), HS_E_INVALID, std::string("cannot open ") + path);
    stream_result(ctx, with_overlap != 0, 0,
                 [&](const uint32_t* r, const uint32_t* s, const uint32_t* ov, uint64_t, uint64_t n) {
                   w.add(r, s, ov, n);
                 });
    HS_REQUIRE(w.close(), HS_E_INVALID, std::string("writing ") + path + " failed");
  });
}

int hs_get_stats(const hs_ctx* ctx, hs_stats* out) {
  if (!ctx || !out) return HS_E_INVALID;
  *out = ctx->stats;
  return HS_OK;
}

int hs_bounds(int32_t sim, double threshold, int32_t bounds_mode, uint32_t l1, uint32_t l2, uint32_t out[5]) {
  return guarded(nullptr, [&] {
    HS_REQUIRE(out, HS_E_INVALID, "NULL output");
    hs::BoundTables T;
    T.build(sim, bounds_mode, threshold, std::max(l1, l2) + 1);
    out[0] = T.minsize[l1];
    out[1] = T.maxsize[l1];
    out[2] = T.minoverlap(l1, l2);
    out[3] = T.maxprefix[l1];
    out[4] = T.midprefix[l1];
  });
}

int hs_popcount_rate(hs_ctx* ctx, int32_t reps, double* gops) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    HS_REQUIRE(gops && reps > 0, HS_E_INVALID, "bad popcount-rate arguments");
    HS_REQUIRE(ctx->device >= 0, HS_E_INVALID, "a host-only context has no device");
    DevBuf out;
    out.ensure(4);
    const uint32_t grid = (uint32_t)ctx->num_cus * 16, iters = 2048;
    hs::launch_popc_peak(ctx->stream, grid, iters, out.as<uint32_t>());  // warm-up
    HS_CHECK_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    for (int i = 0; i < reps; ++i) hs::launch_popc_peak(ctx->stream, grid, iters, out.as<uint32_t>());
    HS_CHECK_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    HS_CHECK_HIP(hipEventSynchronize(ctx->ev[1]));
    const double ms = elapsed(ctx->ev[0], ctx->ev[1]);
    const double ops = (double)grid * 256.0 * iters * 4 * 8 * 2 * reps;
    *gops = ms > 0 ? ops / (ms * 1e-3) / 1e9 : 0.0;
  });
}

int hs_copy_bandwidth(hs_ctx* ctx, uint64_t bytes, int32_t reps, double* gbps) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    HS_REQUIRE(gbps && reps > 0 && bytes >= 16, HS_E_INVALID, "bad copy-bandwidth arguments");
    bytes &= ~(uint64_t)15;
    DevBuf a, b;
    a.ensure(bytes);
    b.ensure(bytes);
    hs::launch_fill_u32(ctx->stream, a.as<uint32_t>(), bytes / 4, 1u);
    hs::launch_copy(ctx->stream, a.ptr, b.ptr, bytes);  // warm-up
    HS_CHECK_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    for (int i = 0; i < reps; ++i) hs::launch_copy(ctx->stream, (i & 1) ? b.ptr : a.ptr, (i & 1) ? a.ptr : b.ptr, bytes);
    HS_CHECK_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    HS_CHECK_HIP(hipEventSynchronize(ctx->ev[1]));
    const double ms = elapsed(ctx->ev[0], ctx->ev[1]);
    *gbps = ms > 0 ? 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9 : 0.0;
  });
}

}  // extern "C"
