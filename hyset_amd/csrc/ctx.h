// Context and collection state shared by the host-side translation units
// of libhyset_amd.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "hs_internal.h"

using hs::DevBuf;
using hs::Error;

// Page-locked host staging buffer (hipHostMalloc; grow-only).
struct PinnedBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void ensure(size_t need) {
    if (need <= bytes && ptr) return;
    release();
    hs::bump_buffer_epoch();
    size_t want = need < 256 ? 256 : need + need / 8;
    void* p = nullptr;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      throw Error{HS_E_OOM, "hipHostMalloc failed"};
    }
    ptr = p;
    bytes = want;
  }
  void release() {
    if (ptr) {
      (void)hipHostFree(ptr);
      hs::bump_buffer_epoch();
    }
    ptr = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(ptr); }
  ~PinnedBuf() { release(); }
};

// ------------------------------------------------------------ structures

struct hs_coll {
  int device = 0;
  uint32_t n = 0;
  uint64_t n_tokens = 0;
  uint32_t universe = 0;
  uint32_t maxlen = 0;
  DevBuf tokens, starts, sizes, size_first, sigs, tiles;
  uint32_t sig_words = 0;
  std::vector<uint32_t> h_tokens, h_starts, h_sizes, h_size_first;
  std::vector<uint32_t> h_tiles;  // tile plan: tile k = [h_tiles[k], h_tiles[k+1])
  // batched collections (hs_upload_batch): join j owns set ids
  // [h_join_first[j], h_join_first[j+1]) and its size table
  // h_size_first[h_sf_base[j] .. h_sf_base[j+1]) (global set ids); tiles of
  // the batched launch in queue order
  uint32_t n_joins = 0;
  std::vector<uint32_t> h_join_first, h_sf_base, h_join_maxlen, h_join_order;
  std::vector<uint32_t> h_tile_lo, h_tile_hi, h_tile_join;
  DevBuf join_first, sf_base, tile_lo, tile_hi, tile_join;
  uint32_t plan_probes = 0, plan_tokens = 0;

  hs::DevColl dev() const {
    return hs::DevColl{tokens.as<uint32_t>(), starts.as<uint32_t>(), sizes.as<uint32_t>(), n};
  }
};

namespace hs {
struct Triples {
  std::vector<uint32_t> r, s, ov;
};

// One device join between its set-up and its launches: hs_join runs one
// launch over its probe range, the CPU/GPU co-processing join one launch per
// probe stripe it claims, all against one index.
struct DevicePlan {
  bool self = false;
  hs_range ir{1, 0};          // indexed range of the index
  uint32_t sig_words = 0;     // fgss signature test width (64-bit words)
  uint32_t bitmap_words = 0;  // signatures computed for the bitmap technique
  uint64_t index_entries = 0;    // entries of the index built (token-major: probe prefixes)
  uint64_t index_side_entries = 0;  // of those, the entries of the index prefixes
  bool counters_ready = false;  // the index build zeroed the work counters
  bool token_major = false;     // fgss self-join of one range: the token-major plan
  TokPack pk{};                 // the token-major plan's packed entry layout
};
}  // namespace hs

// One double-buffered slot of the hybrid candidate pipeline.
struct HybridSlot {
  PinnedBuf h_probes, h_offsets, h_cands;
  DevBuf d_probes, d_offsets, d_cands;
  hipEvent_t done = nullptr;
  bool used = false;
};

struct hs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool have_threshold = false;
  int kind = HS_JACCARD, mode = HS_BOUNDS_REFERENCE;
  double t = 0.0;
  int32_t host_algo = HS_HOST_ALLPAIRS;  // hs_set_host_algorithm
  // co-processing's break-even gate: per join key, the measured wall times
  // of the device alone and of the last co-processed run (< 0: none yet)
  struct CoprocHistory {
    double device_ms = -1.0, shared_ms = -1.0;
  };
  std::map<std::string, CoprocHistory> coproc_hist;
  hs::BoundTables tables;
  bool tables_valid = false;
  DevBuf d_tables;
  hs::DevBounds dev_bounds{};
  hs::IndexBuild ib;
  // the result pairs as three rows of out_cap ids (r, s, overlap) in one
  // allocation, so one 2D copy moves a range of all three
  DevBuf out, counters;
  uint64_t out_cap = 0;
  PinnedBuf h_counters, h_cum;
  DevBuf d_cum;
  // ev[0..5]: join start, index end, kernel end, counters back, -, kernel
  // start; ev[6..9]: the token-major plan's filter end, second pass start
  // and end, verification start (HS_TUNE_STAGE_TIMES)
  hipEvent_t ev[10] = {};
  hs_stats stats{};
  uint64_t last_count = 0;  // pairs of the last result: on the device, then host_tail
  std::vector<hs::Triples> host_tail;  // pairs the host joined (co-processing), after the device's
  HybridSlot slots[2];
  std::mutex submit_mu;
  std::vector<uint32_t> h_offsets, h_set_ids, h_metas;  // host copy of the index (hybrid; metas for PPJoin)
  // CPU/GPU co-processing join: a copy stream for the index, pinned host
  // copies of the index, claim-completion events
  hipStream_t copy_stream = nullptr;
  PinnedBuf h_index_offsets, h_index_ids, h_index_metas;
  hipEvent_t index_built = nullptr, index_copied = nullptr;
  hipEvent_t claim_ev[2] = {};
  PinnedBuf stage[2];            // pinned chunks of the result stream
  hipEvent_t stage_ev[2] = {};   // chunk k has landed in stage[k]
  int num_cus = 256;
  DevBuf batch_queue;  // per-XCD tile counters of the batched launch
  DevBuf feat;         // partition-feature accumulators (hs_features)
  DevBuf feat_hist;    // per-token prefix counts of two partitions (hs_features_pair)
  DevBuf bal_cost, bal_tiles, bal_ntiles;  // work-balanced fgss tile plan
  DevBuf bal_slice_a, bal_slice_n, d_icum;  // item slices and the item table
  DevBuf bal_rec, bal_slot;                 // per-probe records of the plan
  PinnedBuf h_icum;
  hipEvent_t icum_ev = nullptr;  // recorded after the last copy out of h_icum
  bool icum_pending = false;
  // the item table the device holds (d_icum): the upload is skipped while
  // the table and the allocation epoch are unchanged
  std::vector<uint32_t> icum_host, icum_tmp;
  bool icum_valid = false;
  uint64_t icum_epoch = 0;
  // captured batched join: replayed while (collection, bounds, plan) match
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  const hs_coll* graph_coll = nullptr;
  std::string graph_key;
  PinnedBuf graph_cum;
  uint64_t graph_entries = 0;
  // captured single join (hs_join with HS_TUNE_GRAPH): replayed while the
  // join's key and the buffer epoch are those of the capture
  hipGraph_t jgraph = nullptr;
  hipGraphExec_t jgraph_exec = nullptr;
  std::string jgraph_key;
  uint64_t jgraph_epoch = 0;
  hs::DevicePlan jgraph_plan;
  uint32_t jgraph_tile_plan = 0;
  PinnedBuf jgraph_cum;         // the captured bound-entry table
  // hs_join's entry table on the device, and what it holds: the copy is
  // skipped while the table, the buffer and the allocation epoch are as
  // they were (graph replays rewrite it, so they clear jcum_valid)
  DevBuf d_jcum;
  std::vector<uint32_t> jcum_host, jcum_tmp;
  bool jcum_valid = false;
  uint64_t jcum_epoch = 0;
  bool capturing_join = false;  // prepare_device_join stages through jgraph_cum
  bool stage_times = false;     // this join records events between its stages (HS_TUNE_STAGE_TIMES)
  uint64_t graph_epoch = 0;  // buffer epoch the capture saw
  DevBuf tok_queue;          // token-major join: survivors of the filters, verified next
  uint32_t tok_queue_cap = 0;
  DevBuf tok_heavy, tok_aoff;  // token-major join: chunks of heavy blocks and the blocks' layouts
  DevBuf tok_hdrs;             // token-major join: each tile's header (first key and its list start)
  uint32_t tok_heavy_cap = 0;
  DevBuf tok_stamp;            // token-major join with stage times: probes already counted
  uint32_t tok_epoch = 0;
  uint64_t tok_stamp_n = 0;
  // hs_join copying its pairs: a prefix of the result rows fetched with the
  // counters (as many as the last such join found), so the pairs usually
  // reach the host in the counters' round trip
  PinnedBuf spec;
  uint64_t spec_cap = 0, spec_guess = 0;
  // hs_join_noalloc: its host bound tables and the staging copy of tables
  // and entry table (host memory only; the device side is the scratch array)
  hs::BoundTables na_tables;
  bool na_tables_valid = false;
  std::vector<uint32_t> na_stage;
};


namespace hs {

template <typename F>
inline int guarded(hs_ctx* ctx, F&& body) {
  try {
    if (ctx && ctx->device >= 0) {
      hipError_t e = hipSetDevice(ctx->device);
      if (e != hipSuccess) throw Error{HS_E_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e)};
    }
    body();
    if (ctx) ctx->err.clear();
    return HS_OK;
  } catch (const Error& e) {
    if (ctx) ctx->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->err = "host allocation failed";
    return HS_E_OOM;
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return HS_E_INVALID;
  } catch (...) {
    if (ctx) ctx->err = "unknown error";
    return HS_E_INVALID;
  }
}

inline float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
    (void)hipGetLastError();
    return 0.f;
  }
  return ms;
}

// Parity signatures of every set of c (words 64-bit words each) into
// c->sigs, on the stream.
inline void compute_signatures(hipStream_t st, hs_coll* c, uint32_t words) {
  c->sigs.ensure((size_t)std::max<uint32_t>(c->n, 1) * words * 8);
  c->sig_words = words;
  launch_signatures(st, c->dev(), words, c->sigs.as<uint64_t>());
}

// Signature words of the fgss signature test: `bits` when given (64, 128,
// 256, 512 or 1024), else chosen from the mean size of the indexed sets.
inline uint32_t fgss_signature_words(const hs_coll* indexed, uint32_t bits) {
  if (bits == 0) return auto_signature_words(indexed->n ? (double)indexed->n_tokens / indexed->n : 0.0);
  HS_REQUIRE(bits == 64 || bits == 128 || bits == 256 || bits == 512 || bits == 1024, HS_E_INVALID,
             "signature width must be 64, 128, 256, 512 or 1024 bits");
  return bits / 64;
}

// Bound tables covering set sizes up to L, uploaded to the device.
inline void ensure_bounds(hs_ctx* ctx, uint32_t L) {
  HS_REQUIRE(ctx->have_threshold, HS_E_INVALID, "hs_set_threshold has not been called");
  if (ctx->tables_valid && ctx->tables.L >= L) return;
  uint32_t want = std::max<uint32_t>(L, 64);
  ctx->tables.build(ctx->kind, ctx->mode, ctx->t, want);
  const auto& T = ctx->tables;
  const size_t nl = T.L + 2;
  const size_t total = 4 * nl + T.mo.size();
  std::vector<uint32_t> flat(total);
  std::copy(T.minsize.begin(), T.minsize.end(), flat.begin());
  std::copy(T.maxsize.begin(), T.maxsize.end(), flat.begin() + nl);
  std::copy(T.maxprefix.begin(), T.maxprefix.end(), flat.begin() + 2 * nl);
  std::copy(T.midprefix.begin(), T.midprefix.end(), flat.begin() + 3 * nl);
  std::copy(T.mo.begin(), T.mo.end(), flat.begin() + 4 * nl);
  ctx->d_tables.ensure(total * sizeof(uint32_t));
  HS_CHECK_HIP(hipMemcpy(ctx->d_tables.ptr, flat.data(), total * sizeof(uint32_t), hipMemcpyHostToDevice));
  const uint32_t* base = ctx->d_tables.as<uint32_t>();
  ctx->dev_bounds = hs::DevBounds{base, base + nl, base + 2 * nl, base + 3 * nl, base + 4 * nl, T.L, T.kind};
  ctx->tables_valid = true;
}

// The row length follows the allocation (its bytes / 12), so the positions
// of the s and overlap rows move only when the buffer is reallocated, which
// bumps the allocation epoch that guards graph replay.
inline void ensure_output(hs_ctx* ctx, uint64_t cap) {
  if (cap <= ctx->out_cap && ctx->out.ptr) return;
  ctx->out.ensure(cap * 3 * sizeof(uint32_t));
  ctx->out_cap = ctx->out.bytes / (3 * sizeof(uint32_t));
}

inline hs::DevOut dev_out(hs_ctx* ctx) {
  uint32_t* o = ctx->out.as<uint32_t>();
  return hs::DevOut{o, o + ctx->out_cap, o + 2 * ctx->out_cap, ctx->counters.as<unsigned long long>(),
                    ctx->out_cap};
}

inline void reset_counters(hs_ctx* ctx) {
  hs::launch_fill_u64(ctx->stream, ctx->counters.as<unsigned long long>(), hs::CNT_WORDS, 0);
}

// Reads the work counters back (synchronizes the stream); `done`, when
// given, is recorded after the copy, so it completes with the device work.
inline void read_counters(hs_ctx* ctx, unsigned long long* dst, hipEvent_t done = nullptr) {
  HS_CHECK_HIP(hipMemcpyAsync(ctx->h_counters.ptr, ctx->counters.ptr, hs::CNT_WORDS * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, ctx->stream));
  if (done) HS_CHECK_HIP(hipEventRecord(done, ctx->stream));
  HS_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(dst, ctx->h_counters.ptr, hs::CNT_WORDS * sizeof(unsigned long long));
  hs::fold_counters(dst);
}

// Pairs per chunk of the result stream (12 bytes each with the overlap).
constexpr uint64_t kStreamChunk = 1ull << 18;

// Sends pairs [0, count) of the device result buffers to `consume` in chunks
// of at most `chunk` pairs (0: kStreamChunk), staged through two pinned
// buffers: chunk k+1 is in flight on the context stream while consume sees
// chunk k.  consume(r, s, overlap or NULL, first pair index, n).
void stream_pairs(hs_ctx* ctx, uint64_t count, bool with_overlap, uint64_t chunk,
                  const std::function<void(const uint32_t*, const uint32_t*, const uint32_t*, uint64_t, uint64_t)>&
                      consume);

// The last result on the context: its device pairs through stream_pairs,
// then the host-held tail, in chunks of at most `chunk` pairs.
void stream_result(hs_ctx* ctx, bool with_overlap, uint64_t chunk,
                   const std::function<void(const uint32_t*, const uint32_t*, const uint32_t*, uint64_t, uint64_t)>&
                       consume);

// The last result (count pairs) into the caller's arrays of res.
void copy_pairs(hs_ctx* ctx, uint64_t count, hs_result* res);

// A join left `count` pairs in the device result buffers and nothing on the host.
inline void set_device_result(hs_ctx* ctx, uint64_t count) {
  ctx->last_count = count;
  ctx->host_tail.clear();
}

inline hs_range clamp_range(const hs_range* r, uint32_t n) {
  hs_range out{0, 0};
  if (n == 0) return hs_range{1, 0};
  if (!r) return hs_range{0, n - 1};
  out = *r;
  if (out.last >= n) out.last = n - 1;
  return out;
}

inline bool range_empty(const hs_range& r) { return r.first > r.last; }

// Signatures and the prefix index over ir on the context stream (the index
// also for the bitmap technique when with_index).
void prepare_device_join(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, hs_range ir, hs_range pr, int32_t algo,
                         const hs_tune& tune, bool with_index, DevicePlan& P);

// Queues the technique over probe range pr against the prepared join; pairs
// and work counters accumulate in the context's result buffers.  Records
// ev[5] just before the main kernel.  Returns the tile plan (HS_PLAN_*).
uint32_t launch_device_join(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, DevicePlan& P, hs_range pr,
                            int32_t algo, const hs_tune& tune);

// Postings visited inside the length windows and distinct candidates kept
// by a host probe loop.
struct HostCounts {
  uint64_t visits = 0, candidates = 0;
};

// CPU join of probes [first, last] of P against the prefix index (offsets,
// ids: id-ascending lists over S's universe) of S's sets in ir, on the
// calling thread; pairs are appended to out.  metas (NULL: AllPairs) are
// the entries' (size << 16) | min(position, 0xFFFF): PPJoin's positional
// filter at a candidate's first common prefix token (an entry whose
// position saturates is kept).  stamp is the thread's deduplication array
// (S.n entries, allocated on first use; keep it across calls of one
// thread).  counts (may be NULL) are added to.
void host_probe_range(const hs_coll& P, const hs_coll& S, const uint32_t* offsets, const uint32_t* ids,
                      const uint32_t* metas, const BoundTables& T, bool self, const hs_range& ir, uint32_t first,
                      uint32_t last, Triples& out, std::vector<uint32_t>& stamp, HostCounts* counts = nullptr);

// The hybrid technique over probe range pr against the index of ir
// (hybrid::allpairs / hybrid::ppjoin feeding device_handler,
// hybrid_algorithms.hpp:16-387): tune.host_threads host threads filter the
// probes against the host copy of the index (off, ids; metas for PPJoin's
// positional filter, else NULL) and queue candidate batches, double-buffered
// in pinned slots, for verification on the context stream; pairs and
// counters accumulate in the context's result buffers.  Returns when every
// batch is queued (not when verified).  counts (may be NULL) are added to.
void hybrid_filter_verify(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, bool self, const hs_range& ir,
                          const hs_range& pr, const hs_tune& tune, const uint32_t* off, const uint32_t* ids,
                          const uint32_t* metas, HostCounts* counts);

// CPU join (host_join.cpp): pairs of probe range pr against indexed range ir
// (PPJoin filtering when ppjoin); reads only host copies and the given
// tables, so it may run on any thread.  counts (may be NULL) receive the
// visits and candidates summed over the threads.
uint64_t host_join(const hs_coll& indexed, const hs_coll& probe, bool self, hs_range ir, hs_range pr,
                   const BoundTables& T, uint32_t threads, std::vector<Triples>* parts_out, bool ppjoin = false,
                   HostCounts* counts = nullptr);


}  // namespace hs
