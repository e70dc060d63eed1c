// Batched joins: many independent self-joins in one device collection, one
// index build and one persistent prefix-filter launch fed by a device work
// queue.
//
// Reference: src/queue.cu:142-279 keeps a host MPMC queue of block pairs
// that a device thread drains with one kernel sequence per pair; the
// cost_model.cu / dichotomy.cu drivers split partitions the same way.  Here
// the unit of work is a tile of probes of one join, the queue lives in
// device memory (per-XCD counters with stealing, fgss.hip), and every join
// of the batch shares one index build and one launch.  Joins get disjoint
// token ranges (token ids are offset by the universes of the joins before
// them), so one index over the union serves them all, and each tile keeps
// its window inside its own join's set-id range.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <chrono>
#include <numeric>
#include <string>
#include <vector>

#include "ctx.h"

using namespace hs;

namespace {

// Greedy tile plan of the batched launch: runs of at most `probes` probes and
// `tokens` tokens inside one join, joins in descending size so the largest
// join's tiles start first.
void plan_batch_tiles(hs_coll* c, uint32_t probes, uint32_t tokens) {
  c->h_tile_lo.clear();
  c->h_tile_hi.clear();
  c->h_tile_join.clear();
  for (uint32_t j : c->h_join_order) {
    const uint32_t a = c->h_join_first[j], b = c->h_join_first[j + 1];
    for (uint32_t i = a; i < b;) {
      const uint32_t lo = i;
      uint64_t tok = 0;
      uint32_t cnt = 0;
      while (i < b && cnt < probes && (cnt == 0 || tok + c->h_sizes[i] <= tokens)) {
        tok += c->h_sizes[i];
        ++cnt;
        ++i;
      }
      c->h_tile_lo.push_back(lo);
      c->h_tile_hi.push_back(i);
      c->h_tile_join.push_back(j);
    }
  }
  const size_t nt = std::max<size_t>(c->h_tile_lo.size(), 1);
  c->tile_lo.ensure(nt * 4);
  c->tile_hi.ensure(nt * 4);
  c->tile_join.ensure(nt * 4);
  if (!c->h_tile_lo.empty()) {
    HS_CHECK_HIP(hipMemcpy(c->tile_lo.ptr, c->h_tile_lo.data(), c->h_tile_lo.size() * 4, hipMemcpyHostToDevice));
    HS_CHECK_HIP(hipMemcpy(c->tile_hi.ptr, c->h_tile_hi.data(), c->h_tile_hi.size() * 4, hipMemcpyHostToDevice));
    HS_CHECK_HIP(hipMemcpy(c->tile_join.ptr, c->h_tile_join.data(), c->h_tile_join.size() * 4, hipMemcpyHostToDevice));
  }
  c->plan_probes = probes;
  c->plan_tokens = tokens;
}

// Entry table of the batched index: cum[sf_base[j] + l] = entries of all
// earlier joins plus those of join j's sets shorter than l.
uint64_t batch_entry_table(const hs_coll* c, const std::vector<uint32_t>& prefix, uint32_t* cum) {
  uint64_t acc = 0;
  for (uint32_t j = 0; j < c->n_joins; ++j) {
    const uint32_t b = c->h_sf_base[j], e = c->h_sf_base[j + 1];
    for (uint32_t k = b; k < e; ++k) {
      const uint32_t l = k - b;
      cum[k] = (uint32_t)acc;
      if (k + 1 < e) acc += (uint64_t)(c->h_size_first[k + 1] - c->h_size_first[k]) * prefix[l];
      HS_REQUIRE(acc < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "index limited to 2^32-1 entries");
    }
  }
  return acc;
}

std::string graph_key(const hs_ctx* ctx, const hs_coll* c, uint32_t workgroups, uint32_t sig_words,
                      uint32_t flags) {
  return std::to_string(ctx->kind) + "/" + std::to_string(ctx->mode) + "/" + std::to_string(ctx->t) + "/" +
         std::to_string(ctx->tables.L) + "/" + std::to_string(c->plan_probes) + "x" + std::to_string(c->plan_tokens) +
         "/" + std::to_string(workgroups) + "/" + std::to_string(sig_words) + "/" +
         std::to_string(flags & (HS_TUNE_NARROW | HS_TUNE_INDEX_MAXPREFIX));
}

void drop_graph(hs_ctx* ctx) {
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->graph_exec = nullptr;
  ctx->graph = nullptr;
  ctx->graph_coll = nullptr;
  ctx->graph_key.clear();
}

// Enqueues the whole batched join (entry table upload, signatures, index
// build, counter reset, persistent launch) on the context stream.
void enqueue_batch(hs_ctx* ctx, hs_coll* c, const uint32_t* h_cum, uint64_t entries, uint32_t workgroups,
                   uint32_t sig_words, uint32_t flags) {
  hipStream_t st = ctx->stream;
  // 64-bit signatures come out of the index's emit pass (it reads every set)
  if (sig_words == 1) {
    c->sigs.ensure((size_t)std::max<uint32_t>(c->n, 1) * 8);
    c->sig_words = 1;
  } else if (sig_words) {
    compute_signatures(st, c, sig_words);
  }
  const uint32_t ntab = (uint32_t)c->h_size_first.size();
  ctx->d_cum.ensure((size_t)ntab * 4);
  HS_CHECK_HIP(hipMemcpyAsync(ctx->d_cum.ptr, h_cum, (size_t)ntab * 4, hipMemcpyHostToDevice, st));
  const uint32_t* index_prefix =
      (flags & HS_TUNE_INDEX_MAXPREFIX) ? ctx->dev_bounds.maxprefix : ctx->dev_bounds.midprefix;
  build_index_batch(st, c->dev(), c->join_first.as<uint32_t>(), c->n_joins, index_prefix,
                    c->size_first.as<uint32_t>(), c->sf_base.as<uint32_t>(), ctx->d_cum.as<uint32_t>(), c->universe,
                    entries, ctx->ib, sig_words == 1 ? c->sigs.as<uint64_t>() : nullptr,
                    sig_words == 1 ? c->sigs.as<uint64_t>() : nullptr);
  reset_counters(ctx);
  FgssArgs a{};
  a.probe = c->dev();
  a.indexed = c->dev();
  a.index = dev_index(ctx->ib);
  a.bounds = ctx->dev_bounds;
  a.size_first = c->size_first.as<uint32_t>();
  a.self = 1;
  a.probe_first = 0;
  a.probe_last = c->n ? c->n - 1 : 0;
  a.batch = FgssBatch{c->tile_lo.as<uint32_t>(), c->tile_hi.as<uint32_t>(), c->tile_join.as<uint32_t>(),
                      c->join_first.as<uint32_t>(), c->sf_base.as<uint32_t>(), ctx->batch_queue.as<uint32_t>()};
  a.out = dev_out(ctx);
  a.sig_words = sig_words;
  a.probe_sigs = a.indexed_sigs = sig_words ? c->sigs.as<uint64_t>() : nullptr;
  launch_fgss_batch(st, a, (uint32_t)c->h_tile_lo.size(), workgroups, (flags & HS_TUNE_NARROW) != 0);
}

void check_batch(hs_ctx* ctx, hs_coll* c) {
  HS_REQUIRE(ctx && c, HS_E_INVALID, "NULL context or collection");
  HS_REQUIRE(c->device == ctx->device, HS_E_INVALID, "collection lives on another device");
  HS_REQUIRE(c->n_joins > 0, HS_E_INVALID, "not a batched collection (use hs_upload_batch)");
}

}  // namespace

extern "C" {

int hs_upload_batch(hs_ctx* ctx, uint32_t n_joins, const uint32_t* const* tokens, const uint32_t* const* starts,
                    const uint32_t* const* sizes, const uint32_t* n_sets, const uint64_t* n_tokens, hs_coll** out) {
  if (!ctx || !out) return HS_E_INVALID;
  *out = nullptr;
  std::unique_ptr<hs_coll> c;
  return guarded(ctx, [&] {
    HS_REQUIRE(n_joins > 0 && tokens && sizes && n_sets && n_tokens, HS_E_INVALID, "NULL batch arrays");
    c.reset(new hs_coll());
    c->device = ctx->device;
    c->n_joins = n_joins;
    c->h_join_first.assign(1, 0);
    c->h_sf_base.assign(1, 0);
    uint64_t tok_base = 0, pos = 0, total_tok = 0, total_sets = 0;
    uint32_t maxlen_all = 0;
    for (uint32_t j = 0; j < n_joins; ++j) {
      total_tok += n_tokens[j];
      total_sets += n_sets[j];
    }
    HS_REQUIRE(total_tok < 0xFFFFFFFFull && total_sets < 0xFFFFFFFFull, HS_E_UNSUPPORTED,
               "batch limited to 2^32-1 sets and tokens");
    c->h_tokens.reserve(total_tok);
    c->h_starts.reserve(total_sets);
    c->h_sizes.reserve(total_sets);
    for (uint32_t j = 0; j < n_joins; ++j) {
      const uint32_t n = n_sets[j];
      HS_REQUIRE(n == 0 || (tokens[j] && sizes[j]), HS_E_INVALID, "NULL arrays of join " + std::to_string(j));
      const uint32_t first = (uint32_t)c->h_sizes.size();
      uint32_t maxtok = 0, maxlen = 0, prev = 0;
      for (uint32_t i = 0; i < n; ++i) {
        const uint32_t l = sizes[j][i];
        HS_REQUIRE(l >= 1 && l >= prev, HS_E_INVALID, "join " + std::to_string(j) + ": empty or unsorted set");
        prev = l;
        maxlen = std::max(maxlen, l);
      }
      // copy tokens with the join's token offset
      uint64_t local = 0;
      for (uint32_t i = 0; i < n; ++i) {
        const uint32_t l = sizes[j][i];
        const uint64_t src = (starts && starts[j]) ? starts[j][i] : local;
        HS_REQUIRE(src + l <= n_tokens[j], HS_E_INVALID, "join " + std::to_string(j) + ": set exceeds tokens");
        const uint32_t* t = tokens[j] + src;
        for (uint32_t k = 1; k < l; ++k) HS_REQUIRE(t[k - 1] < t[k], HS_E_INVALID, "tokens not ascending");
        maxtok = std::max(maxtok, t[l - 1]);
        c->h_starts.push_back((uint32_t)pos);
        c->h_sizes.push_back(l);
        for (uint32_t k = 0; k < l; ++k) c->h_tokens.push_back((uint32_t)(t[k] + tok_base));
        pos += l;
        local += l;
        HS_REQUIRE(pos < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "batch limited to 2^32-1 tokens");
      }
      const uint32_t universe_j = n ? maxtok + 1 : 1;
      tok_base += universe_j;
      HS_REQUIRE(tok_base < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "batch token space exceeds 2^32-1");
      c->h_join_first.push_back((uint32_t)c->h_sizes.size());
      c->h_join_maxlen.push_back(maxlen);
      maxlen_all = std::max(maxlen_all, maxlen);
      // size table of join j in global set ids
      for (uint32_t l = 0, i = first; l <= maxlen + 1; ++l) {
        while (i < first + n && c->h_sizes[i] < l) ++i;
        c->h_size_first.push_back(i);
      }
      c->h_sf_base.push_back((uint32_t)c->h_size_first.size());
    }
    c->n = (uint32_t)c->h_sizes.size();
    c->n_tokens = pos;
    c->universe = (uint32_t)std::max<uint64_t>(tok_base, 1);
    c->maxlen = maxlen_all;
    c->h_join_order.resize(n_joins);
    std::iota(c->h_join_order.begin(), c->h_join_order.end(), 0u);
    std::stable_sort(c->h_join_order.begin(), c->h_join_order.end(), [&](uint32_t a, uint32_t b) {
      return c->h_join_first[a + 1] - c->h_join_first[a] > c->h_join_first[b + 1] - c->h_join_first[b];
    });
    c->tokens.ensure(std::max<uint64_t>(pos, 1) * 4);
    c->starts.ensure((size_t)std::max<uint32_t>(c->n, 1) * 4);
    c->sizes.ensure((size_t)std::max<uint32_t>(c->n, 1) * 4);
    c->size_first.ensure(c->h_size_first.size() * 4);
    c->join_first.ensure(c->h_join_first.size() * 4);
    c->sf_base.ensure(c->h_sf_base.size() * 4);
    if (pos) HS_CHECK_HIP(hipMemcpy(c->tokens.ptr, c->h_tokens.data(), pos * 4, hipMemcpyHostToDevice));
    if (c->n) {
      HS_CHECK_HIP(hipMemcpy(c->starts.ptr, c->h_starts.data(), (size_t)c->n * 4, hipMemcpyHostToDevice));
      HS_CHECK_HIP(hipMemcpy(c->sizes.ptr, c->h_sizes.data(), (size_t)c->n * 4, hipMemcpyHostToDevice));
    }
    HS_CHECK_HIP(hipMemcpy(c->size_first.ptr, c->h_size_first.data(), c->h_size_first.size() * 4,
                           hipMemcpyHostToDevice));
    HS_CHECK_HIP(hipMemcpy(c->join_first.ptr, c->h_join_first.data(), c->h_join_first.size() * 4,
                           hipMemcpyHostToDevice));
    HS_CHECK_HIP(hipMemcpy(c->sf_base.ptr, c->h_sf_base.data(), c->h_sf_base.size() * 4, hipMemcpyHostToDevice));
    plan_batch_tiles(c.get(), TILE_PROBES, TILE_TOKENS);
    *out = c.release();
  });
}

int hs_join_batch(hs_ctx* ctx, hs_coll* batch, const hs_tune* tune_in, hs_result* result) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    check_batch(ctx, batch);
    HS_REQUIRE(result, HS_E_INVALID, "NULL result");
    hs_tune tune{};
    if (tune_in) tune = *tune_in;
    HS_REQUIRE((tune.flags & ~(uint32_t)(HS_TUNE_GRAPH | HS_TUNE_NO_SIGNATURE | HS_TUNE_NARROW |
                                         HS_TUNE_INDEX_MAXPREFIX)) == 0,
               HS_E_INVALID, "unknown tune flags");
    const bool narrow = (tune.flags & HS_TUNE_NARROW) != 0;
    const uint32_t max_probes = narrow ? TILE_PROBES / 2 : TILE_PROBES;
    const uint32_t max_tokens = narrow ? TILE_TOKENS / 2 : TILE_TOKENS;
    ensure_bounds(ctx, batch->maxlen);
    const uint32_t sig_words =
        (tune.flags & HS_TUNE_NO_SIGNATURE) ? 0u : fgss_signature_words(batch, tune.signature_bits);
    // tile plan of the requested granularity
    const uint32_t probes = tune.probes_per_tile ? std::min<uint32_t>(tune.probes_per_tile, max_probes) : max_probes;
    const uint32_t tokens = tune.tile_tokens ? std::min<uint32_t>(tune.tile_tokens, max_tokens) : max_tokens;
    if (probes != batch->plan_probes || tokens != batch->plan_tokens) plan_batch_tiles(batch, probes, tokens);
    const uint32_t wpc = tune.workgroups_per_cu ? tune.workgroups_per_cu : (narrow ? 4 : 2);
    const uint32_t workgroups = (uint32_t)ctx->num_cus * wpc;
    ctx->batch_queue.ensure(8 * 4);
    ctx->stats = hs_stats{};
    hipStream_t st = ctx->stream;
    const bool use_graph = (tune.flags & HS_TUNE_GRAPH) != 0;
    const std::string key = graph_key(ctx, batch, workgroups, sig_words, tune.flags);
    HS_CHECK_HIP(hipEventRecord(ctx->ev[0], st));
    unsigned long long cnt[CNT_WORDS];
    for (int attempt = 0; attempt < 2; ++attempt) {
      // replay only while every buffer the capture recorded is still where it was
      if (use_graph && ctx->graph_exec && ctx->graph_coll == batch && ctx->graph_key == key &&
          ctx->graph_epoch == buffer_epoch()) {
        HS_CHECK_HIP(hipGraphLaunch(ctx->graph_exec, st));
      } else {
        const uint32_t ntab = (uint32_t)batch->h_size_first.size();
        PinnedBuf& cum = use_graph ? ctx->graph_cum : ctx->h_cum;
        cum.ensure((size_t)ntab * 4);
        const uint64_t entries = batch_entry_table(
            batch, (tune.flags & HS_TUNE_INDEX_MAXPREFIX) ? ctx->tables.maxprefix : ctx->tables.midprefix,
            cum.as<uint32_t>());
        if (use_graph) {
          // an eager run first allocates every work buffer; the capture then
          // records a sequence that allocates nothing and replays as is
          drop_graph(ctx);
          enqueue_batch(ctx, batch, cum.as<uint32_t>(), entries, workgroups, sig_words, tune.flags);
          HS_CHECK_HIP(hipStreamSynchronize(st));
          HS_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
          enqueue_batch(ctx, batch, cum.as<uint32_t>(), entries, workgroups, sig_words, tune.flags);
          HS_CHECK_HIP(hipStreamEndCapture(st, &ctx->graph));
          HS_CHECK_HIP(hipGraphInstantiate(&ctx->graph_exec, ctx->graph, nullptr, nullptr, 0));
          ctx->graph_coll = batch;
          ctx->graph_key = key;
          ctx->graph_entries = entries;
          ctx->graph_epoch = buffer_epoch();
          HS_CHECK_HIP(hipGraphLaunch(ctx->graph_exec, st));
        } else {
          enqueue_batch(ctx, batch, cum.as<uint32_t>(), entries, workgroups, sig_words, tune.flags);
        }
        ctx->stats.index_entries = entries;
      }
      HS_CHECK_HIP(hipEventRecord(ctx->ev[2], st));
      read_counters(ctx, cnt);
      if (cnt[CNT_PAIRS] <= ctx->out_cap) break;
      HS_REQUIRE(attempt == 0, HS_E_DEVICE, "result count changed between runs");
      drop_graph(ctx);  // the output buffers move
      ensure_output(ctx, cnt[CNT_PAIRS] + cnt[CNT_PAIRS] / 4 + 1024);
    }
    if (use_graph) ctx->stats.index_entries = ctx->graph_entries;
    const uint64_t count = cnt[CNT_PAIRS];
    set_device_result(ctx, count);
    result->count = count;
    HS_CHECK_HIP(hipEventRecord(ctx->ev[3], st));
    if (!result->count_only && count > 0) {
      HS_REQUIRE(count <= result->capacity, HS_E_OVERFLOW, "result capacity too small");
      copy_pairs(ctx, count, result);
    }
    HS_CHECK_HIP(hipEventRecord(ctx->ev[4], st));
    HS_CHECK_HIP(hipEventSynchronize(ctx->ev[4]));
    ctx->stats.filter_ms = elapsed(ctx->ev[0], ctx->ev[2]);
    ctx->stats.kernel_ms = ctx->stats.filter_ms;
    ctx->stats.transfer_ms = elapsed(ctx->ev[3], ctx->ev[4]);
    ctx->stats.total_ms = elapsed(ctx->ev[0], ctx->ev[4]);
    ctx->stats.pairs = count;
    ctx->stats.posting_visits = cnt[CNT_VISITS];
    ctx->stats.candidates = cnt[CNT_CANDS];
    ctx->stats.candidate_tokens = cnt[CNT_CAND_TOKENS];
    ctx->stats.probe_prefix_tokens = cnt[CNT_ITEMS];
    ctx->stats.signature_tests = cnt[CNT_SIG_TESTS];
    ctx->stats.signature_bits = sig_words * 64;
  });
}

int hs_autotune_batch(hs_ctx* ctx, hs_coll* batch, int32_t reps, hs_tune* best, double* best_ms) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    check_batch(ctx, batch);
    HS_REQUIRE(best, HS_E_INVALID, "NULL tune output");
    const int r = reps > 0 ? reps : 3;
    // the median of r timed runs after one warm-up
    auto time_of = [&](const hs_tune& t, int runs) {
      std::vector<double> ms;
      for (int k = 0; k <= runs; ++k) {
        hs_result res{};
        res.count_only = 1;
        int rc = hs_join_batch(ctx, batch, &t, &res);
        if (rc != HS_OK) throw Error{rc, ctx->err};
        if (k > 0) ms.push_back(ctx->stats.total_ms);  // the first run warms up
      }
      std::sort(ms.begin(), ms.end());
      return ms[ms.size() / 2];
    };
    // the default settings are the first incumbent; a candidate replaces
    // the incumbent only when 2% faster, so run-to-run noise (a few
    // percent on a 15 ms batch) cannot pick a slower setting
    const hs_tune defaults{};
    hs_tune bt = defaults;
    double best_t = time_of(defaults, r);
    auto trial = [&](const hs_tune& t) {
      const double ms = time_of(t, r);
      if (ms < best_t * 0.98) {
        best_t = ms;
        bt = t;
      }
    };
    // threads per probe (workgroup width over probes per tile) and grid at
    // the default signature width, then the prefix scheme, then the width
    for (const bool narrow : {false, true}) {
      const uint32_t lanes = narrow ? TILE_PROBES / 2 : TILE_PROBES;
      for (uint32_t probes : {64u, 128u, 256u, 512u}) {
        if (probes > lanes) continue;
        for (uint32_t wpc : narrow ? std::initializer_list<uint32_t>{2u, 4u, 8u}
                                   : std::initializer_list<uint32_t>{1u, 2u, 4u}) {
          hs_tune t{};
          t.flags = narrow ? HS_TUNE_NARROW : 0u;
          t.probes_per_tile = probes;
          t.tile_tokens = probes * 16;
          t.workgroups_per_cu = wpc;
          trial(t);
        }
      }
    }
    hs_tune scheme = bt;
    scheme.flags |= HS_TUNE_INDEX_MAXPREFIX;
    trial(scheme);
    const hs_tune base = bt;
    for (uint32_t bits : {64u, 128u, 256u}) {
      hs_tune t = base;
      t.signature_bits = bits;
      trial(t);
    }
    hs_tune off = base;
    off.flags |= HS_TUNE_NO_SIGNATURE;
    trial(off);
    // the winner against the defaults once more, with twice the runs
    const bool changed = bt.flags != defaults.flags || bt.probes_per_tile != defaults.probes_per_tile ||
                         bt.tile_tokens != defaults.tile_tokens || bt.workgroups_per_cu != defaults.workgroups_per_cu ||
                         bt.signature_bits != defaults.signature_bits;
    if (changed) {
      const double d = time_of(defaults, 2 * r), w = time_of(bt, 2 * r);
      if (w < d * 0.98) {
        best_t = w;
      } else {
        bt = defaults;
        best_t = d;
      }
    }
    *best = bt;
    if (best_ms) *best_ms = best_t;
  });
}

}  // extern "C"
