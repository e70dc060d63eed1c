// hs_join_noalloc: the prefix-filter join (fgss) without any device or
// pinned allocation, for callers that cannot allocate (tight serving loops,
// harnesses that forbid hipMalloc).  No counterpart in the reference.
//
// Everything hs_join keeps in buffers it grows -- the bound tables, the
// per-size entry table of the index, the parity signatures, the inverted
// index with its radix-sort work space, and the work counters -- lives in
// one file-scope device array, `scratch`, of NOALLOC_SCRATCH_BYTES per
// device; the result triples go to the caller's device arrays.  The host
// side never names the array in a kernel argument or copy: at context
// creation (hs_create, once per device, with the device current) its device
// address is taken with hipGetSymbolAddress and kept per device, and every
// use goes through that address.  Each device holds the array until the
// process ends, and every call on a device shares it: calls that overlap on
// one device (two contexts on one GPU) race.  Nothing in the array has to
// start at zero except the counters, which the index build (or a fill
// kernel when there is no index) clears on every call.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ctx.h"

namespace hs {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)HS_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t* g_scratch[kMaxDevices] = {};

uint8_t* scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_mu);
  return g_scratch[device];
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
}  // namespace

// Takes the device address of `scratch` on `device` (the current device)
// once; later calls return at once.  Returns false when the runtime refused.
bool noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return false;
  std::lock_guard<std::mutex> g(g_mu);
  if (g_scratch[device]) return true;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return false;
  }
  g_scratch[device] = static_cast<uint8_t*>(p);
  return true;
}

}  // namespace hs

using namespace hs;

extern "C" int hs_join_noalloc(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, const hs_range* indexed_range,
                               const hs_range* probe_range, const hs_tune* tune_in, uint32_t* r, uint32_t* s,
                               uint32_t* overlap, uint64_t capacity, uint64_t* count) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    HS_REQUIRE(indexed && probe && count, HS_E_INVALID, "NULL collection or count");
    HS_REQUIRE(ctx->device >= 0, HS_E_INVALID, "a host-only context has no device");
    HS_REQUIRE(indexed->n_joins == 0 && probe->n_joins == 0, HS_E_INVALID,
               "batched collections join with hs_join_batch");
    HS_REQUIRE(indexed->device == ctx->device && probe->device == ctx->device, HS_E_INVALID,
               "collection lives on another device");
    HS_REQUIRE(capacity == 0 || (r && s && overlap), HS_E_INVALID, "NULL result arrays");
    HS_REQUIRE(ctx->have_threshold, HS_E_INVALID, "hs_set_threshold has not been called");
    uint8_t* base = scratch_of(ctx->device);
    HS_REQUIRE(base, HS_E_DEVICE, "the allocation-free scratch is not available on this device");
    *count = 0;
    hs_tune tune{};
    if (tune_in) tune = *tune_in;
    const bool self = indexed == probe;
    const hs_range ir = clamp_range(indexed_range, indexed->n);
    const hs_range pr = clamp_range(probe_range, probe->n);
    hipStream_t st = ctx->stream;

    // ---- bound tables, on the host (host memory only)
    BoundTables& T = ctx->na_tables;
    const uint32_t L = std::max<uint32_t>(std::max(indexed->maxlen, probe->maxlen), 64);
    if (!ctx->na_tables_valid || T.kind != ctx->kind || T.mode != ctx->mode || T.t != ctx->t || T.L < L) {
      T.build(ctx->kind, ctx->mode, ctx->t, L);
      ctx->na_tables_valid = true;
    }
    const size_t nl = T.L + 2;
    const size_t n_tab = 4 * nl + T.mo.size();
    // entries per size class of the indexed range (sets are sorted by size)
    const auto& ptab = self ? T.midprefix : T.maxprefix;
    const auto& sf = indexed->h_size_first;
    const size_t nlc = sf.size();  // maxlen + 2
    std::vector<uint32_t>& stage = ctx->na_stage;
    stage.resize(n_tab + nlc);
    std::copy(T.minsize.begin(), T.minsize.end(), stage.begin());
    std::copy(T.maxsize.begin(), T.maxsize.end(), stage.begin() + nl);
    std::copy(T.maxprefix.begin(), T.maxprefix.end(), stage.begin() + 2 * nl);
    std::copy(T.midprefix.begin(), T.midprefix.end(), stage.begin() + 3 * nl);
    std::copy(T.mo.begin(), T.mo.end(), stage.begin() + 4 * nl);
    uint64_t entries = 0;
    for (size_t l = 0; l + 1 < nlc; ++l) {
      stage[n_tab + l] = (uint32_t)entries;
      const uint32_t lo = std::max(ir.first, sf[l]);
      const uint32_t hi = std::min(ir.last + 1, sf[l + 1]);
      if (!range_empty(ir) && hi > lo) entries += (uint64_t)(hi - lo) * ptab[std::min<size_t>(l, T.L)];
      HS_REQUIRE(entries < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "index limited to 2^32-1 entries");
    }
    stage[n_tab + nlc - 1] = (uint32_t)entries;

    // ---- signature width of the fgss signature test
    const bool run = !range_empty(ir) && !range_empty(pr);
    uint32_t W = 0;
    if (run && !(tune.flags & HS_TUNE_NO_SIGNATURE)) W = fgss_signature_words(indexed, tune.signature_bits);
    const bool sig_in_emit = W == 1 && self && ir.first == 0 && ir.last + 1 == indexed->n;

    // ---- layout of the scratch array
    size_t off = 0;
    auto take = [&](size_t bytes) {
      const size_t at = off;
      off = align256(off + bytes);
      return at;
    };
    const size_t o_stage = take(stage.size() * 4);
    const size_t o_cnt = take(CNT_WORDS * 8);
    const size_t o_sig_i = take(W ? (size_t)std::max<uint32_t>(indexed->n, 1) * W * 8 : 0);
    const size_t o_sig_p = take(W && !self ? (size_t)std::max<uint32_t>(probe->n, 1) * W * 8 : 0);
    const size_t o_index = off;
    if (run) off += index_region(indexed->universe, entries, nullptr, nullptr);
    HS_REQUIRE(off <= NOALLOC_SCRATCH_BYTES, HS_E_UNSUPPORTED,
               "join needs " + std::to_string(off) + " bytes of scratch; hs_join_noalloc holds " +
                   std::to_string(NOALLOC_SCRATCH_BYTES));

    const uint32_t* dtab = reinterpret_cast<const uint32_t*>(base + o_stage);
    const DevBounds B{dtab, dtab + nl, dtab + 2 * nl, dtab + 3 * nl, dtab + 4 * nl, T.L, T.kind};
    const uint32_t* dcum = dtab + n_tab;
    unsigned long long* cnt_d = reinterpret_cast<unsigned long long*>(base + o_cnt);
    uint64_t* sig_i = W ? reinterpret_cast<uint64_t*>(base + o_sig_i) : nullptr;
    uint64_t* sig_p = W ? (self ? sig_i : reinterpret_cast<uint64_t*>(base + o_sig_p)) : nullptr;

    ctx->stats = hs_stats{};
    HS_CHECK_HIP(hipEventRecord(ctx->ev[0], st));
    // tables and entry table: one copy from host memory
    HS_CHECK_HIP(hipMemcpyAsync(base + o_stage, stage.data(), stage.size() * 4, hipMemcpyHostToDevice, st));
    bool counters_ready = false;
    uint32_t tile_plan = 0;
    if (run) {
      if (W && !sig_in_emit) {
        launch_signatures(st, indexed->dev(), W, sig_i);
        if (!self) launch_signatures(st, probe->dev(), W, sig_p);
      }
      IndexBufs ib{};
      index_region(indexed->universe, entries, base + o_index, &ib);
      build_index_into(st, indexed->dev(), ir.first, ir.last, self ? B.midprefix : B.maxprefix,
                       indexed->size_first.as<uint32_t>(), dcum, indexed->universe, entries, ib,
                       W == 1 ? sig_i : nullptr, sig_in_emit ? sig_i : nullptr, cnt_d);
      counters_ready = entries > 0;
      if (!counters_ready) launch_fill_u64(st, cnt_d, CNT_WORDS, 0);

      FgssArgs a{};
      a.probe = probe->dev();
      a.indexed = indexed->dev();
      a.index = DevIndex{ib.offsets, ib.set_ids, ib.ent, W == 1 ? 4u : 2u, indexed->universe};
      a.bounds = B;
      a.size_first = indexed->size_first.as<uint32_t>();
      a.size_first_len = (uint32_t)sf.size();
      a.probe_first = pr.first;
      a.probe_last = pr.last;
      a.idx_first = ir.first;
      a.idx_last = ir.last;
      a.self = self;
      a.out = DevOut{r, s, overlap, cnt_d, capacity};
      a.sig_words = W;
      a.probe_sigs = sig_p;
      a.indexed_sigs = sig_i;
      const uint32_t ppt = tune.probes_per_tile;
      if (ppt == 0) {
        // the probe collection's upload-time tiles, clipped to the probe range
        const auto& tl = probe->h_tiles;
        const uint32_t ft = (uint32_t)(std::upper_bound(tl.begin(), tl.end(), pr.first) - tl.begin()) - 1;
        const uint32_t lt = (uint32_t)(std::upper_bound(tl.begin(), tl.end(), pr.last) - tl.begin()) - 1;
        a.tile_starts = probe->tiles.as<uint32_t>();
        a.first_tile = ft;
        a.n_plan_tiles = lt - ft + 1;
      }
      tile_plan = ppt ? HS_PLAN_FIXED : HS_PLAN_UPLOAD;
      launch_fgss(st, a, ppt, 0);
    } else {
      launch_fill_u64(st, cnt_d, CNT_WORDS, 0);
    }
    HS_CHECK_HIP(hipMemcpyAsync(ctx->h_counters.ptr, cnt_d, CNT_WORDS * 8, hipMemcpyDeviceToHost, st));
    HS_CHECK_HIP(hipEventRecord(ctx->ev[3], st));
    HS_CHECK_HIP(hipStreamSynchronize(st));
    unsigned long long cnt[CNT_WORDS];
    std::memcpy(cnt, ctx->h_counters.ptr, sizeof(cnt));
    fold_counters(cnt);
    ctx->stats.total_ms = elapsed(ctx->ev[0], ctx->ev[3]);
    ctx->stats.pairs = cnt[CNT_PAIRS];
    ctx->stats.posting_visits = cnt[CNT_VISITS];
    ctx->stats.candidates = cnt[CNT_CANDS];
    ctx->stats.candidate_tokens = cnt[CNT_CAND_TOKENS];
    ctx->stats.probe_prefix_tokens = cnt[CNT_ITEMS];
    ctx->stats.index_entries = run ? entries : 0;
    ctx->stats.signature_tests = cnt[CNT_SIG_TESTS];
    ctx->stats.signature_bits = W * 64;
    ctx->stats.tile_plan = tile_plan;
    *count = cnt[CNT_PAIRS];
    HS_REQUIRE(cnt[CNT_PAIRS] <= capacity, HS_E_OVERFLOW, "result capacity too small");
  });
}
