// One join shared by the device and host cores running concurrently.
//
// Reference: the work-sharing driver src/queue.cu:142-279 (a queue of block
// pairs drained by a device thread and by the host thread; the device
// thread runs fgss, bitmap or, --cooperative, the hybrid technique) and the
// static split of src/dichotomy.cu:143-285 (the small-set partition Rs
// joined with itself by one side, R against the rest Rl by the other, the
// sides named by the --policy pair, the device side fgss, bitmap or the
// hybrid cpu-gpu technique).  Here the unit of work
// is a stripe of consecutive probe ids joined against the whole indexed
// range, so any set of disjoint stripes covering the probe ids gives exactly
// the join.  Sets are sorted by size, so the work per probe grows with its
// id: the device claims from the top, where its wide launches pay, the host
// threads from the bottom, in stripes small enough that the last one ends
// soon after the device's.
//
// The host needs no index of its own: the device builds the prefix index
// once and a copy of it goes to pinned host memory on a second stream while
// the device already runs its first stripe.  Device stripes accumulate pairs
// in the context's result buffers (every launch only adds to the counters);
// the host's pairs stay in host memory as the tail of the context's result,
// which the result sinks deliver after the device's pairs.
#include <sched.h>

#include <cstdio>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "ctx.h"

using namespace hs;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

uint32_t usable_cpus() {
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) return std::max(1, CPU_COUNT(&set));
  return std::max(1u, std::thread::hardware_concurrency());
}

// Work proxy of probe ids [a, b): set size plus a fixed per-probe cost,
// summed over the size classes the ids span (sets are sorted by size, sf[l]
// is the first id of size l or more).
constexpr uint64_t kProbeCost = 8;

uint64_t probe_weight(const std::vector<uint32_t>& sf, uint32_t a, uint32_t b) {
  uint64_t w = 0;
  for (uint32_t l = 0; l + 1 < (uint32_t)sf.size(); ++l) {
    const uint32_t lo = std::max(a, sf[l]), hi = std::min(b, sf[l + 1]);
    if (hi > lo) w += (uint64_t)(hi - lo) * (l + kProbeCost);
  }
  return w;
}

// Smallest x in [a, b] with probe_weight(a, x) >= u.
uint32_t weight_cut(const std::vector<uint32_t>& sf, uint32_t a, uint32_t b, uint64_t u) {
  uint64_t acc = 0;
  for (uint32_t l = 0; l + 1 < (uint32_t)sf.size(); ++l) {
    const uint32_t lo = std::max(a, sf[l]), hi = std::min(b, sf[l + 1]);
    if (hi <= lo) continue;
    const uint64_t c = l + kProbeCost, cls = (uint64_t)(hi - lo) * c;
    if (acc + cls >= u) return lo + (uint32_t)std::min<uint64_t>(hi - lo, (u - acc + c - 1) / c);
    acc += cls;
  }
  return b;
}

// Unclaimed probe ids [lo, hi): host claims move lo up, device claims move
// hi down; the host never claims at or above host_limit.
struct Stripes {
  std::mutex mu;
  uint32_t lo = 0, hi = 0, host_limit = 0;

  // A host claim of at most `want` sets and at most an equal share of what
  // is left between the host threads, so the last stripes are short.
  bool take_low(uint32_t want, uint32_t threads, uint32_t& first, uint32_t& n) {
    std::lock_guard<std::mutex> g(mu);
    const uint32_t lim = std::min(hi, host_limit);
    if (lo >= lim) return false;
    const uint32_t share = std::max(1u, (lim - lo + threads - 1) / threads);
    n = std::min({want, lim - lo, share});
    first = lo;
    lo += n;
    return true;
  }
};

struct HostWorker {
  Triples out;
  std::vector<uint32_t> stamp;
  uint32_t probes = 0, claims = 0;
  Clock::time_point done;
  bool failed = false;
  Error err{HS_OK, ""};
};

// Host stripe size when the caller gives none: about this much work per
// claim at the thread's last measured rate.
constexpr double kHostClaimMs = 0.1;

// The break-even gate's default: a join that takes the device alone less
// than this is not shared.
constexpr double kGateMs = 50.0;

// The gate's key of a join: the collections, technique, threshold and the
// knobs that change either side's speed.
std::string coproc_key(const hs_ctx* ctx, const hs_coll* indexed, const hs_coll* probe, const hs_coprocess& o) {
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%p/%u/%llu|%p/%u/%llu|a%d|s%d|m%d|t%.17g|f%u|b%u|c%u|h%d", (const void*)indexed,
                indexed->n, (unsigned long long)indexed->n_tokens, (const void*)probe, probe->n,
                (unsigned long long)probe->n_tokens, o.algo, ctx->kind, ctx->mode, ctx->t, o.tune.flags,
                o.tune.signature_bits, o.cpu_threads, ctx->host_algo);
  return buf;
}

}  // namespace

extern "C" {

int hs_join_coprocess(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, const hs_coprocess* options,
                      hs_result* result, hs_coprocess_stats* stats) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    HS_REQUIRE(options && result, HS_E_INVALID, "NULL options or result");
    HS_REQUIRE(indexed && probe, HS_E_INVALID, "NULL collection");
    HS_REQUIRE(ctx->device >= 0, HS_E_INVALID, "co-processing needs a device context");
    HS_REQUIRE(indexed->n_joins == 0 && probe->n_joins == 0, HS_E_INVALID, "batched collections are not supported");
    HS_REQUIRE(indexed->device == ctx->device && probe->device == ctx->device, HS_E_INVALID,
               "collection lives on another device");
    const hs_coprocess opt = *options;
    HS_REQUIRE(opt.mode == HS_COPROC_QUEUE || opt.mode == HS_COPROC_STATIC, HS_E_INVALID, "unknown mode");
    HS_REQUIRE(opt.algo == HS_ALGO_FGSS || opt.algo == HS_ALGO_BITMAP || opt.algo == HS_ALGO_HYBRID, HS_E_INVALID,
               "the device technique must be fgss, bitmap or hybrid");
    HS_REQUIRE(opt.policy == HS_POLICY_CPU_GPU || opt.policy == HS_POLICY_GPU_CPU, HS_E_INVALID, "unknown policy");
    HS_REQUIRE(opt.mode == HS_COPROC_STATIC || opt.policy == HS_POLICY_CPU_GPU, HS_E_INVALID,
               "the queue mode's device claims from the large sets (policy cpu,gpu)");
    const bool hybrid = opt.algo == HS_ALGO_HYBRID;
    HS_REQUIRE(opt.mode != HS_COPROC_STATIC || (opt.cpu_share >= 0.0 && opt.cpu_share <= 1.0), HS_E_INVALID,
               "cpu_share must lie in [0, 1]");
    const hs_tune tune = opt.tune;
    const auto t_start = Clock::now();
    hs_coprocess_stats cs{};
    const bool self = indexed == probe;
    const hs_range ir = clamp_range(nullptr, indexed->n);
    const hs_range pr = clamp_range(nullptr, probe->n);
    ensure_bounds(ctx, std::max(indexed->maxlen, probe->maxlen));
    const BoundTables T = ctx->tables;  // the host threads' copy
    ctx->stats = hs_stats{};
    hipStream_t st = ctx->stream;

    if (range_empty(ir) || range_empty(pr)) {
      set_device_result(ctx, 0);
      result->count = 0;
      cs.wall_ms = ms_between(t_start, Clock::now());
      if (stats) *stats = cs;
      return;
    }

    // ---- the break-even gate: the device alone unless sharing measured faster
    hs_ctx::CoprocHistory* hist = nullptr;
    if (opt.mode == HS_COPROC_QUEUE && opt.min_device_ms >= 0.0) {
      hist = &ctx->coproc_hist[coproc_key(ctx, indexed, probe, opt)];
      const double gate_ms = opt.min_device_ms > 0.0 ? opt.min_device_ms : kGateMs;
      const bool alone = hist->device_ms < 0.0 || hist->device_ms < gate_ms ||
                         (hist->shared_ms >= 0.0 && hist->shared_ms >= hist->device_ms);
      if (alone) {
        const auto t0 = Clock::now();
        const int rc = hs_join(ctx, indexed, probe, nullptr, nullptr, opt.algo, &tune, result);
        if (rc != HS_OK) throw Error{rc, ctx->err};
        const double ms = ms_between(t0, Clock::now());
        // keep the faster of the measurements (the first call also pays warm-up)
        hist->device_ms = hist->device_ms < 0.0 ? ms : std::min(hist->device_ms, ms);
        cs.gated = 1;
        cs.device_pairs = result->count;
        cs.device_probes = pr.last - pr.first + 1;
        cs.device_claims = 1;
        cs.split = pr.first;
        cs.device_ms = ms;
        cs.index_ms = ctx->stats.index_ms;
        cs.device_alone_ms = hist->device_ms;
        cs.shared_ms = std::max(hist->shared_ms, 0.0);
        cs.wall_ms = ms_between(t_start, Clock::now());
        if (stats) *stats = cs;
        return;
      }
    }
    if (!ctx->copy_stream) HS_CHECK_HIP(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
    if (!ctx->index_built) HS_CHECK_HIP(hipEventCreate(&ctx->index_built));
    if (!ctx->index_copied) HS_CHECK_HIP(hipEventCreate(&ctx->index_copied));
    for (auto& e : ctx->claim_ev)
      if (!e) HS_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));

    // ---- the index, once, and its copy to the host behind it
    HS_CHECK_HIP(hipEventRecord(ctx->ev[0], st));
    DevicePlan P;
    prepare_device_join(ctx, indexed, probe, ir, pr, opt.algo, tune, true, P);
    if (P.counters_ready) {
      // counters zeroed by the index build; nothing more to prepare
    } else {
      reset_counters(ctx);
    }
    HS_CHECK_HIP(hipEventRecord(ctx->index_built, st));
    const uint32_t U = indexed->universe;
    const uint64_t entries = P.index_entries;
    ctx->h_index_offsets.ensure(((size_t)U + 1) * 4);
    ctx->h_index_ids.ensure(std::max<uint64_t>(entries, 1) * 4);
    HS_CHECK_HIP(hipStreamWaitEvent(ctx->copy_stream, ctx->index_built, 0));
    HS_CHECK_HIP(hipMemcpyAsync(ctx->h_index_offsets.ptr, ctx->ib.offsets.ptr, ((size_t)U + 1) * 4,
                                hipMemcpyDeviceToHost, ctx->copy_stream));
    if (entries)
      HS_CHECK_HIP(hipMemcpyAsync(ctx->h_index_ids.ptr, ctx->ib.set_ids.ptr, entries * 4, hipMemcpyDeviceToHost,
                                  ctx->copy_stream));
    const bool ppjoin = ctx->host_algo == HS_HOST_PPJOIN;  // the host threads filter as PPJoin
    if (ppjoin) {
      ctx->h_index_metas.ensure(std::max<uint64_t>(entries, 1) * 4);
      if (entries)
        HS_CHECK_HIP(hipMemcpyAsync(ctx->h_index_metas.ptr, ctx->ib.metas.ptr, entries * 4, hipMemcpyDeviceToHost,
                                    ctx->copy_stream));
    }
    HS_CHECK_HIP(hipEventRecord(ctx->index_copied, ctx->copy_stream));

    // ---- the split of the probe ids
    Stripes S;
    S.lo = pr.first;
    S.hi = pr.last + 1;
    S.host_limit = S.hi;
    const uint32_t np = pr.last - pr.first + 1;
    // static: [pr.first, cut) is the small part Rs, [cut, n) the large part Rl
    const uint32_t cut = opt.mode == HS_COPROC_STATIC
                             ? pr.first + (uint32_t)std::llround(opt.cpu_share * (double)np) : pr.last + 1;
    const bool dev_small = opt.mode == HS_COPROC_STATIC && opt.policy == HS_POLICY_GPU_CPU;
    if (opt.mode == HS_COPROC_STATIC) {
      if (dev_small) {
        S.lo = cut;  // the host's share is Rl; the device's Rs is launched below
      } else {
        S.host_limit = cut;
        S.hi = cut;  // the device's share Rl is launched below, in one claim
      }
    }
    const uint32_t nthreads = opt.cpu_threads ? opt.cpu_threads : std::max(1u, usable_cpus() - 1);
    cs.host_threads = nthreads;
    const uint32_t min_claim = opt.device_min_claim ? opt.device_min_claim : 16384u;

    // ---- host consumers
    const uint32_t* h_off = ctx->h_index_offsets.as<uint32_t>();
    const uint32_t* h_ids = ctx->h_index_ids.as<uint32_t>();
    const uint32_t* h_metas = ppjoin ? ctx->h_index_metas.as<uint32_t>() : nullptr;
    const int device = ctx->device;
    hipEvent_t copied = ctx->index_copied;
    const std::vector<uint32_t>& psf = probe->h_size_first;
    std::atomic<uint64_t> host_weight_done{0};  // work proxy of the stripes the host has joined
    std::atomic<int64_t> host_start_ns{0};      // when the first host thread had the index
    std::vector<HostWorker> workers(nthreads);
    std::vector<std::thread> pool;
    const auto t_host = Clock::now();
    for (uint32_t k = 0; k < nthreads; ++k) {
      pool.emplace_back([&, k] {
        HostWorker& w = workers[k];
        w.done = Clock::now();
        try {
          HS_CHECK_HIP(hipSetDevice(device));
          HS_CHECK_HIP(hipEventSynchronize(copied));
          int64_t zero = 0;
          host_start_ns.compare_exchange_strong(zero, Clock::now().time_since_epoch().count());
          uint32_t want = opt.cpu_block ? opt.cpu_block : 16u, first = 0, n = 0;
          while (S.take_low(want, nthreads, first, n)) {
            const auto a = Clock::now();
            host_probe_range(*probe, *indexed, h_off, h_ids, h_metas, T, self, ir, first, first + n - 1, w.out,
                             w.stamp);
            const double dt = ms_between(a, Clock::now());
            host_weight_done += probe_weight(psf, first, first + n);
            w.probes += n;
            ++w.claims;
            if (!opt.cpu_block) {
              const double per = std::max(dt / n, 1e-5);
              want = (uint32_t)std::clamp(kHostClaimMs / per, 1.0, 65536.0);
            }
          }
        } catch (const Error& e) {
          w.failed = true;
          w.err = e;
        } catch (...) {
          w.failed = true;
          w.err = Error{HS_E_INVALID, "host stripe failed"};
        }
        w.done = Clock::now();
      });
    }

    // ---- device consumer (this thread): up to two stripes in flight.  Until
    // both sides have a measured rate the device claims half of what is
    // left; then it claims the top part of the unclaimed ids sized so that
    // the host, joining the rest, ends when the device does.  With the
    // hybrid technique a claim is filtered on this thread (and
    // tune.host_threads - 1 more) before its candidates are verified.
    uint32_t dev_lo = pr.last + 1, dev_hi = pr.first;  // the device's share is [dev_lo, dev_hi)
    HostCounts dev_hc;
    uint32_t tile_plan = 0;
    std::string dev_err;
    int dev_rc = HS_OK;
    const auto t_dev = Clock::now();
    try {
      struct Claim {
        uint64_t weight;
        Clock::time_point launched;
        uint32_t ev;
      };
      std::deque<Claim> inflight;
      uint32_t claims = 0;
      uint64_t dev_weight_done = 0;
      double dev_busy_ms = 0.0;
      Clock::time_point prev_done = t_dev;
      auto launch = [&](uint32_t first, uint32_t n) {
        const auto t_launch = Clock::now();
        const hs_range cr{first, first + n - 1};
        if (hybrid) {
          HS_CHECK_HIP(hipEventSynchronize(copied));
          hybrid_filter_verify(ctx, indexed, probe, self, ir, cr, tune, h_off, h_ids, h_metas, &dev_hc);
        } else {
          tile_plan = launch_device_join(ctx, indexed, probe, P, cr, opt.algo, tune);
        }
        HS_CHECK_HIP(hipEventRecord(ctx->claim_ev[claims & 1], st));
        inflight.push_back(Claim{probe_weight(psf, first, first + n), t_launch, claims & 1u});
        ++claims;
        dev_lo = std::min(dev_lo, first);
        dev_hi = std::max(dev_hi, first + n);
        cs.device_probes += n;
      };
      // the next device claim [x, hi) under the lock, or false
      auto decide = [&](uint32_t& first, uint32_t& n) {
        std::lock_guard<std::mutex> g(S.mu);
        if (S.hi <= S.lo) return false;
        const uint32_t rem = S.hi - S.lo;
        const uint64_t hw = host_weight_done.load();
        const int64_t h0 = host_start_ns.load();
        const double host_ms = h0 ? ms_between(Clock::time_point(Clock::duration(h0)), Clock::now()) : 0.0;
        uint32_t x;
        if (dev_busy_ms <= 0.0 || hw == 0 || host_ms <= 0.0) {
          x = S.hi - std::min(rem, std::max(min_claim, rem - rem / 2));
        } else {
          const double rd = dev_weight_done / dev_busy_ms, rh = hw / host_ms;  // weight per ms
          double queued = 0.0;  // device work already launched
          for (const Claim& c : inflight) queued += c.weight;
          const double W = (double)probe_weight(psf, S.lo, S.hi);
          // host share u: u / rh = queued / rd + (W - u) / rd
          const double u = (queued / rd + W / rd) / (1.0 / rh + 1.0 / rd);
          x = u >= W ? S.hi : weight_cut(psf, S.lo, S.hi, (uint64_t)u);
          if (inflight.empty() && x == S.hi) x = S.hi - 1;  // an idle device always takes something
          if (!inflight.empty() && S.hi - x < std::max(1u, min_claim / 4)) return false;  // not worth it yet
        }
        if (x >= S.hi) return false;
        first = x;
        n = S.hi - x;
        S.hi = x;
        return true;
      };
      if (opt.mode == HS_COPROC_STATIC) {
        if (dev_small) {
          if (cut > pr.first) launch(pr.first, cut - pr.first);
        } else if (cut <= pr.last) {
          launch(cut, pr.last + 1 - cut);
        }
      }
      for (;;) {
        uint32_t first = 0, n = 0;
        while (opt.mode == HS_COPROC_QUEUE && inflight.size() < 2 && decide(first, n)) launch(first, n);
        if (inflight.empty()) {
          std::lock_guard<std::mutex> g(S.mu);
          if (S.hi <= S.lo || opt.mode == HS_COPROC_STATIC) break;
          continue;  // unreachable in practice: an idle device with work left always claims
        }
        const Claim c = inflight.front();
        HS_CHECK_HIP(hipEventSynchronize(ctx->claim_ev[c.ev]));
        inflight.pop_front();
        const auto now = Clock::now();
        dev_busy_ms += ms_between(std::max(c.launched, prev_done), now);
        dev_weight_done += c.weight;
        prev_done = now;
      }
      cs.device_claims = claims;
    } catch (const Error& e) {
      dev_rc = e.code;
      dev_err = e.msg;
      std::lock_guard<std::mutex> g(S.mu);
      S.host_limit = 0;  // stop the host threads after their current stripe
    }
    const auto t_dev_end = Clock::now();
    for (auto& th : pool) th.join();
    if (dev_rc != HS_OK) throw Error{dev_rc, dev_err};
    for (auto& w : workers)
      if (w.failed) throw w.err;
    HS_REQUIRE(S.lo == S.hi, HS_E_DEVICE, "probe stripes left unclaimed");
    cs.split = dev_small ? cut : std::max(dev_lo, S.hi);

    // ---- the device's pairs; a second pass over its share if they overflowed
    unsigned long long cnt[CNT_WORDS];
    read_counters(ctx, cnt);
    if (cnt[CNT_PAIRS] > ctx->out_cap) {
      ensure_output(ctx, cnt[CNT_PAIRS] + cnt[CNT_PAIRS] / 4 + 1024);
      reset_counters(ctx);
      if (dev_lo < dev_hi) {
        const hs_range dr{dev_lo, dev_hi - 1};
        if (hybrid)
          hybrid_filter_verify(ctx, indexed, probe, self, ir, dr, tune, h_off, h_ids, h_metas, nullptr);
        else
          launch_device_join(ctx, indexed, probe, P, dr, opt.algo, tune);
      }
      unsigned long long again[CNT_WORDS];
      read_counters(ctx, again);
      HS_REQUIRE(again[CNT_PAIRS] == cnt[CNT_PAIRS], HS_E_DEVICE, "result count changed between runs");
    }
    const uint64_t g = cnt[CNT_PAIRS];

    // ---- host pairs: kept on the host as the tail of the context's result
    uint64_t h = 0;
    set_device_result(ctx, g);
    for (auto& w : workers) {
      h += w.out.r.size();
      cs.host_probes += w.probes;
      cs.host_claims += w.claims;
      cs.host_ms = std::max(cs.host_ms, ms_between(t_host, w.done));
      if (!w.out.r.empty()) ctx->host_tail.push_back(std::move(w.out));
    }
    const uint64_t total = g + h;
    ctx->last_count = total;
    result->count = total;
    if (!result->count_only && total) {
      HS_REQUIRE(total <= result->capacity, HS_E_OVERFLOW, "result capacity too small");
      copy_pairs(ctx, total, result);
    }

    cs.device_pairs = g;
    cs.host_pairs = h;
    cs.index_ms = elapsed(ctx->ev[0], ctx->index_built);
    cs.copy_ms = elapsed(ctx->index_built, ctx->index_copied);
    cs.device_ms = ms_between(t_dev, t_dev_end);
    cs.wall_ms = ms_between(t_start, Clock::now());
    if (hist) {
      hist->shared_ms = cs.wall_ms;
      cs.device_alone_ms = std::max(hist->device_ms, 0.0);
      cs.shared_ms = cs.wall_ms;
    }
    ctx->stats.pairs = total;
    ctx->stats.index_ms = cs.index_ms;
    ctx->stats.total_ms = cs.wall_ms;
    // the hybrid device side counts its visits and candidates on the host
    ctx->stats.posting_visits = hybrid ? dev_hc.visits : cnt[CNT_VISITS];
    ctx->stats.candidates = hybrid ? dev_hc.candidates : cnt[CNT_CANDS];
    ctx->stats.candidate_tokens = cnt[CNT_CAND_TOKENS];
    ctx->stats.probe_prefix_tokens = cnt[CNT_ITEMS];
    ctx->stats.index_entries = entries;
    ctx->stats.tile_plan = tile_plan;
    ctx->stats.signature_tests = cnt[CNT_SIG_TESTS];
    ctx->stats.signature_bits = P.sig_words * 64;
    if (stats) *stats = cs;
  });
}

}  // extern "C"
