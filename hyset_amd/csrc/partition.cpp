// Partition pairs of one self-join: the partitioners, the cost model's
// technique choice per partition pair, and the runner of such a plan.
//
// Reference: include/hyset/partitioner.hpp:11-68 (even and dichotomize
// over blocks of consecutive sets), include/hyset/cost.hpp:28-58 (cost::run:
// per partition pair (i, j), i <= j, the features of statistics.hpp:163-283
// and the technique of least linear cost) and the execution half of
// src/cost_model.cu:150-262 (hybrid pairs, then CPU, bitmap and fgss pairs).
// Partitions are set-id ranges here: sets are sorted by size, so a
// partition of blocks is one range, and a pair (left, right) joins the probe
// sets of right against the sets of left with s < r.  The reference runs
// its CPU pairs on the main thread between the device ones; here they run on
// host threads while the device pairs run on the calling thread.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "ctx.h"

using namespace hs;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// Blocks of block_size consecutive sets over [0, n): block k = [k*B, min(n, (k+1)*B)).
uint32_t n_blocks(uint32_t n, uint32_t block) { return (uint32_t)(((uint64_t)n + block - 1) / block); }

hs_range blocks_range(uint32_t n, uint32_t block, uint64_t b0, uint64_t b1) {
  // set ids of blocks [b0, b1); empty when b1 <= b0
  if (b1 <= b0) return hs_range{1, 0};
  const uint64_t first = b0 * block, end = std::min<uint64_t>(n, b1 * block);
  return hs_range{(uint32_t)first, (uint32_t)(end - 1)};
}

// The technique indices in the order of the reference's std::map keys
// ("bitmap" < "cpu" < "fgssjoin" < "hybrid"): std::min_element keeps the
// first of equal costs (cost.hpp:48).
constexpr int32_t kMapOrder[4] = {HS_TECH_BITMAP, HS_TECH_CPU, HS_TECH_FGSS, HS_TECH_HYBRID};

long double linear_cost(const hs_cost_coefficients& c, const float* f, uint32_t bitmap, bool wrap_square) {
  // statistics::base holds sizes as unsigned and the moments as float
  const uint32_t size = (uint32_t)f[0], universe = (uint32_t)f[4], index = (uint32_t)f[5];
  const long double sq = wrap_square ? (long double)(uint32_t)(size * size) : (long double)size * (long double)size;
  const long double v = (long double)c.set_size_std * f[2] + (long double)c.average_set_size * f[1] +
                        (long double)c.token_freq * f[3] + (long double)c.size * size + (long double)c.size_square * sq +
                        (long double)c.universe_size * universe + (long double)c.index_size * index +
                        (long double)c.bitmap * bitmap + (long double)c.intercept;
  return v < 0 ? -v : v;
}

int32_t algo_of(int32_t tech) {
  return tech == HS_TECH_FGSS ? HS_ALGO_FGSS : tech == HS_TECH_BITMAP ? HS_ALGO_BITMAP : HS_ALGO_HYBRID;
}

}  // namespace

extern "C" {

int hs_partition_even(uint32_t n_sets, uint32_t block_size, uint32_t parts, hs_range* out, uint32_t* n_out) {
  if (!out || !n_out || block_size == 0 || parts == 0) return HS_E_INVALID;
  const uint32_t nb = n_blocks(n_sets, block_size);
  // partitioner.hpp:14: ceil(blocks / partitions) blocks each
  const uint32_t per = (uint32_t)std::ceil((double)nb / (double)parts);
  uint32_t k = 0;
  for (uint64_t b = 0; b < nb; b += per) out[k++] = blocks_range(n_sets, block_size, b, std::min<uint64_t>(nb, b + per));
  *n_out = k;
  return HS_OK;
}

int hs_partition_dichotomize(uint32_t n_sets, uint32_t block_size, double split_pct, hs_range out[3]) {
  if (!out || block_size == 0 || !(split_pct >= 0.0 && split_pct <= 100.0)) return HS_E_INVALID;
  const uint32_t nb = n_blocks(n_sets, block_size);
  // partitioner.hpp:35: ceil((split / 100) * blocks) blocks in the small part
  const uint64_t portion = std::min<uint64_t>(nb, (uint64_t)std::ceil((split_pct / 100.0) * nb));
  out[0] = blocks_range(n_sets, block_size, 0, nb);
  out[1] = blocks_range(n_sets, block_size, 0, portion);
  out[2] = blocks_range(n_sets, block_size, portion, nb);
  return HS_OK;
}

int hs_cost_run(hs_ctx* ctx, hs_coll* coll, const hs_range* parts, uint32_t n_parts,
                const hs_cost_coefficients coef[4], uint32_t bitmap_bits, uint32_t flags, int32_t* technique,
                float* features, double* costs) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    HS_REQUIRE(coll && parts && coef && technique && n_parts > 0, HS_E_INVALID, "NULL argument");
    bool any = false;
    for (int k = 0; k < 4; ++k) any = any || coef[k].enabled;
    HS_REQUIRE(any, HS_E_INVALID, "no technique enabled");
    for (uint32_t i = 0; i < n_parts; ++i)
      HS_REQUIRE(parts[i].first <= parts[i].last && parts[i].last < coll->n, HS_E_INVALID,
                 "partitions must be non-empty ranges of the collection");
    const bool wrap = (flags & HS_COST_WRAP_SQUARE) != 0;
    uint64_t q = 0;
    float f[HS_NFEATURES];
    for (uint32_t i = 0; i < n_parts; ++i)
      for (uint32_t j = i; j < n_parts; ++j, ++q) {
        // cost.hpp:41-46: the partition's own features on the diagonal, the pair's elsewhere
        const int rc = i == j ? hs_features(ctx, coll, &parts[i], f) : hs_features_pair(ctx, coll, &parts[i], &parts[j], f);
        if (rc != HS_OK) throw Error{rc, ctx->err};
        int32_t best = -1;
        long double best_cost = 0;
        for (int32_t k : kMapOrder) {
          if (!coef[k].enabled) {
            if (costs) costs[q * 4 + k] = -1.0;
            continue;
          }
          const long double c = linear_cost(coef[k], f, bitmap_bits, wrap);
          if (costs) costs[q * 4 + k] = (double)c;
          if (best < 0 || c < best_cost) {
            best = k;
            best_cost = c;
          }
        }
        technique[q] = best;
        if (features) std::copy(f, f + HS_NFEATURES, features + q * HS_NFEATURES);
      }
  });
}

int hs_run_partition_pairs(hs_ctx* ctx, hs_coll* coll, hs_partition_pair* pairs, uint32_t n_pairs,
                           uint32_t cpu_threads, const hs_tune* tune_in, hs_result* result, double* wall_ms) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    HS_REQUIRE(coll && result && (pairs || n_pairs == 0), HS_E_INVALID, "NULL argument");
    HS_REQUIRE(coll->n_joins == 0 && coll->device == ctx->device, HS_E_INVALID, "bad collection");
    HS_REQUIRE(ctx->device >= 0, HS_E_INVALID, "the plan needs a device context");
    for (uint32_t p = 0; p < n_pairs; ++p) {
      HS_REQUIRE(pairs[p].technique >= HS_TECH_FGSS && pairs[p].technique <= HS_TECH_CPU, HS_E_INVALID,
                 "bad technique");
      HS_REQUIRE(pairs[p].left.first <= pairs[p].left.last && pairs[p].left.last < coll->n &&
                     pairs[p].right.first <= pairs[p].right.last && pairs[p].right.last < coll->n,
                 HS_E_INVALID, "partition pairs must be non-empty ranges of the collection");
    }
    const hs_tune tune = tune_in ? *tune_in : hs_tune{};
    const bool keep = !result->count_only;
    ensure_bounds(ctx, coll->maxlen);  // the host threads read these tables while device joins run
    const BoundTables T = ctx->tables;
    const bool ppjoin = ctx->host_algo == HS_HOST_PPJOIN;
    std::vector<uint32_t> cpu, dev;
    for (uint32_t p = 0; p < n_pairs; ++p) (pairs[p].technique == HS_TECH_CPU ? cpu : dev).push_back(p);
    // host threads take the pairs with the most probe sets first
    std::stable_sort(cpu.begin(), cpu.end(), [&](uint32_t a, uint32_t b) {
      return pairs[a].right.last - pairs[a].right.first > pairs[b].right.last - pairs[b].right.first;
    });
    // device pairs: hybrid, then bitmap, then fgss (cost_model.cu:150-262)
    auto rank = [](int32_t t) { return t == HS_TECH_HYBRID ? 0 : t == HS_TECH_BITMAP ? 1 : 2; };
    std::stable_sort(dev.begin(), dev.end(),
                     [&](uint32_t a, uint32_t b) { return rank(pairs[a].technique) < rank(pairs[b].technique); });
    const auto t0 = Clock::now();
    std::vector<Triples> host_parts(n_pairs);  // per pair, kept pairs (unless count-only)
    std::atomic<uint32_t> next{0};
    std::atomic<bool> stop{false};
    std::vector<Error> errs(std::max<uint32_t>(1, cpu_threads), Error{HS_OK, ""});
    std::vector<std::thread> pool;
    const uint32_t nt = cpu.empty() ? 0 : std::max<uint32_t>(1, cpu_threads);
    for (uint32_t k = 0; k < nt; ++k)
      pool.emplace_back([&, k] {
        try {
          for (;;) {
            const uint32_t i = next.fetch_add(1);
            if (i >= cpu.size() || stop.load()) break;
            hs_partition_pair& pp = pairs[cpu[i]];
            const auto a = Clock::now();
            std::vector<Triples> parts;
            pp.pairs = host_join(*coll, *coll, true, pp.left, pp.right, T, 1, keep ? &parts : nullptr, ppjoin);
            pp.ms = ms_between(a, Clock::now());
            if (keep && !parts.empty()) host_parts[cpu[i]] = std::move(parts[0]);
          }
        } catch (const Error& e) {
          errs[k] = e;
          stop = true;
        } catch (...) {
          errs[k] = Error{HS_E_INVALID, "host partition join failed"};
          stop = true;
        }
      });
    // the device pairs on this thread
    int rc = HS_OK;
    std::string err;
    std::vector<Triples> dev_parts(n_pairs);
    for (uint32_t p : dev) {
      if (stop.load()) break;
      hs_partition_pair& pp = pairs[p];
      hs_result res{};
      res.count_only = 1;
      const auto a = Clock::now();
      const int r = hs_join(ctx, coll, coll, &pp.left, &pp.right, algo_of(pp.technique), &tune, &res);
      if (r != HS_OK) {
        rc = r;
        err = ctx->err;
        stop = true;
        break;
      }
      pp.pairs = res.count;
      if (keep && res.count) {
        Triples& t = dev_parts[p];
        t.r.resize(res.count);
        t.s.resize(res.count);
        t.ov.resize(res.count);
        stream_pairs(ctx, res.count, true, 0,
                     [&](const uint32_t* r, const uint32_t* s, const uint32_t* ov, uint64_t first, uint64_t n) {
                       std::memcpy(t.r.data() + first, r, n * 4);
                       std::memcpy(t.s.data() + first, s, n * 4);
                       std::memcpy(t.ov.data() + first, ov, n * 4);
                     });
      }
      pp.ms = ms_between(a, Clock::now());
    }
    for (auto& th : pool) th.join();
    if (rc != HS_OK) throw Error{rc, err};
    for (const auto& e : errs)
      if (e.code != HS_OK) throw e;
    uint64_t total = 0;
    for (uint32_t p = 0; p < n_pairs; ++p) total += pairs[p].pairs;
    // the result: device pairs in run order, then the CPU pairs, all on the host
    set_device_result(ctx, 0);
    if (keep) {
      for (uint32_t p : dev)
        if (!dev_parts[p].r.empty()) ctx->host_tail.push_back(std::move(dev_parts[p]));
      for (uint32_t p : cpu)
        if (!host_parts[p].r.empty()) ctx->host_tail.push_back(std::move(host_parts[p]));
      ctx->last_count = total;
    }
    result->count = total;
    ctx->stats = hs_stats{};
    ctx->stats.pairs = total;
    ctx->stats.total_ms = ms_between(t0, Clock::now());
    if (wall_ms) *wall_ms = ctx->stats.total_ms;
    if (keep && total) {
      HS_REQUIRE(total <= result->capacity, HS_E_OVERFLOW, "result capacity too small");
      copy_pairs(ctx, total, result);
    }
  });
}

}  // extern "C"
