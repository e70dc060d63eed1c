// C++ host layer over the C ABI of libhyset_amd.so that restores the
// reference's handler interface, so its drivers' call sites keep their shape.
//
// Reference interface each type mirrors (paths relative to the reference):
//   output::handler / count_handler / pairs_handler   include/hyset/output.hpp:13-37
//   structs::block / structs::partition               include/hyset/structs.hpp:10-54, :105-148
//   collection::device_collection                     include/hyset/collection.hpp:54-71
//   algorithms::device::bitmap::handler               include/hyset/device_algorithms.hpp:120-260
//   algorithms::device::fgssjoin::handler             include/hyset/device_algorithms.hpp:305-446
//   algorithms::hybrid::device_handler                include/hyset/hybrid_algorithms.hpp:16-205
//   algorithms::host::handler (allpairs / ppjoin)     include/hyset/host_algorithms.hpp:16-250,
//                                                     src/cpu_single.cu:118-135, src/cpu_multi.cu:118-207
//   partitioner::even / dichotomize                   include/hyset/partitioner.hpp:11-68
//   cost::run, cost::read_coefficients               include/hyset/cost.hpp:28-58,
//                                                     src/cost_model.cu:296-323
//
// Differences from the reference, all deliberate:
//   * errors throw hyset_amd::error carrying the HS_E_* code and message
//     (the reference lets CUDA exceptions escape);
//   * the output sink is chosen by the `aggregate` flag exactly as in the
//     reference, but pairs_handler also keeps the overlap of every pair;
//   * handlers take a hyset_amd::context (one HIP stream) instead of a
//     device timer; timings are available through context::stats().
#pragma once

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <future>
#include <map>
#include <memory>
#include <sstream>
#include <utility>
#include <stdexcept>
#include <string>
#include <vector>

#include "../hyset_amd.h"

namespace hyset_amd {

struct error : std::runtime_error {
  int code;
  error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, hs_ctx* ctx, const char* what) {
  if (rc != HS_OK) throw error(rc, std::string(what) + ": " + (ctx ? hs_last_error(ctx) : "failed"));
}

// ------------------------------------------------------------------ output

namespace output {
struct handler {
  virtual void addPair(const unsigned int& rec1, const unsigned int& rec2, unsigned int overlap) = 0;
  virtual unsigned long getCount() const = 0;
  virtual ~handler() = default;
};

struct count_handler : handler {
  unsigned long count = 0;
  void addPair(const unsigned int&, const unsigned int&, unsigned int) override { ++count; }
  unsigned long getCount() const override { return count; }
};

struct pair {
  unsigned int first, second, overlap;
};

struct pairs_handler : handler {
  std::vector<pair> pairs;
  void addPair(const unsigned int& a, const unsigned int& b, unsigned int ov) override { pairs.push_back({a, b, ov}); }
  unsigned long getCount() const override { return (unsigned long)pairs.size(); }
};
}  // namespace output

// ----------------------------------------------------------------- structs

namespace structs {
// A run of consecutive set ids (include/hyset/structs.hpp:10-54).
struct block {
  unsigned int id = 0, startID = 0, endID = 0;
  unsigned int size() const { return endID - startID + 1; }
};

// A list of blocks, or a plain id range (include/hyset/structs.hpp:105-148).
struct partition {
  unsigned int id = 0;
  std::vector<block*> blocks;
  bool blockBased = true;
  unsigned int startID = 0, endID = 0;
  explicit partition(unsigned int i) : id(i) {}
  partition(unsigned int i, unsigned int s, unsigned int e) : id(i), blockBased(false), startID(s), endID(e) {}
  unsigned int start_id() const { return blockBased ? blocks.front()->startID : startID; }
  unsigned int end_id() const { return blockBased ? blocks.back()->endID : endID; }
};

// Cuts [0, n) into blocks of blockSize sets (collection.hpp:147-172).
inline std::vector<block> make_blocks(unsigned int n, unsigned int blockSize) {
  std::vector<block> out;
  for (unsigned int s = 0, id = 0; s < n; s += blockSize, ++id)
    out.push_back(block{id, s, (n - s > blockSize ? s + blockSize : n) - 1});
  return out;
}
}  // namespace structs

// ----------------------------------------------------- context, collection

class context {
 public:
  explicit context(int device = 0) { check(hs_create(device, &ctx_), nullptr, "hs_create"); }
  ~context() { hs_destroy(ctx_); }
  context(const context&) = delete;
  context& operator=(const context&) = delete;
  hs_ctx* get() const { return ctx_; }
  hs_stats stats() const {
    hs_stats s{};
    hs_get_stats(ctx_, &s);
    return s;
  }

 private:
  hs_ctx* ctx_ = nullptr;
};

namespace collection {
// Device-resident collection (collection.hpp:54-71).
class device_collection {
 public:
  device_collection(context& ctx, const std::vector<unsigned int>& tokens, const std::vector<unsigned int>& starts,
                    const std::vector<unsigned int>& sizes)
      : sizes_(sizes) {
    check(hs_upload_collection(ctx.get(), tokens.data(), starts.data(), sizes.data(), (uint32_t)sizes.size(),
                               tokens.size(), &coll_),
          ctx.get(), "hs_upload_collection");
  }
  ~device_collection() { hs_free_collection(coll_); }
  device_collection(const device_collection&) = delete;
  device_collection& operator=(const device_collection&) = delete;
  hs_coll* get() const { return coll_; }
  const std::vector<unsigned int>& sizes() const { return sizes_; }

 private:
  hs_coll* coll_ = nullptr;
  std::vector<unsigned int> sizes_;
};
}  // namespace collection

namespace detail {
// Appends the pairs of the last join on ctx to a pairs sink, streamed from
// the device in pinned chunks (hs_stream_last_result).
inline int append_chunk(void* user, const uint32_t* r, const uint32_t* s, const uint32_t* ov, uint64_t n) {
  auto& pairs = static_cast<output::pairs_handler*>(user)->pairs;
  for (uint64_t i = 0; i < n; ++i) pairs.push_back({r[i], s[i], ov[i]});
  return 0;
}
inline void deliver_pairs(hs_ctx* ctx, uint64_t count, output::pairs_handler& p) {
  p.pairs.reserve(p.pairs.size() + count);
  check(hs_stream_last_result(ctx, 0, 1, append_chunk, &p), ctx, "hs_stream_last_result");
}

// Runs one device join and moves its result into the sink the reference
// would have downcast to (device_algorithms.hpp:193-221).
inline void run_join(context& ctx, collection::device_collection& indexed, collection::device_collection& probe,
                     hs_range ir, hs_range pr, int32_t algo, const hs_tune& tune, bool aggregate,
                     const std::shared_ptr<output::handler>& out) {
  hs_result res{};
  res.count_only = 1;
  check(hs_join(ctx.get(), indexed.get(), probe.get(), &ir, &pr, algo, &tune, &res), ctx.get(), "hs_join");
  if (!out) throw error(HS_E_INVALID, "no output handler set");
  if (aggregate) {
    auto* c = dynamic_cast<output::count_handler*>(out.get());
    if (!c) throw error(HS_E_INVALID, "aggregate mode needs a count_handler");
    c->count += res.count;
    return;
  }
  auto* p = dynamic_cast<output::pairs_handler*>(out.get());
  if (!p) throw error(HS_E_INVALID, "pairs mode needs a pairs_handler");
  deliver_pairs(ctx.get(), res.count, *p);
}

// The self-join block-pair rule of length_filter (device_algorithms.hpp:21-53):
// pairs with s < r in [first.start, first.end] x [second.start, second.end].
inline hs_range range_of(const structs::partition* p) { return hs_range{p->start_id(), p->end_id()}; }
inline hs_range range_of(const structs::block* b) { return hs_range{b->startID, b->endID}; }
}  // namespace detail

// ------------------------------------------------------------ techniques

namespace algorithms {
namespace device {

namespace bitmap {
struct handler {
  handler(context& ctx, collection::device_collection& coll, unsigned int bitmapSize, unsigned int blockSize,
          bool aggregate, double threshold)
      : ctx_(ctx), coll_(coll), blockSize(blockSize), aggregate(aggregate), threshold(threshold) {
    tune_.bitmap_bits = bitmapSize;
    tune_.flags = HS_TUNE_STAGE_TIMES;  // per-stage times, as the reference's device timer
    check(hs_set_threshold(ctx.get(), HS_JACCARD, threshold, HS_BOUNDS_REFERENCE), ctx.get(), "hs_set_threshold");
  }
  void setOutputHandler(std::shared_ptr<output::handler>& h) { output_ = h; }
  void join(const structs::block* first, const structs::block* second) { run(detail::range_of(first), detail::range_of(second)); }
  void join(const structs::partition* first, const structs::partition* second) {
    run(detail::range_of(first), detail::range_of(second));
  }
  unsigned int blockSize;
  bool aggregate;
  double threshold;

 private:
  void run(hs_range ir, hs_range pr) {
    check(hs_set_threshold(ctx_.get(), HS_JACCARD, threshold, HS_BOUNDS_REFERENCE), ctx_.get(), "hs_set_threshold");
    detail::run_join(ctx_, coll_, coll_, ir, pr, HS_ALGO_BITMAP, tune_, aggregate, output_);
  }
  context& ctx_;
  collection::device_collection& coll_;
  hs_tune tune_{};
  std::shared_ptr<output::handler> output_;
};
}  // namespace bitmap

namespace fgssjoin {
struct handler {
  handler(context& ctx, collection::device_collection& coll, unsigned int blockSize, bool aggregate, double threshold)
      : ctx_(ctx), coll_(coll), blockSize(blockSize), aggregate(aggregate), threshold(threshold) {
    tune_.flags = HS_TUNE_STAGE_TIMES;  // per-stage times, as the reference's device timer
    check(hs_set_threshold(ctx.get(), HS_JACCARD, threshold, HS_BOUNDS_REFERENCE), ctx.get(), "hs_set_threshold");
  }
  void setOutputHandler(std::shared_ptr<output::handler>& h) { output_ = h; }
  // The reference copies a host-built index here (device_algorithms.hpp:354-361);
  // the device builds its own index inside every join, so this is a no-op kept
  // for call-site compatibility.
  template <typename Index>
  void transferIndex(Index*) const {}
  void join(const structs::block* first, const structs::block* second) { run(detail::range_of(first), detail::range_of(second)); }
  void join(const structs::partition* first, const structs::partition* second) {
    run(detail::range_of(first), detail::range_of(second));
  }
  unsigned int blockSize;
  bool aggregate;
  double threshold;

 private:
  void run(hs_range ir, hs_range pr) {
    check(hs_set_threshold(ctx_.get(), HS_JACCARD, threshold, HS_BOUNDS_REFERENCE), ctx_.get(), "hs_set_threshold");
    detail::run_join(ctx_, coll_, coll_, ir, pr, HS_ALGO_FGSS, tune_, aggregate, output_);
  }
  context& ctx_;
  collection::device_collection& coll_;
  hs_tune tune_{};
  std::shared_ptr<output::handler> output_;
};
}  // namespace fgssjoin

}  // namespace device

namespace hybrid {
// CPU-filtered candidates verified on the device
// (include/hyset/hybrid_algorithms.hpp:16-205).  Candidates accumulate in a
// host batch of maxCandidates; a full batch or join() hands it to a worker
// thread that verifies it (hs_verify_candidates) and delivers its pairs
// while the caller goes on filtering into the other batch -- the overlap of
// hybrid_algorithms.hpp:107-120, one batch in flight.  flush() returns once
// every batch is delivered; an error of a batch surfaces from the next
// join() or flush().
struct device_handler {
  device_handler(context& ctx, collection::device_collection& indexed, collection::device_collection& probe,
                 unsigned int threads, unsigned int scenario, bool aggregate, unsigned int maxProbes,
                 unsigned int maxCandidates, double threshold)
      : threads(threads), scenario(scenario), aggregate(aggregate), threshold(threshold), maxProbes(maxProbes),
        maxCandidates(maxCandidates), ctx_(ctx), indexed_(indexed), probe_(probe) {
    check(hs_set_threshold(ctx.get(), HS_JACCARD, threshold, HS_BOUNDS_REFERENCE), ctx.get(), "hs_set_threshold");
  }
  void setMaxSetSize(unsigned int s) { maxSetSize = s; }
  void setOutputHandler(std::shared_ptr<output::handler>& h) { output_ = h; }
  // hybrid_algorithms.hpp:71-105: a candidate list may be split across two
  // batches when it does not fit.
  void addCandidates(unsigned int probeID, const std::vector<unsigned int>& candidateSet) {
    size_t done = 0;
    while (done < candidateSet.size()) {
      size_t room = maxCandidates > candidates_.size() ? maxCandidates - candidates_.size() : 0;
      if (room == 0 || probes_.size() >= maxProbes) {
        join();
        continue;
      }
      size_t take = std::min(room, candidateSet.size() - done);
      candidates_.insert(candidates_.end(), candidateSet.begin() + done, candidateSet.begin() + done + take);
      probes_.push_back(probeID);
      offsets_.push_back((unsigned int)candidates_.size());
      done += take;
    }
  }
  ~device_handler() {
    try {
      wait();
    } catch (...) {
    }
  }
  device_handler(const device_handler&) = delete;
  device_handler& operator=(const device_handler&) = delete;
  void join() {
    if (probes_.empty()) return;
    wait();  // one verification at a time on the context; it delivers before the next starts
    std::swap(probes_, f_probes_);
    std::swap(offsets_, f_offsets_);
    std::swap(candidates_, f_candidates_);
    probes_.clear();
    offsets_.clear();
    candidates_.clear();
    inflight_ = std::async(std::launch::async, [this] { verify_batch(); });
  }
  void flush() {
    join();
    wait();
  }

  unsigned int threads, scenario;
  bool aggregate;
  double threshold;
  unsigned int maxProbes, maxCandidates, maxSetSize = 0;

 private:
  void wait() {
    if (inflight_.valid()) inflight_.get();  // rethrows the batch's error
  }
  void verify_batch() {
    check(hs_set_threshold(ctx_.get(), HS_JACCARD, threshold, HS_BOUNDS_REFERENCE), ctx_.get(), "hs_set_threshold");
    hs_result res{};
    res.count_only = 1;
    check(hs_verify_candidates(ctx_.get(), indexed_.get(), probe_.get(), f_probes_.data(), f_offsets_.data(),
                               f_candidates_.data(), (uint32_t)f_probes_.size(), (int32_t)scenario, &res),
          ctx_.get(), "hs_verify_candidates");
    deliver(res.count);
  }
  void deliver(uint64_t count) {
    if (!output_) throw error(HS_E_INVALID, "no output handler set");
    if (aggregate) {
      dynamic_cast<output::count_handler&>(*output_).count += count;
      return;
    }
    detail::deliver_pairs(ctx_.get(), count, dynamic_cast<output::pairs_handler&>(*output_));
  }
  context& ctx_;
  collection::device_collection& indexed_;
  collection::device_collection& probe_;
  std::vector<unsigned int> probes_, offsets_, candidates_;        // the batch being filled
  std::vector<unsigned int> f_probes_, f_offsets_, f_candidates_;  // the batch in flight
  std::future<void> inflight_;
  std::shared_ptr<output::handler> output_;
};
}  // namespace hybrid

namespace host {
// The CPU techniques of cpu_single / cpu_multi: host::allpairs or
// host::ppjoin (include/hyset/host_algorithms.hpp:16-250) over `threads`
// host threads (src/cpu_multi.cu:118-207), through hs_join_host.
// algorithm is "allpairs" or "ppjoin" (the reference's --algorithm).
struct handler {
  handler(context& ctx, collection::device_collection& coll, unsigned int threads, const std::string& algorithm,
          bool aggregate, double threshold)
      : threads(threads), aggregate(aggregate), threshold(threshold), ctx_(ctx), coll_(coll) {
    if (algorithm == "allpairs") algo_ = HS_HOST_ALLPAIRS;
    else if (algorithm == "ppjoin") algo_ = HS_HOST_PPJOIN;
    else throw error(HS_E_INVALID, "CPU algorithm must be allpairs or ppjoin");
  }
  void setOutputHandler(std::shared_ptr<output::handler>& h) { output_ = h; }
  void join(const structs::partition* first, const structs::partition* second) {
    run(detail::range_of(first), detail::range_of(second));
  }
  void join(const structs::block* first, const structs::block* second) {
    run(detail::range_of(first), detail::range_of(second));
  }
  unsigned int threads;
  bool aggregate;
  double threshold;

 private:
  void run(hs_range ir, hs_range pr) {
    check(hs_set_threshold(ctx_.get(), HS_JACCARD, threshold, HS_BOUNDS_REFERENCE), ctx_.get(), "hs_set_threshold");
    check(hs_set_host_algorithm(ctx_.get(), algo_), ctx_.get(), "hs_set_host_algorithm");
    if (!output_) throw error(HS_E_INVALID, "no output handler set");
    if (aggregate) {
      auto* c = dynamic_cast<output::count_handler*>(output_.get());
      if (!c) throw error(HS_E_INVALID, "aggregate mode needs a count_handler");
      hs_result res{};
      res.count_only = 1;
      check(hs_join_host(ctx_.get(), coll_.get(), coll_.get(), &ir, &pr, threads, &res), ctx_.get(), "hs_join_host");
      c->count += res.count;
      return;
    }
    auto* p = dynamic_cast<output::pairs_handler*>(output_.get());
    if (!p) throw error(HS_E_INVALID, "pairs mode needs a pairs_handler");
    // arrays for as many pairs as probes; a larger result runs again at its size
    uint64_t cap = std::max<uint64_t>(1u << 16, pr.last >= pr.first ? (uint64_t)pr.last - pr.first + 1 : 0);
    for (;;) {
      std::vector<uint32_t> r(cap), s(cap), ov(cap);
      hs_result res{};
      res.capacity = cap;
      res.r = r.data();
      res.s = s.data();
      res.overlap = ov.data();
      const int rc = hs_join_host(ctx_.get(), coll_.get(), coll_.get(), &ir, &pr, threads, &res);
      if (rc == HS_E_OVERFLOW && res.count > cap) {
        cap = res.count;
        continue;
      }
      check(rc, ctx_.get(), "hs_join_host");
      for (uint64_t i = 0; i < res.count; ++i) p->pairs.push_back({r[i], s[i], ov[i]});
      return;
    }
  }
  context& ctx_;
  collection::device_collection& coll_;
  int32_t algo_ = HS_HOST_ALLPAIRS;
  std::shared_ptr<output::handler> output_;
};
}  // namespace host
}  // namespace algorithms

// ------------------------------------------------ partitioner, cost model

namespace partitioner {
// partitioner::even (partitioner.hpp:11-30): ceil(blocks / n) consecutive
// blocks per partition.
inline std::vector<structs::partition> even(std::vector<structs::block>& blocks, unsigned int numOfPartitions) {
  std::vector<structs::partition> out;
  if (numOfPartitions == 0) throw error(HS_E_INVALID, "no partitions");
  const unsigned int per = (unsigned int)std::ceil((double)blocks.size() / (double)numOfPartitions);
  unsigned int id = 0;
  for (size_t i = 0; i < blocks.size(); i += per) {
    structs::partition p(id++);
    for (size_t b = i; b < std::min(blocks.size(), i + per); ++b) p.blocks.push_back(&blocks[b]);
    out.push_back(p);
  }
  return out;
}

// partitioner::dichotomize (partitioner.hpp:32-68): {complete, the first
// ceil(split / 100 * blocks) blocks, the rest}; an empty part has no blocks.
inline std::vector<structs::partition> dichotomize(std::vector<structs::block>& blocks, double split) {
  const size_t portion = std::min(blocks.size(), (size_t)std::ceil((split / 100.0) * blocks.size()));
  std::vector<structs::partition> out{structs::partition(0), structs::partition(1), structs::partition(2)};
  for (size_t b = 0; b < blocks.size(); ++b) {
    out[0].blocks.push_back(&blocks[b]);
    out[b < portion ? 1 : 2].blocks.push_back(&blocks[b]);
  }
  return out;
}
}  // namespace partitioner

namespace cost {
// The reference's technique names (cost_model.cu:304) and HS_TECH_* values.
inline const std::vector<std::pair<std::string, int32_t>>& techniques() {
  static const std::vector<std::pair<std::string, int32_t>> t{
      {"bitmap", HS_TECH_BITMAP}, {"cpu", HS_TECH_CPU}, {"fgssjoin", HS_TECH_FGSS}, {"hybrid", HS_TECH_HYBRID}};
  return t;
}

namespace detail {
// A JSON value reduced to what a coefficient file holds: objects and numbers.
struct json {
  std::map<std::string, json> obj;
  double num = 0.0;
  bool is_num = false;
};
struct json_reader {
  const std::string& s;
  size_t i = 0;
  void ws() {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
  }
  [[noreturn]] void fail(const char* what) const {
    throw error(HS_E_INVALID, std::string("coefficient file: ") + what + " at offset " + std::to_string(i));
  }
  std::string str() {
    if (s[i] != '"') fail("expected a string");
    std::string out;
    for (++i; i < s.size() && s[i] != '"'; ++i) {
      if (s[i] == '\\' && i + 1 < s.size()) ++i;
      out += s[i];
    }
    if (i >= s.size()) fail("unterminated string");
    ++i;
    return out;
  }
  json value() {
    ws();
    if (i >= s.size()) fail("unexpected end");
    json v;
    if (s[i] == '{') {
      ++i;
      ws();
      if (s[i] == '}') {
        ++i;
        return v;
      }
      for (;;) {
        ws();
        std::string k = str();
        ws();
        if (s[i] != ':') fail("expected ':'");
        ++i;
        v.obj[k] = value();
        ws();
        if (s[i] == ',') {
          ++i;
          continue;
        }
        if (s[i] == '}') {
          ++i;
          return v;
        }
        fail("expected ',' or '}'");
      }
    }
    char* end = nullptr;
    v.num = std::strtod(s.c_str() + i, &end);
    if (end == s.c_str() + i) fail("expected a number or an object");
    i = (size_t)(end - s.c_str());
    v.is_num = true;
    return v;
  }
};
inline double number(const json& j, const std::string& key, const char* technique) {
  auto it = j.obj.find(key);
  if (it == j.obj.end() || !it->second.is_num)
    throw error(HS_E_INVALID, std::string("coefficient file: no number ") + key + " for " + technique);
  return it->second.num;
}
}  // namespace detail

// read_coefficients (src/cost_model.cu:296-323): per technique the entry of
// the threshold printed with two significant digits ("0.8", "0.85"), terms
// Intercept, set_size_std, avg_set_size, token_freq, dataset_size,
// dataset_size_square, universe_size, and index_size (not bitmap) or bitmap
// (bitmap only).
inline std::map<std::string, hs_cost_coefficients> read_coefficients(const std::string& path, double threshold) {
  std::ifstream f(path);
  if (!f) throw error(HS_E_INVALID, "cannot read " + path);
  std::stringstream buf;
  buf << f.rdbuf();
  const std::string text = buf.str();
  detail::json_reader rd{text};
  const detail::json root = rd.value();
  char key[32];
  std::snprintf(key, sizeof(key), "%.2g", threshold);
  std::map<std::string, hs_cost_coefficients> out;
  for (const auto& [name, tech] : techniques()) {
    (void)tech;
    auto t = root.obj.find(name);
    if (t == root.obj.end()) throw error(HS_E_INVALID, "coefficient file: no technique " + name);
    auto e = t->second.obj.find(key);
    if (e == t->second.obj.end()) throw error(HS_E_INVALID, "coefficient file: no threshold " + std::string(key));
    const detail::json& j = e->second;
    hs_cost_coefficients c{};
    c.intercept = detail::number(j, "Intercept", name.c_str());
    c.set_size_std = detail::number(j, "set_size_std", name.c_str());
    c.average_set_size = detail::number(j, "avg_set_size", name.c_str());
    c.token_freq = detail::number(j, "token_freq", name.c_str());
    c.size = detail::number(j, "dataset_size", name.c_str());
    c.size_square = detail::number(j, "dataset_size_square", name.c_str());
    c.universe_size = detail::number(j, "universe_size", name.c_str());
    if (name == "bitmap") c.bitmap = detail::number(j, "bitmap", name.c_str());
    else c.index_size = detail::number(j, "index_size", name.c_str());
    c.enabled = 1;
    out[name] = c;
  }
  return out;
}

using partition_pair = std::pair<structs::partition*, structs::partition*>;

// cost::run (cost.hpp:28-58): every partition pair (left, right), right.id
// >= left.id, under the technique of least modelled cost; techniques
// missing from `coefficients` (as bitmap when --bitmap is 0) are no choice.
inline std::map<std::string, std::vector<partition_pair>> run(std::map<std::string, hs_cost_coefficients>& coefficients,
                                                              std::vector<structs::partition>& partitions,
                                                              context& ctx, collection::device_collection& coll,
                                                              unsigned int bitmapSize) {
  hs_cost_coefficients coef[4] = {};
  std::map<int32_t, std::string> name_of;
  for (const auto& [name, tech] : techniques()) {
    name_of[tech] = name;
    auto it = coefficients.find(name);
    if (it != coefficients.end()) coef[tech] = it->second;
    else coef[tech].enabled = 0;
  }
  std::vector<hs_range> ranges;
  for (auto& p : partitions) ranges.push_back(hs_range{p.start_id(), p.end_id()});
  const size_t n = partitions.size();
  std::vector<int32_t> tech(n * (n + 1) / 2);
  check(hs_cost_run(ctx.get(), coll.get(), ranges.data(), (uint32_t)n, coef, bitmapSize, 0, tech.data(), nullptr,
                    nullptr),
        ctx.get(), "hs_cost_run");
  std::map<std::string, std::vector<partition_pair>> out;
  size_t q = 0;
  for (size_t i = 0; i < n; ++i)
    for (size_t j = i; j < n; ++j, ++q) out[name_of[tech[q]]].push_back({&partitions[i], &partitions[j]});
  return out;
}
}  // namespace cost

}  // namespace hyset_amd
