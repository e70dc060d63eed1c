// Command-line self-join driver over the handler layer, with the flags of
// the reference's GPU drivers (src/bitmap.cu, src/fgss.cu, src/hybrid.cu):
//
//   hyset_join --input FILE [--threshold 0.95] [--block 10000]
//              [--technique fgss|bitmap|hybrid|cpu|queue|dichotomy|cost]
//              [--algorithm allpairs|ppjoin]
//              [--bitmap 64] [--scenario 1] [--threads 32] [--device-memory 512M]
//              [--output FILE] [--overlap]
//              [--prepare words|qgrams|integers [--q 3] [--lowercase]]
//              [--binary] [--save-binary FILE]
//
// --algorithm allpairs|ppjoin is the CPU filtering algorithm, as in the
// reference's src/hybrid.cu:36 and src/cpu_single.cu:35: the hybrid
// technique's host filter (the default technique when only --algorithm
// allpairs|ppjoin is given, as the reference's hybrid driver) or the CPU
// technique (cpu_single / cpu_multi: host::allpairs / host::ppjoin over
// --threads threads).  --algorithm also accepts a technique name
// (fgss|bitmap|hybrid|cpu), the same as --technique.
//
// --prepare reads raw records (one per line) and prepares them with
// hs_preprocess_file (the step the reference leaves to external scripts,
// README.md:14); pair ids are then set ids of the prepared collection, and
// --output also writes FILE.records with the input line of every set id.
// --binary reads the binary CSR cache instead of text; --save-binary writes
// the (prepared) collection to it.
//
// CPU and device together (the reference's queue, dichotomy and cost_model
// drivers), --cpu-threads host threads joining beside the device:
//   --technique queue [--cooperative] [--bitmap B]: src/queue.cu, one work
//     queue drained from both ends (hs_join_coprocess, HS_COPROC_QUEUE); the
//     device side runs bitmap when --bitmap is given, the hybrid technique
//     (host filter, device verification) with --cooperative, else fgss;
//   --technique dichotomy [--policy cpu,gpu] [--split 50]: src/dichotomy.cu,
//     the first ceil(split% of the blocks) (Rs) joined with itself by the
//     first side of the policy pair, R against the rest by the second; sides
//     cpu, gpu (fgss, or bitmap with --bitmap) and cpu-gpu (hybrid);
//   --technique cost --coefficients FILE [--partitions 4] [--bitmap B]:
//     src/cost_model.cu, partitioner::even into --partitions partitions of
//     blocks, cost::run over the partition pairs with the coefficient file's
//     table for the threshold (bitmap left out unless --bitmap is given),
//     each pair run with its technique (hs_run_partition_pairs).
//
// The whole collection is one partition of blocks joined with itself
// (src/fgss.cu:108-116).  With --output the pairs are written as "r s" lines
// (output.hpp:39-76; "r s overlap" with --overlap); otherwise the count is
// printed.  The hybrid technique filters on the host (candidate generation
// of hybrid::allpairs or hybrid::ppjoin, hybrid_algorithms.hpp:209-387) and
// verifies the batches on the device through the device_handler.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hyset_amd/handlers.hpp"

namespace ha = hyset_amd;

namespace {

struct Options {
  std::string input, output, technique, filter = "allpairs", deviceMemory = "512M";
  std::string prepare, saveBinary, policy = "cpu,gpu", coefficients;
  double threshold = 0.95, split = 50.0;
  unsigned block = 10000, bitmap = 64, scenario = 1, threads = 32, q = 3, partitions = 4, cpuThreads = 0;
  bool overlap = false, binary = false, lowercase = false, cooperative = false, bitmapGiven = false;
};

[[noreturn]] void usage(const char* msg) {
  std::fprintf(stderr, "error: %s\nusage: hyset_join --input FILE [--threshold T] [--block B] "
                       "[--technique fgss|bitmap|hybrid|cpu|queue|dichotomy|cost] [--algorithm allpairs|ppjoin] "
                       "[--cooperative] [--policy cpu,gpu|gpu,cpu|cpu,cpu-gpu|cpu-gpu,cpu] [--split PCT] "
                       "[--coefficients FILE] [--partitions N] [--cpu-threads N] "
                       "[--bitmap 64|128|256] [--scenario 1|2] [--threads N] "
                       "[--device-memory 512M] [--output FILE] [--overlap] "
                       "[--prepare words|qgrams|integers [--q N] [--lowercase]] [--binary] [--save-binary FILE]\n",
               msg);
  std::exit(1);
}

Options parse(int argc, char** argv) {
  Options o;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto val = [&]() -> std::string {
      if (i + 1 >= argc) usage(("missing value for " + a).c_str());
      return argv[++i];
    };
    if (a == "--input") o.input = val();
    else if (a == "--output") o.output = val();
    else if (a == "--technique") o.technique = val();
    else if (a == "--algorithm") {
      const std::string v = val();
      if (v == "allpairs" || v == "ppjoin") {
        o.filter = v;
        if (o.technique.empty()) o.technique = "hybrid-default";
      } else {
        o.technique = v;
      }
    }
    else if (a == "--threshold") o.threshold = std::strtod(val().c_str(), nullptr);
    else if (a == "--block") o.block = (unsigned)std::strtoul(val().c_str(), nullptr, 10);
    else if (a == "--bitmap") {
      o.bitmap = (unsigned)std::strtoul(val().c_str(), nullptr, 10);
      o.bitmapGiven = true;
    }
    else if (a == "--cooperative") o.cooperative = true;
    else if (a == "--policy") o.policy = val();
    else if (a == "--split") o.split = std::strtod(val().c_str(), nullptr);
    else if (a == "--coefficients") o.coefficients = val();
    else if (a == "--partitions") o.partitions = (unsigned)std::strtoul(val().c_str(), nullptr, 10);
    else if (a == "--cpu-threads") o.cpuThreads = (unsigned)std::strtoul(val().c_str(), nullptr, 10);
    else if (a == "--scenario") o.scenario = (unsigned)std::strtoul(val().c_str(), nullptr, 10);
    else if (a == "--threads") o.threads = (unsigned)std::strtoul(val().c_str(), nullptr, 10);
    else if (a == "--device-memory") o.deviceMemory = val();
    else if (a == "--overlap") o.overlap = true;
    else if (a == "--prepare") o.prepare = val();
    else if (a == "--q") o.q = (unsigned)std::strtoul(val().c_str(), nullptr, 10);
    else if (a == "--lowercase") o.lowercase = true;
    else if (a == "--binary") o.binary = true;
    else if (a == "--save-binary") o.saveBinary = val();
    else usage(("unknown option " + a).c_str());
  }
  if (o.input.empty()) usage("no input dataset given");
  if (!o.prepare.empty() && o.prepare != "words" && o.prepare != "qgrams" && o.prepare != "integers")
    usage("--prepare takes words, qgrams or integers");
  if (!o.prepare.empty() && o.binary) usage("--prepare reads raw text, --binary a prepared cache");
  if (o.block == 0) usage("--block must be positive");
  if (o.split < 0 || o.split > 100) usage("--split must lie in [0, 100]");  // dichotomy.cu:71-74
  if (o.partitions == 0) usage("--partitions must be positive");
  if (o.technique == "cost" && o.coefficients.empty()) usage("--technique cost needs --coefficients");
  // a CPU algorithm alone selects the hybrid technique, as the reference's hybrid driver
  if (o.technique.empty()) o.technique = "fgss";
  if (o.technique == "hybrid-default") o.technique = "hybrid";
  return o;
}

void check_rc(int rc, ha::context& ctx, const char* what) { ha::check(rc, ctx.get(), what); }

// "512M" -> number of 4-byte candidates (include/hyset/helpers.hpp:17-34).
unsigned long long candidates_for(const std::string& mem) {
  if (mem.empty()) return 1ull << 27;
  char scale = mem.back();
  double v = std::strtod(mem.substr(0, mem.size() - 1).c_str(), nullptr);
  double bytes = v * (scale == 'M' ? 1e6 : 1e9);
  return (unsigned long long)(bytes / 4.0);
}

// Host candidate generation for the hybrid driver: AllPairs or PPJoin
// filtering without verification (hybrid_algorithms.hpp:209-387) over a
// host index of midprefix entries sorted by (token, set id), each with its
// position in its set; PPJoin takes a candidate at a common prefix token
// only if the positional filter passes there (:333-348).
void hybrid_filter(const hs_host_collection& c, double t, bool ppjoin, ha::algorithms::hybrid::device_handler& dh) {
  std::vector<unsigned> offs(c.universe + 1, 0), ids, pos;
  std::vector<uint32_t> b(5);
  auto bnd = [&](uint32_t l1, uint32_t l2) {
    ha::check(hs_bounds(HS_JACCARD, t, HS_BOUNDS_REFERENCE, l1, l2, b.data()), nullptr, "hs_bounds");
    return b;
  };
  const uint32_t maxlen = c.n ? c.sizes[c.n - 1] : 0;
  std::vector<uint32_t> minsize(maxlen + 1), maxsize(maxlen + 1), maxprefix(maxlen + 1), midprefix(maxlen + 1);
  for (uint32_t l = 1; l <= maxlen; ++l) {
    auto v = bnd(l, l);
    minsize[l] = v[0];
    maxsize[l] = v[1];
    maxprefix[l] = v[3];
    midprefix[l] = v[4];
  }
  for (uint32_t i = 0; i < c.n; ++i)
    for (uint32_t k = 0; k < midprefix[c.sizes[i]]; ++k) offs[c.tokens[c.starts[i] + k] + 1]++;
  for (uint32_t u = 0; u < c.universe; ++u) offs[u + 1] += offs[u];
  ids.resize(offs[c.universe]);
  pos.resize(offs[c.universe]);
  std::vector<unsigned> cur(offs.begin(), offs.end() - 1);
  for (uint32_t i = 0; i < c.n; ++i)
    for (uint32_t k = 0; k < midprefix[c.sizes[i]]; ++k) {
      const unsigned at = cur[c.tokens[c.starts[i] + k]]++;
      ids[at] = i;
      pos[at] = k;
    }
  std::vector<unsigned> seen(c.n, 0xFFFFFFFFu), cand;
  for (uint32_t p = 0; p < c.n; ++p) {
    const uint32_t l = c.sizes[p];
    cand.clear();
    for (uint32_t j = 0; j < maxprefix[l]; ++j) {
      const uint32_t tok = c.tokens[c.starts[p] + j];
      for (unsigned it = offs[tok]; it < offs[tok + 1]; ++it) {
        const unsigned s = ids[it];
        if (c.sizes[s] < minsize[l]) continue;
        if (s >= p || c.sizes[s] > maxsize[l]) break;
        if (seen[s] == p) continue;
        if (ppjoin && bnd(l, c.sizes[s])[2] > std::min(l - j, c.sizes[s] - pos[it])) continue;
        seen[s] = p;
        cand.push_back(s);
      }
    }
    if (!cand.empty()) dh.addCandidates(p, cand);
  }
  dh.flush();
}

// The policy pair of src/dichotomy.cu:76-87: two different sides out of
// cpu, gpu and cpu-gpu, never cpu-gpu with gpu.
void parse_policy(const std::string& p, std::string& small, std::string& large) {
  const size_t comma = p.find(',');
  if (comma == std::string::npos) usage("--policy takes two sides, e.g. cpu,gpu");
  small = p.substr(0, comma);
  large = p.substr(comma + 1);
  auto side = [](const std::string& x) { return x == "cpu" || x == "gpu" || x == "cpu-gpu"; };
  if (!side(small) || !side(large) || small == large || (small == "cpu-gpu" && large == "gpu") ||
      (small == "gpu" && large == "cpu-gpu"))
    usage("wrong policy values");
}

// CPU and device together on one context: the queue and dichotomy drivers
// (hs_join_coprocess) and the cost-model driver (hs_run_partition_pairs).
// Returns the pair count; the pairs stay in the context's last result.
uint64_t run_shared(const Options& o, ha::context& ctx, ha::collection::device_collection& dc,
                    std::vector<ha::structs::block>& blocks) {
  check_rc(hs_set_threshold(ctx.get(), HS_JACCARD, o.threshold, HS_BOUNDS_REFERENCE), ctx, "hs_set_threshold");
  check_rc(hs_set_host_algorithm(ctx.get(), o.filter == "ppjoin" ? HS_HOST_PPJOIN : HS_HOST_ALLPAIRS), ctx,
           "hs_set_host_algorithm");
  hs_result res{};
  res.count_only = 1;
  const uint32_t n = (uint32_t)dc.sizes().size();
  if (o.technique == "queue" || o.technique == "dichotomy") {
    hs_coprocess cp{};
    cp.cpu_threads = o.cpuThreads;
    cp.tune.bitmap_bits = o.bitmap;
    cp.tune.verify_scenario = o.scenario;
    if (o.technique == "queue") {
      cp.mode = HS_COPROC_QUEUE;
      cp.algo = o.bitmapGiven && o.bitmap > 0 ? HS_ALGO_BITMAP : o.cooperative ? HS_ALGO_HYBRID : HS_ALGO_FGSS;
      cp.min_device_ms = -1.0;  // the reference's queue always shares
    } else {
      std::string small, large;
      parse_policy(o.policy, small, large);
      const std::string& dev = small == "cpu" ? large : small;
      cp.mode = HS_COPROC_STATIC;
      cp.policy = small == "cpu" ? HS_POLICY_CPU_GPU : HS_POLICY_GPU_CPU;
      cp.algo = dev == "cpu-gpu" ? HS_ALGO_HYBRID : o.bitmapGiven && o.bitmap > 0 ? HS_ALGO_BITMAP : HS_ALGO_FGSS;
      // Rs = the first ceil(split% of the blocks) blocks (partitioner.hpp:35)
      auto parts = ha::partitioner::dichotomize(blocks, o.split);
      const uint32_t rs = parts[1].blocks.empty() ? 0 : parts[1].end_id() + 1;
      cp.cpu_share = n ? (double)rs / (double)n : 0.0;
    }
    hs_coprocess_stats st{};
    check_rc(hs_join_coprocess(ctx.get(), dc.get(), dc.get(), &cp, &res, &st), ctx, "hs_join_coprocess");
    std::printf("Probe sets: device %u, host %u (%u host threads); pairs: device %llu, host %llu\n",
                st.device_probes, st.host_probes, st.host_threads, (unsigned long long)st.device_pairs,
                (unsigned long long)st.host_pairs);
    return res.count;
  }
  // cost model
  auto coef = ha::cost::read_coefficients(o.coefficients, o.threshold);
  if (!o.bitmapGiven || o.bitmap == 0) coef.erase("bitmap");  // cost_model.cu:133-135
  auto parts = ha::partitioner::even(blocks, o.partitions);
  auto runs = ha::cost::run(coef, parts, ctx, dc, o.bitmapGiven ? o.bitmap : 0);
  std::vector<hs_partition_pair> plan;
  for (const auto& [name, pairs] : runs) {
    int32_t tech = HS_TECH_FGSS;
    for (const auto& [nm, t] : ha::cost::techniques())
      if (nm == name) tech = t;
    std::printf("%s: %zu partition pairs\n", name.c_str(), pairs.size());
    for (const auto& pp : pairs) {
      hs_partition_pair x{};
      x.left = hs_range{pp.first->start_id(), pp.first->end_id()};
      x.right = hs_range{pp.second->start_id(), pp.second->end_id()};
      x.technique = tech;
      plan.push_back(x);
    }
  }
  hs_tune tune{};
  tune.bitmap_bits = o.bitmapGiven ? o.bitmap : 64;
  tune.verify_scenario = o.scenario;
  res.count_only = o.output.empty() ? 1 : 0;  // keep the pairs for the writer
  const int rc = hs_run_partition_pairs(ctx.get(), dc.get(), plan.data(), (uint32_t)plan.size(),
                                        o.cpuThreads ? o.cpuThreads : 1, &tune, &res, nullptr);
  if (rc != HS_E_OVERFLOW) check_rc(rc, ctx, "hs_run_partition_pairs");
  return res.count;
}

}  // namespace

int main(int argc, char** argv) {
  Options o = parse(argc, argv);
  hs_host_collection hc{};
  uint32_t* record_of = nullptr;
  int rc;
  if (!o.prepare.empty()) {
    hs_prep_params pp{};
    pp.mode = o.prepare == "words" ? HS_TOKENS_WORDS : o.prepare == "qgrams" ? HS_TOKENS_QGRAMS : HS_TOKENS_INTEGERS;
    pp.q = o.q;
    pp.lowercase = o.lowercase;
    rc = hs_preprocess_file(o.input.c_str(), &pp, &hc, &record_of);
  } else {
    rc = o.binary ? hs_read_collection_binary(o.input.c_str(), &hc) : hs_read_collection_text(o.input.c_str(), &hc);
  }
  if (rc != HS_OK) {
    std::fprintf(stderr, "Error: cannot read %s\n", o.input.c_str());
    return 1;
  }
  if (!o.saveBinary.empty() &&
      hs_write_collection_binary(o.saveBinary.c_str(), hc.tokens, hc.starts, hc.sizes, hc.n) != HS_OK) {
    std::fprintf(stderr, "Error: cannot write %s\n", o.saveBinary.c_str());
    return 1;
  }
  if (record_of && !o.output.empty()) {
    FILE* f = std::fopen((o.output + ".records").c_str(), "w");
    if (!f) {
      std::fprintf(stderr, "Error: cannot write %s.records\n", o.output.c_str());
      return 1;
    }
    for (uint32_t i = 0; i < hc.n; ++i) std::fprintf(f, "%u %u\n", i, record_of[i]);
    std::fclose(f);
  }
  hs_free_array(record_of);
  try {
    ha::context ctx(0);
    std::vector<unsigned> toks(hc.tokens, hc.tokens + hc.n_tokens), starts(hc.starts, hc.starts + hc.n),
        sizes(hc.sizes, hc.sizes + hc.n);
    ha::collection::device_collection dc(ctx, toks, starts, sizes);
    std::vector<ha::structs::block> blocks = ha::structs::make_blocks(hc.n, o.block);
    ha::structs::partition part(0);
    for (auto& b : blocks) part.blocks.push_back(&b);
    const bool aggregate = o.output.empty();
    if (o.technique == "queue" || o.technique == "dichotomy" || o.technique == "cost") {
      auto t0 = std::chrono::steady_clock::now();
      const uint64_t count = hc.n ? run_shared(o, ctx, dc, blocks) : 0;
      auto t1 = std::chrono::steady_clock::now();
      std::printf("Total time without I/O (ms): %.3f\n", std::chrono::duration<double, std::milli>(t1 - t0).count());
      if (aggregate) {
        std::printf("Result: %llu\n", (unsigned long long)count);
      } else {
        check_rc(hs_write_last_result_text(ctx.get(), o.output.c_str(), o.overlap ? 1 : 0), ctx,
                 "hs_write_last_result_text");
        std::printf("Result: %llu pairs written to %s\n", (unsigned long long)count, o.output.c_str());
      }
      hs_host_collection_free(&hc);
      return 0;
    }
    std::shared_ptr<ha::output::handler> out;
    if (aggregate) out = std::make_shared<ha::output::count_handler>();
    else out = std::make_shared<ha::output::pairs_handler>();

    auto t0 = std::chrono::steady_clock::now();
    if (hc.n > 0) {
      if (o.technique == "fgss") {
        ha::algorithms::device::fgssjoin::handler h(ctx, dc, o.block, aggregate, o.threshold);
        h.setOutputHandler(out);
        h.join(&part, &part);
      } else if (o.technique == "bitmap") {
        ha::algorithms::device::bitmap::handler h(ctx, dc, o.bitmap, o.block, aggregate, o.threshold);
        h.setOutputHandler(out);
        h.join(&part, &part);
      } else if (o.technique == "hybrid") {
        ha::algorithms::hybrid::device_handler dh(ctx, dc, dc, o.threads, o.scenario, aggregate, hc.n,
                                                  (unsigned)std::min<unsigned long long>(candidates_for(o.deviceMemory),
                                                                                         0xFFFFFFFFull),
                                                  o.threshold);
        dh.setMaxSetSize(hc.n ? hc.sizes[hc.n - 1] : 0);
        dh.setOutputHandler(out);
        hybrid_filter(hc, o.threshold, o.filter == "ppjoin", dh);
      } else if (o.technique == "cpu") {
        ha::algorithms::host::handler h(ctx, dc, o.threads, o.filter, aggregate, o.threshold);
        h.setOutputHandler(out);
        h.join(&part, &part);
      } else {
        usage("unknown technique");
      }
    }
    auto t1 = std::chrono::steady_clock::now();
    std::printf("Total time without I/O (ms): %.3f\n", std::chrono::duration<double, std::milli>(t1 - t0).count());
    if (aggregate) {
      std::printf("Result: %lu\n", out->getCount());
    } else {
      auto& pairs = static_cast<ha::output::pairs_handler&>(*out).pairs;
      std::vector<uint32_t> r, s, ov;
      for (auto& p : pairs) {
        r.push_back(p.first);
        s.push_back(p.second);
        ov.push_back(p.overlap);
      }
      if (hs_write_pairs_text(o.output.c_str(), r.data(), s.data(), o.overlap ? ov.data() : nullptr, r.size()) !=
          HS_OK) {
        std::fprintf(stderr, "Error writing to file\n");
        hs_host_collection_free(&hc);
        return 1;
      }
      std::printf("Result: %zu pairs written to %s\n", pairs.size(), o.output.c_str());
    }
  } catch (const ha::error& e) {
    std::fprintf(stderr, "hyset_join: %s\n", e.what());
    hs_host_collection_free(&hc);
    return 2;
  }
  hs_host_collection_free(&hc);
  return 0;
}
