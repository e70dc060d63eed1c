// Internal declarations shared by the host code and the HIP kernels of
// libhyset_amd.so.  Nothing here crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "hyset_amd.h"

namespace hs {

// ------------------------------------------------------------------ errors

struct Error {
  int code;
  std::string msg;
};

#define HS_CHECK_HIP(expr)                                                                  \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess)                                                                   \
      throw ::hs::Error{HS_E_DEVICE, std::string(#expr " failed: ") + hipGetErrorString(_e)}; \
  } while (0)

#define HS_REQUIRE(cond, code, text)                 \
  do {                                               \
    if (!(cond)) throw ::hs::Error{(code), (text)};  \
  } while (0)

// ----------------------------------------------------------- device memory

// Allocation epoch of the process: every device or pinned allocation and
// free bumps it.  A captured graph holds raw buffer addresses, so it is
// replayed only while the epoch still reads what it read at capture.
uint64_t buffer_epoch();
void bump_buffer_epoch();

// A growable device allocation.  Memory is only ever grown, never shrunk,
// so steady-state joins allocate nothing.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void ensure(size_t need);
  void release();
  template <typename T>
  T* as() const { return static_cast<T*>(ptr); }
  ~DevBuf() { release(); }
};

// ------------------------------------------------------------ bound tables

// Integer bound tables for one (similarity, threshold, mode).  All entries
// are computed on the host with the same expressions as the CPU join, then
// used as plain integers on the device, so device and host can never round
// differently.
struct BoundTables {
  int kind = HS_JACCARD;
  int mode = HS_BOUNDS_REFERENCE;
  double t = 0.0;
  uint32_t L = 0;  // largest set size covered
  std::vector<uint32_t> minsize, maxsize, maxprefix, midprefix;
  // minoverlap before the min(l1,l2) clamp: keyed by l1+l2 for Jaccard and
  // Dice, by l1*(L+1)+l2 for cosine (clamped already).
  std::vector<uint32_t> mo;
  void build(int kind, int mode, double t, uint32_t L);
  uint32_t minoverlap(uint32_t l1, uint32_t l2) const;
};

// Device view of BoundTables (one allocation, sub-arrays).
struct DevBounds {
  const uint32_t* minsize;
  const uint32_t* maxsize;
  const uint32_t* maxprefix;
  const uint32_t* midprefix;
  const uint32_t* mo;
  uint32_t L;
  int kind;
};

__host__ __device__ inline uint32_t dev_minoverlap(const DevBounds& b, uint32_t l1, uint32_t l2) {
  uint32_t m = l1 < l2 ? l1 : l2;
  uint32_t v = (b.kind == HS_COSINE) ? b.mo[(size_t)l1 * (b.L + 1) + l2] : b.mo[l1 + l2];
  return v < m ? v : m;
}

// Host-side bound evaluation (also used to fill the tables).
uint32_t bound_minsize(int kind, int mode, double t, uint64_t p, uint64_t q, uint32_t l);
uint32_t bound_maxsize(int kind, int mode, double t, uint64_t p, uint64_t q, uint32_t l);
uint32_t bound_minoverlap_raw(int kind, int mode, double t, uint64_t p, uint64_t q, uint32_t l1, uint32_t l2);
bool threshold_rational(double t, uint64_t* p, uint64_t* q);

// --------------------------------------------------------------- kernels

struct DevColl {
  const uint32_t* tokens;
  const uint32_t* starts;
  const uint32_t* sizes;
  uint32_t n;
};

// Index entry e of list(t) = [offsets[t], offsets[t+1]): set_ids[e] (lists
// ascending) and the record ent[e * ent_words ..] = {set id, meta, and for
// an index built with 64-bit signatures the set's signature (low, high
// word)}, meta = (min(|s|, 0xFFFF) << 16) | min(position of t in s,
// 0xFFFF).  Size, position and signature ride along so a posting visit is
// filtered from its own entry, without reading sizes[] or the signature
// table at random; interleaved, a visit reads one record instead of three
// arrays.  The window search reads the compact set_ids (a list is searched
// over more entries than its window visits) in aligned 16-byte quads, so
// set_ids is allocated IDS_PAD ids longer than the index.
constexpr uint32_t IDS_PAD = 4;
struct DevIndex {
  const uint32_t* offsets;  // universe + 1
  const uint32_t* set_ids;  // the ids alone (host copies, the item-slice pass)
  const uint32_t* ent;      // ent_words words per entry
  uint32_t ent_words;       // 4 with 64-bit signatures, else 2
  uint32_t universe;
};

// Output cursor and work counters (64-bit): pairs found, posting visits
// (bitmap: pair tests), distinct candidates verified (bitmap: filter
// survivors), tokens of the verified candidate sets, probe-prefix items,
// index entries, and (fgss) visits that reached the signature test.
// CNT_QUEUE: candidates the token-major filter appended to its queue;
// CNT_HEAVY: chunks of visits it left to the second filter pass;
// CNT_CAND_PROBES: sum of |r| over the distinct probes with candidates (the
// token-major join, when asked to count it); CNT_INDEX_ERR: nonzero when a
// sort pass of the packed token-major index gave up waiting for an earlier
// tile (never in a correct run; the join then fails); CNT_HCLAIM: runs of
// second-pass items claimed so far.
// The first CNT_STATS words hold the counters; CNT_SLOTS slots of
// CNT_SLOT_WORDS words (a 128-byte line each) follow, into which a
// workgroup may add its pure statistics (visits, items, signature tests,
// candidates, their tokens) at the same word index, slot blockIdx.x %
// CNT_SLOTS: a wide grid's end-of-workgroup atomics then spread over 16
// lines instead of meeting on one (same-line device atomics serialize,
// ~8 ns each).  fold_counters adds the slots into the first words.
enum {
  CNT_PAIRS = 0, CNT_VISITS = 1, CNT_CANDS = 2, CNT_CAND_TOKENS = 3, CNT_ITEMS = 4, CNT_ENTRIES = 5,
  CNT_SIG_TESTS = 6, CNT_QUEUE = 7, CNT_HEAVY = 8, CNT_CAND_PROBES = 9, CNT_INDEX_ERR = 10, CNT_HCLAIM = 11,
  CNT_STATS = 16,
  CNT_SLOTS = 16, CNT_SLOT_WORDS = 16, CNT_WORDS = CNT_STATS + CNT_SLOTS * CNT_SLOT_WORDS
};
#ifdef __HIPCC__
// This workgroup's statistics slot of a counters array.
__device__ __forceinline__ unsigned long long* stat_slot(unsigned long long* counters) {
  return counters + CNT_STATS + (blockIdx.x % CNT_SLOTS) * CNT_SLOT_WORDS;
}
#endif
// Adds the statistics slots into the first CNT_STATS words (host copy).
inline void fold_counters(unsigned long long* c) {
  for (int k = 0; k < CNT_SLOTS; ++k)
    for (int w = 0; w < CNT_SLOT_WORDS; ++w) c[w] += c[CNT_STATS + k * CNT_SLOT_WORDS + w];
}

struct DevOut {
  uint32_t* r;
  uint32_t* s;
  uint32_t* ov;
  unsigned long long* counters;
  uint64_t capacity;
};

// Launch helpers (defined in the .hip files).
void launch_fill_u32(hipStream_t st, uint32_t* p, size_t n, uint32_t v);
void launch_fill_u64(hipStream_t st, unsigned long long* p, size_t n, unsigned long long v);
void launch_copy(hipStream_t st, const void* src, void* dst, size_t bytes);
// rows r, s, overlap [0, min(counters[CNT_PAIRS], n)) of the result rows
// (row stride cap) into dst (row stride dst_cap), a device address of
// mapped host memory
void launch_rows_to_host(hipStream_t st, const uint32_t* rows, uint64_t cap, const unsigned long long* counters,
                         uint64_t n, uint32_t* dst, uint64_t dst_cap);
// XOR-popcount chains: grid x 256 lanes x iters x 4 x 8 steps of 2 VALU ops.
void launch_popc_peak(hipStream_t st, uint32_t grid, uint32_t iters, uint32_t* out);
void launch_exclusive_scan(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t* out2, size_t n,
                           DevBuf& scratch);

struct IndexBuild {
  DevBuf offsets, set_ids, metas, ent;  // the index
  DevBuf keys[2], vals, pos, hist, scan_scratch;  // radix-sort work space
  DevBuf packed[2], pk_hist, pk_sync;  // the packed token-major index and its sort's work space
  const uint64_t* packed_sorted = nullptr;  // the packed index after its last sort pass
  uint32_t ent_words = 2;
  uint32_t universe = 0;
  uint64_t entries = 0;
};

// Packed token-major index entry: one 64-bit word holding, from the low
// bits up, the entry's position in its set (pb bits), the set's size (sb
// bits), the set id (ib bits) and the key (token << 1 | position >=
// midprefix, key_bits bits from kshift = pb + sb + ib).  Ordering words by
// their bits above pb + sb orders entries by (key, set id).  The sort runs
// `passes` stable passes of digit_bits key bits each, low digit first.
struct TokPack {
  uint32_t pb, sb, ib, kshift, key_bits, passes, digit_bits;
};
// The packing of a self-join index over sets of at most maxlen tokens, ids
// below n_sets and tokens below universe: false when the fields need more
// than 64 bits or the index more than 2^30 entries.
bool tok_pack_plan(uint32_t n_sets, uint32_t maxlen, uint32_t universe, uint64_t entries, TokPack* out);
// The same buffers as raw device pointers: the index (offsets, set_ids,
// metas, ent) and the radix sort's work space (keys, vals, pos, the
// per-tile digit histograms and the digit totals).  A meta is the entry's
// (min(|s|, 0xFFFF) << 16) | min(position, 0xFFFF), written by the emit
// pass and carried through the sort.
struct IndexBufs {
  uint32_t* offsets;
  uint32_t* set_ids;
  uint32_t* metas;
  uint32_t* ent;
  uint32_t* keys[2];
  uint32_t* vals;
  uint32_t* pos;
  uint32_t* hist;
  uint32_t* totals;
};
// Bytes an index of `entries` entries over `universe` tokens occupies laid
// out in one region (each buffer 256-byte aligned); with base != NULL the
// pointers into the region starting at base are stored in *out.
size_t index_region(uint32_t universe, uint64_t entries, uint8_t* base, IndexBufs* out);
// Builds the inverted index over the prefix entries of sets [first, last]
// of `c` (sets sorted by size).  prefix_len is indexed by set size;
// size_first[l] is the first set id of size >= l; cum[l] is the number of
// prefix entries of the range's sets shorter than l; `entries` is the total.
// sigs (may be NULL): one 64-bit signature per set of c, copied into the
// entries (4-word records).  sigs_out (may be NULL): the emit pass computes the
// 64-bit signatures of the range's sets into it on the way (then pass the
// same array as sigs).  counters (may be NULL): CNT_WORDS work counters the
// last kernel zeroes, with the index size in CNT_ENTRIES (only when the
// index has entries).  done (may be NULL): an event given the end of the
// build's last kernel.
void build_index(hipStream_t st, const DevColl& c, uint32_t first, uint32_t last, const uint32_t* prefix_len,
                 const uint32_t* size_first, const uint32_t* cum, uint32_t universe, uint64_t entries,
                 IndexBuild& ib, const uint64_t* sigs, uint64_t* sigs_out = nullptr,
                 unsigned long long* counters = nullptr, hipEvent_t done = nullptr);
// The packed token-major index (TokPack) over the probe prefixes
// (prefix_len = maxprefix; flag_len = midprefix) of sets [first, last]: a
// pass over the sets that writes the words in set-id order and computes
// the key-digit histograms (and, with sigs_out != NULL, the sets' 64-bit
// signatures), then one stable scatter per key digit.  cum[l] is the first
// entry of size class l.  The sorted words are left in ib.packed_sorted.
// num_cus sizes the emit's grid.  counters are zeroed; done (may be NULL)
// takes the end.
void build_token_index(hipStream_t st, const DevColl& c, uint32_t first, uint32_t last, const uint32_t* prefix_len,
                       const uint32_t* size_first, const uint32_t* cum, uint64_t entries, const TokPack& pk,
                       IndexBuild& ib, const uint32_t* flag_len, uint64_t* sigs_out, unsigned long long* counters,
                       uint32_t num_cus, hipEvent_t done = nullptr);
// The same into buffers the caller laid out (index_region); allocates nothing.
void build_index_into(hipStream_t st, const DevColl& c, uint32_t first, uint32_t last, const uint32_t* prefix_len,
                      const uint32_t* size_first, const uint32_t* cum, uint32_t universe, uint64_t entries,
                      const IndexBufs& b, const uint64_t* sigs, uint64_t* sigs_out,
                      unsigned long long* counters, hipEvent_t done = nullptr);
// The same over every join of a batched collection at once (token ids of
// different joins are disjoint, so one index serves all joins).
void build_index_batch(hipStream_t st, const DevColl& c, const uint32_t* join_first, uint32_t n_joins,
                       const uint32_t* prefix_len, const uint32_t* size_first, const uint32_t* sf_base,
                       const uint32_t* cum, uint32_t universe, uint64_t entries, IndexBuild& ib,
                       const uint64_t* sigs, uint64_t* sigs_out = nullptr);
// The device view of a built index.
inline DevIndex dev_index(const IndexBuild& ib) {
  return DevIndex{ib.offsets.as<uint32_t>(), ib.set_ids.as<uint32_t>(), ib.ent.as<uint32_t>(), ib.ent_words,
                  ib.universe};
}

// Batched launch: tile k covers probes [tile_lo[k], tile_hi[k]) of join
// tile_join[k]; join j owns set ids [join_first[j], join_first[j+1]) and its
// size table size_first[sf_base[j] .. sf_base[j+1]); queue[8] are the
// per-XCD tile counters of the device work queue.
struct FgssBatch {
  const uint32_t* tile_lo;
  const uint32_t* tile_hi;
  const uint32_t* tile_join;
  const uint32_t* join_first;
  const uint32_t* sf_base;
  uint32_t* queue;
};

struct FgssArgs {
  DevColl probe;
  DevColl indexed;
  DevIndex index;
  DevBounds bounds;
  const uint32_t* size_first;  // indexed collection: first id with size >= l, l in [0, maxlen+1]
  uint32_t size_first_len;     // maxlen + 2
  uint32_t probe_first, probe_last;  // inclusive
  uint32_t idx_first, idx_last;      // inclusive
  int self;                          // 1: restrict to s < r
  // tile plan of the probe collection: tile k covers probes
  // [tile_starts[k], tile_starts[k+1]) (<= 512 probes, <= 8192 tokens);
  // NULL selects fixed tiles of probes_per_tile probes
  const uint32_t* tile_starts;
  uint32_t first_tile, n_plan_tiles;
  FgssBatch batch;  // tile_join != NULL: persistent batched launch
  // serving: per-probe thresholds.  probe_tidx[p] selects probe_bounds[...]
  // (at most MAX_THRESHOLDS tables); NULL uses `bounds` for every probe
  const DevBounds* probe_bounds;
  const uint8_t* probe_tidx;
  // non-NULL: a persistent grid claims plan tiles from 8 per-XCD counters
  // (zeroed by launch_fgss unless queue_zeroed: the balanced planner zeroes
  // them on its way), heaviest tiles first; NULL: one workgroup per tile
  uint32_t* tile_queue;
  uint32_t queue_zeroed;
  const uint32_t* dev_ntiles;  // non-NULL: the tile count lives on the device
  // item-slice pass (plain joins): per item slot the posting-list slice of
  // ids in the probe's window
  uint32_t narrow_tiles;  // balanced plan: tiles of 256 lanes and probes (very long posting lists)
  const uint32_t* slice_a;
  const uint32_t* slice_n;
  // per probe of the range (planning pass): {start, size | maxprefix << 16,
  // window lo, window hi} and its first item slot
  const uint4* probe_rec;
  const uint32_t* probe_slot;
  // parity signatures, sig_words 64-bit words per set (probe and indexed
  // collection, indexed by set id); sig_words == 0 skips the signature test
  const uint64_t* probe_sigs;
  const uint64_t* indexed_sigs;
  uint32_t sig_words;
  DevOut out;
};
constexpr uint32_t MAX_THRESHOLDS = 8;
constexpr uint32_t TILE_TOKENS = 8192;  // probe tokens staged in LDS per tile
constexpr uint32_t TILE_PROBES = 512;
// persistent_wgs: grid size when a.tile_queue is set (0: one per tile).
// t0 / t1 (may be NULL): events given the k_fgss dispatch's start and end.
void launch_fgss(hipStream_t st, const FgssArgs& a, uint32_t probes_per_tile, uint32_t persistent_wgs = 0,
                 hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
// The batched join's persistent launch: `workgroups` workgroups of 512
// lanes (narrow: 256, for tiles of at most 256 probes) drain the batch's
// tile queue.
void launch_fgss_batch(hipStream_t st, const FgssArgs& a, uint32_t ntiles, uint32_t workgroups, bool narrow);
// Work buffers and the host-made item table of the item-slice pass.
struct BalancedPlan {
  // device table of 5 * nl entries, per size class l of the probe range:
  // first item slot, first probe, window lo, window hi (before the self-join
  // cut at the probe id), maxprefix
  const uint32_t* size_tab;
  uint32_t nl;  // size classes + 1 (cum[nl - 1] = total_items)
  uint64_t total_items;
  DevBuf *slice_a, *slice_n, *cost, *tiles, *ntiles, *scan_scratch, *rec, *slot;
};
// Runs the item-slice pass and plans work-balanced tiles of a's probe range
// on the device (after the index build), pointing a at both; launch with
// a.tile_queue set.
void plan_balanced_tiles(hipStream_t st, FgssArgs& a, const BalancedPlan& bp);

// Token-major self-join of the id range [first, last] with itself.  The
// index is built over the probe prefixes (maxprefix) of the range's sets,
// keyed (token << 1) | (position >= midprefix(|s|)) and sorted by (key, set
// id): each token's list holds its index entries (the index prefixes)
// first, then its other probe entries, both ascending in id.  Entry e is
// the packed word ent[e] (TokPack: key, set id, size, position); the set's
// signature comes from the per-set table.  The visits of a probe entry are
// the index entries of its token
// with ids in its length window below the probe: one run of the token's
// index part (ids ascend, and so do sizes).  A workgroup takes a tile of
// TOK_BLOCK entries and stages in LDS the window of the index from the
// start of its first token's list to the tile's end (at most TOK_WINDOW
// entries), finds every entry's run by bisection there, and tests the
// tile's visits -- positional filter, signature test -- from LDS; the
// survivors {r, s, position of the token in r} go to `queue`, verified by a
// last kernel.  A tile's first TOK_CHUNK visits are tested by its own
// workgroup, the others in items of TOK_HCHUNK through `heavy` ({tile,
// item >= 1, visits of the tile}) by a persistent second pass, `aoff`
// ({run start, first visit} per entry) carrying the tile's layout there.
// Items smaller than the first chunk spread a few heavy tiles over the
// whole grid (on the 1M-set index at Jaccard 0.8, 240 items of 8192
// visits kept 272 of 512 workgroups idle, and a tile that passes nearly
// every visit held its workgroup four times as long as the others; the
// second pass took 25.6 us with 8192-visit items, 21.9 / 19.9 / 23.0 us
// with 4096 / 2048 / 1024).
// 1024-entry tiles: the token filter measured 78 us on the 1M-set index
// at Jaccard 0.8 (512-entry tiles: 86 us; either with 8192-visit chunks:
// 93 / 84 us at 512 / 1024 threads)
constexpr uint32_t TOK_BLOCK = 1024;
constexpr uint32_t TOK_WINDOW = 2048;
constexpr uint32_t TOK_CHUNK = 8192;
constexpr uint32_t TOK_HCHUNK = 2048;
struct TokArgs {
  DevColl c;
  const uint64_t* ent;  // the packed index, sorted by (key, set id)
  TokPack pk;           // its word layout
  uint32_t n_ent;
  DevBounds bounds;
  const uint32_t* size_first;  // first id of size >= l, l in [0, maxlen + 1]
  uint32_t size_first_len;     // maxlen + 2
  uint32_t first, last;
  uint32_t sig_words;          // 0: no signature test, else 1-16 words
  const uint64_t* sigs;        // per-set signatures, sig_words words each
  uint4* queue;
  uint32_t queue_cap;
  uint4* heavy;
  uint32_t heavy_cap;
  uint2* aoff;  // n_ent entries
  uint2* hdrs;  // one per tile of TOK_BLOCK entries: {first key, first entry of its index part}
  // optional: stamp[r] == epoch marks a probe already counted in
  // CNT_CAND_PROBES (NULL: not counted)
  uint32_t* stamp;
  uint32_t epoch;
  DevOut out;
};
// Filter, then verify the queue (a persistent grid of num_cus * 8
// workgroups); t0 / t1 (may be NULL) take the filter's start and the
// verify's end.
// t0 / t1 (may be NULL) take the filter's start and the verification's
// end; mid (may be NULL) four events: the filter's end, the second pass's
// start and end, the verification's start, each taken by its dispatch.
void launch_tok_join(hipStream_t st, const TokArgs& a, uint32_t num_cus, hipEvent_t t0 = nullptr,
                     hipEvent_t t1 = nullptr, const hipEvent_t* mid = nullptr);

struct BitmapArgs {
  DevColl probe;
  DevColl indexed;
  DevBounds bounds;
  const uint64_t* probe_sigs;
  const uint64_t* indexed_sigs;
  const uint32_t* size_first;
  uint32_t size_first_len;
  uint32_t words;  // 1, 2, 4, 8 or 16 (64 to 1024 bits)
  uint32_t probe_first, probe_last;
  uint32_t idx_first, idx_last;
  int self;
  DevOut out;
  uint32_t parts;  // waves per 64 probes, each taking a share of their window (0: 1)
};
// words: 1, 2, 4, 8 or 16 (64 ... 1024 bits)
void launch_signatures(hipStream_t st, const DevColl& c, uint32_t words, uint64_t* sigs);
// Signature width (64-bit words) for the fgss signature test of sets of
// mean size `mean`: the smallest power of two of at least 2 * mean bits,
// within [64, 1024] -- wide enough that unrelated sets differ in about half
// of the bits, which rules them out.
uint32_t auto_signature_words(double mean);
void launch_bitmap(hipStream_t st, const BitmapArgs& a);

struct VerifyArgs {
  DevColl probe;
  DevColl indexed;
  DevBounds bounds;
  const uint32_t* probes;      // probe id per batch row
  const uint32_t* offsets;     // inclusive end of each row's candidates
  const uint32_t* candidates;  // indexed ids
  uint32_t n_probes;
  uint64_t n_candidates;
  DevOut out;
};
void launch_verify(hipStream_t st, const VerifyArgs& a, int scenario);

// hs_join_noalloc's scratch array: takes its device address on `device`
// (the current device) once per process; false when the runtime refused.
bool noalloc_init(int device);

}  // namespace hs
