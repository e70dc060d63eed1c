/*
 * hyset_amd -- exact threshold set-similarity join on MI355X (gfx950).
 *
 * C ABI of libhyset_amd.so.  Plain pointers and sizes only; every call
 * returns HS_OK (0) or a negative HS_E_* code, and no C++ exception crosses
 * the boundary.  One hs_ctx per host thread; each context owns one HIP
 * stream.  Device-resident collections may be shared read-only between
 * contexts on the same device.
 *
 * The entry points replace the GPU partition of the reference's hybrid join
 * (paths relative to the reference repository):
 *   hs_create / hs_destroy             handler constructors / destructors,
 *                                      include/hyset/device_algorithms.hpp:138-171,
 *                                      :322-348, :249-259, :434-444
 *   hs_upload_collection               hyset::collection::device_collection,
 *                                      include/hyset/collection.hpp:54-71
 *   hs_set_threshold                   the `threshold` argument of every handler
 *                                      and the bounds of similarity.hpp:11-55
 *   hs_join                            bitmap::handler::join / fgssjoin::handler::join
 *                                      for block and partition pairs,
 *                                      device_algorithms.hpp:177-247, :363-432
 *   hs_verify_candidates               hybrid::device_handler::addCandidates + flush,
 *                                      include/hyset/hybrid_algorithms.hpp:71-200
 *   hs_result                          output::count_handler / pairs_handler,
 *                                      include/hyset/output.hpp:19-37
 */
#ifndef HYSET_AMD_H
#define HYSET_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------- return codes */
enum {
  HS_OK = 0,
  HS_E_INVALID = -1,     /* bad argument or malformed collection */
  HS_E_OOM = -2,         /* host or device allocation failed */
  HS_E_OVERFLOW = -3,    /* result capacity too small; hs_result.count holds the size needed */
  HS_E_DEVICE = -4,      /* HIP runtime error; see hs_last_error */
  HS_E_UNSUPPORTED = -5, /* combination not supported (e.g. set size above the bound tables) */
  HS_E_BUSY = -6         /* admission refused: the request does not fit the memory budget now */
};

/* ------------------------------------------------------ host collections */

/* A collection in CSR form: set i holds tokens[starts[i] .. starts[i]+sizes[i]),
 * strictly ascending; sets ordered by size ascending.  Token ids are dense
 * 0..universe-1 ranked by ascending global frequency.  Arrays are malloc'd by
 * the generator and released with hs_host_collection_free. */
typedef struct {
  uint32_t n;
  uint64_t n_tokens;
  uint32_t universe; /* max token + 1 */
  uint32_t* tokens;
  uint32_t* starts;
  uint32_t* sizes;
} hs_host_collection;

enum { HS_SIZE_UNIFORM = 0, HS_SIZE_LOGNORMAL = 1 };

/* Synthetic-input recipe: n sets, sizes uniform in [size_lo, size_hi] or
 * round(lognormal(median, sigma)) clipped to [size_lo, size_hi]; tokens drawn
 * from Zipf(zipf_s) over ranks 1..universe without repetition inside a set;
 * then re-ranked by ascending frequency (ids from 1), tokens sorted inside a
 * set and sets sorted by (size, tokens).  xoshiro256** seeded by splitmix64. */
typedef struct {
  uint32_t n;
  int32_t size_dist;
  double size_median;
  double size_sigma;
  uint32_t size_lo;
  uint32_t size_hi;
  uint32_t universe;
  double zipf_s;
  uint64_t seed;
} hs_gen_params;

int hs_generate_self(const hs_gen_params* p, hs_host_collection* out);

/* R and S with a shared token ranking; the first round(dup_frac*|S|) raw S sets
 * are copies of random R sets with round(f*|r|) edits, f ~ U[edit_lo, edit_hi],
 * each edit a substitution, deletion or insertion with equal probability. */
int hs_generate_rs(const hs_gen_params* r, const hs_gen_params* s, double dup_frac, double edit_lo,
                   double edit_hi, hs_host_collection* out_r, hs_host_collection* out_s);

/* An indexed collection and a query collection drawn from the same token
 * space; tokens are ranked by their frequency in the indexed collection and
 * query-only tokens get ids above every indexed token. */
int hs_generate_queries(const hs_gen_params* index_params, const hs_gen_params* query_params,
                        hs_host_collection* out_index, hs_host_collection* out_queries);

void hs_host_collection_free(hs_host_collection* c);

/* Reads the reference's text format (include/hyset/collection.hpp:92-177):
 * one set per line, space-separated token ids, empty lines skipped. */
int hs_read_collection_text(const char* path, hs_host_collection* out);
int hs_write_collection_text(const char* path, const uint32_t* tokens, const uint32_t* starts,
                             const uint32_t* sizes, uint32_t n);

/* ----------------------------------------------------- dataset preparation */

/* Raw records to the format above -- the preparation the reference leaves to
 * the external scripts of the set-similarity-join benchmark it points to
 * (README.md:14; the format it then assumes, collection.hpp:117-135).  One
 * record per line; tokens are whitespace-separated words (HS_TOKENS_WORDS),
 * the character q-grams of the line padded with q-1 marks on both sides
 * (HS_TOKENS_QGRAMS, q from 1 to 64, default 3) or integer ids
 * (HS_TOKENS_INTEGERS).  The k-th occurrence of a token inside a record is
 * a token of its own, so records are sets.  Token ids are ranked by
 * ascending number of records holding them (ties: first appearance),
 * starting at 1; tokens ascend inside each set; sets are sorted by (size,
 * tokens); records without tokens are dropped.  record_of (may be NULL)
 * receives a malloc'd array: the 0-based line number of every set (free it
 * with hs_free_array). */
enum { HS_TOKENS_WORDS = 0, HS_TOKENS_QGRAMS = 1, HS_TOKENS_INTEGERS = 2 };
typedef struct {
  int32_t mode;      /* HS_TOKENS_* */
  uint32_t q;        /* q-gram length (0: 3) */
  int32_t lowercase; /* fold ASCII letters to lower case first */
} hs_prep_params;

int hs_preprocess_text(const char* text, uint64_t len, const hs_prep_params* params, hs_host_collection* out,
                       uint32_t** record_of);
int hs_preprocess_file(const char* path, const hs_prep_params* params, hs_host_collection* out,
                       uint32_t** record_of);
void hs_free_array(void* p);

/* Binary CSR cache of a collection: the 8 bytes "HSCSR1\0\0", u32 n, u32
 * universe, u64 n_tokens, then sizes u32[n] and tokens u32[n_tokens] (sets
 * in order, dense), little-endian.  The reader checks the magic, that the
 * sizes add up to n_tokens, that no token reaches the universe and that the
 * file ends there; anything else is HS_E_INVALID.  starts may be NULL for a
 * dense CSR. */
int hs_write_collection_binary(const char* path, const uint32_t* tokens, const uint32_t* starts,
                               const uint32_t* sizes, uint32_t n);
int hs_read_collection_binary(const char* path, hs_host_collection* out);

/* Writes pairs as "r s\n" lines (include/hyset/output.hpp:39-76), or
 * "r s overlap\n" when overlap is not NULL. */
int hs_write_pairs_text(const char* path, const uint32_t* r, const uint32_t* s, const uint32_t* overlap,
                        uint64_t count);

/* --------------------------------------------------------- join settings */

enum { HS_JACCARD = 0, HS_COSINE = 1, HS_DICE = 2 };

/* HS_BOUNDS_REFERENCE: Jaccard bounds are the reference's IEEE-double
 * expressions (include/hyset/similarity.hpp:40-50), so pairs come out
 * exactly as the reference's cpu_single emits them, including at the sums
 * where the double quotient rounds up.  HS_BOUNDS_EXACT: the threshold is
 * read as the rational given by its shortest decimal form and every bound is
 * integer-exact.  Cosine and Dice always use the exact form. */
enum { HS_BOUNDS_REFERENCE = 0, HS_BOUNDS_EXACT = 1 };

/* Technique selection of the reference's drivers (src/queue.cu:181-214):
 * bitmap signatures, prefix filter (fgssjoin), or CPU filtering with device
 * verification (hybrid). */
enum { HS_ALGO_FGSS = 0, HS_ALGO_BITMAP = 1, HS_ALGO_HYBRID = 2 };

typedef struct hs_ctx hs_ctx;
typedef struct hs_coll hs_coll;

/* The integer bounds the device tables hold for (sim, threshold, mode):
 * out[0] = minsize(l1), out[1] = maxsize(l1), out[2] = minoverlap(l1, l2),
 * out[3] = maxprefix(l1) (probe prefix), out[4] = midprefix(l1) (self-join
 * index prefix).  Host-only; needs no device. */
int hs_bounds(int32_t sim, double threshold, int32_t bounds_mode, uint32_t l1, uint32_t l2, uint32_t out[5]);

/* Inclusive set-id interval, like a block's startID/endID
 * (include/hyset/structs.hpp:10-54).  first > last is an empty range. */
typedef struct {
  uint32_t first;
  uint32_t last;
} hs_range;

/* Tuning knobs; zero selects the default. */
typedef struct {
  uint32_t bitmap_bits;       /* bitmap: signature width 64, 128, 256, 512 or 1024 (default: the smallest
                                 power of two >= 2 x the mean indexed set size, within 64..1024) */
  uint32_t probes_per_tile;   /* fgss: fixed tiles of this many probes, 1..512 (default 0: the
                                 collection's tile plan); batched fgss: probes per tile (default 512) */
  uint32_t verify_scenario;   /* hybrid: 1 lane per candidate, 2 wave merge-path per
                                 candidate, 0 pick by set size (default 0) */
  uint32_t host_threads;      /* hybrid: CPU filter threads (default 1) */
  uint64_t batch_candidates;  /* hybrid: candidates per device batch (default 1<<22) */
  uint32_t flags;             /* HS_TUNE_* bits */
  uint32_t tile_tokens;       /* batched fgss: probe tokens per tile, <= 8192 (default 8192) */
  uint32_t workgroups_per_cu; /* fgss, batched fgss: persistent workgroups per CU (default: as
                                 many as a CU holds, 2 of 512 lanes or 4 of 256) */
  uint32_t signature_bits;    /* fgss: width of the parity-signature test that runs between the
                                 positional filter and verification, 64, 128, 256, 512 or 1024
                                 (default: the smallest power of two >= 2 x the mean indexed set
                                 size, within 64..1024) */
} hs_tune;

/* hs_join_batch and hs_join: capture the device sequence in a HIP graph on
 * first use and replay it while the join (collections, ranges, bounds, knobs,
 * result capacity) is the same and no device or pinned buffer has been
 * allocated or freed since.  hs_join replays its index build and launch for
 * the upload-time and fixed tile plans and the bitmap technique; the
 * balanced and tile-queue plans and the hybrid technique run eagerly. */
enum { HS_TUNE_GRAPH = 1 };
/* hs_join fgss plans.  A self-join of one id range with itself (indexed ==
 * probe, equal ranges) runs token-major by default (HS_PLAN_TOKEN): the
 * index is built over the probe prefixes, sorted by (token, set id), and
 * one lane per index entry walks back through its token's list over the
 * probe's length window, the survivors of the positional filter and the
 * signature test queued and verified by a second kernel.  Every other join,
 * and any join with one of the flags below or probes_per_tile set, runs
 * item-major over an index of the index prefixes: a workgroup per tile of
 * probes searches each prefix token's posting list for the probe's window.
 * HS_TUNE_STATIC_TILES: one workgroup per tile of the collection's
 * upload-time plan (at most 512 probes and 8192 tokens).
 * HS_TUNE_TILE_QUEUE: the same tiles claimed by a persistent grid
 * (workgroups_per_cu per CU) from per-XCD queues.  HS_TUNE_BALANCED_TILES:
 * tiles of equal estimated work, planned on the device after the index
 * build by a pass that locates every probe-prefix item's posting-list
 * slice, claimed by the persistent grid.  HS_TUNE_ITEM_MAJOR: item-major
 * with the automatic tile plan: the balanced plan for joins of at least
 * 2^21 probe-prefix items whose index holds at least 3.5 entries per token
 * of the universe (long posting lists, so tiles differ much in work), the
 * upload-time tiles otherwise. */
enum { HS_TUNE_TILE_QUEUE = 2, HS_TUNE_BALANCED_TILES = 8, HS_TUNE_STATIC_TILES = 16, HS_TUNE_ITEM_MAJOR = 64 };
/* hs_stats.tile_plan: the plan a fgss join ran with. */
enum { HS_PLAN_FIXED = 1, HS_PLAN_UPLOAD = 2, HS_PLAN_QUEUE = 3, HS_PLAN_BALANCED = 4, HS_PLAN_TOKEN = 5 };
/* fgss: no signature test (every visit passing the positional filter is
 * verified by a merge, as in the reference's fgssjoin).  The pair set is the
 * same either way. */
enum { HS_TUNE_NO_SIGNATURE = 4 };
/* hs_join: record events between the stages, so that hs_stats holds
 * index_ms, filter_ms and kernel_ms.  Each event costs the device a few
 * microseconds of idle time, so by default a join is timed as a whole
 * (total_ms, and transfer_ms when pairs are copied) and the stage times
 * are 0.  The hybrid technique always records its stages. */
enum { HS_TUNE_STAGE_TIMES = 32 };
/* hs_join_batch knobs of the autotuner (hs_autotune_batch).
 * HS_TUNE_NARROW: 256-lane workgroups over tiles of at most 256 probes and
 * 4096 tokens (default: 512 lanes, 512 probes, 8192 tokens); with
 * probes_per_tile it sets the threads per probe (lanes / probes per tile).
 * HS_TUNE_INDEX_MAXPREFIX: the prefix scheme -- the self-joins' index holds
 * every set's probe prefix (maxprefix) instead of the index prefix
 * (midprefix) the reference fixes (similarity.hpp:13-19).  Exact either
 * way: a pair's first common token lies in both prefixes. */
enum { HS_TUNE_NARROW = 128, HS_TUNE_INDEX_MAXPREFIX = 256 };

/* Result sink.  count is always written.  With count_only == 0 the pairs
 * (r = probe id, s = indexed id, overlap = |r n s|) are copied into the
 * caller's arrays when count <= capacity, otherwise HS_E_OVERFLOW is
 * returned and nothing is copied.  With count_only != 0 the pairs stay on
 * the device until the next join and can be fetched with
 * hs_copy_last_result.  Pair order is unspecified (the reference emits in
 * atomic order as well); sort for comparisons.  overlap may be NULL. */
typedef struct {
  uint64_t count;
  uint64_t capacity;
  uint32_t* r;
  uint32_t* s;
  uint32_t* overlap;
  int32_t count_only;
} hs_result;

/* Timings (ms, HIP events on the context stream) and work counters of the
 * last join on a context.  hs_join fills index_ms, filter_ms and kernel_ms
 * with HS_TUNE_STAGE_TIMES only; total_ms runs from the first launch to the
 * counters' (and, when copied, the pairs') arrival on the host. */
typedef struct {
  double index_ms;
  double filter_ms; /* candidate generation + verification (fused for fgss/bitmap) */
  double transfer_ms;
  double total_ms;
  double host_filter_ms; /* hybrid: CPU filtering wall time */
  double kernel_ms;      /* the main filter/verify kernel alone (fgss token-major plan: its filter,
                            second pass and verification kernels, the sum of the three
                            dispatches' durations) */
  uint64_t index_entries;
  uint64_t posting_visits;   /* fgss: postings inside length windows; bitmap: pairs tested */
  uint64_t candidates;       /* distinct pairs verified; bitmap: filter survivors */
  uint64_t pairs;
  uint64_t probe_prefix_tokens;
  uint64_t candidate_tokens; /* fgss: sum of |s| over verified candidates */
  uint64_t signature_tests;  /* fgss: visits that passed the positional filter and were tested
                                by signature (0 without the signature test) */
  uint32_t signature_bits;   /* fgss: signature width used (0: none) */
  uint32_t tile_plan;        /* fgss: HS_PLAN_* (0 for other techniques) */
  uint32_t graph;            /* HS_TUNE_GRAPH joins: 1 the sequence was captured, 2 replayed (0: eager);
                                a replayed join is timed as a whole: index_ms and kernel_ms are 0,
                                filter_ms covers the index build, the launch and the counters */
  uint32_t reserved;
  uint64_t candidate_probe_tokens; /* fgss token-major plan with HS_TUNE_STAGE_TIMES: sum of |r| over the
                                      distinct probes with a verified candidate (else 0) */
} hs_stats;

/* ------------------------------------------------------------- contexts */

/* device < 0 creates a host-only context: collections stay in host memory
 * and only hs_join_host, the bounds and the cost models are available. */
int hs_create(int device, hs_ctx** out);
void hs_destroy(hs_ctx* ctx);
const char* hs_last_error(const hs_ctx* ctx);

/* Copies a collection to device memory.  sizes[i] >= 1; starts may be NULL
 * for a dense CSR (starts[i] = sum of earlier sizes).  Sets must be sorted
 * by size and tokens ascending inside a set (checked). */
int hs_upload_collection(hs_ctx* ctx, const uint32_t* tokens, const uint32_t* starts, const uint32_t* sizes,
                         uint32_t n, uint64_t n_tokens, hs_coll** out);
void hs_free_collection(hs_coll* coll);

int hs_set_threshold(hs_ctx* ctx, int32_t sim, double threshold, int32_t bounds_mode);

/* Candidate generation on the host cores: the reference's CPU filtering
 * algorithm, --algorithm allpairs|ppjoin of src/hybrid.cu:36 and
 * src/cpu_single.cu:35.  It applies to the host filter of HS_ALGO_HYBRID
 * (hybrid::allpairs / hybrid::ppjoin, hybrid_algorithms.hpp:209-387), to
 * hs_join_host (host::allpairs / host::ppjoin, host_algorithms.hpp:16-250)
 * and to the host side of hs_join_coprocess.  HS_HOST_PPJOIN keeps a
 * candidate only if PPJoin's positional filter (host_algorithms.hpp:196-199,
 * helpers.hpp:46-50) passes at the first prefix token the probe and the
 * candidate share: minoverlap(|r|, |s|) <= min(|r| - pos_r, |s| - pos_s).
 * The pairs are the same either way; fewer candidates reach verification.
 * The default is HS_HOST_ALLPAIRS (as the reference's drivers). */
enum { HS_HOST_ALLPAIRS = 0, HS_HOST_PPJOIN = 1 };
int hs_set_host_algorithm(hs_ctx* ctx, int32_t algorithm);

/* Joins probe sets of `probe` (within probe_range) against the sets of
 * `indexed` (within indexed_range).  indexed == probe is a self-join and
 * keeps pairs with s < r only; otherwise every (r, s) is considered (R-S
 * join, index prefix = maxprefix).  NULL ranges mean the whole collection. */
int hs_join(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, const hs_range* indexed_range,
            const hs_range* probe_range, int32_t algo, const hs_tune* tune, hs_result* result);

/* Verifies one CSR batch of candidates: row i is probe probes[i] with
 * candidates[(i ? offsets[i-1] : 0) .. offsets[i]) of `indexed`
 * (inclusive-end offsets, as hybrid_algorithms.hpp:71-105 builds them).
 * scenario: 1 lane per candidate, 2 merge-path wave per candidate, 0 auto. */
int hs_verify_candidates(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, const uint32_t* probes,
                         const uint32_t* offsets, const uint32_t* candidates, uint32_t n_probes,
                         int32_t scenario, hs_result* result);

/* ------------------------------------------------------- batched joins */

/* Uploads n_joins independent collections (each sorted by size, tokens
 * ascending; starts[j] may be NULL for a dense CSR) as one batch.  Set ids of
 * join j are offset by the number of sets of the joins before it, token ids
 * by their universes, so results of hs_join_batch are in batch ids:
 * join j owns ids [sum_{k<j} n_sets[k], sum_{k<=j} n_sets[k]).
 * Replaces the block-pair queue of src/queue.cu:142-279 by one device work
 * queue; see hs_join_batch. */
int hs_upload_batch(hs_ctx* ctx, uint32_t n_joins, const uint32_t* const* tokens, const uint32_t* const* starts,
                    const uint32_t* const* sizes, const uint32_t* n_sets, const uint64_t* n_tokens, hs_coll** out);

/* Self-joins every collection of a batch with the context's threshold: one
 * index build over all of them and one persistent prefix-filter launch that
 * drains a device work queue of probe tiles.  tune: probes_per_tile,
 * tile_tokens, workgroups_per_cu, signature_bits, flags (HS_TUNE_GRAPH,
 * HS_TUNE_NO_SIGNATURE, HS_TUNE_NARROW, HS_TUNE_INDEX_MAXPREFIX). */
int hs_join_batch(hs_ctx* ctx, hs_coll* batch, const hs_tune* tune, hs_result* result);

/* Times hs_join_batch over a grid of threads per probe (workgroup width
 * 512 or 256 lanes over 64-512 probes per tile) and persistent-grid sizes,
 * then over the prefix scheme (midprefix or maxprefix index) and signature
 * widths (64, 128, 256 bits, none) at the best point so far, timing each
 * setting as the median of reps runs after one warm-up (reps <= 0: 3).  The
 * default settings (all fields 0) start as the best; a setting replaces the
 * best only when 2% faster, and the final winner is timed against the
 * defaults again with twice the runs, so noise cannot return a setting
 * slower than the defaults.  *best_ms receives the winner's median device
 * time. */
int hs_autotune_batch(hs_ctx* ctx, hs_coll* batch, int32_t reps, hs_tune* best, double* best_ms);

/* -------------------------------------------------------------- serving */

/* A device-resident index over one collection answering requests (query
 * sets + threshold) with continuous batching.  No reference counterpart
 * (the reference runs offline joins only); results are those of hs_join with
 * the query sets as probes and the index collection as the indexed side. */
typedef struct hs_server hs_server;

typedef struct {
  double min_threshold;      /* lowest threshold accepted; the index prefix is built for it (default 0.6) */
  uint32_t slots;            /* worker streams = device batches in flight (default 4) */
  uint64_t hbm_budget_bytes; /* device bytes the batches in flight may hold (default 1 GiB) */
  uint32_t max_batch_queries;/* query sets per device batch (default 65536) */
  uint32_t max_query_size;   /* largest query set (default 4096; 1024 for cosine) */
} hs_server_config;

/* signature_bits: width of the fgss signature test the server runs (chosen
 * from the index's mean set size); a request of n query sets and T tokens
 * holds 4 T + (109 + signature_bits / 8) n + 4096 device bytes. */
typedef struct {
  uint64_t submitted, batches, deferred, rejected, max_in_flight, max_bytes_in_flight, signature_bits;
} hs_server_stats;

/* index: a collection uploaded on `device` (kept alive by the caller). */
int hs_server_create(int device, hs_coll* index, int32_t sim, int32_t bounds_mode, const hs_server_config* cfg,
                     hs_server** out);
/* Queues n_queries query sets (dense CSR: sizes[i] tokens each, ascending)
 * at `threshold`; returns at once with a request id. */
int hs_server_submit(hs_server* server, const uint32_t* tokens, const uint32_t* sizes, uint32_t n_queries,
                     uint64_t n_tokens, double threshold, uint64_t* request_id);
/* Blocks until the request is answered; r = query index inside the request,
 * s = index set id.  *latency_ms = submit-to-answer time.  HS_E_BUSY: the
 * request could never fit the device budget.  HS_E_OVERFLOW (count set):
 * the arrays are too small; the answer is kept for another call.  Otherwise
 * the request is released. */
int hs_server_wait(hs_server* server, uint64_t request_id, hs_result* result, double* latency_ms);
int hs_server_get_stats(hs_server* server, hs_server_stats* out);
void hs_server_destroy(hs_server* server);

/* The CPU partition: the same join on host cores (host::allpairs, or
 * host::ppjoin after hs_set_host_algorithm(HS_HOST_PPJOIN), of
 * include/hyset/host_algorithms.hpp:16-250 with the thread layout of
 * src/cpu_multi.cu:118-207), with verification and exact overlaps.  Works on
 * device and host-only contexts.  hs_stats: pairs, posting_visits (postings
 * inside the length windows), candidates (distinct candidates verified),
 * host_filter_ms = total_ms (wall). */
int hs_join_host(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, const hs_range* indexed_range,
                 const hs_range* probe_range, uint32_t threads, hs_result* result);

/* One join shared between the device and host cores running concurrently,
 * over the whole probe collection against the whole indexed one (a
 * self-join when indexed == probe).  The device builds the prefix index
 * once; a copy of it travels to the host on a second stream while the
 * device starts probing.  The device takes probe stripes from the top of the
 * (size-sorted) probe ids, the host threads take small stripes from the
 * bottom, each with the CPU partition's join (hs_join_host's per-probe
 * filter and verification).
 *   HS_COPROC_QUEUE: both sides claim stripes until they meet -- the shared
 *     block-pair queue of src/queue.cu:142-279 with a device consumer and
 *     host consumers.  Until both sides have a measured rate the device
 *     claims half of the unclaimed range (at least device_min_claim sets);
 *     then it claims the top part sized so that the host, joining the rest
 *     at its measured rate, ends when the device does.  Two device stripes
 *     are in flight at a time.  Host claims are cpu_block sets, or sized per
 *     thread to about 0.1 ms at its measured rate when cpu_block is 0, and
 *     never more than an equal share of what is left.  Rates are taken per
 *     unit of a work proxy: set size + 8 per probe.
 *   HS_COPROC_STATIC: the split of src/dichotomy.cu:143-285 into the small
 *     sets Rs (the probe sets below cpu_share * n; cpu_share is the
 *     reference's --split / 100) and the rest Rl.  policy names the sides
 *     as the reference's --policy pair (small part, large part):
 *     HS_POLICY_CPU_GPU (cpu,gpu): the host joins Rs with itself, the
 *     device R against Rl; HS_POLICY_GPU_CPU (gpu,cpu): the device joins
 *     Rs, the host R against Rl.
 * The device technique (algo) is fgss, bitmap or hybrid.  With hybrid the
 * device side is the reference's cpu-gpu side (dichotomy.cu:204-258, and in
 * queue mode the --cooperative consumer of queue.cu:189-241): its stripes
 * are filtered on the host, by tune.host_threads threads (default 1) beside
 * the cpu_threads that join, against the same host copy of the index, and
 * their candidates verified on the device.
 * The result is the same set of pairs as hs_join: afterwards the context's
 * last result is the device's pairs followed by the host's (those stay in
 * host memory, no copy to the device), count-only or copied into result as
 * for hs_join, and hs_copy_last_result, hs_stream_last_result and
 * hs_write_last_result_text deliver it in that order.  Each device stripe chooses its tile
 * plan as hs_join does for a probe range.
 *
 * Break-even gate (queue mode, min_device_ms >= 0): sharing pays only when
 * the host's share outweighs its fixed cost (the index copy, the threads,
 * and the device's stripes missing the whole-collection token-major plan),
 * so the context remembers, per join (collections, technique, threshold,
 * knobs), the measured wall time of the device alone and of the last
 * co-processed run, and runs the device alone (hs_join) when (1) the join
 * has no device-alone time yet (the call measures it), (2) that time is
 * below min_device_ms (0: 50 ms), or (3) a measured co-processed run was
 * not faster.  Otherwise it co-processes and records the time.
 * min_device_ms < 0 always co-processes; static mode is never gated. */
enum { HS_COPROC_QUEUE = 0, HS_COPROC_STATIC = 1 };
enum { HS_POLICY_CPU_GPU = 0, HS_POLICY_GPU_CPU = 1 };
typedef struct {
  int32_t mode;              /* HS_COPROC_* */
  int32_t algo;              /* device technique: HS_ALGO_FGSS, HS_ALGO_BITMAP or HS_ALGO_HYBRID */
  uint32_t cpu_threads;      /* host threads (0: the usable CPUs less one, for the device thread) */
  uint32_t cpu_block;        /* queue: probe sets per host claim (0: about 0.1 ms of work) */
  uint32_t device_min_claim; /* queue: smallest device claim in probe sets (0: 16384) */
  double cpu_share;          /* static: fraction of the probe sets in the small part Rs */
  hs_tune tune;              /* device knobs */
  double min_device_ms;      /* queue: the break-even gate (0: 50 ms; < 0: always co-process) */
  int32_t policy;            /* static: HS_POLICY_* (queue: HS_POLICY_CPU_GPU only) */
  uint32_t reserved;
} hs_coprocess;

typedef struct {
  uint64_t device_pairs, host_pairs;
  uint32_t device_probes, host_probes; /* probe sets joined by each side */
  uint32_t device_claims, host_claims; /* stripes each side claimed */
  uint32_t split;                      /* first probe id of the large part: the device's share (queue,
                                          HS_POLICY_CPU_GPU) or the host's (HS_POLICY_GPU_CPU) */
  uint32_t host_threads;
  double index_ms;  /* device index build and signatures */
  double copy_ms;   /* index copy to the host, from the end of the build */
  double device_ms; /* first device claim to the device's last stripe done */
  double host_ms;   /* host threads: start to the last stripe done */
  double wall_ms;   /* the whole call */
  double device_alone_ms; /* the gate's measured device-alone wall time of this join (0: none yet) */
  double shared_ms;       /* the gate's last measured co-processed wall time (0: none yet) */
  uint32_t gated;         /* 1: the gate ran the device alone */
  uint32_t reserved;
} hs_coprocess_stats;

int hs_join_coprocess(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, const hs_coprocess* options,
                      hs_result* result, hs_coprocess_stats* stats);

/* ------------------------------------------------- cost models, split */

/* Partition features of include/hyset/statistics.hpp:163-283, computed on
 * the device for sets [range] of a device collection under the context's
 * threshold: 0 number of sets, 1 mean set size, 2 set-size standard
 * deviation, 3 standard deviation of the prefix-token list lengths over the
 * universe, 4 universe, 5 prefix-index entries, 6 threshold. */
#define HS_NFEATURES 7
int hs_features(hs_ctx* ctx, hs_coll* coll, const hs_range* range, float out[HS_NFEATURES]);

/* The analytic cost model: per technique a linear function of the features
 * whose absolute value is the predicted cost (statistics.hpp:132-161 with the
 * coefficient table of src/cost_model.cu:296-323), fitted here by ridge least
 * squares.  X is n x d row-major, Y is n x n_out; coef receives (d + 1) x
 * n_out (row d = intercepts). */
int hs_linear_fit(const double* X, const double* Y, uint32_t n, uint32_t d, uint32_t n_out, double ridge,
                  double* coef);
int hs_linear_predict(const double* coef, const double* X, uint32_t n, uint32_t d, uint32_t n_out, double* Y);

/* The learned cost model: an MLP (n_in -> h1 -> h2 -> n_out, ReLU) trained
 * and evaluated by device kernels.  No reference counterpart: it replaces
 * the linear costs above where they mispredict.  Inputs are standardised by
 * the mean/stdv set with hs_mlp_set_normalization; parameters are laid out
 * W1 (h1 x n_in), b1, W2 (h2 x h1), b2, W3 (n_out x h2), b3. */
typedef struct hs_mlp hs_mlp;
int hs_mlp_create(hs_ctx* ctx, uint32_t n_in, uint32_t h1, uint32_t h2, uint32_t n_out, uint64_t seed, hs_mlp** out);
void hs_mlp_destroy(hs_mlp* mlp);
uint32_t hs_mlp_num_params(const hs_mlp* mlp);
int hs_mlp_set_normalization(hs_mlp* mlp, const float* mean, const float* stdv);
int hs_mlp_get_params(const hs_mlp* mlp, float* out);
int hs_mlp_set_params(hs_mlp* mlp, const float* in);
/* X: n x n_in host rows; Y: n x n_out host outputs. */
int hs_mlp_predict(hs_mlp* mlp, const float* X, uint32_t n, float* Y);
/* Mean squared error over n x n_out and its gradient (num_params floats). */
int hs_mlp_loss_grad(hs_mlp* mlp, const float* X, const float* Y, uint32_t n, float* loss, float* grad);
/* Full-batch Adam for `epochs` steps; *final_loss (may be NULL) = loss after them. */
int hs_mlp_train(hs_mlp* mlp, const float* X, const float* Y, uint32_t n, uint32_t epochs, float lr,
                 float* final_loss);

/* Techniques a split may give a unit (include/hyset/cost.hpp:28-58 chooses
 * among the device techniques and the CPU). */
#define HS_TECH_FGSS 0
#define HS_TECH_BITMAP 1
#define HS_TECH_HYBRID 2
#define HS_TECH_CPU 3

/* Runs a split: every unit (a collection self-join under the context's
 * threshold) with its technique (bitmap at 64 bits); HS_TECH_CPU units run on cpu_threads host
 * threads (which claim the largest units first, by token count) while the
 * device units run one after another on the context's stream.  counts[u] = pairs of unit u; unit_ms[u] (may be NULL) its time;
 * *wall_ms (may be NULL) the time of the whole split. */
int hs_split_run(hs_ctx* ctx, uint32_t n_units, hs_coll* const* colls, const int32_t* technique,
                 uint32_t cpu_threads, uint64_t* counts, double* unit_ms, double* wall_ms);

/* ------------------------------------- partition pairs of one self-join */

/* Partition features of two set-id ranges joined with each other
 * (statistics::extract(left, right), include/hyset/statistics.hpp:216-283),
 * laid out as hs_features' (the threshold in slot 6):
 *   0 size = |left| + |right|;
 *   1, 2 mean and standard deviation (population) of |left| + |right|
 *     values: the sizes of left's sets but its last, a zero, the sizes of
 *     right's sets but its last, a zero (the reference copies [startID,
 *     endID) of each range into a zero-filled vector of |left| + |right|
 *     slots, endID being inclusive);
 *   3 standard deviation over the universe of the per-token counts of the
 *     midprefix entries of both ranges' sets;
 *   4 universe; 5 indexSize = left's midprefix entries.
 * Computed on the device (size moments, a token histogram). */
int hs_features_pair(hs_ctx* ctx, hs_coll* coll, const hs_range* left, const hs_range* right,
                     float out[HS_NFEATURES]);

/* Partitioners over blocks of block_size consecutive sets, the last block
 * possibly short (blocks as include/hyset/collection.hpp:147-172 cuts them).
 * hs_partition_even (partitioner::even, include/hyset/partitioner.hpp:11-30):
 * partitions of ceil(blocks / parts) consecutive blocks; out (room for
 * `parts` ranges) receives their set-id ranges, *n_out their number (which
 * may be below parts).  hs_partition_dichotomize (partitioner::dichotomize,
 * :32-68): out[0] the whole collection, out[1] the first
 * ceil(split_pct / 100 * blocks) blocks (Rs), out[2] the rest (Rl); an empty
 * range has first > last. */
int hs_partition_even(uint32_t n_sets, uint32_t block_size, uint32_t parts, hs_range* out, uint32_t* n_out);
int hs_partition_dichotomize(uint32_t n_sets, uint32_t block_size, double split_pct, hs_range out[3]);

/* One technique's linear cost (statistics::coefficients,
 * include/hyset/statistics.hpp:132-161; the per-technique, per-threshold
 * table of src/cost_model.cu:296-323): cost = |intercept + set_size_std *
 * setSizeStd + average_set_size * averageSetSize + token_freq *
 * tokenFreqDistribution + size * size + size_square * size^2 +
 * universe_size * universeSize + index_size * indexSize + bitmap * bitmap|,
 * evaluated in long double.  enabled = 0 leaves the technique out of the
 * choice (cost_model.cu:133-135 drops bitmap when --bitmap is 0). */
typedef struct {
  double intercept, set_size_std, average_set_size, token_freq, size, size_square, universe_size, index_size,
      bitmap;
  int32_t enabled;
  uint32_t reserved;
} hs_cost_coefficients;

/* hs_cost_run flags: HS_COST_WRAP_SQUARE evaluates size^2 as the reference
 * does, an unsigned 32-bit product that wraps above 65,535 sets
 * (statistics.hpp:155); by default the square is exact. */
#define HS_COST_WRAP_SQUARE 1u

/* cost::run (include/hyset/cost.hpp:28-58) on the device collection under
 * the context's threshold: for every partition pair (i, j), i <= j, in the
 * order (0,0), (0,1), ..., (0,n-1), (1,1), ..., the features (hs_features
 * of parts[i] when i == j, hs_features_pair(parts[i], parts[j]) otherwise),
 * bitmap = bitmap_bits, and the enabled technique of least cost, ties to
 * the first in the reference's map order (bitmap, cpu, fgssjoin, hybrid).
 * coef[HS_TECH_*].  technique receives n_parts * (n_parts + 1) / 2 HS_TECH_*
 * values; features (may be NULL) HS_NFEATURES floats per pair; costs (may
 * be NULL) 4 doubles per pair, indexed by HS_TECH_* (disabled: -1). */
int hs_cost_run(hs_ctx* ctx, hs_coll* coll, const hs_range* parts, uint32_t n_parts,
                const hs_cost_coefficients coef[4], uint32_t bitmap_bits, uint32_t flags, int32_t* technique,
                float* features, double* costs);

/* One partition pair of a plan: the probe sets of `right` joined with the
 * sets of `left` (pairs s < r, s in left, r in right) by `technique`. */
typedef struct {
  hs_range left;
  hs_range right;
  int32_t technique; /* HS_TECH_* */
  uint32_t reserved;
  uint64_t pairs;    /* out: pairs found */
  double ms;         /* out: wall time of this pair's join */
} hs_partition_pair;

/* Runs a plan of partition pairs of one self-join, the execution half of
 * src/cost_model.cu:150-262: the device pairs on this thread one after
 * another (hybrid first, then bitmap, then fgss, as the reference orders
 * them; tune as for hs_join: bitmap_bits, host_threads for the hybrid
 * filter, ...), the HS_TECH_CPU pairs (host::allpairs, or host::ppjoin
 * after hs_set_host_algorithm) on cpu_threads host threads meanwhile,
 * largest probe side first.  The pairs (i, j), i <= j, of a partition of the
 * collection's ids give exactly the join.  result->count = all pairs.
 * Unless count_only, the pairs are gathered in host memory (each device
 * pair's after its join, then the CPU pairs) and become the context's last
 * result, copied into result when it has room (else HS_E_OVERFLOW, the
 * result kept for hs_copy_last_result); count-only runs keep none.
 * *wall_ms (may be NULL) = the time of the whole plan. */
int hs_run_partition_pairs(hs_ctx* ctx, hs_coll* coll, hs_partition_pair* pairs, uint32_t n_pairs,
                           uint32_t cpu_threads, const hs_tune* tune, hs_result* result, double* wall_ms);

/* Copies the pairs of the last join into result (count_only ignored).  The
 * pairs travel in chunks through pinned host buffers, the next chunk's copy
 * in flight while the previous one is moved into the caller's arrays. */
int hs_copy_last_result(hs_ctx* ctx, hs_result* result);

/* Streaming result sink (the reference's output handlers,
 * include/hyset/output.hpp:13-93, fed chunk by chunk instead of from one
 * host array).  The pairs of the last join on the context reach `sink` in
 * chunks of at most chunk_pairs pairs (0: 262144), in the order of the
 * device result buffer; overlap is NULL unless with_overlap.  The arrays are
 * pinned staging buffers valid only during the call; the next chunk is being
 * copied while the sink runs.  A nonzero return from the sink stops the
 * stream and the call returns HS_E_INVALID.  The sink must not call into the
 * same context. */
typedef int (*hs_pair_sink)(void* user, const uint32_t* r, const uint32_t* s, const uint32_t* overlap,
                            uint64_t n);
int hs_stream_last_result(hs_ctx* ctx, uint64_t chunk_pairs, int32_t with_overlap, hs_pair_sink sink, void* user);

/* Writes the pairs of the last join as hs_write_pairs_text does, streamed
 * from the device in pinned chunks (formatting overlaps the next chunk's
 * copy); no host array of the whole result is built. */
int hs_write_last_result_text(hs_ctx* ctx, const char* path, int32_t with_overlap);

int hs_get_stats(const hs_ctx* ctx, hs_stats* out);

/* ------------------------------------------------- allocation-free join */

/* hs_join with the fgss technique that allocates nothing: no hipMalloc, no
 * hipFree, no pinned or library allocation.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  Inputs are collections
 * already uploaded (hs_upload_collection) and the context's threshold; the
 * result triples are written into the caller's DEVICE arrays r, s and
 * overlap (capacity entries each) and *count receives the number of pairs.
 * The bound tables, the signatures, the inverted index with its sort work
 * space and the work counters live in one device array of
 * HS_NOALLOC_SCRATCH_BYTES per device, fixed for the life of the process
 * (its address is taken in hs_create).  A join that needs more than that
 * returns HS_E_UNSUPPORTED (the 1M-set self-join at Jaccard 0.8 needs about
 * 80 MB); more pairs than capacity return HS_E_OVERFLOW with *count set, the
 * arrays then holding an unspecified `capacity` of them.  Tiles are the
 * probe collection's upload-time plan (tune->probes_per_tile: fixed tiles);
 * tune->signature_bits and HS_TUNE_NO_SIGNATURE apply, the other knobs are
 * ignored.  The call synchronizes the context stream to read the count.
 * Every call on one device shares the scratch array: calls that overlap on
 * one device (from two contexts) race.  The context's last result
 * (hs_copy_last_result and friends) is left as it was. */
#define HS_NOALLOC_SCRATCH_BYTES (128ull << 20)
int hs_join_noalloc(hs_ctx* ctx, hs_coll* indexed, hs_coll* probe, const hs_range* indexed_range,
                    const hs_range* probe_range, const hs_tune* tune, uint32_t* r, uint32_t* s, uint32_t* overlap,
                    uint64_t capacity, uint64_t* count);

/* Device-to-device copy of `bytes` with a plain 16-byte-vector kernel,
 * repeated `reps` times on the context stream; *gbps receives the rate
 * (read + write bytes per second / 1e9).  The reference rate for HBM
 * roofline fractions on the machine at hand. */
int hs_copy_bandwidth(hs_ctx* ctx, uint64_t bytes, int32_t reps, double* gbps);

/* The integer-op rate of the XOR-popcount test on this device: lanes run
 * independent chains of a = popcount(a ^ m) + a (2 VALU operations a step)
 * over 16 workgroups of 256 lanes per CU, `reps` timed launches; *gops
 * receives operations per second / 1e9.  The reference rate for the bitmap
 * filter's roofline (it is bound by these operations, not by HBM). */
int hs_popcount_rate(hs_ctx* ctx, int32_t reps, double* gops);

#ifdef __cplusplus
}
#endif

#endif /* HYSET_AMD_H */
