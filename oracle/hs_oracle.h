/*
 * CPU oracle for the exact set-similarity join -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C, the reference's CPU join so that the
 * GPU path can be checked against it.  Only the test-suite, the smoke check
 * and the CPU-baseline leg of the benchmark load it; the product library
 * never links, loads or falls back to it.
 *
 * Followed reference code (paths relative to the reference repository):
 *   bounds      include/hyset/similarity.hpp:11-55 (double formulas, min clamps)
 *   index       include/hyset/index.hpp:46-125, :217-282 (lists sorted by
 *               (token, setID); list(t) = [offsets[t-1], offsets[t]))
 *   allpairs    include/hyset/host_algorithms.hpp:16-129
 *   ppjoin      include/hyset/host_algorithms.hpp:131-250, helpers.hpp:46-50
 *   verify      include/hyset/verification.hpp:9-31
 *   cpu_single  src/cpu_single.cu:80-144
 *   cpu_multi   src/cpu_multi.cu:118-207 (sqrt-spaced probe partitions,
 *               per-thread index over [0, end_i])
 *
 * Extensions that the reference does not have (no reference oracle exists
 * for them; they are pinned by brute force and closed forms in the tests):
 *   cosine and Dice bounds, exact-rational bounds, R-S (foreign) joins and
 *   the integer overlap of every result pair.
 */
#ifndef HS_ORACLE_H
#define HS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_JACCARD = 0, OR_COSINE = 1, OR_DICE = 2 };
enum { OR_BOUNDS_REFERENCE = 0, OR_BOUNDS_EXACT = 1 };
enum { OR_ALLPAIRS = 0, OR_PPJOIN = 1 };

typedef struct {
  uint64_t count;
  uint32_t* r; /* probe set id */
  uint32_t* s; /* indexed set id */
  uint32_t* overlap;
  /* filter statistics of the last run */
  uint64_t candidates;      /* distinct (r,s) candidates verified */
  uint64_t posting_visits;  /* posting entries visited inside length windows */
  uint64_t probe_tokens;    /* probe prefix tokens looked up */
  uint64_t index_entries;   /* entries in the inverted index */
  /* steady-clock intervals (ns) of the reference's cpu_single timers
     (src/cpu_single.cu:80-144): the index build, then the join proper
     (probing and verification, src/cpu_single.cu:115-138).  cpu_multi:
     index_ns is the slowest thread's index build, join_ns the wall time of
     the whole call. */
  uint64_t index_ns;
  uint64_t join_ns;
} or_pairs;

/* Bounds.  kind: OR_JACCARD / OR_COSINE / OR_DICE.  mode: OR_BOUNDS_REFERENCE
 * evaluates the reference's IEEE-double expressions (Jaccard only; cosine and
 * Dice always use the exact rational form).  Returns 0 on success. */
int or_bounds(int kind, int mode, double t, uint32_t l1, uint32_t l2, uint32_t* minsize,
              uint32_t* maxsize, uint32_t* minoverlap, uint32_t* maxprefix, uint32_t* midprefix);

/* The rational p/q the exact mode uses for t. */
int or_threshold_rational(double t, uint64_t* p, uint64_t* q);

/* Self-join over sets [first, last] of one collection (cpu_single).  The
 * index holds the midprefix entries of all sets in [first, last]; probes are
 * the same range; a pair is (r, s) with s < r. */
int or_self_join(const uint32_t* tokens, const uint32_t* starts, const uint32_t* sizes, uint32_t n,
                 uint32_t universe, int kind, int mode, double t, int algo, uint32_t first,
                 uint32_t last, or_pairs* out);

/* Block-pair self-join: index over sets [i_first, i_last], probes [p_first,
 * p_last] (the reference's join(block*, block*)).  s < r still applies. */
int or_self_join_blocks(const uint32_t* tokens, const uint32_t* starts, const uint32_t* sizes,
                        uint32_t n, uint32_t universe, int kind, int mode, double t, int algo,
                        uint32_t i_first, uint32_t i_last, uint32_t p_first, uint32_t p_last,
                        or_pairs* out);

/* R-S join: every probe of R against an index over all of S (S must be
 * size-sorted); no id restriction; the index prefix is maxprefix. */
int or_rs_join(const uint32_t* r_tokens, const uint32_t* r_starts, const uint32_t* r_sizes,
               uint32_t r_n, const uint32_t* s_tokens, const uint32_t* s_starts,
               const uint32_t* s_sizes, uint32_t s_n, uint32_t universe, int kind, int mode,
               double t, or_pairs* out);

/* Brute force over all pairs with an exact intersection (small inputs). */
int or_bruteforce_self(const uint32_t* tokens, const uint32_t* starts, const uint32_t* sizes,
                       uint32_t n, int kind, int mode, double t, or_pairs* out);
int or_bruteforce_rs(const uint32_t* r_tokens, const uint32_t* r_starts, const uint32_t* r_sizes,
                     uint32_t r_n, const uint32_t* s_tokens, const uint32_t* s_starts,
                     const uint32_t* s_sizes, uint32_t s_n, int kind, int mode, double t,
                     or_pairs* out);

/* cpu_multi: `threads` sqrt-spaced probe partitions run concurrently, each
 * with its own index over [0, end_i]. */
int or_cpu_multi(const uint32_t* tokens, const uint32_t* starts, const uint32_t* sizes, uint32_t n,
                 uint32_t universe, int kind, int mode, double t, int algo, int threads,
                 or_pairs* out);

/* Sort pairs by (r, s) in place. */
void or_pairs_sort(or_pairs* p);
void or_pairs_free(or_pairs* p);

/* Merge-path split (helpers.hpp:96-108): number of elements taken from a on
 * diagonal diag of the merge of a and b. */
uint32_t or_merge_path(const uint32_t* a, uint32_t a_n, const uint32_t* b, uint32_t b_n, uint32_t diag);

/* Size of the intersection of two ascending arrays. */
uint32_t or_intersect(const uint32_t* a, uint32_t a_n, const uint32_t* b, uint32_t b_n);

#ifdef __cplusplus
}
#endif
#endif
