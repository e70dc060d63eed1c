/*
 * CPU oracle for the exact set-similarity join -- TEST INFRASTRUCTURE ONLY.
 * Loaded by the tests, the smoke check and the CPU-baseline leg of the
 * benchmark, never by the product library.  The header lists the reference
 * functions each routine restates.
 */
#define _POSIX_C_SOURCE 199309L
#include "hs_oracle.h"

#include <math.h>
#include <time.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ bounds */

/* hyset::min / hyset::max operate on int (include/hyset/helpers.hpp:36-44). */
static int imin(int a, int b) { return (a > b) ? b : a; }

static uint64_t gcd_u64(uint64_t a, uint64_t b) {
  while (b) {
    uint64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

int or_threshold_rational(double t, uint64_t* p, uint64_t* q) {
  if (!(t > 0.0) || t > 1.0) return -1;
  uint64_t den = 1;
  for (int k = 0; k <= 9; ++k, den *= 10) {
    double num = t * (double)den;
    uint64_t pn = (uint64_t)llround(num);
    if ((double)pn / (double)den == t) {
      uint64_t g = gcd_u64(pn, den);
      *p = pn / g;
      *q = den / g;
      return 0;
    }
  }
  return -1;
}

typedef unsigned __int128 u128;

static uint64_t ceil_div128(u128 a, u128 b) { return (uint64_t)((a + b - 1) / b); }

/* The reference's Jaccard expressions in IEEE double
 * (include/hyset/similarity.hpp:40-50). */
static uint32_t ref_jac_minoverlap(uint32_t l1, uint32_t l2, double t) {
  return (uint32_t)(ceil((l1 + l2) * t / (1 + t)));
}
static uint32_t ref_jac_minsize(uint32_t l, double t) { return (uint32_t)(ceil(t * l)); }
static uint32_t ref_jac_maxsize(uint32_t l, double t) { return (uint32_t)((l / t)); }

/* Exact rational forms with t = p/q. */
static uint32_t ex_minoverlap(int kind, uint64_t p, uint64_t q, uint32_t l1, uint32_t l2) {
  switch (kind) {
    case OR_JACCARD: return (uint32_t)ceil_div128((u128)p * (l1 + (uint64_t)l2), (u128)(p + q));
    case OR_DICE: return (uint32_t)ceil_div128((u128)p * (l1 + (uint64_t)l2), (u128)(2 * q));
    default: {
      /* smallest o with o^2 q^2 >= p^2 l1 l2 */
      u128 rhs = (u128)p * p * l1 * l2;
      u128 qq = (u128)q * q;
      double approx = (double)p / (double)q * sqrt((double)l1 * (double)l2);
      uint64_t o = approx > 2.0 ? (uint64_t)approx - 2 : 0;
      while ((u128)o * o * qq < rhs) ++o;
      while (o > 0 && (u128)(o - 1) * (o - 1) * qq >= rhs) --o;
      return (uint32_t)o;
    }
  }
}
static uint32_t ex_minsize(int kind, uint64_t p, uint64_t q, uint32_t l) {
  switch (kind) {
    case OR_JACCARD: return (uint32_t)ceil_div128((u128)p * l, q);
    case OR_DICE: return (uint32_t)ceil_div128((u128)p * l, 2 * q - p);
    default: return (uint32_t)ceil_div128((u128)p * p * l, (u128)q * q);
  }
}
static uint32_t ex_maxsize(int kind, uint64_t p, uint64_t q, uint32_t l) {
  u128 v;
  switch (kind) {
    case OR_JACCARD: v = (u128)q * l / p; break;
    case OR_DICE: v = (u128)(2 * q - p) * l / p; break;
    default: v = (u128)q * q * l / ((u128)p * p); break;
  }
  return v > 0xFFFFFFFFu ? 0xFFFFFFFFu : (uint32_t)v;
}

typedef struct {
  int kind, exact;
  double t;
  uint64_t p, q;
} bounds_t;

static int bounds_init(bounds_t* b, int kind, int mode, double t) {
  if (kind < OR_JACCARD || kind > OR_DICE) return -1;
  if (!(t > 0.0) || t > 1.0) return -1;
  b->kind = kind;
  b->t = t;
  b->exact = (kind != OR_JACCARD) || (mode == OR_BOUNDS_EXACT);
  if (or_threshold_rational(t, &b->p, &b->q) != 0) {
    if (b->exact) return -1;
    b->p = b->q = 1;
  }
  return 0;
}

/* generic_similarity::minoverlap clamps with min(l1, l2) (similarity.hpp:21-23) */
static uint32_t b_minoverlap(const bounds_t* b, uint32_t l1, uint32_t l2) {
  uint32_t raw = b->exact ? ex_minoverlap(b->kind, b->p, b->q, l1, l2) : ref_jac_minoverlap(l1, l2, b->t);
  return (uint32_t)imin((int)l2, imin((int)l1, (int)raw));
}
static uint32_t b_minsize(const bounds_t* b, uint32_t l) {
  return b->exact ? ex_minsize(b->kind, b->p, b->q, l) : ref_jac_minsize(l, b->t);
}
static uint32_t b_maxsize(const bounds_t* b, uint32_t l) {
  return b->exact ? ex_maxsize(b->kind, b->p, b->q, l) : ref_jac_maxsize(l, b->t);
}
/* similarity.hpp:13-15 */
static uint32_t b_maxprefix(const bounds_t* b, uint32_t l) {
  return (uint32_t)imin((int)l, (int)(l - b_minsize(b, l) + 1));
}
/* similarity.hpp:17-19 */
static uint32_t b_midprefix(const bounds_t* b, uint32_t l) {
  return (uint32_t)imin((int)l, (int)(l - b_minoverlap(b, l, l) + 1));
}

int or_bounds(int kind, int mode, double t, uint32_t l1, uint32_t l2, uint32_t* minsize,
              uint32_t* maxsize, uint32_t* minoverlap, uint32_t* maxprefix, uint32_t* midprefix) {
  bounds_t b;
  if (bounds_init(&b, kind, mode, t) != 0) return -1;
  if (minsize) *minsize = b_minsize(&b, l1);
  if (maxsize) *maxsize = b_maxsize(&b, l1);
  if (minoverlap) *minoverlap = b_minoverlap(&b, l1, l2);
  if (maxprefix) *maxprefix = b_maxprefix(&b, l1);
  if (midprefix) *midprefix = b_midprefix(&b, l1);
  return 0;
}

/* ------------------------------------------------------------ pair output */

static int pairs_push(or_pairs* o, uint64_t* cap, uint32_t r, uint32_t s, uint32_t ov) {
  if (o->count == *cap) {
    uint64_t nc = *cap ? *cap * 2 : 1024;
    uint32_t* nr = (uint32_t*)realloc(o->r, nc * sizeof(uint32_t));
    if (!nr) return -1;
    o->r = nr;
    uint32_t* ns = (uint32_t*)realloc(o->s, nc * sizeof(uint32_t));
    if (!ns) return -1;
    o->s = ns;
    uint32_t* no = (uint32_t*)realloc(o->overlap, nc * sizeof(uint32_t));
    if (!no) return -1;
    o->overlap = no;
    *cap = nc;
  }
  o->r[o->count] = r;
  o->s[o->count] = s;
  o->overlap[o->count] = ov;
  o->count++;
  return 0;
}

void or_pairs_free(or_pairs* p) {
  if (!p) return;
  free(p->r);
  free(p->s);
  free(p->overlap);
  memset(p, 0, sizeof(*p));
}

static int cmp_idx(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : x > y ? 1 : 0;
}

void or_pairs_sort(or_pairs* p) {
  if (!p || p->count < 2) return;
  /* key = r<<32|s, payload = overlap; sort 128-bit records via two arrays */
  uint64_t n = p->count;
  uint64_t* keys = (uint64_t*)malloc(n * 2 * sizeof(uint64_t));
  if (!keys) return;
  for (uint64_t i = 0; i < n; ++i) {
    keys[2 * i] = ((uint64_t)p->r[i] << 32) | p->s[i];
    keys[2 * i + 1] = p->overlap[i];
  }
  qsort(keys, n, 2 * sizeof(uint64_t), cmp_idx);
  for (uint64_t i = 0; i < n; ++i) {
    p->r[i] = (uint32_t)(keys[2 * i] >> 32);
    p->s[i] = (uint32_t)keys[2 * i];
    p->overlap[i] = (uint32_t)keys[2 * i + 1];
  }
  free(keys);
}

/* ------------------------------------------------------------ primitives */

uint32_t or_intersect(const uint32_t* a, uint32_t a_n, const uint32_t* b, uint32_t b_n) {
  uint32_t i = 0, j = 0, c = 0;
  while (i < a_n && j < b_n) {
    if (a[i] == b[j]) {
      ++c;
      ++i;
      ++j;
    } else if (a[i] < b[j]) {
      ++i;
    } else {
      ++j;
    }
  }
  return c;
}

/* include/hyset/helpers.hpp:96-108 */
uint32_t or_merge_path(const uint32_t* a, uint32_t a_n, const uint32_t* b, uint32_t b_n, uint32_t diag) {
  int begin = (int)diag - (int)b_n;
  if (begin < 0) begin = 0;
  int end = imin((int)diag, (int)a_n);
  while (begin < end) {
    int mid = (begin + end) / 2;
    if (a[mid] < b[diag - 1 - mid]) begin = mid + 1;
    else end = mid;
  }
  return (uint32_t)begin;
}

/* include/hyset/verification.hpp:9-31 (early-terminating merge) */
static int verify_pair(const uint32_t* r1, uint32_t n1, const uint32_t* r2, uint32_t n2,
                       uint32_t thres, uint32_t pos1, uint32_t pos2, uint32_t found) {
  uint32_t max1 = n1 - pos1 + found;
  uint32_t max2 = n2 - pos2 + found;
  while (max1 >= thres && max2 >= thres && found < thres) {
    if (r1[pos1] == r2[pos2]) {
      ++pos1;
      ++pos2;
      ++found;
    } else if (r1[pos1] < r2[pos2]) {
      ++pos1;
      --max1;
    } else {
      ++pos2;
      --max2;
    }
  }
  return found >= thres;
}

/* ---------------------------------------------------------------- index */

/* Inverted index over the prefix entries of sets [first, last]; lists sorted
 * by (token, setID) as after thrust::sort with entry_comparator
 * (index.hpp:278-279).  A counting sort over set-id order gives that order. */
typedef struct {
  uint32_t universe;
  uint32_t* offsets; /* universe + 1 */
  uint32_t* set_ids;
  uint32_t* positions;
  uint32_t* prefix_len; /* indexed prefix length per set id (0 if not indexed) */
  uint64_t entries;
} index_t;

static void index_free(index_t* x) {
  free(x->offsets);
  free(x->set_ids);
  free(x->positions);
  free(x->prefix_len);
  memset(x, 0, sizeof(*x));
}

/* use_maxprefix: 0 -> midprefix (self-join, collection.hpp:128),
 *                1 -> maxprefix (R-S join). */
static int index_build(index_t* x, const uint32_t* tokens, const uint32_t* starts, const uint32_t* sizes,
                       uint32_t n, uint32_t universe, const bounds_t* b, uint32_t first, uint32_t last,
                       int use_maxprefix) {
  memset(x, 0, sizeof(*x));
  x->universe = universe;
  x->offsets = (uint32_t*)calloc((size_t)universe + 1, sizeof(uint32_t));
  x->prefix_len = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
  if (!x->offsets || !x->prefix_len) return -1;
  uint64_t total = 0;
  if (n > 0 && first <= last) {
    for (uint32_t i = first; i <= last; ++i) {
      uint32_t l = sizes[i];
      uint32_t pl = use_maxprefix ? b_maxprefix(b, l) : b_midprefix(b, l);
      x->prefix_len[i] = pl;
      for (uint32_t k = 0; k < pl; ++k) x->offsets[tokens[starts[i] + k] + 1]++;
      total += pl;
    }
  }
  for (uint32_t t = 0; t < universe; ++t) x->offsets[t + 1] += x->offsets[t];
  x->entries = total;
  x->set_ids = (uint32_t*)malloc((total ? total : 1) * sizeof(uint32_t));
  x->positions = (uint32_t*)malloc((total ? total : 1) * sizeof(uint32_t));
  uint32_t* cursor = (uint32_t*)malloc(((size_t)universe + 1) * sizeof(uint32_t));
  if (!x->set_ids || !x->positions || !cursor) {
    free(cursor);
    return -1;
  }
  memcpy(cursor, x->offsets, ((size_t)universe + 1) * sizeof(uint32_t));
  if (n > 0 && first <= last) {
    for (uint32_t i = first; i <= last; ++i) {
      for (uint32_t k = 0; k < x->prefix_len[i]; ++k) {
        uint32_t tk = tokens[starts[i] + k];
        uint32_t at = cursor[tk]++;
        x->set_ids[at] = i;
        x->positions[at] = k;
      }
    }
  }
  free(cursor);
  return 0;
}

/* ------------------------------------------------------------ the joins */

typedef struct {
  const uint32_t *tokens, *starts, *sizes;
  uint32_t n;
} coll_t;

/* One probe range against one index: host::allpairs (host_algorithms.hpp:16-129)
 * or host::ppjoin (:131-250).  self != 0 adds the s < r restriction. */
static int probe_range(const coll_t* R, const coll_t* S, const index_t* idx, const bounds_t* b,
                       int algo, int self, uint32_t p_first, uint32_t p_last, or_pairs* out,
                       uint64_t* cap) {
  uint32_t* count = (uint32_t*)calloc(S->n ? S->n : 1, sizeof(uint32_t));
  uint32_t* ppos = (uint32_t*)calloc(S->n ? S->n : 1, sizeof(uint32_t));
  uint32_t* ipos = (uint32_t*)calloc(S->n ? S->n : 1, sizeof(uint32_t));
  uint32_t* mo_of = (uint32_t*)calloc(S->n ? S->n : 1, sizeof(uint32_t));
  uint32_t* cands = (uint32_t*)malloc((S->n ? S->n : 1) * sizeof(uint32_t));
  int rc = 0;
  if (!count || !ppos || !ipos || !mo_of || !cands) {
    rc = -1;
    goto done;
  }
  if (R->n == 0 || p_first > p_last) goto done;
  for (uint32_t probe = p_first; probe <= p_last; ++probe) {
    const uint32_t* ptok = R->tokens + R->starts[probe];
    uint32_t psize = R->sizes[probe];
    uint32_t minsize = b_minsize(b, psize);
    uint32_t prefix = b_maxprefix(b, psize);
    uint32_t maxsize = b_maxsize(b, psize);
    uint32_t ncand = 0;
    out->probe_tokens += prefix;
    for (uint32_t pos = 0; pos < prefix; ++pos) {
      uint32_t token = ptok[pos];
      if (token >= idx->universe) continue;
      uint32_t it = idx->offsets[token], end = idx->offsets[token + 1];
      /* min-length filter: skip entries shorter than minsize */
      while (it != end && S->sizes[idx->set_ids[it]] < minsize) ++it;
      for (; it != end; ++it) {
        uint32_t sid = idx->set_ids[it];
        if (self && sid >= probe) break;
        uint32_t ssize = S->sizes[sid];
        if (ssize > maxsize) break;
        out->posting_visits++;
        if (algo == OR_PPJOIN) {
          uint32_t ip = idx->positions[it];
          if (count[sid] == 0) {
            uint32_t mo = b_minoverlap(b, psize, ssize);
            /* position_filter, helpers.hpp:46-50 */
            if (!(mo <= (uint32_t)imin((int)(psize - pos), (int)(ssize - ip)))) continue;
            cands[ncand++] = sid;
            mo_of[sid] = mo;
          }
          count[sid]++;
          ppos[sid] = pos;
          ipos[sid] = ip;
        } else {
          if (count[sid] == 0) cands[ncand++] = sid;
          count[sid]++;
        }
      }
    }
    out->candidates += ncand;
    for (uint32_t c = 0; c < ncand; ++c) {
      uint32_t sid = cands[c];
      const uint32_t* stok = S->tokens + S->starts[sid];
      uint32_t ssize = S->sizes[sid];
      uint32_t mo = (algo == OR_PPJOIN) ? mo_of[sid] : b_minoverlap(b, psize, ssize);
      uint32_t lastposind = idx->prefix_len[sid];
      uint32_t lastposprobe = prefix;
      uint32_t rlast = ptok[lastposprobe - 1];
      uint32_t slast = stok[lastposind - 1];
      uint32_t recpos, indrecpos;
      if (algo == OR_PPJOIN) {
        recpos = ppos[sid];
        indrecpos = ipos[sid];
        if (rlast > slast) {
          recpos += 1;
          indrecpos = lastposind;
        } else {
          recpos = lastposprobe;
          indrecpos += 1;
        }
      } else {
        if (rlast > slast) {
          recpos = count[sid];
          indrecpos = lastposind;
        } else {
          recpos = lastposprobe;
          indrecpos = count[sid];
        }
      }
      if (verify_pair(ptok, psize, stok, ssize, mo, recpos, indrecpos, count[sid])) {
        uint32_t ov = or_intersect(ptok, psize, stok, ssize);
        if (pairs_push(out, cap, probe, sid, ov) != 0) {
          rc = -1;
          goto done;
        }
      }
      count[sid] = 0;
    }
  }
done:
  free(count);
  free(ppos);
  free(ipos);
  free(mo_of);
  free(cands);
  return rc;
}

static int check_coll(const uint32_t* tokens, const uint32_t* starts, const uint32_t* sizes, uint32_t n) {
  if (n && (!tokens || !starts || !sizes)) return -1;
  for (uint32_t i = 0; i < n; ++i)
    if (sizes[i] == 0) return -1;
  return 0;
}

static uint64_t now_ns(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

int or_self_join_blocks(const uint32_t* tokens, const uint32_t* starts, const uint32_t* sizes,
                        uint32_t n, uint32_t universe, int kind, int mode, double t, int algo,
                        uint32_t i_first, uint32_t i_last, uint32_t p_first, uint32_t p_last,
                        or_pairs* out) {
  bounds_t b;
  if (!out || bounds_init(&b, kind, mode, t) != 0 || check_coll(tokens, starts, sizes, n)) return -1;
  memset(out, 0, sizeof(*out));
  if (n == 0) return 0;
  if (i_last >= n) i_last = n - 1;
  if (p_last >= n) p_last = n - 1;
  coll_t C = {tokens, starts, sizes, n};
  index_t idx;
  const uint64_t t0 = now_ns();
  if (index_build(&idx, tokens, starts, sizes, n, universe, &b, i_first, i_last, 0) != 0) {
    index_free(&idx);
    return -1;
  }
  const uint64_t t1 = now_ns();
  out->index_entries = idx.entries;
  uint64_t cap = 0;
  int rc = probe_range(&C, &C, &idx, &b, algo, 1, p_first, p_last, out, &cap);
  out->index_ns = t1 - t0;
  out->join_ns = now_ns() - t1;
  index_free(&idx);
  return rc;
}

int or_self_join(const uint32_t* tokens, const uint32_t* starts, const uint32_t* sizes, uint32_t n,
                 uint32_t universe, int kind, int mode, double t, int algo, uint32_t first,
                 uint32_t last, or_pairs* out) {
  return or_self_join_blocks(tokens, starts, sizes, n, universe, kind, mode, t, algo, first, last,
                             first, last, out);
}

int or_rs_join(const uint32_t* r_tokens, const uint32_t* r_starts, const uint32_t* r_sizes,
               uint32_t r_n, const uint32_t* s_tokens, const uint32_t* s_starts,
               const uint32_t* s_sizes, uint32_t s_n, uint32_t universe, int kind, int mode,
               double t, or_pairs* out) {
  bounds_t b;
  if (!out || bounds_init(&b, kind, mode, t) != 0) return -1;
  if (check_coll(r_tokens, r_starts, r_sizes, r_n) || check_coll(s_tokens, s_starts, s_sizes, s_n)) return -1;
  memset(out, 0, sizeof(*out));
  if (r_n == 0 || s_n == 0) return 0;
  coll_t R = {r_tokens, r_starts, r_sizes, r_n};
  coll_t S = {s_tokens, s_starts, s_sizes, s_n};
  index_t idx;
  if (index_build(&idx, s_tokens, s_starts, s_sizes, s_n, universe, &b, 0, s_n - 1, 1) != 0) {
    index_free(&idx);
    return -1;
  }
  out->index_entries = idx.entries;
  uint64_t cap = 0;
  int rc = probe_range(&R, &S, &idx, &b, OR_ALLPAIRS, 0, 0, r_n - 1, out, &cap);
  index_free(&idx);
  return rc;
}

static int pair_passes(const bounds_t* b, uint32_t lr, uint32_t ls, uint32_t ov) {
  if (ls < b_minsize(b, lr) || ls > b_maxsize(b, lr)) return 0;
  return ov >= b_minoverlap(b, lr, ls);
}

int or_bruteforce_self(const uint32_t* tokens, const uint32_t* starts, const uint32_t* sizes,
                       uint32_t n, int kind, int mode, double t, or_pairs* out) {
  bounds_t b;
  if (!out || bounds_init(&b, kind, mode, t) != 0 || check_coll(tokens, starts, sizes, n)) return -1;
  memset(out, 0, sizeof(*out));
  uint64_t cap = 0;
  for (uint32_t r = 0; r < n; ++r) {
    for (uint32_t s = 0; s < r; ++s) {
      uint32_t ov = or_intersect(tokens + starts[r], sizes[r], tokens + starts[s], sizes[s]);
      if (pair_passes(&b, sizes[r], sizes[s], ov))
        if (pairs_push(out, &cap, r, s, ov) != 0) return -1;
    }
  }
  return 0;
}

int or_bruteforce_rs(const uint32_t* r_tokens, const uint32_t* r_starts, const uint32_t* r_sizes,
                     uint32_t r_n, const uint32_t* s_tokens, const uint32_t* s_starts,
                     const uint32_t* s_sizes, uint32_t s_n, int kind, int mode, double t,
                     or_pairs* out) {
  bounds_t b;
  if (!out || bounds_init(&b, kind, mode, t) != 0) return -1;
  if (check_coll(r_tokens, r_starts, r_sizes, r_n) || check_coll(s_tokens, s_starts, s_sizes, s_n)) return -1;
  memset(out, 0, sizeof(*out));
  uint64_t cap = 0;
  for (uint32_t r = 0; r < r_n; ++r) {
    for (uint32_t s = 0; s < s_n; ++s) {
      uint32_t ov = or_intersect(r_tokens + r_starts[r], r_sizes[r], s_tokens + s_starts[s], s_sizes[s]);
      if (pair_passes(&b, r_sizes[r], s_sizes[s], ov))
        if (pairs_push(out, &cap, r, s, ov) != 0) return -1;
    }
  }
  return 0;
}

/* --------------------------------------------------------------- cpu_multi */

typedef struct {
  const uint32_t *tokens, *starts, *sizes;
  uint32_t n, universe;
  const bounds_t* b;
  int algo;
  uint32_t first, last;
  or_pairs out;
  int rc;
} worker_t;

static void* worker_run(void* arg) {
  worker_t* w = (worker_t*)arg;
  coll_t C = {w->tokens, w->starts, w->sizes, w->n};
  index_t idx;
  memset(&w->out, 0, sizeof(w->out));
  /* per-thread index over [0, end_i] (cpu_multi.cu:143-158) */
  const uint64_t t0 = now_ns();
  w->rc = index_build(&idx, w->tokens, w->starts, w->sizes, w->n, w->universe, w->b, 0, w->last, 0);
  w->out.index_ns = now_ns() - t0;
  if (w->rc == 0) {
    uint64_t cap = 0;
    w->out.index_entries = idx.entries;
    w->rc = probe_range(&C, &C, &idx, w->b, w->algo, 1, w->first, w->last, &w->out, &cap);
  }
  index_free(&idx);
  return NULL;
}

int or_cpu_multi(const uint32_t* tokens, const uint32_t* starts, const uint32_t* sizes, uint32_t n,
                 uint32_t universe, int kind, int mode, double t, int algo, int threads,
                 or_pairs* out) {
  bounds_t b;
  if (!out || threads < 1 || bounds_init(&b, kind, mode, t) != 0 || check_coll(tokens, starts, sizes, n))
    return -1;
  memset(out, 0, sizeof(*out));
  if (n == 0) return 0;
  /* x_i = k * sqrt((i + 1) / T), truncated (cpu_multi.cu:118-125) */
  uint32_t* points = (uint32_t*)malloc(sizeof(uint32_t) * (threads + 1));
  worker_t* ws = (worker_t*)calloc(threads, sizeof(worker_t));
  pthread_t* th = (pthread_t*)calloc(threads, sizeof(pthread_t));
  int* started = (int*)calloc(threads, sizeof(int));
  int rc = 0;
  if (!points || !ws || !th || !started) {
    rc = -1;
    goto done;
  }
  points[0] = 0;
  const uint64_t t_start = now_ns();
  for (int i = 0; i < threads; ++i) points[i + 1] = (uint32_t)(n * sqrt((double)(i + 1) / (double)threads));
  for (int i = 0; i < threads; ++i) {
    if (points[i + 1] <= points[i]) continue;
    worker_t* w = &ws[i];
    w->tokens = tokens;
    w->starts = starts;
    w->sizes = sizes;
    w->n = n;
    w->universe = universe;
    w->b = &b;
    w->algo = algo;
    w->first = points[i];
    w->last = points[i + 1] - 1;
    if (pthread_create(&th[i], NULL, worker_run, w) == 0) started[i] = 1;
    else worker_run(w);
  }
  uint64_t cap = 0;
  for (int i = 0; i < threads; ++i) {
    if (started[i]) pthread_join(th[i], NULL);
    if (points[i + 1] <= points[i]) continue;
    worker_t* w = &ws[i];
    if (w->rc != 0) rc = -1;
    for (uint64_t k = 0; rc == 0 && k < w->out.count; ++k)
      if (pairs_push(out, &cap, w->out.r[k], w->out.s[k], w->out.overlap[k]) != 0) rc = -1;
    out->candidates += w->out.candidates;
    out->posting_visits += w->out.posting_visits;
    out->probe_tokens += w->out.probe_tokens;
    out->index_entries += w->out.index_entries;
    if (w->out.index_ns > out->index_ns) out->index_ns = w->out.index_ns;
    or_pairs_free(&w->out);
  }
  out->join_ns = now_ns() - t_start;
done:
  free(points);
  free(ws);
  free(th);
  free(started);
  return rc;
}
