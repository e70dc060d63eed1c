// Prefix-filter candidate generation fused with verification (the fgss
// technique) for gfx950.
//
// Reference: include/hyset/device_algorithms.hpp:265-303 (prefix_filter:
// one block per probe, global atomicAdd of shared-prefix counts into a dense
// blockSize^2 matrix) followed by verification.hpp:60-138 (scan of the whole
// matrix, sequential merge per cell).  The result set is the reference's:
// pairs (r, s) inside the length window of r that share a probe-prefix token
// of r with the index prefix of s and reach minoverlap(|r|, |s|).
//
// Design here:
//   * a workgroup of 512 lanes owns a tile of consecutive probes (at most
//     512 probes and 8192 tokens, planned once per collection); their
//     tokens are one contiguous CSR range and are staged in LDS with
//     coalesced loads.  A tile's work is a chain of dependent loads and
//     barriers whose length hardly depends on the tile's size, so the
//     kernel's time follows the number of tiles: 512-probe tiles in two
//     workgroups per CU ran 8-18% faster than 256-probe tiles in four (the
//     same 16 waves per CU), 2048-token tiles 1.5x slower;
//   * the length window of a probe is an id interval (sets are sorted by
//     size), so each (probe, prefix token) "item" binary-searches its slice
//     of the id-sorted posting list -- no length checks per posting;
//   * items, then the posting entries of each batch of items ("visits"),
//     are flattened with workgroup scans so every lane gets one visit;
//   * a (r, s) pair reached through several prefix tokens is verified once:
//     only the visit whose token is the first common token of r and s
//     continues (that token always lies in both prefixes), which replaces
//     the reference's dense count matrix by a merge that starts anyway;
//   * PPJoin's positional filter (host_algorithms.hpp:196-199) runs first,
//     from the token positions stored in the index;
//   * then the bitmap filter of the reference's bitmap technique
//     (device_algorithms.hpp:76-118): |r n s| <= (|r| + |s| -
//     popcount(sig_r ^ sig_s)) / 2 for XOR-parity signatures of 64-1024
//     bits, so a visit with |r| + |s| - popcount < 2 minoverlap cannot be a
//     result.  Both filters only drop visits that cannot pass, so the result
//     set is unchanged.  64-bit signatures of indexed sets are stored in the
//     index entries and those of the tile's probes in LDS, so the test reads
//     nothing beyond the entry; wider ones are read from the per-set table;
//   * visits that pass both are pushed onto an LDS stack (one LDS atomic per
//     wave) and verified 512 at a time, one per lane, so the merges run with
//     full waves instead of in the few lanes whose visit passed;
//   * the merge is driven by s, whose tokens arrive 12 at a time in
//     registers, against r in LDS; it keeps the early-termination bound of
//     verify_pair (verification.hpp:9-31) for failing pairs and runs to the
//     end for passing ones, which yields the exact overlap;
//   * passing pairs leave through one atomic per wave.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "hs_internal.h"

namespace hs {

constexpr int FG_TPB = 512;         // lanes per workgroup of the usual launch, also max probes per tile
constexpr int FG_TPB_NARROW = 256;  // lanes of the narrow launch (balanced tiles over very long lists)
static_assert(TILE_PROBES == FG_TPB && TILE_TOKENS == 16u * FG_TPB, "one probe per lane in phase A");

// Geometry of a launch of TPB lanes: one probe and 16 staged tokens per
// lane, 3 items per lane and batch, a candidate stack of 2 TPB (fewer than
// TPB entries wait between rounds and a round of TPB visits adds at most
// TPB).
template <int TPB>
struct FgGeom {
  static constexpr uint32_t TOKCAP = 16u * TPB;
  static constexpr uint32_t ITEMCAP = 3u * TPB;
  static constexpr uint32_t QCAP = 2u * TPB;
};
constexpr uint32_t NO_LDS = 0xFFFFFFFFu;
constexpr uint32_t FG_MOCAP = 1024;   // minoverlap entries kept in LDS (|r| + |s| < 1024)

constexpr uint32_t SREG = 12;  // tokens of s held in registers per refill

// Verify (r, s) from the visit reached through r[j]: an s-driven merge.  The
// tokens of s are loaded SREG at a time into registers with independent
// loads (one memory latency per chunk instead of one per merge step); r is
// read from LDS (kRInLds) or global memory.  The first common token must be
// r[j] -- otherwise another visit of the same pair owns it and this one
// stops.  The bound found + min(|r| - i, |s| - q) >= minoverlap of
// verify_pair (verification.hpp:9-31) ends failing pairs early; passing
// pairs run to the end, which gives the exact overlap.
template <bool kRInLds>
__device__ __forceinline__ bool verify_visit(const uint32_t* __restrict__ r, uint32_t lr,
                                             const uint32_t* __restrict__ s, uint32_t ls, uint32_t j, uint32_t mo,
                                             uint32_t* ov) {
  uint32_t i = 0, found = 0;
  for (uint32_t base = 0; base < ls; base += SREG) {
    uint32_t sv[SREG];
#pragma unroll
    for (uint32_t q = 0; q < SREG; ++q) sv[q] = (base + q < ls) ? s[base + q] : 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t q = 0; q < SREG; ++q) {
      const uint32_t pos = base + q;
      if (pos >= ls) break;
      if (found + min(lr - i, ls - pos) < mo) return false;
      const uint32_t x = sv[q];
      while (i < lr && r[i] < x) ++i;
      if (i >= lr) {
        *ov = found;
        return found >= mo;
      }
      if (r[i] == x) {
        if (found == 0 && i != j) return false;  // not the first-common visit
        ++found;
        ++i;
      }
    }
  }
  *ov = found;
  return found >= mo;
}

// Workgroup-shared state of one tile.
template <int TPB>
struct FgssLds {
  uint32_t tok[FgGeom<TPB>::TOKCAP];
  uint32_t pstart[TPB];
  uint32_t psize[TPB];
  uint32_t plds[TPB];
  uint32_t lo[TPB];
  uint32_t hi[TPB];
  uint32_t item_off[TPB + 1];
  uint32_t it_a[FgGeom<TPB>::ITEMCAP];
  uint32_t it_pj[FgGeom<TPB>::ITEMCAP];
  uint32_t vis_off[FgGeom<TPB>::ITEMCAP + 1];
  uint32_t tmp[TPB / WAVE + 1];
  unsigned long long stats[5];
  uint32_t next;
  uint32_t nxt[2];                  // claimed tiles, current and next (persistent single join)
  uint32_t slot0;                   // first item slot of the tile (item-slice pass)
  uint32_t mo_lds;                  // 1: mo[] holds the minoverlap table (keyed by |r| + |s|)
  uint16_t mo[FG_MOCAP];
  // candidate stack: probe slot | position in r << 9, indexed set id; the
  // entries of round k are counted in qadd[k % 3] (cleared two rounds ahead)
  uint32_t q_pj[FgGeom<TPB>::QCAP];
  uint32_t q_sid[FgGeom<TPB>::QCAP];
  uint32_t qadd[3];
  uint64_t psig[TPB];               // probe signatures (64-bit signature tests)
  uint8_t ptidx[TPB];               // per-probe threshold table (serving)
  DevBounds bounds[MAX_THRESHOLDS];  // the tables of a multi-threshold launch
};

// The probes of a tile and the part of the indexed collection they join.
struct TileScope {
  uint32_t p0, p1;                // probes [p0, p1)
  uint32_t idx_first, idx_last;   // indexed ids, inclusive
  const uint32_t* size_first;     // first indexed id of size >= l
  uint32_t size_first_len;
};

// Window slices of K posting lists at once: for each k, a = lower_bound of
// k_lo[k] and b = lower_bound of k_hi[k] >= k_lo[k] in the sorted ids
// ids[f_lo[k], f_lo[k] + n_lo[k]).  Both bounds are searched together, and
// a round reads whole aligned quads of ids (one 16-byte load each): every
// id x of the list that a round reads tightens both answers at once
// (ids[x] < key puts the answer above x, otherwise at or below x).  A round
// gives an item WQ quads: all the quads of the interval still open when
// they fit, else evenly spaced pivot quads, split between the two bounds
// once their intervals part.  Lists of up to 13 ids resolve in one round
// and every round shrinks each open interval, so an interval of n ids takes
// about log_{WQ + 1}(n / 4) rounds instead of the 2 log_5(n) of a
// 4-ary search per bound over single ids.  All K searches advance in
// lockstep, so their loads overlap.  The quads may read up to 3 ids past
// the end of the last list: the id array carries IDS_PAD ids of padding.
// On return f_lo[k] = start of the slice and n_hi[k] = b - a.
constexpr uint32_t NO_QUAD = 0xFFFFFFFFu;
// quads per item and round in k_fgss, whose tile state already holds most
// of the registers
constexpr uint32_t FG_SEARCH_QUADS = 2;

// Index record e: set id, meta and (64-bit signature tests) the signature,
// in one load.
template <int kW>
__device__ __forceinline__ void load_entry(const DevIndex& X, uint32_t e, uint32_t& sid, uint32_t& meta,
                                           uint64_t& sig) {
  if constexpr (kW == 1) {
    const uint4 v = reinterpret_cast<const uint4*>(X.ent)[e];
    sid = v.x;
    meta = v.y;
    sig = ((uint64_t)v.w << 32) | v.z;
  } else {
    const uint2 v = *reinterpret_cast<const uint2*>(X.ent + (size_t)e * X.ent_words);
    sid = v.x;
    meta = v.y;
  }
}

// Plans m quads over the open interval [lo, hi) of a list that starts at
// global position f: every quad of the interval when at most m of them hold
// it, else m pivot quads, the i-th at position lo + (i + 1) (hi - lo) / (m + 1).
template <uint32_t M>
__device__ __forceinline__ void plan_quads(uint32_t f, uint32_t lo, uint32_t hi, uint32_t (&q)[M], uint32_t o,
                                           uint32_t m) {
  const uint32_t first = (f + lo) >> 2, last = (f + hi - 1) >> 2;
  const uint32_t span = last - first + 1;
#pragma unroll
  for (uint32_t i = 0; i < M; ++i) {
    if (i < o || i >= o + m) continue;
    const uint32_t r = i - o;
    if (span <= m) q[i] = r < span ? first + r : NO_QUAD;
    else q[i] = (f + lo + (uint32_t)((uint64_t)(r + 1) * (hi - lo) / (m + 1))) >> 2;
  }
}

template <uint32_t K, uint32_t WQ>
__device__ __forceinline__ void lockstep_bounds(const uint32_t* __restrict__ ids, uint32_t (&f_lo)[K],
                                                const uint32_t (&n_lo)[K], uint32_t (&n_hi)[K],
                                                const uint32_t (&k_lo)[K], const uint32_t (&k_hi)[K]) {
  // a lies in [alo, ahi], b in [blo, bhi] (counts of ids below the key)
  uint32_t alo[K], ahi[K], blo[K], bhi[K];
#pragma unroll
  for (uint32_t k = 0; k < K; ++k) {
    alo[k] = blo[k] = 0;
    ahi[k] = bhi[k] = n_lo[k];
  }
  for (;;) {
    bool any = false;
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) any |= alo[k] < ahi[k] || blo[k] < bhi[k];
    if (!any) break;
    uint32_t q[K][WQ];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
#pragma unroll
      for (uint32_t i = 0; i < WQ; ++i) q[k][i] = NO_QUAD;
      const bool oa = alo[k] < ahi[k], ob = blo[k] < bhi[k];
      const uint32_t f = f_lo[k];
      if (oa && ob) {
        const uint32_t span = ((f + bhi[k] - 1) >> 2) - ((f + alo[k]) >> 2) + 1;
        if (span <= WQ || (alo[k] == blo[k] && ahi[k] == bhi[k])) {
          plan_quads(f, alo[k], bhi[k], q[k], 0, WQ);  // one interval for both bounds
        } else {
          plan_quads(f, alo[k], ahi[k], q[k], 0, WQ / 2);
          plan_quads(f, blo[k], bhi[k], q[k], WQ / 2, WQ - WQ / 2);
        }
      } else if (oa) {
        plan_quads(f, alo[k], ahi[k], q[k], 0, WQ);
      } else if (ob) {
        plan_quads(f, blo[k], bhi[k], q[k], 0, WQ);
      }
    }
    uint4 v[K][WQ];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k)
#pragma unroll
      for (uint32_t i = 0; i < WQ; ++i)
        v[k][i] = q[k][i] != NO_QUAD ? reinterpret_cast<const uint4*>(ids)[q[k][i]] : make_uint4(0, 0, 0, 0);
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
#pragma unroll
      for (uint32_t i = 0; i < WQ; ++i) {
        if (q[k][i] == NO_QUAD) continue;
        const uint32_t e[4] = {v[k][i].x, v[k][i].y, v[k][i].z, v[k][i].w};
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) {
          const uint32_t x = 4u * q[k][i] + c - f_lo[k];  // list position (wraps below the list)
          if (x >= n_lo[k]) continue;
          if (e[c] < k_lo[k]) alo[k] = max(alo[k], x + 1);
          else ahi[k] = min(ahi[k], x);
          if (e[c] < k_hi[k]) blo[k] = max(blo[k], x + 1);
          else bhi[k] = min(bhi[k], x);
        }
      }
      // a <= b
      blo[k] = max(blo[k], alo[k]);
      ahi[k] = min(ahi[k], bhi[k]);
    }
  }
#pragma unroll
  for (uint32_t k = 0; k < K; ++k) {
    f_lo[k] += alo[k];
    n_hi[k] = bhi[k] > alo[k] ? bhi[k] - alo[k] : 0u;
  }
}

// Indexed ids [lo, hi) that probe p of size l joins: its length window
// (an id interval, the indexed sets being sorted by size) cut to the
// indexed range, and to ids below p in a self-join.
__device__ __forceinline__ void probe_window(const DevBounds& B, uint32_t l, uint32_t p, const uint32_t* size_first,
                                             uint32_t size_first_len, uint32_t idx_first, uint32_t idx_last,
                                             int self, uint32_t& lo, uint32_t& hi) {
  const uint32_t last = size_first_len - 1;
  const uint32_t mins = min(B.minsize[l], last);
  const uint32_t maxs1 = min(B.maxsize[l], last - 1) + 1;
  lo = max(size_first[mins], idx_first);
  hi = min(size_first[maxs1], idx_last + 1);
  if (self) hi = min(hi, p);
}

// Work counters of one workgroup's lanes.
struct FgssCounts {
  uint32_t cands = 0, cand_tokens = 0, sig_tests = 0;
};

template <int TPB, bool kMulti>
__device__ __forceinline__ uint32_t tile_minoverlap(const FgssArgs& A, const FgssLds<TPB>& L, uint32_t pl, uint32_t lr,
                                                    uint32_t ls) {
  if (!kMulti && L.mo_lds) return min((uint32_t)L.mo[lr + ls], min(lr, ls));
  return dev_minoverlap(kMulti ? L.bounds[L.ptidx[pl]] : A.bounds, lr, ls);
}

// Verifies one candidate of the stack (valid lanes) and appends the passing
// pairs; called by every lane of the wave.  pj = probe slot | j << 9, where
// r[j] is the token through which the candidate was reached.
template <int TPB, bool kMulti>
__device__ __forceinline__ void verify_candidate(const FgssArgs& A, const FgssLds<TPB>& L, uint32_t p0, bool valid,
                                                 uint32_t pj, uint32_t sid, FgssCounts& C) {
  bool pass = false;
  uint32_t ov = 0, rid = 0;
  if (valid) {
    const uint32_t pl = pj & 0x1FFu;
    const uint32_t j = pj >> 9;
    rid = p0 + pl;
    const uint32_t sst = A.indexed.starts[sid];
    const uint32_t ls = A.indexed.sizes[sid];
    const uint32_t lr = L.psize[pl];
    const uint32_t mo = tile_minoverlap<TPB, kMulti>(A, L, pl, lr, ls);
    const uint32_t* sp = A.indexed.tokens + sst;
    const uint32_t off = L.plds[pl];
    if (off != NO_LDS) pass = verify_visit<true>(L.tok + off, lr, sp, ls, j, mo, &ov);
    else pass = verify_visit<false>(A.probe.tokens + L.pstart[pl], lr, sp, ls, j, mo, &ov);
    ++C.cands;
    C.cand_tokens += ls;
  }
  emit_pairs(pass, rid, sid, ov, A.out.r, A.out.s, A.out.ov, &A.out.counters[CNT_PAIRS], A.out.capacity);
}

// kW: signature words (0: no signature test); kSlices: item slices come
// from the item-slice pass; kMulti: per-probe threshold tables (serving).
// rnd counts the workgroup's visit rounds (it selects the stack counter).
template <int TPB, int kW, bool kSlices, bool kMulti>
__device__ __forceinline__ void fgss_tile(const FgssArgs& A, const TileScope& T, FgssLds<TPB>& L, FgssCounts& C,
                                          uint32_t& rnd) {
  const uint32_t tid = threadIdx.x;
  const uint32_t p0 = T.p0;
  const uint32_t np = T.p1 - T.p0;

  // ---- phase A: probe metadata, windows, LDS staging of probe tokens
  const uint32_t base = A.probe.starts[p0];
  const uint32_t tok_end = A.probe.starts[T.p1 - 1] + A.probe.sizes[T.p1 - 1];
  const uint32_t ntok = tok_end - base;
  uint32_t n_items = 0, n_active = 0;
  if (tid < np) {
    const uint32_t p = p0 + tid;
    uint32_t l, st, lo, hi, mp;
    if constexpr (kSlices) {
      // one record per probe from the planning pass: start, size | maxprefix << 16, window
      const uint4 rec = A.probe_rec[p - A.probe_first];
      st = rec.x;
      l = rec.y & 0xFFFFu;
      mp = rec.y >> 16;
      lo = rec.z;
      hi = rec.w;
      if (tid == 0) L.slot0 = A.probe_slot[p - A.probe_first];
    } else {
      l = A.probe.sizes[p];
      st = A.probe.starts[p];
      const uint32_t ti = kMulti ? A.probe_tidx[p] : 0u;
      const DevBounds& B = kMulti ? L.bounds[ti] : A.bounds;
      L.ptidx[tid] = (uint8_t)ti;
      probe_window(B, l, p, T.size_first, T.size_first_len, T.idx_first, T.idx_last, A.self, lo, hi);
      mp = B.maxprefix[l];
    }
    L.pstart[tid] = st;
    L.psize[tid] = l;
    L.lo[tid] = lo;
    L.hi[tid] = hi;
    if constexpr (kW == 1) L.psig[tid] = A.probe_sigs[p];
    const uint32_t off = st - base;
    L.plds[tid] = (off + l <= FgGeom<TPB>::TOKCAP) ? off : NO_LDS;
    n_active = (lo < hi) ? mp : 0;
    // the slice pass lays out maxprefix(l) items for every probe, in probe order
    n_items = kSlices ? mp : n_active;
  }
  {
    // every load of the staging is issued before the first LDS store
    constexpr uint32_t R = FgGeom<TPB>::TOKCAP / TPB;
    const uint32_t m = min(ntok, FgGeom<TPB>::TOKCAP);
    uint32_t tv[R];
#pragma unroll
    for (uint32_t k = 0; k < R; ++k) tv[k] = tid + k * TPB < m ? A.probe.tokens[base + tid + k * TPB] : 0u;
#pragma unroll
    for (uint32_t k = 0; k < R; ++k)
      if (tid + k * TPB < m) L.tok[tid + k * TPB] = tv[k];
  }
  n_active = wave_sum(n_active);
  if (lane_id() == 0 && n_active) atomicAdd(&L.stats[3], (unsigned long long)n_active);
  uint32_t total_items;
  uint32_t ioff = block_exclusive_scan<TPB>(n_items, L.tmp, &total_items);
  if (tid < np) L.item_off[tid] = ioff;
  if (tid == 0) L.item_off[np] = total_items;
  __syncthreads();

  // ---- phase B: batches of items
  uint32_t qn = 0;  // entries on the candidate stack (the same in every lane)
  for (uint32_t ib = 0; ib < total_items; ib += FgGeom<TPB>::ITEMCAP) {
    const uint32_t nb = min(FgGeom<TPB>::ITEMCAP, total_items - ib);
    // slice of the posting list for every item of the batch: read from the
    // item-slice pass when there is one, otherwise searched here, PER items
    // per thread in lockstep
    constexpr uint32_t PER = FgGeom<TPB>::ITEMCAP / TPB;
    if constexpr (kSlices) {
#pragma unroll
      for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t it = tid + k * TPB;
        uint32_t len = 0;
        if (it < nb) {
          const uint32_t item = ib + it;
          const uint32_t pl = segment_of(L.item_off, np, item);
          const uint32_t j = item - L.item_off[pl];
          L.it_pj[it] = pl | (j << 9);
          L.it_a[it] = A.slice_a[L.slot0 + item];
          len = A.slice_n[L.slot0 + item];
        }
        L.vis_off[it] = len;
      }
    } else {
      uint32_t f_lo[PER], n_lo[PER], n_hi[PER], k_lo[PER], k_hi[PER];
#pragma unroll
      for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t it = tid + k * TPB;
        f_lo[k] = n_lo[k] = n_hi[k] = k_lo[k] = k_hi[k] = 0;
        if (it < nb) {
          const uint32_t item = ib + it;
          const uint32_t pl = segment_of(L.item_off, np, item);
          const uint32_t j = item - L.item_off[pl];
          const uint32_t* r = (L.plds[pl] != NO_LDS) ? (L.tok + L.plds[pl]) : (A.probe.tokens + L.pstart[pl]);
          const uint32_t token = r[j];
          L.it_pj[it] = pl | (j << 9);
          k_lo[k] = L.lo[pl];
          k_hi[k] = L.hi[pl];
          if (token < A.index.universe) {
            const uint32_t lb = A.index.offsets[token];
            const uint32_t le = A.index.offsets[token + 1];
            f_lo[k] = lb;
            n_lo[k] = n_hi[k] = le - lb;
          }
        }
      }
      // on the compact ids: a list is searched over more entries than its
      // window visits, so 4-byte ids touch fewer lines than the records
      lockstep_bounds<PER, FG_SEARCH_QUADS>(A.index.set_ids, f_lo, n_lo, n_hi, k_lo, k_hi);
#pragma unroll
      for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t it = tid + k * TPB;
        L.it_a[it] = f_lo[k];
        L.vis_off[it] = (it < nb) ? n_hi[k] : 0u;
      }
    }
    __syncthreads();
    // exclusive scan of the slice lengths (PER per thread)
    uint32_t v[PER];
    uint32_t local = 0;
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
      v[k] = L.vis_off[tid * PER + k];
      local += v[k];
    }
    uint32_t total_vis;
    uint32_t run = block_exclusive_scan<TPB>(local, L.tmp, &total_vis);
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
      L.vis_off[tid * PER + k] = run;
      run += v[k];
    }
    if (tid == 0) {
      L.vis_off[FgGeom<TPB>::ITEMCAP] = total_vis;
      L.stats[0] += total_vis;
    }
    __syncthreads();

    // one visit per lane and sub-round; the index entry of the lane's next
    // visit is fetched before this one is filtered, so its latency overlaps
    // the filter's loads.  Visits passing both filters go onto the LDS
    // stack; whenever it holds TPB they are verified, one per lane.  The
    // last batch of the tile ends with one sub-round without visits that
    // verifies the rest of the stack (one verification site keeps the
    // register budget).
    const bool last_batch = ib + FgGeom<TPB>::ITEMCAP >= total_items;
    // sub-rounds of TPB visits: a lane takes two consecutive visits, the
    // even one in sub-round 2k and the odd one in 2k + 1, so one search
    // finds the items of both (the odd visit's item is the even one's or a
    // later one a step or two on); the last batch adds a sub-round without
    // visits that drains the stack
    const uint32_t n_sub = 2u * ((total_vis + 2u * TPB - 1u) / (2u * TPB)) + (last_batch ? 1u : 0u);
    // (64-bit signatures ride in the index entry and arrive with it).  After
    // its odd visit a lane's next visit lies 2 TPB - 1 further on: with long
    // slices (long sets) it is in the same item or a near one, found by
    // galloping from the current item; with short slices a bisection over
    // the batch's items is cheaper.  The search was a fifth of the kernel's
    // time before pairs halved the number of searches.
    const bool gallop = total_vis >= 64u * nb;
    uint32_t n_it = 0, n_sid = 0, n_meta = 0;
    uint64_t n_sig = 0;
    if (2u * tid < total_vis) {
      n_it = segment_of(L.vis_off, nb, 2u * tid);
      load_entry<kW>(A.index, L.it_a[n_it] + (2u * tid - L.vis_off[n_it]), n_sid, n_meta, n_sig);
    }
    for (uint32_t sr = 0; sr < n_sub; ++sr, ++rnd) {
      const uint32_t vi = (sr >> 1) * (2u * TPB) + 2u * tid + (sr & 1u);
      const uint32_t it = n_it, sid = n_sid, meta = n_meta;
      const uint64_t ssig = n_sig;
      const uint32_t vn = (sr & 1u) ? vi - 1u + 2u * TPB : vi + 1u;
      if (vn < total_vis) {
        if (sr & 1u) {
          n_it = gallop ? segment_from(L.vis_off, nb, vn, n_it) : segment_of(L.vis_off, nb, vn);
        } else {
          while (L.vis_off[n_it + 1] <= vn) ++n_it;  // vis_off[nb] = total_vis > vn ends it
        }
        load_entry<kW>(A.index, L.it_a[n_it] + (vn - L.vis_off[n_it]), n_sid, n_meta, n_sig);
      }
      if (tid == 0) L.qadd[(rnd + 1) % 3] = 0;  // its last reader passed the previous barrier
      bool keep = false;
      uint32_t pj = 0;
      if (vi < total_vis) {
        const uint32_t kp = meta & 0xFFFFu;
        pj = L.it_pj[it];
        const uint32_t pl = pj & 0x1FFu;
        const uint32_t j = pj >> 9;
        const uint32_t lr = L.psize[pl];
        const uint32_t ls = (meta >> 16) == 0xFFFFu ? A.indexed.sizes[sid] : (meta >> 16);
        const uint32_t mo = tile_minoverlap<TPB, kMulti>(A, L, pl, lr, ls);
        // Positional filter (helpers.hpp:46-50) at this visit's token: if it is
        // the first common token, |r n s| <= 1 + min(|r|-1-j, |s|-1-kp).  Only
        // the first-common visit can emit the pair, so a visit failing the
        // bound is dropped without reading s.
        keep = kp == 0xFFFFu || 1u + min(lr - 1u - j, ls - 1u - kp) >= mo;
        if constexpr (kW == 1) {
          // both signatures at hand: the probe's in LDS, the set's from its entry
          if (keep) {
            ++C.sig_tests;
            keep = lr + ls >= 2u * mo + (uint32_t)__popcll(L.psig[pl] ^ ssig);
          }
        } else if constexpr (kW > 1) {
          // the popcount of a prefix of the words is a lower bound of the
          // whole one, so the test stops at the first pair of words (128 bits)
          // that already rules the visit out: unrelated long sets differ in
          // about half of their bits and leave after a few words
          if (keep) {
            ++C.sig_tests;
            const uint64_t* sr = A.probe_sigs + (size_t)(T.p0 + pl) * kW;
            const uint64_t* ss = A.indexed_sigs + (size_t)sid * kW;
            const uint32_t slack = lr + ls - 2u * mo;  // mo <= min(lr, ls)
            uint32_t pc = 0;
            for (int q = 0; q < kW && pc <= slack; q += 2)
              pc += (uint32_t)__popcll(sr[q] ^ ss[q]) + (uint32_t)__popcll(sr[q + 1] ^ ss[q + 1]);
            keep = pc <= slack;
          }
        }
      }
      const uint64_t m = __ballot(keep);
      if (m) {
        const uint32_t leader = __ffsll((unsigned long long)m) - 1;
        uint32_t b = 0;
        if (lane_id() == leader) b = atomicAdd(&L.qadd[rnd % 3], (uint32_t)__popcll(m));
        b = __shfl(b, leader);
        if (keep) {
          const uint32_t slot = qn + b + lanes_below(m);
          L.q_pj[slot] = pj;
          L.q_sid[slot] = sid;
        }
      }
      __syncthreads();
      qn += L.qadd[rnd % 3];
      // TPB entries, or in the final round (no visits, fewer left) the rest
      const uint32_t take = qn >= TPB ? TPB : (sr + 1u == n_sub && last_batch ? qn : 0u);
      if (take) {
        const bool valid = tid < take;
        const uint32_t e = qn - take + tid;
        const uint32_t qpj = valid ? L.q_pj[e] : 0u, qsid = valid ? L.q_sid[e] : 0u;
        __syncthreads();  // every lane holds its entry before the next round reuses the slots
        qn -= take;
        verify_candidate<TPB, kMulti>(A, L, p0, valid, qpj, qsid, C);
      }
    }
    __syncthreads();
  }
}

// Pulls the next tile of a batched launch: each workgroup drains the queue
// of its own XCD (a contiguous run of tiles, so neighbouring tiles share an
// L2), then steals from the other XCDs' queues.
__device__ __forceinline__ uint32_t next_batch_tile(const FgssBatch& B, uint32_t ntiles, uint32_t home) {
  constexpr uint32_t NXCD = 8;
  const uint32_t per = (ntiles + NXCD - 1) / NXCD;
  for (uint32_t q = 0; q < NXCD; ++q) {
    const uint32_t x = (home + q) % NXCD;
    const uint32_t lo = x * per, hi = min(lo + per, ntiles);
    if (lo >= hi) continue;
    if (B.queue[x] >= hi - lo) continue;  // drained (read without claiming)
    const uint32_t k = atomicAdd(&B.queue[x], 1u);
    if (k < hi - lo) return lo + k;
  }
  return 0xFFFFFFFFu;
}

// Pulls the next tile of a persistent single-join launch: XCD x drains the
// contiguous run [x * per, (x + 1) * per) of the plan (neighbouring tiles
// read neighbouring sets, which then stay in that XCD's L2), then steals
// from the other runs.
__device__ __forceinline__ uint32_t next_plain_tile(uint32_t* queue, uint32_t ntiles, uint32_t home) {
  constexpr uint32_t NXCD = 8;
  const uint32_t per = (ntiles + NXCD - 1) / NXCD;
  for (uint32_t q = 0; q < NXCD; ++q) {
    const uint32_t x = (home + q) % NXCD;
    const uint32_t lo = x * per, hi = min(lo + per, ntiles);
    if (lo >= hi || queue[x] >= hi - lo) continue;
    const uint32_t k = atomicAdd(&queue[x], 1u);
    if (k < hi - lo) return lo + k;
  }
  return 0xFFFFFFFFu;
}

__device__ __forceinline__ void plain_tile_range(const FgssArgs& A, uint32_t tile, uint32_t ppt, uint32_t& p0,
                                                 uint32_t& p1) {
  if (A.tile_starts) {
    const uint32_t t = A.first_tile + tile;
    p0 = max(A.tile_starts[t], A.probe_first);
    p1 = min(A.tile_starts[t + 1], A.probe_last + 1);
  } else {
    p0 = A.probe_first + tile * ppt;
    p1 = min(p0 + ppt, A.probe_last + 1);
  }
}

// Launch modes: one workgroup per tile, a persistent grid over the tiles of
// one join, a persistent grid over the tiles of a batch of joins.
enum FgssMode { FG_TILES = 0, FG_QUEUE = 1, FG_BATCH = 2 };

template <int TPB, int kMode, bool kSlices, bool kMulti, int kW>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(4))) void k_fgss(FgssArgs A, uint32_t ppt,
                                                                                     uint32_t ntiles) {
  __shared__ FgssLds<TPB> L;
  const uint32_t tid = threadIdx.x;
  if constexpr (kMulti) {
    if (tid < MAX_THRESHOLDS) L.bounds[tid] = A.probe_bounds[tid];
    __syncthreads();
  }
  FgssCounts C;
  uint32_t rnd = 0;
  if (tid < 5) L.stats[tid] = 0;
  if (tid < 3) L.qadd[tid] = 0;
  if constexpr (!kMulti) {
    // the 1-D minoverlap table (Jaccard, Dice) when it fits: one LDS read per visit
    const uint32_t nmo = 2 * A.bounds.L + 1;
    const bool fits = A.bounds.kind != HS_COSINE && nmo <= FG_MOCAP;
    if (fits)
      for (uint32_t i = tid; i < nmo; i += TPB) L.mo[i] = (uint16_t)A.bounds.mo[i];
    if (tid == 0) L.mo_lds = fits;
  }
  __syncthreads();
  if constexpr (kMode == FG_BATCH) {
    // persistent: one launch for every tile of every join of the batch
    for (;;) {
      __syncthreads();
      if (tid == 0) L.next = next_batch_tile(A.batch, ntiles, blockIdx.x % 8);
      __syncthreads();
      const uint32_t k = L.next;
      if (k == 0xFFFFFFFFu) break;
      const uint32_t j = A.batch.tile_join[k];
      TileScope T{A.batch.tile_lo[k], A.batch.tile_hi[k], A.batch.join_first[j], A.batch.join_first[j + 1] - 1,
                  A.size_first + A.batch.sf_base[j], A.batch.sf_base[j + 1] - A.batch.sf_base[j]};
      fgss_tile<TPB, kW, kSlices, kMulti>(A, T, L, C, rnd);
    }
  } else if constexpr (kMode == FG_QUEUE) {
    // persistent over the tiles of one join (tile count on the device when
    // the plan was made there)
    if (A.dev_ntiles) ntiles = *A.dev_ntiles;
    // the claim of the next tile is issued when a tile starts; its latency
    // overlaps the tile's own loads
    if (tid == 0) L.nxt[0] = next_plain_tile(A.tile_queue, ntiles, blockIdx.x % 8);
    __syncthreads();
    for (uint32_t it = 0;; ++it) {
      const uint32_t tile = L.nxt[it & 1];
      if (tile == 0xFFFFFFFFu) break;
      uint32_t claimed = 0;
      if (tid == 0) claimed = next_plain_tile(A.tile_queue, ntiles, blockIdx.x % 8);
      uint32_t p0, p1;
      plain_tile_range(A, tile, ppt, p0, p1);
      TileScope T{p0, p1, A.idx_first, A.idx_last, A.size_first, A.size_first_len};
      fgss_tile<TPB, kW, kSlices, kMulti>(A, T, L, C, rnd);
      if (tid == 0) L.nxt[(it + 1) & 1] = claimed;
      __syncthreads();
    }
  } else {
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x, ntiles);
    if (tile == 0xFFFFFFFFu) return;
    uint32_t p0, p1;
    plain_tile_range(A, tile, ppt, p0, p1);
    TileScope T{p0, p1, A.idx_first, A.idx_last, A.size_first, A.size_first_len};
    fgss_tile<TPB, kW, kSlices, kMulti>(A, T, L, C, rnd);
  }

  // ---- statistics: one global atomic per workgroup and counter
  const uint32_t cands = wave_sum(C.cands);
  const uint32_t cand_tokens = wave_sum(C.cand_tokens);
  const uint32_t sig_tests = wave_sum(C.sig_tests);
  if (lane_id() == 0) {
    if (cands) {
      atomicAdd(&L.stats[1], (unsigned long long)cands);
      atomicAdd(&L.stats[2], (unsigned long long)cand_tokens);
    }
    if (sig_tests) atomicAdd(&L.stats[4], (unsigned long long)sig_tests);
  }
  __syncthreads();
  if (tid == 0) {
    unsigned long long* slot = stat_slot(A.out.counters);
    if (L.stats[0]) atomicAdd(&slot[CNT_VISITS], L.stats[0]);
    if (L.stats[1]) atomicAdd(&slot[CNT_CANDS], L.stats[1]);
    if (L.stats[2]) atomicAdd(&slot[CNT_CAND_TOKENS], L.stats[2]);
    if (L.stats[3]) atomicAdd(&slot[CNT_ITEMS], L.stats[3]);
    if (L.stats[4]) atomicAdd(&slot[CNT_SIG_TESTS], L.stats[4]);
  }
}

// ---- item slices and the work-balanced tile plan
//
// Item slot s of the probe range belongs to probe p and prefix position j
// with s = cum[l] + (p - first[l]) * maxprefix(l) + j, where first[l] is the
// first probe of size l in the range (sets are sorted by size, so the slots
// of one size are one run).  The host makes, per size class l, the table
// {cum, first, window lo, window hi, maxprefix} (the window of a size is an
// id interval of the indexed range; a self-join further cuts it at p).  One
// thread per slot finds its posting-list slice [a, a + n) of ids inside the
// window; probes whose window is empty get n = 0.
constexpr uint32_t SZ_CAP = 1024;  // size classes staged in LDS

struct SizeTab {
  const uint32_t *cum, *first, *lo, *hi, *mp;
};

__device__ __forceinline__ SizeTab size_tab(const uint32_t* t, uint32_t nl) {
  return SizeTab{t, t + nl, t + 2 * nl, t + 3 * nl, t + 4 * nl};
}

constexpr uint32_t IS_PER = 2;           // item slots per thread, searched in lockstep
constexpr uint32_t IS_SEARCH_QUADS = 4;  // quads per item and search round

__global__ __launch_bounds__(256) void k_item_slices(FgssArgs A, const uint32_t* __restrict__ gtab, uint32_t nl,
                                                     uint32_t total, uint32_t* __restrict__ slice_a,
                                                     uint32_t* __restrict__ slice_n) {
  __shared__ uint32_t stab[5 * SZ_CAP];
  const uint32_t* tab = gtab;
  if (nl <= SZ_CAP) {
    for (uint32_t i = threadIdx.x; i < 5 * nl; i += blockDim.x) stab[i] = gtab[i];
    __syncthreads();
    tab = stab;
  }
  const SizeTab T = size_tab(tab, nl);
  // slots base + threadIdx.x + k * 256: coalesced stores
  const uint32_t base = blockIdx.x * (256 * IS_PER) + threadIdx.x;
  uint32_t f_lo[IS_PER], n_lo[IS_PER], n_hi[IS_PER], k_lo[IS_PER], k_hi[IS_PER];
#pragma unroll
  for (uint32_t k = 0; k < IS_PER; ++k) {
    f_lo[k] = n_lo[k] = n_hi[k] = k_lo[k] = k_hi[k] = 0;
    const uint32_t s = base + k * 256;
    if (s >= total) continue;
    // size class: the last l with cum[l] <= s (its run is non-empty)
    uint32_t lo_l = 0, hi_l = nl - 1;  // cum[nl - 1] = total > s
    while (hi_l - lo_l > 1) {
      const uint32_t mid = (lo_l + hi_l) >> 1;
      if (T.cum[mid] <= s) lo_l = mid;
      else hi_l = mid;
    }
    const uint32_t l = lo_l;
    const uint32_t mp = T.mp[l];
    const uint32_t q = s - T.cum[l];
    const uint32_t p = T.first[l] + q / mp;
    const uint32_t j = q % mp;
    k_lo[k] = T.lo[l];
    k_hi[k] = A.self ? min(T.hi[l], p) : T.hi[l];
    if (k_lo[k] < k_hi[k]) {
      const uint32_t token = A.probe.tokens[A.probe.starts[p] + j];
      if (token < A.index.universe) {
        const uint32_t lb = A.index.offsets[token];
        f_lo[k] = lb;
        n_lo[k] = n_hi[k] = A.index.offsets[token + 1] - lb;
      }
    }
  }
  lockstep_bounds<IS_PER, IS_SEARCH_QUADS>(A.index.set_ids, f_lo, n_lo, n_hi, k_lo, k_hi);
#pragma unroll
  for (uint32_t k = 0; k < IS_PER; ++k) {
    const uint32_t s = base + k * 256;
    if (s < total) {
      slice_a[s] = f_lo[k];
      slice_n[s] = (k_lo[k] < k_hi[k]) ? n_hi[k] : 0u;
    }
  }
}

// Per probe a cost in units of COST_W-th of a tile of `lanes` lanes: the
// largest of COST_W / lanes (so a tile holds at most `lanes` probes),
// COST_W / (16 lanes) per token (at most 16 lanes tokens besides the last
// probe's) and COST_W / (32 lanes) per posting it visits (about 32 lanes
// visits per tile), capped at one tile.  Taking the largest term keeps all
// three limits while a tile of light probes still takes `lanes` of them.  Tile
// k takes the probes whose exclusive cost prefix lies in [k W, (k+1) W).  A
// probe costs at most W, so consecutive probes never skip a tile: every
// tile is non-empty.
constexpr uint32_t COST_W = 1024;

__global__ __launch_bounds__(256) void k_probe_cost(FgssArgs A, const uint32_t* __restrict__ gtab, uint32_t nl,
                                                    uint32_t* __restrict__ cost, uint4* __restrict__ rec,
                                                    uint32_t* __restrict__ slot, uint32_t lanes) {
  __shared__ uint32_t stab[5 * SZ_CAP];
  const uint32_t* tab = gtab;
  if (nl <= SZ_CAP) {
    for (uint32_t k = threadIdx.x; k < 5 * nl; k += blockDim.x) stab[k] = gtab[k];
    __syncthreads();
    tab = stab;
  }
  const SizeTab T = size_tab(tab, nl);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t p = A.probe_first + i;
  if (p > A.probe_last) return;
  const uint32_t l = A.probe.sizes[p];
  const uint32_t mp = T.mp[l];
  const uint32_t base = T.cum[l] + (p - T.first[l]) * mp;
  const uint32_t lo = T.lo[l];
  const uint32_t hi = A.self ? min(T.hi[l], p) : T.hi[l];
  rec[i] = make_uint4(A.probe.starts[p], l | (mp << 16), lo, hi);
  slot[i] = base;
  const uint32_t vcap = 32u * lanes;  // visits of a full tile
  uint32_t v = 0;
  for (uint32_t j = 0; j < mp && v < vcap; ++j) v += A.slice_n[base + j];
  const uint32_t c_tok = (uint32_t)(((uint64_t)l * COST_W + 16u * lanes - 1) / (16u * lanes));
  const uint32_t c_vis = (uint32_t)(((uint64_t)min(v, vcap) * COST_W + vcap - 1) / vcap);
  const uint32_t c = max(COST_W / lanes, max(c_tok, c_vis));
  cost[i] = min(c, COST_W);
}

// queue (may be NULL): the persistent grid's 8 tile counters, zeroed here.
__global__ void k_balanced_tiles(const uint32_t* __restrict__ cum, uint32_t n, uint32_t probe_first,
                                 uint32_t* __restrict__ tile_starts, uint32_t* __restrict__ ntiles,
                                 uint32_t* __restrict__ queue) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (queue && i < 8) queue[i] = 0;
  if (i >= n) return;
  const uint32_t k = cum[i] / COST_W;
  if (i == 0 || cum[i - 1] / COST_W != k) tile_starts[k] = probe_first + i;
  if (i == n - 1) {
    tile_starts[k + 1] = probe_first + n;
    *ntiles = k + 1;
  }
}

void plan_balanced_tiles(hipStream_t st, FgssArgs& a, const BalancedPlan& bp) {
  const uint32_t n = a.probe_last - a.probe_first + 1;
  HS_REQUIRE((uint64_t)n * COST_W < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "probe range too large for a balanced plan");
  bp.slice_a->ensure((size_t)std::max<uint64_t>(bp.total_items, 1) * 4);
  bp.slice_n->ensure((size_t)std::max<uint64_t>(bp.total_items, 1) * 4);
  bp.cost->ensure((size_t)n * 4);
  bp.rec->ensure((size_t)n * 16);
  bp.slot->ensure((size_t)n * 4);
  bp.tiles->ensure(((size_t)n + 1) * 4);
  bp.ntiles->ensure(4);
  a.slice_a = bp.slice_a->as<uint32_t>();
  a.slice_n = bp.slice_n->as<uint32_t>();
  if (bp.total_items)
    hipLaunchKernelGGL(k_item_slices, dim3((unsigned)((bp.total_items + 256 * IS_PER - 1) / (256 * IS_PER))),
                       dim3(256), 0, st, a,
                       bp.size_tab, bp.nl, (uint32_t)bp.total_items, bp.slice_a->as<uint32_t>(),
                       bp.slice_n->as<uint32_t>());
  uint32_t* c = bp.cost->as<uint32_t>();
  hipLaunchKernelGGL(k_probe_cost, dim3((n + 255) / 256), dim3(256), 0, st, a, bp.size_tab, bp.nl, c,
                     bp.rec->as<uint4>(), bp.slot->as<uint32_t>(),
                     (uint32_t)(a.narrow_tiles ? FG_TPB_NARROW : FG_TPB));
  a.probe_rec = bp.rec->as<uint4>();
  a.probe_slot = bp.slot->as<uint32_t>();
  launch_exclusive_scan(st, c, c, nullptr, n, *bp.scan_scratch);
  hipLaunchKernelGGL(k_balanced_tiles, dim3((n + 255) / 256), dim3(256), 0, st, c, n, a.probe_first,
                     bp.tiles->as<uint32_t>(), bp.ntiles->as<uint32_t>(), a.tile_queue);
  a.queue_zeroed = a.tile_queue != nullptr;
  HS_CHECK_HIP(hipGetLastError());
  a.tile_starts = bp.tiles->as<uint32_t>();
  a.first_tile = 0;
  a.n_plan_tiles = n;  // an upper bound; the kernel reads the count
  a.dev_ntiles = bp.ntiles->as<uint32_t>();
}

// One k_fgss instance per signature width.  t0 / t1 (may be NULL) take the
// dispatch's own start and end times: no marker packets between kernels.
template <int TPB, int kMode, bool kSlices, bool kMulti>
static void launch_width(hipStream_t st, const FgssArgs& a, dim3 grid, uint32_t ppt, uint32_t ntiles,
                         hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr) {
  const dim3 block(TPB);
  switch (a.sig_words) {
    case 0: hipExtLaunchKernelGGL((k_fgss<TPB, kMode, kSlices, kMulti, 0>), grid, block, 0, st, t0, t1, 0, a, ppt, ntiles); break;
    case 1: hipExtLaunchKernelGGL((k_fgss<TPB, kMode, kSlices, kMulti, 1>), grid, block, 0, st, t0, t1, 0, a, ppt, ntiles); break;
    case 2: hipExtLaunchKernelGGL((k_fgss<TPB, kMode, kSlices, kMulti, 2>), grid, block, 0, st, t0, t1, 0, a, ppt, ntiles); break;
    case 4: hipExtLaunchKernelGGL((k_fgss<TPB, kMode, kSlices, kMulti, 4>), grid, block, 0, st, t0, t1, 0, a, ppt, ntiles); break;
    case 8: hipExtLaunchKernelGGL((k_fgss<TPB, kMode, kSlices, kMulti, 8>), grid, block, 0, st, t0, t1, 0, a, ppt, ntiles); break;
    case 16: hipExtLaunchKernelGGL((k_fgss<TPB, kMode, kSlices, kMulti, 16>), grid, block, 0, st, t0, t1, 0, a, ppt, ntiles); break;
    default: throw Error{HS_E_INVALID, "signature width must be 0, 64, 128, 256, 512 or 1024 bits"};
  }
  HS_CHECK_HIP(hipGetLastError());
}

static void check_sigs(const FgssArgs& a) {
  HS_REQUIRE(a.sig_words == 0 || (a.probe_sigs && a.indexed_sigs), HS_E_INVALID, "signature arrays missing");
  HS_REQUIRE(a.sig_words != 1 || a.index.ent_words == 4, HS_E_INVALID,
             "64-bit signature test needs an index with signatures");
}

void launch_fgss(hipStream_t st, const FgssArgs& a, uint32_t ppt, uint32_t persistent_wgs, hipEvent_t t0,
                 hipEvent_t t1) {
  // no launch: the events still complete, as markers
  auto none = [&] {
    if (t0) HS_CHECK_HIP(hipEventRecord(t0, st));
    if (t1) HS_CHECK_HIP(hipEventRecord(t1, st));
  };
  if (a.probe.n == 0 || a.probe_first > a.probe_last) return none();
  check_sigs(a);
  if (ppt == 0 || ppt > (uint32_t)FG_TPB) ppt = FG_TPB;
  const uint32_t nprobes = a.probe_last - a.probe_first + 1;
  const uint32_t ntiles = a.tile_starts ? a.n_plan_tiles : (nprobes + ppt - 1) / ppt;
  if (ntiles == 0) return none();
  uint32_t grid = (ntiles + 7) / 8 * 8;
  if (a.tile_queue) {
    if (!a.queue_zeroed) launch_fill_u32(st, a.tile_queue, 8, 0);
    if (persistent_wgs) grid = std::min(grid, (persistent_wgs + 7) / 8 * 8);
  }
  const dim3 g(grid);
  if (a.tile_queue) {
    if (a.slice_a && a.narrow_tiles) launch_width<FG_TPB_NARROW, FG_QUEUE, true, false>(st, a, g, ppt, ntiles, t0, t1);
    else if (a.slice_a) launch_width<FG_TPB, FG_QUEUE, true, false>(st, a, g, ppt, ntiles, t0, t1);
    else launch_width<FG_TPB, FG_QUEUE, false, false>(st, a, g, ppt, ntiles, t0, t1);
  } else if (a.probe_tidx) {
    launch_width<FG_TPB, FG_TILES, false, true>(st, a, g, ppt, ntiles, t0, t1);
  } else {
    launch_width<FG_TPB, FG_TILES, false, false>(st, a, g, ppt, ntiles, t0, t1);
  }
}

void launch_fgss_batch(hipStream_t st, const FgssArgs& a, uint32_t ntiles, uint32_t workgroups, bool narrow) {
  if (ntiles == 0) return;
  check_sigs(a);
  launch_fill_u32(st, a.batch.queue, 8, 0);
  if (narrow) launch_width<FG_TPB_NARROW, FG_BATCH, false, false>(st, a, dim3(workgroups), 0u, ntiles);
  else launch_width<FG_TPB, FG_BATCH, false, false>(st, a, dim3(workgroups), 0u, ntiles);
}

uint32_t auto_signature_words(double mean) {
  uint32_t w = 1;
  while (w < 16 && 64.0 * w < 2.0 * mean) w <<= 1;
  return w;
}

// ---- token-major self-join
//
// The lists of the probe-prefix index are short (at 1M sets and Jaccard
// 0.8 the longest of 647K lists holds 558 entries, the median 2), so a lane
// that owns one entry finds all its visits by walking back through its own
// list: the entries of one list lie next to each other, ascending in set
// id, so the ids below the probe and inside its length window are the run
// just before it.  Lanes of a wave own consecutive entries and read
// overlapping runs: the records stream in with coalesced loads, and the only
// random reads left are those of verification (the survivors of both
// filters, ~1% of the visits).
namespace {
constexpr int TK_TPB = 512;
constexpr uint32_t TK_EPL = TOK_BLOCK / TK_TPB;  // entries per lane
constexpr uint32_t TK_LUT = 512;  // set sizes whose window start and minoverlap sit in LDS
constexpr uint32_t TK_QCAP = 512;  // survivors held in LDS before one flush to the global queue
constexpr uint32_t TK_HDR = 64;    // tile headers a filter workgroup computes at a time
constexpr uint32_t TK_HMERGE = 8;  // second-pass items of one tile visited in one pass
constexpr uint32_t TK_HCLAIM = 16;  // second-pass items a workgroup claims at a time, at most
constexpr uint32_t TK_NONE = 0xFFFFFFFFu;
static_assert(TOK_BLOCK % TK_TPB == 0 && TOK_WINDOW >= TOK_BLOCK, "tile geometry");

struct TokCounts {
  uint32_t visits = 0, items = 0, tests = 0;
};

// Fields of a packed index word (TokPack): key, set id, meta (size << 16
// | position), and (key, set id) as one ordered value.
__device__ __forceinline__ uint32_t pk_key(const TokArgs& A, uint64_t w) { return (uint32_t)(w >> A.pk.kshift); }
__device__ __forceinline__ uint32_t pk_id(const TokArgs& A, uint64_t w) {
  return (uint32_t)((w >> (A.pk.pb + A.pk.sb)) & ((1ull << A.pk.ib) - 1));
}
__device__ __forceinline__ uint32_t pk_meta(const TokArgs& A, uint64_t w) {
  const uint32_t pos = (uint32_t)w & ((1u << A.pk.pb) - 1);
  const uint32_t size = (uint32_t)(w >> A.pk.pb) & ((1u << A.pk.sb) - 1);
  return (size << 16) | pos;
}
__device__ __forceinline__ uint64_t pk_keyid(const TokArgs& A, uint64_t w) { return w >> (A.pk.pb + A.pk.sb); }

// Entry e of the sorted index: its set id, meta and (64-bit signature
// tests) the set's signature, gathered from the per-set table.
template <int kW>
__device__ __forceinline__ void tok_record(const TokArgs& A, uint32_t e, uint32_t& sid, uint32_t& meta,
                                           uint64_t& sig) {
  const uint64_t w = A.ent[e];
  sid = pk_id(A, w);
  meta = pk_meta(A, w);
  sig = kW == 1 ? A.sigs[sid] : 0;
}

// Workgroup state of a token-major workgroup: the bound tables (window
// start and minoverlap per size), the staged window of the index
// (entries [ws, ws + wn) of keys, set ids, metas and 64-bit signatures),
// the tile's runs and first visits, and the survivors not yet flushed.
struct TokLds {
  uint32_t lo_of[TK_LUT + 2];  // first id of the window of a probe of size l
  uint16_t mo[2 * TK_LUT + 1];
  uint32_t wk[TOK_WINDOW], wid[TOK_WINDOW], wm[TOK_WINDOW];
  uint32_t ks[TOK_WINDOW];  // first staged position of each position's key
  uint64_t wsig[TOK_WINDOW];
  uint32_t a[TOK_BLOCK], off[TOK_BLOCK + 1];
  uint32_t tmp[TK_TPB / WAVE + 1];
  uint32_t q_r[TK_QCAP], q_s[TK_QCAP], q_j[TK_QCAP];
  uint32_t qn, qbase, ws, wn, trunc;
  uint32_t hq;  // the tile's first second-pass item
  uint32_t next_ws;  // the window start of the workgroup's next tile
  uint32_t stats[3];
  uint2 hdr[TK_HDR];  // headers of the workgroup's tiles k0 .. k0 + TK_HDR - 1 (its k-th tile: g + k * grid)
};
}  // namespace

__device__ __forceinline__ bool tok_lut(const TokArgs& A) { return A.bounds.kind != HS_COSINE && A.bounds.L <= TK_LUT; }

__device__ __forceinline__ uint32_t tok_window_lo(const TokArgs& A, uint32_t lr) {
  return max(A.size_first[min(A.bounds.minsize[lr], A.size_first_len - 1)], A.first);
}

__device__ __forceinline__ void tok_tables(const TokArgs& A, TokLds& L) {
  if (tok_lut(A)) {
    for (uint32_t l = threadIdx.x; l < A.bounds.L + 2; l += TK_TPB) L.lo_of[l] = tok_window_lo(A, min(l, A.bounds.L));
    for (uint32_t k = threadIdx.x; k < 2 * A.bounds.L + 1; k += TK_TPB) L.mo[k] = (uint16_t)A.bounds.mo[k];
  }
  if (threadIdx.x < 3) L.stats[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    L.qn = 0;
    L.ws = 0;
    L.wn = 0;
  }
}

// Moves the survivors held in LDS to the global queue: one atomic for all
// of them.  Called by every lane, between barriers.
__device__ __forceinline__ void tok_flush_queue(const TokArgs& A, TokLds& L) {
  const uint32_t n = min(L.qn, TK_QCAP);  // the others went to the global queue directly
  if (n == 0) return;
  if (threadIdx.x == 0) L.qbase = (uint32_t)min(atomicAdd(&A.out.counters[CNT_QUEUE], (unsigned long long)n),
                                                 (unsigned long long)0xFFFFFFFFu);
  __syncthreads();
  const uint32_t base = L.qbase;
  for (uint32_t k = threadIdx.x; k < n; k += TK_TPB) {
    const uint64_t slot = (uint64_t)base + k;
    if (slot < A.queue_cap) A.queue[slot] = make_uint4(L.q_r[k], L.q_s[k], L.q_j[k], 0u);
  }
  __syncthreads();
  if (threadIdx.x == 0) L.qn = 0;
  __syncthreads();
}

// Entry e's set id, meta and 64-bit signature: from the staged window
// (kMixed == false: the caller knows the window holds e), else from the
// window when it holds e and the records otherwise.  The two forms are
// separate code: a choice between an LDS and a global address compiles to
// one generic (flat) load, which takes the vector memory path even for LDS.
template <int kW, bool kMixed>
__device__ __forceinline__ void tok_entry(const TokArgs& A, const TokLds& L, uint32_t e, uint32_t& sid,
                                          uint32_t& meta, uint64_t& sig) {
  const uint32_t w = e - L.ws;
  if (!kMixed || w < L.wn) {
    sid = L.wid[w];
    meta = L.wm[w];
    sig = kW == 1 ? L.wsig[w] : 0;
  } else {
    tok_record<kW>(A, e, sid, meta, sig);
  }
}

// A visit's test: the probe entry (r, meta mr = |r| << 16 | position,
// 64-bit signature) against the list entry (s, ms, sig_s).
template <int kW>
__device__ __forceinline__ bool tok_keep(const TokArgs& A, const TokLds& L, bool lut, uint32_t r, uint32_t mr,
                                         uint64_t sig_r, uint32_t s, uint32_t ms, uint64_t sig_s, TokCounts& C) {
  const uint32_t lr = mr >> 16, jr = mr & 0xFFFFu, ls = ms >> 16, ps = ms & 0xFFFFu;
  ++C.visits;
  const uint32_t mo = lut ? min((uint32_t)L.mo[lr + ls], min(lr, ls)) : dev_minoverlap(A.bounds, lr, ls);
  // positional filter (helpers.hpp:46-50) at this token, which only the
  // first common token of r and s may be
  bool keep = 1u + min(lr - 1u - jr, ls - 1u - ps) >= mo;
  if (keep && kW == 1) {
    ++C.tests;
    keep = lr + ls >= 2u * mo + (uint32_t)__popcll(sig_r ^ sig_s);
  } else if (keep && kW > 1) {
    ++C.tests;
    const uint64_t* sr = A.sigs + (size_t)r * kW;
    const uint64_t* ss = A.sigs + (size_t)s * kW;
    const uint32_t slack = lr + ls - 2u * mo;
    uint32_t pc = 0;
    for (int q = 0; q < kW && pc <= slack; q += 2)
      pc += (uint32_t)__popcll(sr[q] ^ ss[q]) + (uint32_t)__popcll(sr[q + 1] ^ ss[q + 1]);
    keep = pc <= slack;
  }
  return keep;
}

constexpr uint32_t TK_G = 4;  // visits a lane loads at once outside the staged window

// Visits [p0, p1) of tile t (entries t TOK_BLOCK + k, their runs L.a and
// first visits L.off): each lane takes a contiguous range of them, finds
// its first entry once and walks on through the runs, reading entries from
// the staged window.  Survivors go to the LDS queue; those past its room
// to the global queue, a wave at a time (below); call tok_flush_queue
// between tiles.
template <int kW, bool kMixed>
__device__ __forceinline__ void tok_visits_(const TokArgs& A, TokLds& L, uint32_t t, uint32_t p0, uint32_t p1,
                                            TokCounts& C) {
  static_assert(TOK_CHUNK <= 32 * TK_TPB && TK_HMERGE * TOK_HCHUNK <= 32 * TK_TPB,
                "a lane's visits fit one 32-bit mask");
  if (p1 <= p0) return;
  const bool lut = tok_lut(A);
  const uint32_t per = (p1 - p0 + TK_TPB - 1) / TK_TPB;
  const uint32_t pb = p0 + threadIdx.x * per, pe = min(pb + per, p1);
  uint32_t spill = 0, k0 = 0;  // bit i: visit pb + i survives past the LDS queue's room; k0: its first entry
  // a survivor of visit p: one LDS atomic (rare), or past the queue's room a spill bit
  auto enqueue = [&](uint32_t p, uint32_t r, uint32_t s, uint32_t jr) {
    const uint32_t slot = atomicAdd(&L.qn, 1u);
    if (slot < TK_QCAP) {
      L.q_r[slot] = r;
      L.q_s[slot] = s;
      L.q_j[slot] = jr;
    } else {
      spill |= 1u << (p - pb);
    }
  };
  if (pb < pe && !kMixed) {
    uint32_t k = k0 = segment_of(L.off, TOK_BLOCK, pb);
    uint32_t r, mr;
    uint64_t sig_r;
    tok_entry<kW, false>(A, L, t * TOK_BLOCK + k, r, mr, sig_r);
    for (uint32_t p = pb; p < pe; ++p) {
      if (p >= L.off[k + 1]) {  // the next entry with visits
        do ++k;
        while (p >= L.off[k + 1]);
        tok_entry<kW, false>(A, L, t * TOK_BLOCK + k, r, mr, sig_r);
      }
      const uint32_t j = L.a[k] + (p - L.off[k]);
      uint32_t s, ms;
      uint64_t sig_s;
      tok_entry<kW, false>(A, L, j, s, ms, sig_s);
      if (tok_keep<kW>(A, L, lut, r, mr, sig_r, s, ms, sig_s, C)) enqueue(p, r, s, mr & 0xFFFFu);
    }
  } else if (pb < pe) {
    // Visits of a tile whose window misses the start of its first token's
    // list: those before the window read the index words and signatures
    // from global memory.  TK_G visits a lane at a time, with every load of
    // the group issued unconditionally (inside the window as well, where
    // the data is the same): two memory latencies a group instead of two a
    // visit; loads under per-visit branches each waited for the one before.
    uint32_t k = k0 = segment_of(L.off, TOK_BLOCK, pb), kr = k;
    uint32_t r, mr;
    uint64_t sig_r;
    tok_entry<kW, false>(A, L, t * TOK_BLOCK + k, r, mr, sig_r);  // (a tile's own entries lie in its window)
    for (uint32_t g0 = pb; g0 < pe; g0 += TK_G) {
      uint32_t kq[TK_G], jq[TK_G];
#pragma unroll
      for (uint32_t q = 0; q < TK_G; ++q) {
        const uint32_t p = g0 + q;
        if (p < pe)
          while (p >= L.off[k + 1]) ++k;
        kq[q] = k;
        jq[q] = min(L.a[k] + (p < pe ? p - L.off[k] : 0u), A.n_ent - 1u);
      }
      uint64_t w[TK_G], sg[TK_G];
#pragma unroll
      for (uint32_t q = 0; q < TK_G; ++q) w[q] = A.ent[jq[q]];
#pragma unroll
      for (uint32_t q = 0; q < TK_G; ++q) sg[q] = kW == 1 ? A.sigs[pk_id(A, w[q])] : 0ull;
#pragma unroll
      for (uint32_t q = 0; q < TK_G; ++q) {
        const uint32_t p = g0 + q;
        if (p >= pe) break;
        if (kq[q] != kr) {
          kr = kq[q];
          tok_entry<kW, false>(A, L, t * TOK_BLOCK + kr, r, mr, sig_r);
        }
        const uint32_t s = pk_id(A, w[q]), ms = pk_meta(A, w[q]);
        if (tok_keep<kW>(A, L, lut, r, mr, sig_r, s, ms, sg[q], C)) enqueue(p, r, s, mr & 0xFFFFu);
      }
    }
  }
  // Survivors past the LDS queue's room (a dense chunk passes nearly all
  // of its 8192 visits): the wave reserves global queue slots for all of
  // its own with one atomic -- an atomic per survivor on the one queue
  // counter, from every dense tile at once, serialized -- and walks its
  // ranges again to write them, one step of every lane at a time.
  if (__ballot(spill != 0) == 0) return;
  const uint32_t n = wave_sum((uint32_t)__popc(spill));
  uint32_t g = 0;
  if (lane_id() == 0)
    g = (uint32_t)min(atomicAdd(&A.out.counters[CNT_QUEUE], (unsigned long long)n), (unsigned long long)0xFFFFFFFFu);
  g = __builtin_amdgcn_readfirstlane(g);
  uint32_t k = k0, r = 0, mr = 0;
  uint64_t sig_r;
  if (spill) tok_entry<kW, kMixed>(A, L, t * TOK_BLOCK + k, r, mr, sig_r);
  for (uint32_t b = 0; b < per; ++b) {
    const bool bit = (spill >> b) & 1u;
    const uint64_t mb = __ballot(bit);
    if (bit) {
      const uint32_t p = pb + b;
      if (p >= L.off[k + 1]) {
        do ++k;
        while (p >= L.off[k + 1]);
        tok_entry<kW, kMixed>(A, L, t * TOK_BLOCK + k, r, mr, sig_r);
      }
      uint32_t s, ms;
      uint64_t sig_s;
      tok_entry<kW, kMixed>(A, L, L.a[k] + (p - L.off[k]), s, ms, sig_s);
      const uint64_t slot = (uint64_t)g + lanes_below(mb);
      if (slot < A.queue_cap) A.queue[slot] = make_uint4(r, s, mr & 0xFFFFu, 0u);
    }
    g += (uint32_t)__popcll(mb);
  }
}

// The visits of a tile whose window holds every entry they read (no list
// of the tile starts before its staged window) read LDS only.
template <int kW>
__device__ __forceinline__ void tok_visits(const TokArgs& A, TokLds& L, uint32_t t, uint32_t p0, uint32_t p1,
                                           TokCounts& C) {
  if (L.trunc == TK_NONE) tok_visits_<kW, false>(A, L, t, p0, p1, C);
  else tok_visits_<kW, true>(A, L, t, p0, p1, C);
}

// Work counters: registers, then LDS, then one global atomic per
// workgroup and counter, once at the end of a persistent workgroup.
__device__ __forceinline__ void tok_flush(const TokArgs& A, TokLds& L, const TokCounts& C) {
  const uint32_t v = wave_sum(C.visits), it = wave_sum(C.items), te = wave_sum(C.tests);
  if (lane_id() == 0) {
    if (v) atomicAdd(&L.stats[0], v);
    if (it) atomicAdd(&L.stats[1], it);
    if (te) atomicAdd(&L.stats[2], te);
  }
  __syncthreads();
  tok_flush_queue(A, L);
  if (threadIdx.x == 0) {
    unsigned long long* slot = stat_slot(A.out.counters);
    if (L.stats[0]) atomicAdd(&slot[CNT_VISITS], (unsigned long long)L.stats[0]);
    if (L.stats[1]) atomicAdd(&slot[CNT_ITEMS], (unsigned long long)L.stats[1]);
    if (L.stats[2]) atomicAdd(&slot[CNT_SIG_TESTS], (unsigned long long)L.stats[2]);
  }
}

// First position p <= hi of the index with (key, id) >= (k, v), hi
// qualifying: gallop back from hi, then bisect the last step (the entries
// whose lists start before the staged window).
__device__ __forceinline__ uint32_t tok_lb_back(const TokArgs& A, uint32_t hi, uint32_t k, uint32_t v) {
  const uint64_t kv = ((uint64_t)k << A.pk.ib) | v;
  uint32_t step = 1, lo;
  for (;;) {
    const uint32_t j = hi >= step ? hi - step : 0u;
    if (pk_keyid(A, A.ent[j]) < kv) {
      lo = j + 1;
      break;
    }
    hi = j;
    if (j == 0) return 0;
    step <<= 1;
  }
  while (lo < hi) {
    const uint32_t m = (lo + hi) >> 1;
    if (pk_keyid(A, A.ent[m]) < kv) lo = m + 1;
    else hi = m;
  }
  return lo;
}


// The staged window of tile t: from the start of its first token's list
// (its index part) to the tile's end, at most TOK_WINDOW entries; every
// other token of the tile lies inside it whole.  hdr = {first key of the
// tile, first entry of that key's index part}.
// Found by a whole wave (every lane calls): 64 probes at i, i - 1, i - 3,
// ..., i - (2^l - 1) bracket the first position with key >= k0, then each
// round 64 evenly spaced probes narrow the bracket 64-fold -- a handful of
// memory latencies where one lane's gallop and bisection take two per
// halving (a tile inside a list of 40K entries: ~32 dependent loads).
__device__ __forceinline__ uint2 tok_header_wave(const TokArgs& A, uint32_t t) {
  const uint32_t lane = lane_id();
  const uint32_t i = t * TOK_BLOCK;
  const uint32_t k0 = pk_key(A, A.ent[i]) & ~1u;
  const uint32_t d = lane < 32 ? (1u << lane) - 1u : 0xFFFFFFFFu;
  const uint32_t pos = d >= i ? 0u : i - d;
  const uint64_t m = __ballot(pk_key(A, A.ent[pos]) < k0);  // a suffix of the lanes (positions descend)
  if (m == 0) return make_uint2(k0, 0u);
  // key(lo) < k0 <= key(hi); lane 0 probed i itself, whose key is >= k0
  const uint32_t l = (uint32_t)__ffsll((unsigned long long)m) - 1u;
  uint32_t lo = __builtin_amdgcn_readlane(pos, l), hi = __builtin_amdgcn_readlane(pos, l - 1u);
  while (hi - lo > 1) {
    const uint32_t step = (hi - lo + WAVE - 1) / WAVE;
    const uint32_t p = min(lo + (lane + 1) * step, hi);  // lane 63's probe is hi
    const uint32_t nb = (uint32_t)__popcll(__ballot(pk_key(A, A.ent[p]) < k0));  // a prefix of the lanes
    const uint32_t lo0 = lo;
    lo = lo0 + nb * step;  // the last probe below k0 (lo itself when none)
    hi = min(lo0 + (nb + 1) * step, hi);
  }
  return make_uint2(k0, hi);
}

// Tile t's window: from its first token's index part (hdr.y) to the tile's
// end t1, at most TOK_WINDOW entries back; every window of the tile lies in
// the span [tok_span_lo(t1), t1).
__device__ __forceinline__ uint32_t tok_t1(const TokArgs& A, uint32_t t) {
  return min(t * TOK_BLOCK + TOK_BLOCK, A.n_ent);
}
__device__ __forceinline__ uint32_t tok_span_lo(uint32_t t1) { return t1 > TOK_WINDOW ? t1 - TOK_WINDOW : 0u; }
__device__ __forceinline__ uint32_t tok_ws(const TokArgs& A, uint32_t t, uint2 hdr) {
  return max(hdr.y, tok_span_lo(tok_t1(A, t)));
}

// Thread 0 sets the window of tile t from its header; then a barrier.
__device__ __forceinline__ void tok_window(const TokArgs& A, TokLds& L, uint32_t t, uint2 hdr) {
  if (threadIdx.x == 0) {
    const uint32_t ws = tok_ws(A, t, hdr);
    L.ws = ws;
    L.wn = tok_t1(A, t) - ws;
    L.trunc = ws > hdr.y ? hdr.x : TK_NONE;
  }
  __syncthreads();
}

// After the staged keys are in LDS: the first position of every
// position's key, a running maximum of the positions where the key
// changes (R consecutive positions per lane); then a barrier.
__device__ __forceinline__ void tok_key_starts(TokLds& L) {
  constexpr uint32_t R = TOK_WINDOW / TK_TPB;
  const uint32_t tid = threadIdx.x, wn = L.wn;
  uint32_t head[R], m = 0;
#pragma unroll
  for (uint32_t q = 0; q < R; ++q) {
    const uint32_t w = tid * R + q;
    head[q] = (w < wn && (w == 0 || L.wk[w] != L.wk[w - 1])) ? w : 0u;
    m = max(m, head[q]);
  }
  uint32_t run = block_exclusive_max<TK_TPB>(m, L.tmp);
#pragma unroll
  for (uint32_t q = 0; q < R; ++q) {
    run = max(run, head[q]);
    L.ks[tid * R + q] = run;
  }
  __syncthreads();
}

// Stages tile t's window from its header; called by every lane between
// barriers (thread 0's hdr counts).  All loads of a lane are issued before
// its first LDS store.
template <int kW>
__device__ __forceinline__ void tok_stage(const TokArgs& A, TokLds& L, uint32_t t, uint2 hdr) {
  const uint32_t tid = threadIdx.x;
  tok_window(A, L, t, hdr);
  const uint32_t ws = L.ws, wn = L.wn;
  constexpr uint32_t R = TOK_WINDOW / TK_TPB;
  uint32_t key[R], sid[R], meta[R];
  uint64_t sig[R];
#pragma unroll
  for (uint32_t q = 0; q < R; ++q) {
    const uint32_t w = tid + q * TK_TPB;
    key[q] = sid[q] = meta[q] = 0;
    sig[q] = 0;
    if (w < wn) {
      const uint64_t e = A.ent[ws + w];
      key[q] = pk_key(A, e);
      sid[q] = pk_id(A, e);
      meta[q] = pk_meta(A, e);
      if (kW == 1) sig[q] = A.sigs[sid[q]];
    }
  }
#pragma unroll
  for (uint32_t q = 0; q < R; ++q) {
    const uint32_t w = tid + q * TK_TPB;
    if (w < wn) {
      L.wk[w] = key[q];
      L.wid[w] = sid[q];
      L.wm[w] = meta[q];
      if (kW == 1) L.wsig[w] = sig[q];
    }
  }
  __syncthreads();
  tok_key_starts(L);
}

// A lane's share of a tile's span, loaded ahead of the tile's staging:
// the words, then (once the window start is known) the signatures of
// those inside the window.
struct TokAhead {
  uint64_t w[TOK_WINDOW / TK_TPB];
  uint64_t sig[TOK_WINDOW / TK_TPB];
};

__device__ __forceinline__ void tok_ahead_words(const TokArgs& A, uint32_t t, TokAhead& P) {
  const uint32_t t1 = tok_t1(A, t), lo = tok_span_lo(t1);
#pragma unroll
  for (uint32_t q = 0; q < TOK_WINDOW / TK_TPB; ++q) {
    const uint32_t g = lo + threadIdx.x + q * TK_TPB;
    P.w[q] = g < t1 ? A.ent[g] : 0ull;
  }
}

template <int kW>
__device__ __forceinline__ void tok_ahead_sigs(const TokArgs& A, uint32_t t, uint32_t ws, TokAhead& P) {
  if (kW != 1) return;
  const uint32_t t1 = tok_t1(A, t), lo = tok_span_lo(t1);
#pragma unroll
  for (uint32_t q = 0; q < TOK_WINDOW / TK_TPB; ++q) {
    const uint32_t g = lo + threadIdx.x + q * TK_TPB;
    P.sig[q] = g >= ws && g < t1 ? A.sigs[pk_id(A, P.w[q])] : 0ull;
  }
}

// tok_stage from the words (and signatures) loaded ahead.
template <int kW>
__device__ __forceinline__ void tok_stage_ahead(const TokArgs& A, TokLds& L, uint32_t t, uint2 hdr,
                                                const TokAhead& P) {
  tok_window(A, L, t, hdr);
  const uint32_t ws = L.ws;
  const uint32_t t1 = tok_t1(A, t), lo = tok_span_lo(t1);
#pragma unroll
  for (uint32_t q = 0; q < TOK_WINDOW / TK_TPB; ++q) {
    const uint32_t g = lo + threadIdx.x + q * TK_TPB;
    if (g >= ws && g < t1) {
      const uint32_t w = g - ws;
      L.wk[w] = pk_key(A, P.w[q]);
      L.wid[w] = pk_id(A, P.w[q]);
      L.wm[w] = pk_meta(A, P.w[q]);
      if (kW == 1) L.wsig[w] = P.sig[q];
    }
  }
  __syncthreads();
  tok_key_starts(L);
}

// First staged position in [lo, hi) with set id >= v (ids ascend there); hi if none.
__device__ __forceinline__ uint32_t tok_lb_ids(const TokLds& L, uint32_t lo, uint32_t hi, uint32_t v) {
  while (lo < hi) {
    const uint32_t m = (lo + hi) >> 1;
    if (L.wid[m] < v) lo = m + 1;
    else hi = m;
  }
  return lo;
}

// The tile of workgroup b's k-th turn (TK_NONE past its last turn) in a
// grid of G: the index is walked from its end, where the tokens with the
// longest lists sit (token ids ascend with frequency) and the tiles with
// the most visits, and in the first round the G heaviest tiles go first to
// the workgroups that get one turn fewer (the last round fills the first
// `last` workgroups only).  Every tile has one turn: round 0 rotates the
// workgroups over tiles [0, G) from the end, round k >= 1 takes
// [k G, k G + G) in order.  On the 1M-set index at Jaccard 0.8 the tiles
// with the last ~60 keys take ~20 us against ~8.5 us on average; dealt
// g, g + G, ... they fell to workgroups already five tiles in (the filter
// took 60-63 us, 57.7 us dealt this way).
__device__ __forceinline__ uint32_t tok_tile_of(uint32_t b, uint32_t k, uint32_t G, uint32_t n_tiles) {
  const uint32_t last = n_tiles - ((n_tiles - 1) / G) * G;  // in [1, G]: workgroups with one turn more
  const uint32_t r = k == 0 ? (b + G - last) % G : k * G + b;
  return r < n_tiles ? n_tiles - 1 - r : TK_NONE;
}

// Persistent: each workgroup takes its turns of tiles (tok_tile_of) of TOK_BLOCK
// entries and stages each tile's window.  Each entry's visits are the run
// [a, b) of its token's index entries with ids in [window start, probe):
// bisections over the staged keys and ids (over the global arrays for the
// first token when its list starts before the staged window).  A scan
// numbers the tile's visits, the workgroup tests the first TOK_CHUNK of
// them, and a heavier tile queues its other chunks for k_tok_heavy.
// Headers of the workgroup's tiles k0 .. k0 + TK_HDR - 1 (tile g + k * grid),
// one lane each, into L.hdr and A.hdrs (for k_tok_heavy); then a barrier.
// Their searches run side by side instead of one per tile on the tile's
// critical path.
__device__ __forceinline__ void tok_headers(const TokArgs& A, TokLds& L, uint32_t k0, uint32_t n_tiles) {
  for (uint32_t q = threadIdx.x / WAVE; q < TK_HDR; q += TK_TPB / WAVE) {
    const uint32_t t = tok_tile_of(blockIdx.x, k0 + q, gridDim.x, n_tiles);
    if (t == TK_NONE) break;  // (wave-uniform; the later turns have no tile either)
    const uint2 h = tok_header_wave(A, t);
    if (lane_id() == 0) {
      L.hdr[q] = h;
      A.hdrs[t] = h;
    }
  }
  __syncthreads();
}

template <int kW>
__global__ __launch_bounds__(TK_TPB) void k_tok_filter(TokArgs A) {
  __shared__ TokLds L;
  const uint32_t tid = threadIdx.x;
  tok_tables(A, L);
  const bool lut = tok_lut(A);
  TokCounts C;
  const uint32_t n_tiles = (A.n_ent + TOK_BLOCK - 1) / TOK_BLOCK;
  // The next tile is loaded while the current one runs: its words right
  // after the current staging, and the signatures of its window once its
  // window start is out (after the scan).  Headers come in batches of TK_HDR.
  TokAhead P;
  uint32_t t = tok_tile_of(blockIdx.x, 0, gridDim.x, n_tiles);
  if (t != TK_NONE) tok_ahead_words(A, t, P);
  tok_headers(A, L, 0, n_tiles);
  if (tid == 0 && t != TK_NONE) L.next_ws = tok_ws(A, t, L.hdr[0]);
  __syncthreads();
  if (t != TK_NONE) tok_ahead_sigs<kW>(A, t, L.next_ws, P);
  for (uint32_t k = 0; t != TK_NONE; ++k) {  // k: the workgroup's turn
    __syncthreads();  // the previous tile's readers are done with L
    tok_stage_ahead<kW>(A, L, t, L.hdr[k % TK_HDR], P);
    const uint32_t tn = tok_tile_of(blockIdx.x, k + 1, gridDim.x, n_tiles);
    if (tn != TK_NONE) {
      if ((k + 1) % TK_HDR == 0) tok_headers(A, L, k + 1, n_tiles);  // (block-uniform; L.hdr no longer read)
      tok_ahead_words(A, tn, P);
      if (tid == 0) L.next_ws = tok_ws(A, tn, L.hdr[(k + 1) % TK_HDR]);
    }
    const uint32_t t0 = t * TOK_BLOCK, t1 = min(t0 + TOK_BLOCK, A.n_ent);
    const uint32_t ws = L.ws;
    uint32_t cnt[TK_EPL], sum = 0;
#pragma unroll
    for (uint32_t q = 0; q < TK_EPL; ++q) {
      const uint32_t k = tid * TK_EPL + q;  // consecutive entries per lane: one scan value each
      const uint32_t i = t0 + k;
      cnt[q] = 0;
      uint32_t x = i;
      if (i < t1) {
        const uint32_t w = i - ws;
        const uint32_t r = L.wid[w], lr = L.wm[w] >> 16;
        const uint32_t lo = lut ? L.lo_of[lr] : tok_window_lo(A, lr);
        if (lo < r) {
          ++C.items;
          const uint32_t key = L.wk[w], k_idx = key & ~1u;
          uint32_t b = i;
          if (k_idx != L.trunc) {
            // the token's index part lies in the window: it starts at ts and
            // ends where the key's run of other probe entries starts
            const uint32_t kst = L.ks[w];
            uint32_t ts = kst, bw = w;
            if (key & 1u) {
              if (kst > 0 && L.wk[kst - 1] == k_idx) ts = L.ks[kst - 1];
              bw = tok_lb_ids(L, ts, kst, r);  // first index entry at or above r
            }
            x = ws + tok_lb_ids(L, ts, bw, lo);
            b = ws + bw;
          } else {
            if (key & 1u) b = tok_lb_back(A, tok_lb_back(A, i, key, 0u), k_idx, r);
            x = tok_lb_back(A, b, k_idx, lo);
          }
          cnt[q] = b - x;
        }
      }
      L.a[k] = x;
      sum += cnt[q];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan<TK_TPB>(sum, L.tmp, &total);
    if (tn != TK_NONE) tok_ahead_sigs<kW>(A, tn, L.next_ws, P);  // (the scan's barriers published next_ws)
#pragma unroll
    for (uint32_t q = 0; q < TK_EPL; ++q) {
      const uint32_t k = tid * TK_EPL + q;
      L.off[k] = run;
      if (total > TOK_CHUNK && t0 + k < t1) A.aoff[t0 + k] = make_uint2(L.a[k], run);
      run += cnt[q];
    }
    // the tile's other visits, for the second pass: items 1 .. n of
    // TOK_HCHUNK visits, reserved by one lane and written by all (a tile
    // of a long list at Jaccard 0.6 has over a hundred)
    const uint32_t n_heavy = total > TOK_CHUNK ? (total - TOK_CHUNK + TOK_HCHUNK - 1) / TOK_HCHUNK : 0u;
    if (tid == 0) {
      L.off[TOK_BLOCK] = total;
      if (n_heavy)
        L.hq = (uint32_t)min(atomicAdd(&A.out.counters[CNT_HEAVY], (unsigned long long)n_heavy),
                             (unsigned long long)0xFFFFFFFFu);
    }
    __syncthreads();
    if (n_heavy) {
      const uint64_t hq = L.hq;
      for (uint32_t k = tid; k < n_heavy; k += TK_TPB)
        if (hq + k < A.heavy_cap) A.heavy[hq + k] = make_uint4(t, k + 1, total, 0u);
    }
    tok_visits<kW>(A, L, t, 0, min(total, TOK_CHUNK), C);
    __syncthreads();
    if (L.qn >= TK_QCAP / 2) tok_flush_queue(A, L);  // (a fuller queue spills to the global one)
    t = tn;
  }
  tok_flush(A, L, C);
}

// The queued chunks of heavy tiles, claimed in order by a persistent grid;
// each restages its tile's window and reloads its layout.
template <int kW>
__global__ __launch_bounds__(TK_TPB) void k_tok_heavy(TokArgs A) {
  __shared__ TokLds L;
  const uint32_t tid = threadIdx.x;
  tok_tables(A, L);
  TokCounts C;
  const unsigned long long queued = A.out.counters[CNT_HEAVY];
  const uint32_t n_items = (uint32_t)min(queued, (unsigned long long)A.heavy_cap);
  // runs of up to TK_HCLAIM items claimed in order by the workgroups as they
  // free up (items differ in cost: those of a truncated window read global
  // memory; equal shares of items left the slowest workgroup 30% behind
  // the average at Jaccard 0.6); a tile's items sit next to each other in
  // the queue, so a claim mostly needs one staging of the window
  // A few items in all (Jaccard 0.8: ~960) are shared out instead, an
  // equal contiguous run per workgroup: claims that small spread the
  // dense items thinner but restaged more windows (25.8 against 20.5 us).
  __shared__ uint32_t s_claim;
  const uint32_t claim = min(TK_HCLAIM, n_items / (4u * gridDim.x));
  uint32_t staged = TK_NONE, q = 0, q1 = 0;
  bool shared_out = false;
  for (;;) {
    if (q >= q1) {  // (block-uniform)
      if (claim == 0) {
        if (shared_out) break;
        const uint32_t per = (n_items + gridDim.x - 1) / gridDim.x;
        q = min(blockIdx.x * per, n_items);
        q1 = min(q + per, n_items);
        shared_out = true;
        if (q >= q1) break;
      } else {
        __syncthreads();
        if (tid == 0)
          s_claim = (uint32_t)min(atomicAdd(&A.out.counters[CNT_HCLAIM], 1ull), (unsigned long long)0xFFFFFFFFu);
        __syncthreads();
        const uint64_t c0 = (uint64_t)s_claim * claim;
        if (c0 >= n_items) break;
        q = (uint32_t)c0;
        q1 = min(q + claim, n_items);
      }
    }
    __syncthreads();  // the previous item's readers are done with L
    const uint4 it = A.heavy[q];
    // the following items of the same tile (a tile's items sit together,
    // in order) join this one, up to TK_HMERGE: one visit pass over their
    // contiguous visits instead of one each -- a pass's fixed cost (the
    // lanes' bisection into the runs, barriers) outweighed 4 visits a lane
    // at Jaccard 0.6, where the second pass takes ~116K items
    const uint32_t n_items_t = (it.z - TOK_CHUNK + TOK_HCHUNK - 1) / TOK_HCHUNK;
    const uint32_t q_end = min(q1, q + min(n_items_t - it.y + 1, TK_HMERGE));
    if (it.x != staged) {
      // the span's words, the header and the layout load together (none
      // waits on another), then the signatures of the window's words: two
      // memory latencies a staging instead of four
      TokAhead P;
      tok_ahead_words(A, it.x, P);
      uint2 hdr = make_uint2(0, 0);
      if (tid == 0) hdr = A.hdrs[it.x];  // written by k_tok_filter
      const uint32_t t0 = it.x * TOK_BLOCK;
      for (uint32_t k = tid; k < TOK_BLOCK; k += TK_TPB) {
        if (t0 + k < A.n_ent) {
          const uint2 ao = A.aoff[t0 + k];
          L.a[k] = ao.x;
          L.off[k] = ao.y;
        } else {
          L.off[k] = it.z;  // past the last entry: no visits
        }
      }
      if (tid == 0) {
        L.off[TOK_BLOCK] = it.z;
        L.next_ws = tok_ws(A, it.x, hdr);
      }
      __syncthreads();
      tok_ahead_sigs<kW>(A, it.x, L.next_ws, P);
      tok_stage_ahead<kW>(A, L, it.x, hdr, P);
      staged = it.x;
    }
    const uint32_t y_last = it.y + (q_end - q) - 1;
    tok_visits<kW>(A, L, it.x, TOK_CHUNK + (it.y - 1) * TOK_HCHUNK, min(it.z, TOK_CHUNK + y_last * TOK_HCHUNK), C);
    __syncthreads();
    if (L.qn >= TK_QCAP / 2) tok_flush_queue(A, L);
    q = q_end;
  }
  __syncthreads();
  tok_flush(A, L, C);
}

// One lane per queued candidate: the s-driven merge of k_fgss's verification
// from the token through which the candidate was reached (a pair reached
// through several tokens is verified by the visit of its first common one).
__global__ __launch_bounds__(TK_TPB) void k_tok_verify(TokArgs A) {
  constexpr uint32_t NW = TK_TPB / WAVE;
  __shared__ unsigned long long s_stats[3];
  __shared__ uint32_t s_pre[NW + 1];
  __shared__ unsigned long long s_base;
  if (threadIdx.x < 3) s_stats[threadIdx.x] = 0;
  __syncthreads();
  const unsigned long long queued = A.out.counters[CNT_QUEUE];
  const uint32_t total = (uint32_t)min(queued, (unsigned long long)A.queue_cap);
  const uint32_t stride = gridDim.x * TK_TPB;
  const uint32_t wave = threadIdx.x / WAVE;
  uint32_t cands = 0, cand_tokens = 0, probe_tokens = 0;
  // block-uniform trip count (barriers inside): a workgroup takes 512
  // consecutive candidates a round and reserves the result rows of all
  // its passing pairs with one atomic -- one per wave put ~860 device-scope
  // atomics on one counter, which serialized (4-8 us of a 16 us kernel)
  for (uint32_t base = blockIdx.x * TK_TPB; base < total; base += stride) {
    const uint32_t q = base + threadIdx.x;
    bool pass = false;
    uint32_t ov = 0, r = 0, s = 0;
    if (q < total) {
      const uint4 e = A.queue[q];
      r = e.x;
      s = e.y;
      const uint32_t lr = A.c.sizes[r], ls = A.c.sizes[s];
      const uint32_t mo = dev_minoverlap(A.bounds, lr, ls);
      pass = verify_visit<false>(A.c.tokens + A.c.starts[r], lr, A.c.tokens + A.c.starts[s], ls, e.z, mo, &ov);
      ++cands;
      cand_tokens += ls;
      if (A.stamp && atomicExch(&A.stamp[r], A.epoch) != A.epoch) probe_tokens += lr;
    }
    const uint64_t m = __ballot(pass);
    if (lane_id() == 0) s_pre[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t run = 0;
#pragma unroll
      for (uint32_t w = 0; w < NW; ++w) {
        const uint32_t c = s_pre[w];
        s_pre[w] = run;
        run += c;
      }
      s_base = run ? atomicAdd(&A.out.counters[CNT_PAIRS], (unsigned long long)run) : 0ull;
    }
    __syncthreads();
    if (pass) {
      const unsigned long long idx = s_base + s_pre[wave] + lanes_below(m);
      if (idx < A.out.capacity) {
        A.out.r[idx] = r;
        A.out.s[idx] = s;
        A.out.ov[idx] = ov;
      }
    }
    __syncthreads();  // s_pre and s_base are rewritten next round
  }
  cands = wave_sum(cands);
  cand_tokens = wave_sum(cand_tokens);
  probe_tokens = wave_sum(probe_tokens);
  if (lane_id() == 0 && probe_tokens) atomicAdd(&s_stats[2], (unsigned long long)probe_tokens);
  if (lane_id() == 0 && cands) {
    atomicAdd(&s_stats[0], (unsigned long long)cands);
    atomicAdd(&s_stats[1], (unsigned long long)cand_tokens);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long* slot = stat_slot(A.out.counters);
    if (s_stats[0]) {
      atomicAdd(&slot[CNT_CANDS], s_stats[0]);
      atomicAdd(&slot[CNT_CAND_TOKENS], s_stats[1]);
    }
    if (s_stats[2]) atomicAdd(&slot[CNT_CAND_PROBES], s_stats[2]);
  }
}

void launch_tok_join(hipStream_t st, const TokArgs& a, uint32_t num_cus, hipEvent_t t0, hipEvent_t t1,
                     const hipEvent_t* mid) {
  HS_REQUIRE(a.sig_words == 0 || a.sigs, HS_E_INVALID, "token-major join: signature table missing");
  const dim3 b(TK_TPB);
  const uint32_t n_tiles = std::max<uint32_t>((a.n_ent + TOK_BLOCK - 1) / TOK_BLOCK, 1);
  // persistent passes: as many workgroups as the CUs hold at once
  static int per_cu = 0;
  if (per_cu == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(k_tok_filter<1>), TK_TPB, 0) !=
            hipSuccess ||
        nb < 1) {
      (void)hipGetLastError();
      nb = 4;
    }
    // at most 3 workgroups per CU (an earlier, smaller layout measured 88 us
    // at 3 and 102 us at 4 on the 1M-set index); the 68 KB layout fits 2,
    // and a 51 KB one launched 3 a CU took 76 against 56 us
    per_cu = std::min(nb, 3);
  }
  const dim3 gp(std::max<uint32_t>(num_cus, 1) * (uint32_t)per_cu);
  const dim3 g(std::min<uint32_t>(n_tiles, gp.x));
  const hipEvent_t f1 = mid ? mid[0] : nullptr, h0 = mid ? mid[1] : nullptr, h1 = mid ? mid[2] : nullptr,
                   v0 = mid ? mid[3] : nullptr;
#define HS_TOK_FILTER(W)                                                  \
  hipExtLaunchKernelGGL(k_tok_filter<W>, g, b, 0, st, t0, f1, 0, a);      \
  hipExtLaunchKernelGGL(k_tok_heavy<W>, gp, b, 0, st, h0, h1, 0, a);
  switch (a.sig_words) {
    case 0: HS_TOK_FILTER(0) break;
    case 1: HS_TOK_FILTER(1) break;
    case 2: HS_TOK_FILTER(2) break;
    case 4: HS_TOK_FILTER(4) break;
    case 8: HS_TOK_FILTER(8) break;
    case 16: HS_TOK_FILTER(16) break;
    default: throw Error{HS_E_INVALID, "signature width must be 0, 64, 128, 256, 512 or 1024 bits"};
  }
#undef HS_TOK_FILTER
  HS_CHECK_HIP(hipGetLastError());
  // verification: no LDS to speak of, four workgroups a CU; each lane's
  // candidates are a chain of dependent loads, so more lanes, fewer rounds
  const dim3 gv(std::max<uint32_t>(num_cus, 1) * 4);
  hipExtLaunchKernelGGL(k_tok_verify, gv, b, 0, st, v0, t1, 0, a);
  HS_CHECK_HIP(hipGetLastError());
}

}  // namespace hs
