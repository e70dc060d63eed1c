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
//   * a workgroup owns a tile of consecutive probes (at most 256 probes and
//     4096 tokens, planned once per collection); their tokens are one
//     contiguous CSR range and are staged in LDS with coalesced loads;
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
//   * the merge is driven by s, whose tokens arrive 16 at a time in
//     registers, against r in LDS; it keeps the early-termination bound of
//     verify_pair (verification.hpp:9-31) for failing pairs and runs to the
//     end for passing ones, which yields the exact overlap;
//   * passing pairs leave through one atomic per wave.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "hs_internal.h"

namespace hs {

constexpr int FG_TPB = 256;                   // threads per workgroup, also max probes per tile
constexpr uint32_t FG_TOKCAP = TILE_TOKENS;    // staged probe tokens (16 KiB)
static_assert(TILE_PROBES == FG_TPB, "one probe per thread in phase A");
constexpr uint32_t FG_ITEMCAP = 1024; // items per batch
constexpr uint32_t NO_LDS = 0xFFFFFFFFu;

constexpr uint32_t SREG = 16;  // tokens of s held in registers per refill

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
struct FgssLds {
  uint32_t tok[FG_TOKCAP];
  uint32_t pstart[FG_TPB];
  uint32_t psize[FG_TPB];
  uint32_t plds[FG_TPB];
  uint32_t lo[FG_TPB];
  uint32_t hi[FG_TPB];
  uint32_t item_off[FG_TPB + 1];
  uint32_t it_a[FG_ITEMCAP];
  uint32_t it_pj[FG_ITEMCAP];
  uint32_t vis_off[FG_ITEMCAP + 1];
  uint32_t tmp[FG_TPB / WAVE + 1];
  unsigned long long stats[4];
  uint32_t next;
  uint8_t ptidx[FG_TPB];            // per-probe threshold table (serving)
  DevBounds bounds[MAX_THRESHOLDS];  // the tables of a multi-threshold launch
};

// The probes of a tile and the part of the indexed collection they join.
struct TileScope {
  uint32_t p0, p1;                // probes [p0, p1)
  uint32_t idx_first, idx_last;   // indexed ids, inclusive
  const uint32_t* size_first;     // first indexed id of size >= l
  uint32_t size_first_len;
};

__device__ __forceinline__ void fgss_tile(const FgssArgs& A, const TileScope& T, FgssLds& L, uint32_t& my_cands,
                                          uint32_t& my_cand_tokens) {
  const uint32_t tid = threadIdx.x;
  const uint32_t p0 = T.p0;
  const uint32_t np = T.p1 - T.p0;

  // ---- phase A: probe metadata, windows, LDS staging of probe tokens
  const uint32_t base = A.probe.starts[p0];
  const uint32_t tok_end = A.probe.starts[T.p1 - 1] + A.probe.sizes[T.p1 - 1];
  const uint32_t ntok = tok_end - base;
  uint32_t n_items = 0;
  if (tid < np) {
    const uint32_t p = p0 + tid;
    const uint32_t l = A.probe.sizes[p];
    const uint32_t st = A.probe.starts[p];
    const uint32_t ti = A.probe_tidx ? A.probe_tidx[p] : 0u;
    const DevBounds& B = A.probe_tidx ? L.bounds[ti] : A.bounds;
    L.ptidx[tid] = (uint8_t)ti;
    const uint32_t last = T.size_first_len - 1;
    const uint32_t mins = min(B.minsize[l], last);
    const uint32_t maxs1 = min(B.maxsize[l], last - 1) + 1;
    uint32_t lo = max(T.size_first[mins], T.idx_first);
    uint32_t hi = min(T.size_first[maxs1], T.idx_last + 1);
    if (A.self) hi = min(hi, p);
    L.pstart[tid] = st;
    L.psize[tid] = l;
    L.lo[tid] = lo;
    L.hi[tid] = hi;
    const uint32_t off = st - base;
    L.plds[tid] = (off + l <= FG_TOKCAP) ? off : NO_LDS;
    n_items = (lo < hi) ? B.maxprefix[l] : 0;
  }
  for (uint32_t i = tid; i < ntok && i < FG_TOKCAP; i += FG_TPB) L.tok[i] = A.probe.tokens[base + i];
  uint32_t total_items;
  uint32_t ioff = block_exclusive_scan<FG_TPB>(n_items, L.tmp, &total_items);
  if (tid < np) L.item_off[tid] = ioff;
  if (tid == 0) {
    L.item_off[np] = total_items;
    L.stats[3] += total_items;
  }
  __syncthreads();

  // ---- phase B: batches of items
  for (uint32_t ib = 0; ib < total_items; ib += FG_ITEMCAP) {
    const uint32_t nb = min(FG_ITEMCAP, total_items - ib);
    // slice of the posting list for every item of the batch
    for (uint32_t it = tid; it < FG_ITEMCAP; it += FG_TPB) {
      uint32_t len = 0;
      if (it < nb) {
        const uint32_t item = ib + it;
        const uint32_t pl = segment_of(L.item_off, np, item);
        const uint32_t j = item - L.item_off[pl];
        const uint32_t* r = (L.plds[pl] != NO_LDS) ? (L.tok + L.plds[pl]) : (A.probe.tokens + L.pstart[pl]);
        const uint32_t token = r[j];
        uint32_t a0 = 0;
        if (token < A.index.universe) {
          const uint32_t lb = A.index.offsets[token];
          const uint32_t le = A.index.offsets[token + 1];
          const uint32_t lo = L.lo[pl], hi = L.hi[pl];
          if (le - lb <= 8) {
            // short list: independent loads, count ids below each bound
            uint32_t below_lo = 0, below_hi = 0;
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) {
              if (lb + q < le) {
                const uint32_t id = A.index.set_ids[lb + q];
                below_lo += id < lo;
                below_hi += id < hi;
              }
            }
            a0 = lb + below_lo;
            len = below_hi - below_lo;
          } else {
            a0 = lower_bound_u32(A.index.set_ids, lb, le, lo);
            len = lower_bound_u32(A.index.set_ids, a0, le, hi) - a0;
          }
        }
        L.it_a[it] = a0;
        L.it_pj[it] = pl | (j << 8);
      }
      L.vis_off[it] = len;
    }
    __syncthreads();
    // exclusive scan of the slice lengths (4 per thread)
    constexpr uint32_t PER = FG_ITEMCAP / FG_TPB;
    uint32_t v[PER];
    uint32_t local = 0;
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
      v[k] = L.vis_off[tid * PER + k];
      local += v[k];
    }
    uint32_t total_vis;
    uint32_t run = block_exclusive_scan<FG_TPB>(local, L.tmp, &total_vis);
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
      L.vis_off[tid * PER + k] = run;
      run += v[k];
    }
    if (tid == 0) {
      L.vis_off[FG_ITEMCAP] = total_vis;
      L.stats[0] += total_vis;
    }
    __syncthreads();

    // one visit per lane
    for (uint32_t vb = 0; vb < total_vis; vb += FG_TPB) {
      const uint32_t vi = vb + tid;
      bool pass = false;
      uint32_t rid = 0, sid = 0, ov = 0, checked_len = 0;
      if (vi < total_vis) {
        const uint32_t it = segment_of(L.vis_off, nb, vi);
        const uint32_t k = vi - L.vis_off[it];
        const uint32_t e = L.it_a[it] + k;
        sid = A.index.set_ids[e];
        const uint32_t sst = A.index.sstart[e];
        const uint32_t meta = A.index.meta[e];
        const uint32_t kp = meta & 0xFFFFu;
        const uint32_t pj = L.it_pj[it];
        const uint32_t pl = pj & 0xFFu;
        const uint32_t j = pj >> 8;
        rid = p0 + pl;
        const uint32_t lr = L.psize[pl];
        const uint32_t ls = (meta >> 16) == 0xFFFFu ? A.indexed.sizes[sid] : (meta >> 16);
        const uint32_t mo = dev_minoverlap(A.probe_tidx ? L.bounds[L.ptidx[pl]] : A.bounds, lr, ls);
        // Positional filter (helpers.hpp:46-50) at this visit's token: if it is
        // the first common token, |r n s| <= 1 + min(|r|-1-j, |s|-1-kp).  Only
        // the first-common visit can emit the pair, so a visit failing the
        // bound is dropped without reading s.
        if (kp == 0xFFFFu || 1u + min(lr - 1u - j, ls - 1u - kp) >= mo) {
          checked_len = ls;  // a verified candidate
          const uint32_t* sp = A.indexed.tokens + sst;
          const uint32_t off = L.plds[pl];
          if (off != NO_LDS) pass = verify_visit<true>(L.tok + off, lr, sp, ls, j, mo, &ov);
          else pass = verify_visit<false>(A.probe.tokens + L.pstart[pl], lr, sp, ls, j, mo, &ov);
        }
      }
      my_cands += checked_len != 0;
      my_cand_tokens += checked_len;
      emit_pairs(pass, rid, sid, ov, A.out.r, A.out.s, A.out.ov, &A.out.counters[CNT_PAIRS], A.out.capacity);
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

__global__ __launch_bounds__(FG_TPB) void k_fgss(FgssArgs A, uint32_t ppt, uint32_t ntiles) {
  __shared__ FgssLds L;
  const uint32_t tid = threadIdx.x;
  if (A.probe_tidx) {
    if (tid < MAX_THRESHOLDS) L.bounds[tid] = A.probe_bounds[tid];
    __syncthreads();
  }
  uint32_t my_cands = 0, my_cand_tokens = 0;
  if (A.batch.tile_join) {
    // persistent: one launch for every tile of every join of the batch
    if (tid < 4) L.stats[tid] = 0;
    for (;;) {
      __syncthreads();
      if (tid == 0) L.next = next_batch_tile(A.batch, ntiles, blockIdx.x % 8);
      __syncthreads();
      const uint32_t k = L.next;
      if (k == 0xFFFFFFFFu) break;
      const uint32_t j = A.batch.tile_join[k];
      TileScope T{A.batch.tile_lo[k], A.batch.tile_hi[k], A.batch.join_first[j], A.batch.join_first[j + 1] - 1,
                  A.size_first + A.batch.sf_base[j], A.batch.sf_base[j + 1] - A.batch.sf_base[j]};
      fgss_tile(A, T, L, my_cands, my_cand_tokens);
    }
  } else {
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x, ntiles);
    if (tile == 0xFFFFFFFFu) return;
    if (tid < 4) L.stats[tid] = 0;
    uint32_t p0, p1;
    if (A.tile_starts) {
      const uint32_t t = A.first_tile + tile;
      p0 = max(A.tile_starts[t], A.probe_first);
      p1 = min(A.tile_starts[t + 1], A.probe_last + 1);
    } else {
      p0 = A.probe_first + tile * ppt;
      p1 = min(p0 + ppt, A.probe_last + 1);
    }
    TileScope T{p0, p1, A.idx_first, A.idx_last, A.size_first, A.size_first_len};
    fgss_tile(A, T, L, my_cands, my_cand_tokens);
  }

  // ---- statistics: one global atomic per workgroup and counter
  my_cands = wave_sum(my_cands);
  my_cand_tokens = wave_sum(my_cand_tokens);
  if (lane_id() == 0 && my_cands) {
    atomicAdd(&L.stats[1], (unsigned long long)my_cands);
    atomicAdd(&L.stats[2], (unsigned long long)my_cand_tokens);
  }
  __syncthreads();
  if (tid == 0) {
    if (L.stats[0]) atomicAdd(&A.out.counters[CNT_VISITS], L.stats[0]);
    if (L.stats[1]) atomicAdd(&A.out.counters[CNT_CANDS], L.stats[1]);
    if (L.stats[2]) atomicAdd(&A.out.counters[CNT_CAND_TOKENS], L.stats[2]);
    if (L.stats[3]) atomicAdd(&A.out.counters[CNT_ITEMS], L.stats[3]);
  }
}

void launch_fgss(hipStream_t st, const FgssArgs& a, uint32_t ppt) {
  if (a.probe.n == 0 || a.probe_first > a.probe_last) return;
  if (ppt == 0 || ppt > (uint32_t)FG_TPB) ppt = FG_TPB;
  const uint32_t nprobes = a.probe_last - a.probe_first + 1;
  const uint32_t ntiles = a.tile_starts ? a.n_plan_tiles : (nprobes + ppt - 1) / ppt;
  const uint32_t grid = (ntiles + 7) / 8 * 8;
  hipLaunchKernelGGL(k_fgss, dim3(grid), dim3(FG_TPB), 0, st, a, ppt, ntiles);
  HS_CHECK_HIP(hipGetLastError());
}

void launch_fgss_batch(hipStream_t st, const FgssArgs& a, uint32_t ntiles, uint32_t workgroups) {
  if (ntiles == 0) return;
  launch_fill_u32(st, a.batch.queue, 8, 0);
  hipLaunchKernelGGL(k_fgss, dim3(workgroups), dim3(FG_TPB), 0, st, a, 0u, ntiles);
  HS_CHECK_HIP(hipGetLastError());
}

}  // namespace hs
