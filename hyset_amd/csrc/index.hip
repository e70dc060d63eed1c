// Inverted-index build over set prefixes for the item-major joins (R-S,
// ranges, batches, serving; the self-join's token-major index is built in
// tokindex.hip), the fills, copies and scans around it.
//
// Reference: include/hyset/index.hpp:217-313 (histogram, thrust::exclusive_scan,
// create_index with atomic scatter, thrust::sort by (token, setID)).
// Here: the prefix entries are emitted in set order, sorted by token with a
// stable LSD radix sort (so each list stays in set-id order), and the
// offsets written at the token boundaries of the sorted keys: list(t) =
// set_ids[offsets[t] .. offsets[t+1]) ascending -- the order of the
// reference's sorted host index (index.hpp:278-279).  Each entry's record
// carries its set's size, its position inside the set (the `position` of
// the reference's 12-byte entry, structs.hpp:56-63, used by the positional
// filter) and, for 64-bit signature tests, the set's signature.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "hs_internal.h"

namespace hs {

// ------------------------------------------------------------------ fills

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, size_t n, unsigned long long v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

static unsigned grid_for(size_t n, unsigned tpb, unsigned cap = 4096) {
  size_t g = (n + tpb - 1) / tpb;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

void launch_fill_u32(hipStream_t st, uint32_t* p, size_t n, uint32_t v) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n, v);
  HS_CHECK_HIP(hipGetLastError());
}
void launch_fill_u64(hipStream_t st, unsigned long long* p, size_t n, unsigned long long v) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n, v);
  HS_CHECK_HIP(hipGetLastError());
}

// The first min(count, n) result rows (r, s and overlap: three rows of `cap`
// words from `rows`; count read from the join's counters) stored straight
// into mapped host memory laid out as three rows of `dst_cap` words: the
// pairs reach the host with the join's own completion, without a separate
// DMA copy.
__global__ void k_rows_to_host(const uint32_t* __restrict__ rows, uint64_t cap,
                               const unsigned long long* __restrict__ counters, uint64_t n, uint32_t* dst,
                               uint64_t dst_cap) {
  const uint64_t m = min((uint64_t)counters[CNT_PAIRS], n);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
    dst[i] = rows[i];
    dst[dst_cap + i] = rows[cap + i];
    dst[2 * dst_cap + i] = rows[2 * cap + i];
  }
}

void launch_rows_to_host(hipStream_t st, const uint32_t* rows, uint64_t cap, const unsigned long long* counters,
                         uint64_t n, uint32_t* dst, uint64_t dst_cap) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_rows_to_host, dim3(grid_for(n, 256, 1024)), dim3(256), 0, st, rows, cap, counters, n, dst,
                     dst_cap);
  HS_CHECK_HIP(hipGetLastError());
}

// Plain device copy in 16-byte vectors: the reference rate for HBM traffic.
__global__ void k_copy_u128(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

void launch_copy(hipStream_t st, const void* src, void* dst, size_t bytes) {
  const size_t n = bytes / sizeof(uint4);
  if (n == 0) return;
  hipLaunchKernelGGL(k_copy_u128, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, (const uint4*)src, (uint4*)dst,
                     n);
  HS_CHECK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------- scan
// Two launches: per-tile sums, then per-tile scans, each workgroup adding
// up the sums of the tiles before its own.  A tile is 256 threads x 16
// elements; at most 4096 tiles (2^24 elements).

constexpr int SC_TPB = 256;
constexpr int SC_IPT = 16;
constexpr int SC_TILE = SC_TPB * SC_IPT;

__global__ __launch_bounds__(SC_TPB) void k_scan_tile_sums(const uint32_t* __restrict__ in, size_t n,
                                                           uint32_t* __restrict__ sums) {
  __shared__ uint32_t red[SC_TPB / WAVE];
  size_t base = (size_t)blockIdx.x * SC_TILE;
  uint32_t acc = 0;
#pragma unroll
  for (int k = 0; k < SC_IPT; ++k) {
    size_t i = base + (size_t)k * SC_TPB + threadIdx.x;
    if (i < n) acc += in[i];
  }
  acc = wave_sum(acc);
  if (lane_id() == 0) red[threadIdx.x / WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < SC_TPB / WAVE; ++w) t += red[w];
    sums[blockIdx.x] = t;
  }
}

// in, out and out2 may alias (in-place scans): each tile is read into LDS
// before any of it is written.  The tile sums are not scanned: each
// workgroup adds up those before its own (at most SC_TILE of them, one
// coalesced read each), so a scan is two launches for up to SC_TILE tiles.
__global__ __launch_bounds__(SC_TPB) void k_scan_apply(const uint32_t* in, size_t n, const uint32_t* __restrict__ sums,
                                                       uint32_t* out, uint32_t* out2) {
  __shared__ uint32_t tile[SC_TILE];
  __shared__ uint32_t tmp[SC_TPB / WAVE + 1];
  size_t base = (size_t)blockIdx.x * SC_TILE;
  uint32_t acc = 0;
  for (uint32_t j = threadIdx.x; j < blockIdx.x; j += SC_TPB) acc += sums[j];
  uint32_t offset;
  (void)block_exclusive_scan<SC_TPB>(acc, tmp, &offset);
  // coalesced load into LDS
#pragma unroll
  for (int k = 0; k < SC_IPT; ++k) {
    size_t i = base + (size_t)k * SC_TPB + threadIdx.x;
    tile[k * SC_TPB + threadIdx.x] = i < n ? in[i] : 0;
  }
  __syncthreads();
  uint32_t v[SC_IPT];
  uint32_t local = 0;
#pragma unroll
  for (int k = 0; k < SC_IPT; ++k) {
    v[k] = tile[threadIdx.x * SC_IPT + k];
    local += v[k];
  }
  uint32_t total;
  uint32_t run = offset + block_exclusive_scan<SC_TPB>(local, tmp, &total);
#pragma unroll
  for (int k = 0; k < SC_IPT; ++k) {
    tile[threadIdx.x * SC_IPT + k] = run;
    run += v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SC_IPT; ++k) {
    size_t i = base + (size_t)k * SC_TPB + threadIdx.x;
    if (i < n) {
      uint32_t x = tile[k * SC_TPB + threadIdx.x];
      out[i] = x;
      if (out2) out2[i] = x;
    }
  }
}

void launch_exclusive_scan(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t* out2, size_t n,
                           DevBuf& scratch) {
  if (n == 0) return;
  size_t nb = (n + SC_TILE - 1) / SC_TILE;
  HS_REQUIRE(nb <= (size_t)SC_TILE, HS_E_UNSUPPORTED, "scan limited to 2^24 elements");
  scratch.ensure(nb * sizeof(uint32_t));
  uint32_t* sums = scratch.as<uint32_t>();
  hipLaunchKernelGGL(k_scan_tile_sums, dim3((unsigned)nb), dim3(SC_TPB), 0, st, in, n, sums);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SC_TPB), 0, st, in, n, sums, out, out2);
  HS_CHECK_HIP(hipGetLastError());
}

// ------------------------------------------------------------ index build
//
// 1. emit: one thread per set writes its prefix entries (token, set id,
//    position) in set-id order.  The entry offset of set i comes from a
//    per-size table (sets are sorted by size, so all sets of one size are
//    contiguous): off(i) = cum[l] + (i - first id of size l) * prefix_len[l].
// 2. a stable LSD radix sort by token, 8 bits per pass, only as many passes
//    as the universe needs; stability keeps every list in set-id order.  Per pass:
//    per-tile digit histograms (LDS atomics), a scan of each digit's row of
//    the histogram table (one workgroup per digit), and a scatter that
//    ranks each element inside its tile with wave ballots, reorders the
//    tile by digit in LDS and writes contiguous runs (no global atomics
//    anywhere: device-scope atomics are slow across the eight XCDs).
// 3. offsets[t] = first entry with token >= t, written at the token
//    boundaries of the sorted keys, and the size (and, for 64-bit signature
//    tests, the signature) of each entry's set.

constexpr int RS_TPB = 256;
constexpr int RS_ROUNDS = 16;
constexpr uint32_t RS_TILE = RS_TPB * RS_ROUNDS;  // 4096 entries per tile
constexpr int RS_WAVES = RS_TPB / WAVE;

// One set's prefix entries at entry offset o, read IX_CHUNK tokens at a
// time with independent loads; with sig != NULL all of the set's tokens are
// read and its 64-bit XOR-parity signature (bit token mod 64) is stored too,
// so a join that indexes the whole collection needs no separate signature
// pass.
constexpr uint32_t IX_CHUNK = 16;

__device__ __forceinline__ void emit_set(const DevColl& c, uint32_t i, uint32_t l, uint32_t pl, uint32_t o,
                                         uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                         uint32_t* __restrict__ meta, uint64_t* __restrict__ sig) {
  const uint32_t* t = c.tokens + c.starts[i];
  const uint32_t len = sig ? l : pl;
  const uint32_t size_bits = min(l, 0xFFFFu) << 16;
  uint64_t w = 0;
  for (uint32_t k0 = 0; k0 < len; k0 += IX_CHUNK) {
    uint32_t v[IX_CHUNK];
#pragma unroll
    for (uint32_t q = 0; q < IX_CHUNK; ++q) v[q] = k0 + q < len ? t[k0 + q] : 0u;
#pragma unroll
    for (uint32_t q = 0; q < IX_CHUNK; ++q) {
      const uint32_t k = k0 + q;
      if (k < pl) {
        keys[o + k] = v[q];
        vals[o + k] = i;
        meta[o + k] = size_bits | min(k, 0xFFFFu);
      }
      if (k < len) w ^= 1ull << (v[q] & 63u);
    }
  }
  if (sig) sig[i] = w;
}

__global__ void k_index_emit(DevColl c, uint32_t first, uint32_t last, const uint32_t* __restrict__ prefix_len,
                             const uint32_t* __restrict__ size_first, const uint32_t* __restrict__ cum,
                             uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, uint32_t* __restrict__ meta,
                             uint64_t* __restrict__ sig) {
  const uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > last) return;
  const uint32_t l = c.sizes[i];
  const uint32_t pl = prefix_len[l];
  const uint32_t o = cum[l] + (i - max(size_first[l], first)) * pl;
  emit_set(c, i, l, pl, o, keys, vals, meta, sig);
}

// Batched collections: set i belongs to join j = last j with join_first[j]
// <= i; its sizes table and entry table start at sf_base[j].
__global__ void k_index_emit_batch(DevColl c, const uint32_t* __restrict__ join_first, uint32_t n_joins,
                                   const uint32_t* __restrict__ prefix_len, const uint32_t* __restrict__ size_first,
                                   const uint32_t* __restrict__ sf_base, const uint32_t* __restrict__ cum,
                                   uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                   uint32_t* __restrict__ meta, uint64_t* __restrict__ sig) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.n) return;
  const uint32_t j = segment_of(join_first, n_joins, i);
  const uint32_t l = c.sizes[i];
  const uint32_t pl = prefix_len[l];
  const uint32_t b = sf_base[j];
  const uint32_t o = cum[b + l] + (i - size_first[b + l]) * pl;
  emit_set(c, i, l, pl, o, keys, vals, meta, sig);
}

// Digit histogram of one tile: LDS atomics, then one column of the
// digit-major table hist[digit * ntiles + tile].
template <int kBits>
__global__ __launch_bounds__(RS_TPB) void k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n, uint32_t shift,
                                                       uint32_t* __restrict__ hist, uint32_t ntiles) {
  constexpr uint32_t NB = 1u << kBits;
  __shared__ uint32_t h[NB];
  for (uint32_t d = threadIdx.x; d < NB; d += RS_TPB) h[d] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * RS_TILE;
#pragma unroll 4
  for (int r = 0; r < RS_ROUNDS; ++r) {
    const uint32_t idx = base + r * RS_TPB + threadIdx.x;
    if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & (NB - 1)], 1u);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < NB; d += RS_TPB) hist[d * ntiles + blockIdx.x] = h[d];
}

// The scan of the digit-major histogram table in one launch: workgroup d
// turns row d (the counts of digit d in every tile) into its exclusive
// prefix in place and stores the row total; the scatter adds the totals of
// the smaller digits itself.  One launch instead of a three-kernel scan of
// the whole table, whose launches cost more than the work at this size.
__global__ __launch_bounds__(RS_TPB) void k_radix_rows(uint32_t* __restrict__ hist, uint32_t ntiles,
                                                      uint32_t* __restrict__ totals) {
  __shared__ uint32_t tmp[RS_TPB / WAVE + 1];
  uint32_t* row = hist + (size_t)blockIdx.x * ntiles;
  constexpr uint32_t PT = 8;  // counts per thread and chunk
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < ntiles; c0 += RS_TPB * PT) {
    uint32_t v[PT], local = 0;
#pragma unroll
    for (uint32_t k = 0; k < PT; ++k) {
      const uint32_t i = c0 + threadIdx.x * PT + k;
      v[k] = i < ntiles ? row[i] : 0u;
      local += v[k];
    }
    uint32_t total;
    uint32_t run = carry + block_exclusive_scan<RS_TPB>(local, tmp, &total);
#pragma unroll
    for (uint32_t k = 0; k < PT; ++k) {
      const uint32_t i = c0 + threadIdx.x * PT + k;
      if (i < ntiles) row[i] = run;
      run += v[k];
    }
    carry += total;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Stable scatter of one tile.  Each wave owns a contiguous quarter of the
// tile (RS_ROUNDS rounds of 64 entries, loaded into registers up front), so
// the tile's order is wave by wave.  The waves count their digits (LDS
// atomics into wave-private rows); one barrier later every wave knows where
// its entries of digit d start inside the tile sorted by digit, ranks its
// rounds on its own (lanes of equal digit found with kBits ballots, the
// leader of each digit advancing the wave's cursor) and places them in an
// LDS copy of the tile in digit order.  The sorted tile is then written out
// with consecutive lanes on consecutive slots, so each digit's run of the
// tile leaves as contiguous stores instead of one scattered store per lane.
template <int kBits>
__global__ __launch_bounds__(RS_TPB) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                          const uint32_t* __restrict__ vin,
                                                          const uint32_t* __restrict__ pin, uint32_t n,
                                                          uint32_t shift, const uint32_t* __restrict__ offs,
                                                          const uint32_t* __restrict__ totals, uint32_t ntiles,
                                                          uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                          uint32_t* __restrict__ pout) {
  constexpr uint32_t NB = 1u << kBits;
  static_assert(NB <= RS_TPB, "one digit total per thread");
  __shared__ uint32_t s_cnt[RS_WAVES][NB];  // per-wave digit counts, then the wave's cursors
  __shared__ uint32_t s_lo[NB];             // first slot of digit d in the sorted tile
  __shared__ uint32_t s_base[NB];           // output position of that slot
  __shared__ uint32_t s_key[RS_TILE];
  __shared__ uint32_t s_val[RS_TILE];
  __shared__ uint32_t s_pos[RS_TILE];
  __shared__ uint32_t s_tmp[RS_TPB / WAVE + 1];
  const uint32_t tid = threadIdx.x;
  const uint32_t wave = tid / WAVE, lane = lane_id();
  const uint32_t tbase = blockIdx.x * RS_TILE;
  const uint32_t wbase = tbase + wave * (RS_ROUNDS * WAVE);
  const uint32_t tn = min(RS_TILE, n - tbase);  // entries of this tile
  uint32_t key[RS_ROUNDS], val[RS_ROUNDS], ps[RS_ROUNDS];
#pragma unroll
  for (int r = 0; r < RS_ROUNDS; ++r) {
    const uint32_t idx = wbase + r * WAVE + lane;
    const bool valid = idx < n;
    key[r] = valid ? kin[idx] : 0u;
    val[r] = valid ? vin[idx] : 0u;
    ps[r] = valid ? pin[idx] : 0u;
  }
  // output base of digit d in this tile: the entries of smaller digits (a
  // scan of the digit totals) plus those of digit d in earlier tiles
  {
    uint32_t all;
    const uint32_t below = block_exclusive_scan<RS_TPB>(tid < NB ? totals[tid] : 0u, s_tmp, &all);
    if (tid < NB) {
#pragma unroll
      for (int w = 0; w < RS_WAVES; ++w) s_cnt[w][tid] = 0;
      s_base[tid] = below + offs[tid * ntiles + blockIdx.x];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ROUNDS; ++r)
    if (wbase + r * WAVE + lane < n) atomicAdd(&s_cnt[wave][(key[r] >> shift) & (NB - 1)], 1u);
  __syncthreads();
  // the tile sorted by digit: digit d occupies [s_lo[d], s_lo[d] + total(d)),
  // wave w's entries of d start after those of the earlier waves
  static_assert(NB % RS_TPB == 0 || NB < RS_TPB, "digits per thread");
  constexpr uint32_t DPT = NB > RS_TPB ? NB / RS_TPB : 1;
  uint32_t tot[DPT], sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < DPT; ++k) {
    const uint32_t d = tid * DPT + k;
    tot[k] = 0;
    if (d < NB)
#pragma unroll
      for (int w = 0; w < RS_WAVES; ++w) tot[k] += s_cnt[w][d];
    sum += tot[k];
  }
  uint32_t all;
  uint32_t run = block_exclusive_scan<RS_TPB>(sum, s_tmp, &all);
#pragma unroll
  for (uint32_t k = 0; k < DPT; ++k) {
    const uint32_t d = tid * DPT + k;
    if (d < NB) {
      s_lo[d] = run;
      uint32_t c = run;
#pragma unroll
      for (int w = 0; w < RS_WAVES; ++w) {
        const uint32_t x = s_cnt[w][d];
        s_cnt[w][d] = c;
        c += x;
      }
    }
    run += tot[k];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ROUNDS; ++r) {
    const bool valid = wbase + r * WAVE + lane < n;
    const uint32_t d = (key[r] >> shift) & (NB - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kBits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    if (valid) {
      const uint32_t o = s_cnt[wave][d] + lanes_below(peers);
      s_key[o] = key[r];
      s_val[o] = val[r];
      s_pos[o] = ps[r];
    }
    // the leader of each digit advances the wave's cursor; the next round's
    // reads come after it in this wave's program order
    __builtin_amdgcn_wave_barrier();
    if (valid && lanes_below(peers) == 0) s_cnt[wave][d] += (uint32_t)__popcll(peers);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // slot i of the sorted tile goes to base(d) + (i - lo(d))
  for (uint32_t i = tid; i < tn; i += RS_TPB) {
    const uint32_t k = s_key[i];
    const uint32_t d = (k >> shift) & (NB - 1);
    const uint32_t o = s_base[d] + (i - s_lo[d]);
    kout[o] = k;
    vout[o] = s_val[i];
    pout[o] = s_pos[i];
  }
}

// offsets[t] = number of entries with token < t, for t in [0, universe],
// from the sorted keys: entry i writes the offsets of the tokens in
// (keys[i-1], keys[i]] (every t is written once); and, for every entry, its
// record: set id, meta (size and position, carried through the sort) and,
// when given, the set's signature (one random read per entry here instead
// of per visit).  With counters != NULL the thread of entry n also prepares
// the join's work counters (all zero, the index size in CNT_ENTRIES), which
// saves the join a fill and a copy launch.
__global__ void k_index_finish(const uint32_t* __restrict__ keys, uint32_t n, uint32_t universe,
                               uint32_t* __restrict__ offsets, const uint32_t* __restrict__ set_ids,
                               const uint32_t* __restrict__ metas, const uint64_t* __restrict__ sigs,
                               uint32_t* __restrict__ ent, unsigned long long* __restrict__ counters) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  if (i == n && counters) {
    for (int k = 0; k < CNT_WORDS; ++k) counters[k] = 0;
    counters[CNT_ENTRIES] = n;
  }
  const uint32_t prev = i == 0 ? 0u : keys[i - 1] + 1;   // first token not yet written
  const uint32_t cur = i == n ? universe : keys[i];      // tokens prev..cur start at entry i
  for (uint32_t t = prev; t <= cur && t <= universe; ++t) offsets[t] = i;
  if (i < n) {
    const uint32_t sid = set_ids[i];
    const uint32_t m = metas[i];
    if (sigs) {
      const uint64_t g = sigs[sid];
      reinterpret_cast<uint4*>(ent)[i] = make_uint4(sid, m, (uint32_t)g, (uint32_t)(g >> 32));
    } else {
      reinterpret_cast<uint2*>(ent)[i] = make_uint2(sid, m);
    }
  }
}

// Digit plan of the LSD sort for keys below `universe`: 8-bit digits, as
// many passes as the key bits need.  (Wider digits halve the passes but
// leave ~4 entries per digit and tile, and the scattered stores then cost
// more than the saved pass: 10-bit digits measured slower on the 1M-set
// index.)
struct RadixPlan {
  uint32_t passes, bits;
};
static RadixPlan radix_plan(uint32_t universe) {
  uint32_t need = 1;
  while (need < 32 && ((uint64_t)(universe - 1) >> need)) ++need;
  return RadixPlan{(need + 7) / 8, 8u};
}
constexpr uint32_t RS_DIGITS = 256;

static uint32_t radix_tiles(uint64_t entries) { return (uint32_t)((entries + RS_TILE - 1) / RS_TILE); }

size_t index_region(uint32_t universe, uint64_t entries, uint8_t* base, IndexBufs* out) {
  const size_t cap = entries ? entries : 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = (off + bytes + 255) & ~size_t(255);
    return base ? base + at : nullptr;
  };
  IndexBufs b{};
  b.offsets = reinterpret_cast<uint32_t*>(take(((size_t)universe + 1) * 4));
  b.set_ids = reinterpret_cast<uint32_t*>(take((cap + IDS_PAD) * 4));  // the window search reads whole quads
  b.metas = reinterpret_cast<uint32_t*>(take(cap * 4));
  b.ent = reinterpret_cast<uint32_t*>(take(cap * 16));
  b.keys[0] = reinterpret_cast<uint32_t*>(take(cap * 4));
  b.keys[1] = reinterpret_cast<uint32_t*>(take(cap * 4));
  b.vals = reinterpret_cast<uint32_t*>(take(cap * 4));
  b.pos = reinterpret_cast<uint32_t*>(take(cap * 4));
  b.hist = reinterpret_cast<uint32_t*>(take((size_t)std::max<uint32_t>(radix_tiles(entries), 1) * RS_DIGITS * 4));
  b.totals = reinterpret_cast<uint32_t*>(take(RS_DIGITS * 4));
  if (out) *out = b;
  return off;
}

static IndexBufs prepare(IndexBuild& ib, uint32_t universe, uint64_t entries) {
  HS_REQUIRE(entries < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "index limited to 2^32-1 entries");
  ib.universe = universe;
  ib.entries = entries;
  const size_t cap = entries ? entries : 1;
  ib.offsets.ensure(((size_t)universe + 1) * sizeof(uint32_t));
  ib.set_ids.ensure((cap + IDS_PAD) * sizeof(uint32_t));  // the window search reads whole quads
  ib.metas.ensure(cap * sizeof(uint32_t));
  ib.ent.ensure(cap * 4 * sizeof(uint32_t));
  ib.keys[0].ensure(cap * sizeof(uint32_t));
  ib.keys[1].ensure(cap * sizeof(uint32_t));
  ib.vals.ensure(cap * sizeof(uint32_t));
  ib.pos.ensure(cap * sizeof(uint32_t));
  ib.hist.ensure((size_t)std::max<uint32_t>(radix_tiles(entries), 1) * RS_DIGITS * sizeof(uint32_t));
  ib.scan_scratch.ensure(RS_DIGITS * sizeof(uint32_t));
  return IndexBufs{ib.offsets.as<uint32_t>(), ib.set_ids.as<uint32_t>(), ib.metas.as<uint32_t>(),
                   ib.ent.as<uint32_t>(), {ib.keys[0].as<uint32_t>(), ib.keys[1].as<uint32_t>()},
                   ib.vals.as<uint32_t>(), ib.pos.as<uint32_t>(), ib.hist.as<uint32_t>(),
                   ib.scan_scratch.as<uint32_t>()};
}

// One LSD pass over the entries in buffer `cur`: histograms, row scans, scatter.
template <int kBits>
static void radix_pass(hipStream_t st, uint32_t n, uint32_t shift, uint32_t ntiles, const IndexBufs& b,
                       uint32_t* const kbuf[2], uint32_t* const vbuf[2], uint32_t* const pbuf[2], int cur) {
  constexpr uint32_t NB = 1u << kBits;
  static_assert(NB <= RS_DIGITS, "digit totals");
  hipLaunchKernelGGL((k_radix_hist<kBits>), dim3(ntiles), dim3(RS_TPB), 0, st, kbuf[cur], n, shift, b.hist, ntiles);
  hipLaunchKernelGGL(k_radix_rows, dim3(NB), dim3(RS_TPB), 0, st, b.hist, ntiles, b.totals);
  hipLaunchKernelGGL((k_radix_scatter<kBits>), dim3(ntiles), dim3(RS_TPB), 0, st, kbuf[cur], vbuf[cur], pbuf[cur], n,
                     shift, b.hist, b.totals, ntiles, kbuf[cur ^ 1], vbuf[cur ^ 1], pbuf[cur ^ 1]);
}

// Sorts the emitted entries (in buffer `cur`) by token and writes the index.
static void sort_and_finish(hipStream_t st, const DevColl& c, uint32_t n, uint32_t universe, const IndexBufs& b,
                            int cur, RadixPlan plan, const uint64_t* sigs, unsigned long long* counters,
                            hipEvent_t done = nullptr) {
  const uint32_t ntiles = radix_tiles(n);
  uint32_t* const kbuf[2] = {b.keys[0], b.keys[1]};
  uint32_t* const vbuf[2] = {b.vals, b.set_ids};
  uint32_t* const pbuf[2] = {b.pos, b.metas};
  for (uint32_t p = 0; p < plan.passes; ++p) {
    const uint32_t shift = plan.bits * p;
    radix_pass<8>(st, n, shift, ntiles, b, kbuf, vbuf, pbuf, cur);
    cur ^= 1;
  }
  hipExtLaunchKernelGGL(k_index_finish, dim3((n + 1 + 255) / 256), dim3(256), 0, st, nullptr, done, 0, kbuf[cur], n,
                        universe, b.offsets, b.set_ids, b.metas, sigs, b.ent, counters);
  HS_CHECK_HIP(hipGetLastError());
}

// The emit writes into the buffer pair that makes the last radix pass land in
// (keys[1], set_ids, positions): buffer 0 for an odd number of passes.
static int emit_buffer(uint32_t passes) { return (passes % 2 == 1) ? 0 : 1; }

void build_index_into(hipStream_t st, const DevColl& c, uint32_t first, uint32_t last, const uint32_t* prefix_len,
                      const uint32_t* size_first, const uint32_t* cum, uint32_t universe, uint64_t entries,
                      const IndexBufs& b, const uint64_t* sigs, uint64_t* sigs_out, unsigned long long* counters,
                      hipEvent_t done) {
  HS_REQUIRE(entries < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "index limited to 2^32-1 entries");
  const uint32_t n = (uint32_t)entries;
  if (c.n > 0 && first <= last && n > 0) {
    const RadixPlan plan = radix_plan(universe);
    const int cur = emit_buffer(plan.passes);
    uint32_t* kb = cur ? b.keys[1] : b.keys[0];
    uint32_t* vb = cur ? b.set_ids : b.vals;
    uint32_t* pb = cur ? b.metas : b.pos;
    const uint32_t nsets = last - first + 1;
    hipLaunchKernelGGL(k_index_emit, dim3((nsets + 255) / 256), dim3(256), 0, st, c, first, last, prefix_len,
                       size_first, cum, kb, vb, pb, sigs_out);
    sort_and_finish(st, c, n, universe, b, cur, plan, sigs, counters, done);
  } else {
    launch_fill_u32(st, b.offsets, (size_t)universe + 1, 0);
    if (done) HS_CHECK_HIP(hipEventRecord(done, st));
  }
  HS_CHECK_HIP(hipGetLastError());
}

void build_index(hipStream_t st, const DevColl& c, uint32_t first, uint32_t last, const uint32_t* prefix_len,
                 const uint32_t* size_first, const uint32_t* cum, uint32_t universe, uint64_t entries,
                 IndexBuild& ib, const uint64_t* sigs, uint64_t* sigs_out, unsigned long long* counters,
                 hipEvent_t done) {
  const IndexBufs b = prepare(ib, universe, entries);
  ib.ent_words = sigs ? 4u : 2u;
  build_index_into(st, c, first, last, prefix_len, size_first, cum, universe, entries, b, sigs, sigs_out,
                   counters, done);
}

void build_index_batch(hipStream_t st, const DevColl& c, const uint32_t* join_first, uint32_t n_joins,
                       const uint32_t* prefix_len, const uint32_t* size_first, const uint32_t* sf_base,
                       const uint32_t* cum, uint32_t universe, uint64_t entries, IndexBuild& ib,
                       const uint64_t* sigs, uint64_t* sigs_out) {
  unsigned long long* counters = nullptr;
  const IndexBufs b = prepare(ib, universe, entries);
  ib.ent_words = sigs ? 4u : 2u;
  const uint32_t n = (uint32_t)entries;
  if (c.n > 0 && n > 0) {
    const RadixPlan plan = radix_plan(universe);
    const int cur = emit_buffer(plan.passes);
    uint32_t* kb = cur ? b.keys[1] : b.keys[0];
    uint32_t* vb = cur ? b.set_ids : b.vals;
    uint32_t* pb = cur ? b.metas : b.pos;
    hipLaunchKernelGGL(k_index_emit_batch, dim3((c.n + 255) / 256), dim3(256), 0, st, c, join_first, n_joins,
                       prefix_len, size_first, sf_base, cum, kb, vb, pb, sigs_out);
    sort_and_finish(st, c, n, universe, b, cur, plan, sigs, counters);
  } else {
    launch_fill_u32(st, b.offsets, (size_t)universe + 1, 0);
  }
  HS_CHECK_HIP(hipGetLastError());
}

}  // namespace hs
