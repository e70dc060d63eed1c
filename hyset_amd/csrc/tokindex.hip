// The packed token-major index of a self-join: every set's probe-prefix
// entries, one 64-bit word each (position, size, set id and key, TokPack),
// sorted by (key, set id).
//
// Reference: include/hyset/index.hpp:217-313 builds its inverted index by a
// token histogram, thrust::exclusive_scan, an atomic scatter of the
// entries and thrust::sort by (token, setID).  Here the entries are never
// counted per token: the join only needs them grouped by key in set-id
// order, which a stable LSD sort over the key digits gives when the
// entries start in set-id order.
//
// 1. k_pack_emit: one thread per set (a grid-stride loop over a grid of a
//    few workgroups per CU) reads the set's prefix (all of it when it also
//    computes the set's 64-bit signature), writes its entries' words in
//    set-id order and counts every key digit of every pass in workgroup
//    histograms; each workgroup stores its histograms as one column of a
//    digit-major table.  It also clears the look-back flags and tile
//    tickets of the passes and the join counters.
// 2. k_pack_hist_sum: one workgroup per (pass, digit) row sums the row.
// 3. k_pack_pass, once per digit: a workgroup takes the next tile of
//    PK_TILE entries by ticket (so every tile it waits for has a running
//    workgroup), loads it, counts its digits, publishes the counts, ranks
//    the tile by digit in LDS, and finds where each digit's run goes by
//    decoupled look-back over the earlier tiles' published counts: a
//    tile's flag word per digit carries either its own count or the
//    inclusive count of all tiles up to it.  The sorted tile leaves as
//    contiguous runs per digit.  No per-pass histogram or scan launches.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "hs_internal.h"

namespace hs {

namespace {
// 8192-entry tiles, 16 waves of 8 rounds: per pass on the 3.0M-entry
// index of the 1M-set self-join at Jaccard 0.8, 23.5 us (256 x 16: 33.8,
// 512 x 8: 30.6, 1024 x 4: 29.1, 512 x 16: 25.9, 512 x 24: 24.9)
constexpr int PK_TPB = 1024;
constexpr int PK_ROUNDS = 8;
constexpr uint32_t PK_TILE = PK_TPB * PK_ROUNDS;  // entries per tile
constexpr int PK_WAVES = PK_TPB / WAVE;
constexpr int PK_EMIT_TPB = 256;
constexpr uint32_t PK_MAX_PASSES = 5;
constexpr uint32_t PK_CHUNK = 16;  // tokens per load round of the emit
// look-back flag words: 2 flag bits over a 30-bit count
constexpr uint32_t PK_AGG = 1u << 30, PK_INC = 2u << 30, PK_COUNT = (1u << 30) - 1;
// a waiting lane gives up (and flags the join as failed) after this many
// polls; a correct run waits a few microseconds at most
constexpr uint32_t PK_SPIN_LIMIT = 1u << 22;
constexpr uint32_t PK_LOOK = 8;  // look-back flags a digit's lane reads at once

uint32_t bits_for(uint64_t v) {  // bits to hold every value <= v (at least 1)
  uint32_t b = 1;
  while (b < 64 && (v >> b)) ++b;
  return b;
}

// The words of the prefix entries of the sets [first, first + nsets) in
// set-id order (set i's at cum[l] + (i - first id of size l in the range)
// * prefix_len[l]), and histograms of every key digit of every pass; hist
// is digit-major: row p * NB + d, one column per workgroup.
template <int D>
__global__ __launch_bounds__(PK_EMIT_TPB) void k_pack_emit(
    DevColl c, uint32_t first, uint32_t nsets, const uint32_t* __restrict__ prefix_len,
    const uint32_t* __restrict__ flag_len, const uint32_t* __restrict__ size_first, const uint32_t* __restrict__ cum,
    TokPack pk, uint64_t* __restrict__ words, uint64_t* __restrict__ sig_out, uint32_t* __restrict__ hist,
    uint32_t* __restrict__ zero, uint32_t n_zero, unsigned long long* __restrict__ counters) {
  constexpr uint32_t NB = 1u << D;
  __shared__ uint32_t h[PK_MAX_PASSES * NB];
  for (uint32_t k = threadIdx.x; k < PK_MAX_PASSES * NB; k += PK_EMIT_TPB) h[k] = 0;
  const uint32_t gtid = blockIdx.x * PK_EMIT_TPB + threadIdx.x, stride = gridDim.x * PK_EMIT_TPB;
  for (uint32_t k = gtid; k < n_zero; k += stride) zero[k] = 0;
  for (uint32_t k = gtid; k < CNT_WORDS; k += stride) counters[k] = 0;  // (a grid of one workgroup included)
  __syncthreads();
  for (uint32_t q = gtid; q < nsets; q += stride) {
    const uint32_t i = first + q;
    const uint32_t l = c.sizes[i];
    const uint32_t pl = prefix_len[l], fl = flag_len[l];
    const uint32_t* t = c.tokens + c.starts[i];
    const uint32_t len = sig_out ? l : pl;
    uint64_t* o = words + cum[l] + (size_t)(i - max(size_first[l], first)) * pl;
    const uint64_t fixed = ((uint64_t)i << (pk.pb + pk.sb)) | ((uint64_t)l << pk.pb);
    uint64_t w = 0;
    for (uint32_t k0 = 0; k0 < len; k0 += PK_CHUNK) {
      uint32_t v[PK_CHUNK];
#pragma unroll
      for (uint32_t j = 0; j < PK_CHUNK; ++j) v[j] = k0 + j < len ? t[k0 + j] : 0u;
#pragma unroll
      for (uint32_t j = 0; j < PK_CHUNK; ++j) {
        const uint32_t k = k0 + j;
        if (k < pl) {
          const uint32_t key = (v[j] << 1) | (k >= fl ? 1u : 0u);
          o[k] = ((uint64_t)key << pk.kshift) | fixed | k;
          for (uint32_t p = 0; p < pk.passes; ++p) atomicAdd(&h[p * NB + ((key >> (p * D)) & (NB - 1))], 1u);
        }
        if (k < len) w ^= 1ull << (v[j] & 63u);
      }
    }
    if (sig_out) sig_out[i] = w;
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < pk.passes * NB; k += PK_EMIT_TPB)
    hist[(size_t)k * gridDim.x + blockIdx.x] = h[k];
}

// totals[row] = sum of the row's `cols` workgroup counts.
__global__ __launch_bounds__(256) void k_pack_hist_sum(const uint32_t* __restrict__ hist, uint32_t cols,
                                                       uint32_t* __restrict__ totals) {
  __shared__ uint32_t part[256 / WAVE];
  const uint32_t* row = hist + (size_t)blockIdx.x * cols;
  uint32_t s = 0;
  for (uint32_t k = threadIdx.x; k < cols; k += 256) s += row[k];
  s = wave_sum(s);
  if (lane_id() == 0) part[threadIdx.x / WAVE] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < 256 / WAVE; ++w) t += part[w];
    totals[blockIdx.x] = t;
  }
}

struct PackPass {
  const uint64_t* in;             // words in the previous pass's order
  uint64_t* out;
  uint32_t n;                     // entries
  uint32_t shift;                 // of this pass's digit in the word
  const uint32_t* totals;         // this pass's digit totals
  uint32_t* flags;                // this pass's look-back words, ntiles x NB, zeroed
  uint32_t* ticket;               // this pass's tile counter, zeroed
  unsigned long long* counters;   // CNT_INDEX_ERR
};

// One stable scatter pass by the digit at A.shift (see the file comment).
template <int D>
__global__ __launch_bounds__(PK_TPB) void k_pack_pass(PackPass A) {
  constexpr uint32_t NB = 1u << D;
  static_assert(NB <= PK_TPB, "one digit per thread");
  __shared__ uint64_t s_w[PK_TILE];
  __shared__ uint32_t s_cnt[PK_WAVES][NB];  // per-wave digit counts, then the wave's cursors
  __shared__ uint32_t s_lo[NB];             // first slot of digit d in the sorted tile
  __shared__ uint32_t s_base[NB];           // output position of that slot
  __shared__ uint32_t s_tmp[PK_TPB / WAVE + 1];
  __shared__ uint32_t s_tile;
  const uint32_t tid = threadIdx.x;
  const uint32_t wave = tid / WAVE, lane = lane_id();
  if (tid == 0) s_tile = atomicAdd(A.ticket, 1u);
  if (tid < NB) {
#pragma unroll
    for (int w = 0; w < PK_WAVES; ++w) s_cnt[w][tid] = 0;
  }
  // the digit's first output position over all tiles
  uint32_t all;
  const uint32_t gbase = block_exclusive_scan<PK_TPB>(tid < NB ? A.totals[tid] : 0u, s_tmp, &all);
  // (the scan's barriers publish s_tile and the zeroed counts)
  const uint32_t t = s_tile;
  if ((uint64_t)t * PK_TILE >= A.n) {  // more workgroups than tiles: never in a correct launch
    if (tid == 0) atomicOr(&A.counters[CNT_INDEX_ERR], 2ull);
    return;
  }
  const uint32_t tbase = t * PK_TILE;
  const uint32_t wbase = tbase + wave * (PK_ROUNDS * WAVE);
  const uint32_t tn = min(PK_TILE, A.n - tbase);
  uint64_t w[PK_ROUNDS];
#pragma unroll
  for (int r = 0; r < PK_ROUNDS; ++r) {
    const uint32_t e = wbase + r * WAVE + lane;
    w[r] = e < A.n ? A.in[e] : 0ull;
  }
#pragma unroll
  for (int r = 0; r < PK_ROUNDS; ++r)
    if (wbase + r * WAVE + lane < A.n) atomicAdd(&s_cnt[wave][(uint32_t)(w[r] >> A.shift) & (NB - 1)], 1u);
  __syncthreads();
  // this tile's count of digit tid, published at once: later tiles read it
  uint32_t tot = 0;
  if (tid < NB) {
#pragma unroll
    for (int v = 0; v < PK_WAVES; ++v) tot += s_cnt[v][tid];
    __hip_atomic_store(&A.flags[(size_t)t * NB + tid], (t == 0 ? PK_INC : PK_AGG) | tot, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  // the tile sorted by digit: digit d occupies [s_lo[d], s_lo[d] + tot(d)),
  // wave v's entries of d start after those of the earlier waves
  uint32_t run = block_exclusive_scan<PK_TPB>(tot, s_tmp, &all);
  if (tid < NB) {
    s_lo[tid] = run;
#pragma unroll
    for (int v = 0; v < PK_WAVES; ++v) {
      const uint32_t x = s_cnt[v][tid];
      s_cnt[v][tid] = run;
      run += x;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PK_ROUNDS; ++r) {
    const bool valid = wbase + r * WAVE + lane < A.n;
    const uint32_t d = (uint32_t)(w[r] >> A.shift) & (NB - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < D; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    if (valid) s_w[s_cnt[wave][d] + lanes_below(peers)] = w[r];
    // the leader of each digit advances the wave's cursor; the next round's
    // reads come after it in this wave's program order
    __builtin_amdgcn_wave_barrier();
    if (valid && lanes_below(peers) == 0) s_cnt[wave][d] += (uint32_t)__popcll(peers);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  // decoupled look-back: digit tid's entries in the tiles before t, the
  // flags of PK_LOOK earlier tiles read at once (one memory latency where a
  // walk of one flag at a time waited for each: 7 us of a 20 us tile)
  if (tid < NB) {
    uint32_t excl = 0;
    if (t > 0) {
      uint32_t p = t, spins = 0;  // the tiles p-1, p-2, ... are still to be read
      for (;;) {
        uint32_t f[PK_LOOK];
#pragma unroll
        for (uint32_t j = 0; j < PK_LOOK; ++j)
          f[j] = p > j ? __hip_atomic_load(&A.flags[(size_t)(p - 1 - j) * NB + tid], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT)
                       : PK_INC;  // (before tile 0, never reached: tile 0's flag is inclusive)
        bool stop = false, fin = false;
        uint32_t used = 0;
#pragma unroll
        for (uint32_t j = 0; j < PK_LOOK; ++j) {
          if (!stop) {
            if ((f[j] & ~PK_COUNT) == 0) {
              stop = true;  // not published yet: read again from this tile
            } else {
              excl += f[j] & PK_COUNT;
              ++used;
              if (f[j] & PK_INC) stop = fin = true;
            }
          }
        }
        if (fin) break;
        p -= used;
        if (used < PK_LOOK) {
          if (++spins > PK_SPIN_LIMIT) {
            atomicOr(&A.counters[CNT_INDEX_ERR], 1ull);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __hip_atomic_store(&A.flags[(size_t)t * NB + tid], PK_INC | (excl + tot), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    s_base[tid] = gbase + excl;
  }
  __syncthreads();
  // slot i of the sorted tile goes to base(d) + (i - lo(d))
  for (uint32_t i = tid; i < tn; i += PK_TPB) {
    const uint64_t v = s_w[i];
    const uint32_t d = (uint32_t)(v >> A.shift) & (NB - 1);
    const uint32_t o = s_base[d] + (i - s_lo[d]);
    if (o < A.n) A.out[o] = v;
    else atomicOr(&A.counters[CNT_INDEX_ERR], 4ull);  // digit totals disagree with the tiles
  }
}

template <int D>
void launch_pack(hipStream_t st, const DevColl& c, uint32_t first, uint32_t nsets, const uint32_t* prefix_len,
                 const uint32_t* flag_len, const uint32_t* size_first, const uint32_t* cum, uint32_t n,
                 const TokPack& pk, IndexBuild& ib, uint64_t* sigs_out, unsigned long long* counters,
                 uint32_t num_cus, hipEvent_t done) {
  constexpr uint32_t NB = 1u << D;
  const uint32_t ntiles = (n + PK_TILE - 1) / PK_TILE;
  const uint32_t cols =
      std::max<uint32_t>(1, std::min<uint32_t>((nsets + PK_EMIT_TPB - 1) / PK_EMIT_TPB, 4 * std::max(num_cus, 1u)));
  ib.packed[0].ensure((size_t)n * 8);
  ib.packed[1].ensure((size_t)n * 8);
  ib.pk_hist.ensure(((size_t)pk.passes * NB * cols + (size_t)pk.passes * NB) * 4);
  const size_t n_flags = (size_t)pk.passes * ntiles * NB;
  HS_REQUIRE(n_flags + pk.passes < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "packed index too large");
  ib.pk_sync.ensure((n_flags + pk.passes) * 4);
  uint32_t* hist = ib.pk_hist.as<uint32_t>();
  uint32_t* totals = hist + (size_t)pk.passes * NB * cols;
  uint32_t* flags = ib.pk_sync.as<uint32_t>();
  uint32_t* tickets = flags + n_flags;
  uint64_t* buf[2] = {ib.packed[0].as<uint64_t>(), ib.packed[1].as<uint64_t>()};
  hipLaunchKernelGGL((k_pack_emit<D>), dim3(cols), dim3(PK_EMIT_TPB), 0, st, c, first, nsets, prefix_len, flag_len,
                     size_first, cum, pk, buf[0], sigs_out, hist, flags, (uint32_t)(n_flags + pk.passes), counters);
  hipLaunchKernelGGL(k_pack_hist_sum, dim3(pk.passes * NB), dim3(256), 0, st, hist, cols, totals);
  PackPass a{};
  a.n = n;
  a.counters = counters;
  for (uint32_t p = 0; p < pk.passes; ++p) {
    a.in = buf[p & 1];
    a.out = buf[(p + 1) & 1];
    a.shift = pk.kshift + p * pk.digit_bits;
    a.totals = totals + (size_t)p * NB;
    a.flags = flags + (size_t)p * ntiles * NB;
    a.ticket = tickets + p;
    hipEvent_t ev = p + 1 == pk.passes ? done : nullptr;
    hipExtLaunchKernelGGL((k_pack_pass<D>), dim3(ntiles), dim3(PK_TPB), 0, st, nullptr, ev, 0, a);
  }
  ib.packed_sorted = buf[pk.passes & 1];
}
}  // namespace

bool tok_pack_plan(uint32_t n_sets, uint32_t maxlen, uint32_t universe, uint64_t entries, TokPack* out) {
  TokPack p{};
  p.pb = bits_for(maxlen);  // positions < maxlen
  p.sb = bits_for(maxlen);
  p.ib = bits_for(n_sets > 0 ? n_sets - 1 : 0);
  p.kshift = p.pb + p.sb + p.ib;
  p.key_bits = bits_for(universe > 0 ? 2ull * universe - 1 : 1);
  if (universe >= 0x80000000u || p.kshift + p.key_bits > 64 || entries >= (1ull << 30) || p.ib > 32) return false;
  p.passes = (p.key_bits + 7) / 8;
  p.digit_bits = std::max<uint32_t>(4, (p.key_bits + p.passes - 1) / p.passes);
  if (p.passes > PK_MAX_PASSES || p.digit_bits > 8) return false;
  if (out) *out = p;
  return true;
}

void build_token_index(hipStream_t st, const DevColl& c, uint32_t first, uint32_t last, const uint32_t* prefix_len,
                       const uint32_t* size_first, const uint32_t* cum, uint64_t entries, const TokPack& pk,
                       IndexBuild& ib, const uint32_t* flag_len, uint64_t* sigs_out, unsigned long long* counters,
                       uint32_t num_cus, hipEvent_t done) {
  HS_REQUIRE(entries < (1ull << 30), HS_E_UNSUPPORTED, "packed index limited to 2^30 entries");
  HS_REQUIRE(counters, HS_E_INVALID, "packed index: counters required");
  HS_REQUIRE(first <= last && last < c.n, HS_E_INVALID, "packed index: empty range");
  ib.ent_words = 0;
  ib.universe = 0;
  ib.entries = entries;
  ib.packed_sorted = nullptr;
  const uint32_t n = (uint32_t)entries;
  const uint32_t nsets = last - first + 1;
  if (n == 0) {
    // no prefix entries: the counters still start at zero, the signatures
    // are still wanted
    launch_fill_u64(st, counters, CNT_WORDS, 0);
    if (sigs_out) launch_signatures(st, c, 1, sigs_out);  // (sigs_out: the whole collection's)
    if (done) HS_CHECK_HIP(hipEventRecord(done, st));
    return;
  }
#define HS_PACK(D)                                                                                             \
  case D:                                                                                                      \
    launch_pack<D>(st, c, first, nsets, prefix_len, flag_len, size_first, cum, n, pk, ib, sigs_out, counters, \
                   num_cus, done);                                                                             \
    break;
  switch (pk.digit_bits) {
    HS_PACK(4)
    HS_PACK(5)
    HS_PACK(6)
    HS_PACK(7)
    HS_PACK(8)
    default: throw Error{HS_E_INVALID, "packed index: digit width out of range"};
  }
#undef HS_PACK
  HS_CHECK_HIP(hipGetLastError());
}

}  // namespace hs
