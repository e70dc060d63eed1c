// Bitmap-signature filter (the bitmap technique) for gfx950.
//
// Reference: include/hyset/device_algorithms.hpp:58-118 (generate: one
// thread per set toggling bit token % b; filter: one thread per right set
// looping over every left set of the block, writing a dense blockSize^2
// matrix) and helpers.hpp:138-150 (change_bit, hamming_distance).
//
// A set's signature is the XOR-parity of its tokens hashed to b bits
// (b = 64, 128, 256), so popcount(sig_r ^ sig_s) <= |r xor s| and
// |r n s| <= (|r| + |s| - popcount) / 2.  The reference keeps a pair when
// (|r|+|s|-pop)/2.0 >= minoverlap, i.e. exactly when the integer test
// |r| + |s| - pop >= 2 * minoverlap holds; that integer form is used here.
// The bit layout inside a signature differs from the reference's
// (bit token % b sits in word (token % b) / 64), which leaves every popcount
// unchanged.
//
// Filter design: a workgroup takes 256 consecutive probes (one per lane,
// signature in registers).  Their windows overlap almost entirely (sizes
// are sorted), so the union window is streamed through LDS in chunks of
// signatures + sizes and every lane tests its own window slice against the
// broadcast LDS entries.  Survivors are verified on the spot by a merge.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "hs_internal.h"

namespace hs {

// One thread per set; the set's tokens are read SIG_CHUNK at a time with
// independent loads (one memory latency per chunk, not per token).
constexpr uint32_t SIG_CHUNK = 16;

template <int W>
__global__ void k_signatures(DevColl c, uint64_t* __restrict__ sigs) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.n) return;
  uint64_t w[W];
#pragma unroll
  for (int k = 0; k < W; ++k) w[k] = 0;
  const uint32_t* t = c.tokens + c.starts[i];
  const uint32_t l = c.sizes[i];
  for (uint32_t k0 = 0; k0 < l; k0 += SIG_CHUNK) {
    uint32_t v[SIG_CHUNK];
#pragma unroll
    for (uint32_t q = 0; q < SIG_CHUNK; ++q) v[q] = k0 + q < l ? t[k0 + q] : 0u;
#pragma unroll
    for (uint32_t q = 0; q < SIG_CHUNK; ++q) {
      if (k0 + q >= l) break;
      const uint32_t bit = v[q] & (64u * W - 1u);
#pragma unroll
      for (int b = 0; b < W; ++b)
        if ((bit >> 6) == (uint32_t)b) w[b] ^= 1ull << (bit & 63u);
    }
  }
#pragma unroll
  for (int k = 0; k < W; ++k) sigs[(size_t)i * W + k] = w[k];
}

// Wide signatures (512, 1024 bits) of long sets: one wave per set reads the
// set's tokens coalesced and toggles bits of an LDS copy of the signature
// with LDS atomic XOR (32-bit words; word k of the u64 layout is dwords 2k,
// 2k+1), then 2W lanes store it.
template <int W>
__global__ __launch_bounds__(256) void k_signatures_wave(DevColl c, uint64_t* __restrict__ sigs) {
  __shared__ uint32_t s[256 / WAVE][2 * W];
  const uint32_t wave = threadIdx.x / WAVE, lane = lane_id();
  const uint32_t i = blockIdx.x * (256 / WAVE) + wave;
  if (lane < 2 * W) s[wave][lane] = 0;
  __syncthreads();
  if (i < c.n) {
    const uint32_t* t = c.tokens + c.starts[i];
    const uint32_t l = c.sizes[i];
    for (uint32_t k = lane; k < l; k += WAVE) {
      const uint32_t bit = t[k] & (64u * W - 1u);
      atomicXor(&s[wave][bit >> 5], 1u << (bit & 31u));
    }
  }
  __syncthreads();
  if (i < c.n && lane < 2 * W) reinterpret_cast<uint32_t*>(sigs)[(size_t)i * 2 * W + lane] = s[wave][lane];
}

void launch_signatures(hipStream_t st, const DevColl& c, uint32_t words, uint64_t* sigs) {
  if (c.n == 0) return;
  dim3 g((c.n + 255) / 256), b(256), gw((c.n + 3) / 4);
  switch (words) {
    case 1: hipLaunchKernelGGL(k_signatures<1>, g, b, 0, st, c, sigs); break;
    case 2: hipLaunchKernelGGL(k_signatures<2>, g, b, 0, st, c, sigs); break;
    case 4: hipLaunchKernelGGL(k_signatures<4>, g, b, 0, st, c, sigs); break;
    case 8: hipLaunchKernelGGL(k_signatures_wave<8>, gw, b, 0, st, c, sigs); break;
    case 16: hipLaunchKernelGGL(k_signatures_wave<16>, gw, b, 0, st, c, sigs); break;
    default: throw Error{HS_E_INVALID, "signature width must be 64, 128, 256, 512 or 1024 bits"};
  }
  HS_CHECK_HIP(hipGetLastError());
}

// Full merge with the early-termination bound; exact overlap for passes.
__device__ __forceinline__ bool verify_merge(const uint32_t* r, uint32_t lr, const uint32_t* s, uint32_t ls,
                                             uint32_t mo, uint32_t* ov) {
  uint32_t i = 0, k = 0, found = 0;
  uint32_t a = r[0], b = s[0];
  while (i < lr && k < ls) {
    if (found + min(lr - i, ls - k) < mo) return false;
    bool eq = a == b;
    bool adv_a = a <= b;
    bool adv_b = a >= b;
    found += eq;
    i += adv_a;
    k += adv_b;
    if (adv_a && i < lr) a = r[i];
    if (adv_b && k < ls) b = s[k];
  }
  *ov = found;
  return found >= mo;
}

constexpr int BM_TPB = 256;

// The indexed signatures through the constant address space: every lane
// of a wave reads the same words, so they arrive by scalar loads into SGPRs
// (the kernel never writes them).
typedef const __attribute__((address_space(4))) uint64_t ConstU64;

// XOR-popcount of words [K0, K1) of the lane's signature against a
// wave-uniform one, added to pop: one popcount-with-accumulate per 32-bit
// half, chained.
template <int K0, int K1, int W>
__device__ __forceinline__ uint32_t bm_pop(const uint64_t (&sig)[W], const uint64_t (&ss)[W], uint32_t pop = 0) {
#pragma unroll
  for (int k = K0; k < K1; ++k) {
    const uint64_t x = sig[k] ^ ss[k];
    pop = (uint32_t)__popc((uint32_t)x) + pop;
    pop = (uint32_t)__popc((uint32_t)(x >> 32)) + pop;
  }
  return pop;
}

// Tests per loop round: the first halves of the round's signatures sit in
// SGPRs, at most 32 of them (8 tests of up to 256 bits, 4 of 512, 2 of 1024).
template <int W>
constexpr uint32_t bm_unroll() {
  constexpr int H = W > 1 ? W / 2 : 1;
  return H >= 8 ? 2u : H >= 4 ? 4u : 8u;
}

// One wave per 64 consecutive probes (sets are sorted by size, so their
// length windows nearly coincide): the wave walks the union of its lanes'
// windows once, in segments over which the indexed sets' size and the set of
// lanes whose window holds the id are both constant, so a pair test is the
// XOR-popcount of the lane's signature against a wave-uniform one and one
// compare with a per-lane slack fixed for the segment (-1 for lanes outside
// their window).  Survivors (rare) are merged by their lanes.  With few
// probes, A.parts waves share one group of 64 probes, each walking an equal
// share of the union window, so the device still holds enough waves.
template <int W>
__global__ __launch_bounds__(BM_TPB) void k_bitmap(BitmapArgs A, uint32_t ntiles) {
  __shared__ unsigned long long s_stats[2];
  const uint32_t parts = A.parts ? A.parts : 1u;
  const uint32_t job = xcd_tile(blockIdx.x, gridDim.x, ntiles * parts);
  if (job == 0xFFFFFFFFu) return;
  const uint32_t tile = job / parts, part = job % parts;
  const uint32_t tid = threadIdx.x;
  if (tid < 2) s_stats[tid] = 0;
  const uint32_t p = A.probe_first + tile * BM_TPB + tid;
  const bool valid = p <= A.probe_last;
  uint64_t sig[W];
  uint32_t lr = 0, lo = 0xFFFFFFFFu, hi = 0;
  if (valid) {
    lr = A.probe.sizes[p];
    const uint32_t last = A.size_first_len - 1;
    const uint32_t mins = min(A.bounds.minsize[lr], last);
    const uint32_t maxs1 = min(A.bounds.maxsize[lr], last - 1) + 1;
    lo = max(A.size_first[mins], A.idx_first);
    hi = min(A.size_first[maxs1], A.idx_last + 1);
    if (A.self) hi = min(hi, p);
    if (lo >= hi) {
      lo = 0xFFFFFFFFu;
      hi = 0;
    }
#pragma unroll
    for (int k = 0; k < W; ++k) sig[k] = A.probe_sigs[(size_t)p * W + k];
  } else {
#pragma unroll
    for (int k = 0; k < W; ++k) sig[k] = 0;
  }
  __syncthreads();  // s_stats zeroed
  const uint32_t* rtok = valid ? A.probe.tokens + A.probe.starts[p] : nullptr;
  // this wave's share of its lanes' union window
  const uint32_t ulo = __builtin_amdgcn_readfirstlane(wave_min(lo));
  const uint32_t uhi = __builtin_amdgcn_readfirstlane(wave_max(hi));
  const uint32_t span = uhi > ulo ? uhi - ulo : 0u;
  const uint32_t wlo = ulo + (uint32_t)((uint64_t)span * part / parts);
  const uint32_t whi = ulo + (uint32_t)((uint64_t)span * (part + 1) / parts);
  const uint32_t sf_last = A.size_first_len - 1;
  uint32_t survivors = 0, tests = 0;
  for (uint32_t sidx = wlo; sidx < whi;) {
    // the size class of sidx and where it ends
    const uint32_t ls = __builtin_amdgcn_readfirstlane(A.indexed.sizes[sidx]);
    const uint32_t cend = min(whi, A.size_first[min(ls + 1, sf_last)]);
    // the next id at which some lane's window opens or closes
    const uint32_t nb = lo > sidx ? lo : hi > sidx ? hi : 0xFFFFFFFFu;
    const uint32_t send = min(cend, __builtin_amdgcn_readfirstlane(wave_min(nb)));
    const bool in = lo <= sidx && sidx < hi;  // constant over [sidx, send)
    int32_t slack = -1;
    uint32_t mo = 0;
    if (in) {
      mo = dev_minoverlap(A.bounds, lr, ls);
      slack = (int32_t)(lr + ls) - 2 * (int32_t)mo;
      tests += send - sidx;
    }
    ConstU64* cs = (ConstU64*)(A.indexed_sigs) + (size_t)sidx * W;
    // a survivor of id q: its lane merges the two sets
    auto survive = [&](uint32_t q) {
      ++survivors;
      uint32_t ov = 0;
      const uint32_t* s = A.indexed.tokens + A.indexed.starts[q];
      if (verify_merge(rtok, lr, s, ls, mo, &ov)) {
        unsigned long long idx = atomicAdd(&A.out.counters[CNT_PAIRS], 1ull);
        if (idx < A.out.capacity) {
          A.out.r[idx] = p;
          A.out.s[idx] = q;
          A.out.ov[idx] = ov;
        }
      }
    };
    // a round tests the first H words of BM_UNROLL signatures (a lower
    // bound of the whole popcount, so a valid filter), and only a round
    // where some lane passes that reads and tests the rest
    constexpr int H = W > 1 ? W / 2 : 1;
    constexpr uint32_t BM_UNROLL = bm_unroll<W>();
    for (; sidx + BM_UNROLL <= send; sidx += BM_UNROLL, cs += BM_UNROLL * W) {
      uint64_t t[BM_UNROLL][W];
#pragma unroll
      for (uint32_t u = 0; u < BM_UNROLL; ++u)
#pragma unroll
        for (int k = 0; k < H; ++k) t[u][k] = cs[u * W + k];
      uint64_t any = 0;  // lanes with a passing test among the round's
      uint32_t part[BM_UNROLL];
#pragma unroll
      for (uint32_t u = 0; u < BM_UNROLL; ++u) {
        part[u] = bm_pop<0, H>(sig, t[u]);
        any |= __ballot((int32_t)part[u] <= slack);
      }
      if (any == 0) continue;  // wave-uniform; survivors are rare
      if (H < W) {  // the rest of the round's signatures, straight-line
#pragma unroll
        for (uint32_t u = 0; u < BM_UNROLL; ++u)
#pragma unroll
          for (int k = H; k < W; ++k) t[u][k] = cs[u * W + k];
        any = 0;
#pragma unroll
        for (uint32_t u = 0; u < BM_UNROLL; ++u) {
          part[u] = bm_pop<H, W>(sig, t[u], part[u]);
          any |= __ballot((int32_t)part[u] <= slack);
        }
        if (any == 0) continue;
      }
      for (uint32_t u = 0; u < BM_UNROLL; ++u)
        if ((int32_t)part[u] <= slack) survive(sidx + u);
    }
    for (; sidx < send; ++sidx, cs += W) {
      uint64_t t[W];
#pragma unroll
      for (int k = 0; k < W; ++k) t[k] = cs[k];
      const bool pass = (int32_t)bm_pop<0, W>(sig, t) <= slack;
      if (__ballot(pass) == 0) continue;
      if (pass) survive(sidx);
    }
  }
  survivors = wave_sum(survivors);
  tests = wave_sum(tests);
  if (lane_id() == 0) {
    if (survivors) atomicAdd(&s_stats[0], (unsigned long long)survivors);
    if (tests) atomicAdd(&s_stats[1], (unsigned long long)tests);
  }
  __syncthreads();
  if (tid == 0) {
    unsigned long long* slot = stat_slot(A.out.counters);
    if (s_stats[0]) atomicAdd(&slot[CNT_CANDS], s_stats[0]);
    if (s_stats[1]) atomicAdd(&slot[CNT_VISITS], s_stats[1]);
  }
}

void launch_bitmap(hipStream_t st, const BitmapArgs& a) {
  if (a.probe.n == 0 || a.probe_first > a.probe_last) return;
  const uint32_t nprobes = a.probe_last - a.probe_first + 1;
  const uint32_t ntiles = (nprobes + BM_TPB - 1) / BM_TPB;
  const uint64_t jobs = (uint64_t)ntiles * (a.parts ? a.parts : 1u);
  HS_REQUIRE(jobs < 0x7FFFFFF0ull, HS_E_UNSUPPORTED, "bitmap launch too large");
  const uint32_t grid = (uint32_t)((jobs + 7) / 8 * 8);
  switch (a.words) {
    case 1: hipLaunchKernelGGL(k_bitmap<1>, dim3(grid), dim3(BM_TPB), 0, st, a, ntiles); break;
    case 2: hipLaunchKernelGGL(k_bitmap<2>, dim3(grid), dim3(BM_TPB), 0, st, a, ntiles); break;
    case 4: hipLaunchKernelGGL(k_bitmap<4>, dim3(grid), dim3(BM_TPB), 0, st, a, ntiles); break;
    case 8: hipLaunchKernelGGL(k_bitmap<8>, dim3(grid), dim3(BM_TPB), 0, st, a, ntiles); break;
    case 16: hipLaunchKernelGGL(k_bitmap<16>, dim3(grid), dim3(BM_TPB), 0, st, a, ntiles); break;
    default: throw Error{HS_E_INVALID, "bitmap width must be 64, 128, 256, 512 or 1024 bits"};
  }
  HS_CHECK_HIP(hipGetLastError());
}

// ---- integer-op peak of the XOR-popcount test, for the bitmap roofline.
// Every lane runs 8 independent chains a = popcount(a ^ m) + a (one XOR and
// one popcount-with-accumulate per step, 2 VALU operations); the result is
// stored only when it equals a value it cannot take, which keeps the
// chains live without per-lane stores.
__global__ __launch_bounds__(256) void k_popc_peak(uint32_t seed, uint32_t iters, uint32_t* out) {
  uint32_t a[8];
  const uint32_t m = seed * 0x9E3779B9u + 1u;
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = seed + (blockIdx.x * blockDim.x + threadIdx.x) * 8u + (uint32_t)k;
  for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] = (uint32_t)__popc(a[k] ^ m) + a[k];
    }
  }
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += a[k];
  if (s == 0xFFFFFFFFu) out[0] = s;
}

void launch_popc_peak(hipStream_t st, uint32_t grid, uint32_t iters, uint32_t* out) {
  hipLaunchKernelGGL(k_popc_peak, dim3(grid), dim3(256), 0, st, 12345u, iters, out);
  HS_CHECK_HIP(hipGetLastError());
}

}  // namespace hs
