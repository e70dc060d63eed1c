// Wave64 / workgroup building blocks shared by the kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hs {

constexpr int WAVE = 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    uint32_t o = __shfl_up(v, d);
    if (lane >= (uint32_t)d) v += o;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) v = min(v, (uint32_t)__shfl_xor(v, d));
  return v;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) v = max(v, (uint32_t)__shfl_xor(v, d));
  return v;
}

// Exclusive scan of one value per thread over a workgroup of NT threads.
// `tmp` must hold NT/64 + 1 words of LDS.  Returns the exclusive prefix;
// *total receives the workgroup sum.  Contains barriers: call uniformly.
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* tmp, uint32_t* total) {
  constexpr int NW = NT / WAVE;
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / WAVE;
  uint32_t inc = wave_inclusive_scan(v);
  if (lane == WAVE - 1) tmp[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      uint32_t x = tmp[w];
      tmp[w] = run;
      run += x;
    }
    tmp[NW] = run;
  }
  __syncthreads();
  uint32_t res = tmp[wave] + inc - v;
  *total = tmp[NW];
  __syncthreads();
  return res;
}

// Exclusive running maximum over the workgroup's threads (0 for thread 0);
// tmp holds NT / WAVE + 1 words.  Contains barriers: call uniformly.
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_max(uint32_t v, uint32_t* tmp) {
  constexpr int NW = NT / WAVE;
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x / WAVE;
  uint32_t inc = v;
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d);
    if (lane >= (uint32_t)d) inc = max(inc, o);
  }
  const uint32_t exc_w = __shfl_up(inc, 1);
  if (lane == WAVE - 1) tmp[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const uint32_t x = tmp[w];
      tmp[w] = run;
      run = max(run, x);
    }
  }
  __syncthreads();
  const uint32_t res = max(tmp[wave], lane ? exc_w : 0u);
  __syncthreads();
  return res;
}

// First index i in [lo, hi) with a[i] >= key (a ascending); hi if none.
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t key) {
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Largest i in [0, n] with a[i] <= key, for ascending a[0..n] with a[0] <= key.
__device__ __forceinline__ uint32_t segment_of(const uint32_t* a, uint32_t n, uint32_t key) {
  uint32_t lo = 0, hi = n;  // answer in [lo, hi)
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= key) lo = mid;
    else hi = mid;
  }
  return lo;
}

// segment_of searched forward from a known lower position `from` (a[from]
// <= key): gallop from it, then bisect inside the last gap -- about
// 2 log2(distance) reads instead of log2(n) when the answer is near.
__device__ __forceinline__ uint32_t segment_from(const uint32_t* a, uint32_t n, uint32_t key, uint32_t from) {
  uint32_t lo = from, step = 1, hi = from + 1;
  while (hi < n && a[hi] <= key) {
    lo = hi;
    step <<= 1;
    hi = lo + step;
  }
  if (hi > n) hi = n;
  while (hi - lo > 1) {  // answer in [lo, hi)
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= key) lo = mid;
    else hi = mid;
  }
  return lo;
}

// blockIdx.x -> tile such that each XCD (workgroups are dealt round-robin
// over the 8 XCDs) receives a contiguous run of tiles, keeping neighbouring
// tiles -- which touch the same posting lists and sets -- in one L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t bid, uint32_t grid, uint32_t ntiles) {
  constexpr uint32_t NXCD = 8;
  uint32_t per = (grid + NXCD - 1) / NXCD;
  uint32_t t = (bid % NXCD) * per + bid / NXCD;
  return t < ntiles ? t : 0xFFFFFFFFu;
}

// Append one (r, s, overlap) per passing lane: one atomic per wave.
__device__ __forceinline__ void emit_pairs(bool pass, uint32_t r, uint32_t s, uint32_t ov, uint32_t* out_r,
                                           uint32_t* out_s, uint32_t* out_ov, unsigned long long* counter,
                                           uint64_t capacity) {
  uint64_t m = __ballot(pass);
  if (m == 0) return;
  uint32_t leader = __ffsll((unsigned long long)m) - 1;
  unsigned long long base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
  uint32_t lo = __shfl((uint32_t)base, leader);
  uint32_t hi = __shfl((uint32_t)(base >> 32), leader);
  base = ((unsigned long long)hi << 32) | lo;
  if (pass) {
    unsigned long long idx = base + lanes_below(m);
    if (idx < capacity) {
      out_r[idx] = r;
      out_s[idx] = s;
      out_ov[idx] = ov;
    }
  }
}

}  // namespace hs
