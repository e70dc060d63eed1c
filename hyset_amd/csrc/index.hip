// Inverted-index build over set prefixes, and the device-wide scan it uses.
//
// Reference: include/hyset/index.hpp:217-313 (histogram, thrust::exclusive_scan,
// create_index with atomic scatter, thrust::sort by (token, setID)).
// Here: count -> exclusive scan -> atomic scatter of set ids -> per-list sort
// by set id, so list(t) = set_ids[offsets[t] .. offsets[t+1]) ascending --
// the order of the reference's sorted host index (index.hpp:278-279).  Lists
// hold 4-byte set ids only; positions are not needed by the device filters.
#include <hip/hip_runtime.h>

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
// Three phases: per-tile sums, one-workgroup scan of the tile sums, per-tile
// scan with the tile offset.  A tile is 256 threads x 16 elements.

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

// Exclusive scan of `sums[0..nb)` in place by a single workgroup.
__global__ __launch_bounds__(SC_TPB) void k_scan_sums(uint32_t* sums, uint32_t nb) {
  __shared__ uint32_t tmp[SC_TPB / WAVE + 1];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < nb; base += SC_TPB * SC_IPT) {
    uint32_t v[SC_IPT];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < SC_IPT; ++k) {
      uint32_t i = base + threadIdx.x * SC_IPT + k;
      v[k] = i < nb ? sums[i] : 0;
      local += v[k];
    }
    uint32_t total;
    uint32_t pre = block_exclusive_scan<SC_TPB>(local, tmp, &total);
    uint32_t run = carry + pre;
#pragma unroll
    for (int k = 0; k < SC_IPT; ++k) {
      uint32_t i = base + threadIdx.x * SC_IPT + k;
      if (i < nb) sums[i] = run;
      run += v[k];
    }
    carry += total;
    __syncthreads();
  }
}

__global__ __launch_bounds__(SC_TPB) void k_scan_apply(const uint32_t* __restrict__ in, size_t n,
                                                       const uint32_t* __restrict__ sums, uint32_t* __restrict__ out,
                                                       uint32_t* __restrict__ out2) {
  __shared__ uint32_t tile[SC_TILE];
  __shared__ uint32_t tmp[SC_TPB / WAVE + 1];
  size_t base = (size_t)blockIdx.x * SC_TILE;
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
  uint32_t run = sums[blockIdx.x] + block_exclusive_scan<SC_TPB>(local, tmp, &total);
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
  HS_REQUIRE(nb < 0xFFFFFFFFull, HS_E_UNSUPPORTED, "scan too large");
  scratch.ensure(nb * sizeof(uint32_t));
  uint32_t* sums = scratch.as<uint32_t>();
  hipLaunchKernelGGL(k_scan_tile_sums, dim3((unsigned)nb), dim3(SC_TPB), 0, st, in, n, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SC_TPB), 0, st, sums, (uint32_t)nb);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SC_TPB), 0, st, in, n, sums, out, out2);
  HS_CHECK_HIP(hipGetLastError());
}

// ------------------------------------------------------------ index build

// One thread per set: count every prefix token.  Sets are contiguous in the
// CSR, so a wave reads ~64 neighbouring prefixes.
__global__ void k_index_count(DevColl c, uint32_t first, uint32_t last, const uint32_t* __restrict__ prefix_len,
                              uint32_t* __restrict__ counts) {
  uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > last) return;
  uint32_t st = c.starts[i];
  uint32_t pl = prefix_len[c.sizes[i]];
  for (uint32_t k = 0; k < pl; ++k) atomicAdd(&counts[c.tokens[st + k]], 1u);
}

__global__ void k_index_scatter(DevColl c, uint32_t first, uint32_t last, const uint32_t* __restrict__ prefix_len,
                                uint32_t* __restrict__ cursor, uint32_t* __restrict__ set_ids) {
  uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x;
  if (i > last) return;
  uint32_t st = c.starts[i];
  uint32_t pl = prefix_len[c.sizes[i]];
  for (uint32_t k = 0; k < pl; ++k) {
    uint32_t pos = atomicAdd(&cursor[c.tokens[st + k]], 1u);
    set_ids[pos] = i;
  }
}

constexpr uint32_t SHORT_LIST = 32;

// One thread per token: insertion-sort lists of up to SHORT_LIST ids in
// place; queue longer lists for k_sort_long_lists.
__global__ void k_sort_short_lists(const uint32_t* __restrict__ offsets, uint32_t universe, uint32_t* set_ids,
                                   uint32_t* long_queue) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= universe) return;
  uint32_t a = offsets[t], b = offsets[t + 1];
  uint32_t len = b - a;
  if (len < 2) return;
  if (len > SHORT_LIST) {
    uint32_t q = atomicAdd(&long_queue[0], 1u);
    long_queue[1 + q] = t;
    return;
  }
  uint32_t* p = set_ids + a;
  for (uint32_t i = 1; i < len; ++i) {
    uint32_t x = p[i];
    uint32_t j = i;
    while (j > 0 && p[j - 1] > x) {
      p[j] = p[j - 1];
      --j;
    }
    p[j] = x;
  }
}

constexpr int LS_TPB = 256;
constexpr uint32_t LS_LDS = 8192;  // ids sorted in LDS (32 KiB)

// Ascending bitonic network in the "mirror" form: the first step of every
// stage compares i with i ^ (k-1), the later steps i with i ^ j, and every
// comparator puts the smaller id at the lower index.  Positions >= len act
// as +infinity; since no comparator ever moves a real element upward past a
// larger one, a pair with a virtual partner never swaps and the network can
// run on exactly `len` stored elements (LDS or global memory).
__device__ void mirror_bitonic_sort(uint32_t* p, uint32_t len) {
  uint32_t n2 = 1;
  while (n2 < len) n2 <<= 1;
  for (uint32_t k = 2; k <= n2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      const uint32_t flip = (j == (k >> 1)) ? (k - 1) : j;
      for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) {
        uint32_t l = i ^ flip;
        if (l > i && l < len) {
          uint32_t x = p[i], y = p[l];
          if (x > y) {
            p[i] = y;
            p[l] = x;
          }
        }
      }
      __syncthreads();
    }
  }
}

// One workgroup per queued long list (grid-stride over the queue).
__global__ __launch_bounds__(LS_TPB) void k_sort_long_lists(const uint32_t* __restrict__ offsets, uint32_t* set_ids,
                                                            const uint32_t* __restrict__ long_queue) {
  __shared__ uint32_t buf[LS_LDS];
  const uint32_t nq = long_queue[0];
  for (uint32_t qi = blockIdx.x; qi < nq; qi += gridDim.x) {
    uint32_t t = long_queue[1 + qi];
    uint32_t a = offsets[t], len = offsets[t + 1] - a;
    uint32_t* p = set_ids + a;
    if (len <= LS_LDS) {
      for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) buf[i] = p[i];
      __syncthreads();
      mirror_bitonic_sort(buf, len);
      for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) p[i] = buf[i];
      __syncthreads();
    } else {
      mirror_bitonic_sort(p, len);  // rare: very long lists sort in global memory
    }
  }
}

void build_index(hipStream_t st, const DevColl& c, uint32_t first, uint32_t last, const uint32_t* prefix_len,
                 uint32_t universe, uint64_t entries_bound, IndexBuild& ib) {
  ib.universe = universe;
  ib.offsets.ensure(((size_t)universe + 1) * sizeof(uint32_t));
  ib.cursor.ensure(((size_t)universe + 1) * sizeof(uint32_t));
  ib.set_ids.ensure((entries_bound ? entries_bound : 1) * sizeof(uint32_t));
  ib.long_queue.ensure(((size_t)universe + 1) * sizeof(uint32_t));
  uint32_t* counts = ib.cursor.as<uint32_t>();
  launch_fill_u32(st, counts, (size_t)universe + 1, 0);
  launch_fill_u32(st, ib.long_queue.as<uint32_t>(), 1, 0);
  if (c.n > 0 && first <= last) {
    uint32_t nsets = last - first + 1;
    hipLaunchKernelGGL(k_index_count, dim3((nsets + 255) / 256), dim3(256), 0, st, c, first, last, prefix_len, counts);
  }
  // offsets = exclusive scan of counts; the scan also refreshes the cursor.
  launch_exclusive_scan(st, counts, ib.offsets.as<uint32_t>(), counts, (size_t)universe + 1, ib.scan_scratch);
  if (c.n > 0 && first <= last) {
    uint32_t nsets = last - first + 1;
    hipLaunchKernelGGL(k_index_scatter, dim3((nsets + 255) / 256), dim3(256), 0, st, c, first, last, prefix_len,
                       counts, ib.set_ids.as<uint32_t>());
  }
  hipLaunchKernelGGL(k_sort_short_lists, dim3((universe + 255) / 256), dim3(256), 0, st, ib.offsets.as<uint32_t>(),
                     universe, ib.set_ids.as<uint32_t>(), ib.long_queue.as<uint32_t>());
  hipLaunchKernelGGL(k_sort_long_lists, dim3(512), dim3(LS_TPB), 0, st, ib.offsets.as<uint32_t>(),
                     ib.set_ids.as<uint32_t>(), ib.long_queue.as<uint32_t>());
  HS_CHECK_HIP(hipGetLastError());
}

}  // namespace hs
