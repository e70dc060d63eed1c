// Verification of candidate batches (the device half of the hybrid
// technique) for gfx950.
//
// Reference: include/hyset/hybrid_kernels.hpp:14-165 and
// hybrid_algorithms.hpp:122-193.  A batch is a CSR of candidates:
// probes[i] is the probe of row i, its candidates are
// candidates[offsets[i-1] .. offsets[i]) (offsets are inclusive ends, row 0
// starts at 0).
//   scenario 1 (hybrid_kernels.hpp:14-74): one lane per candidate, sequential
//     early-terminating merge.  Here rows are flattened, so a wave is always
//     full regardless of how many candidates a probe has.
//   scenario 2 (hybrid_kernels.hpp:76-165): merge-path intersection.  Here a
//     64-lane wave takes one candidate, each lane binary-searches its split
//     of the |r|+|s| merge diagonal and counts matches over its ceil(
//     (|r|+|s|)/64) steps; a wave reduction gives |r n s|.  (The reference's
//     block version drops per-thread counts for more than 32 threads; this
//     one counts every match exactly once.)
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "hs_internal.h"

namespace hs {

// Row of candidate c: first i with offsets[i] > c.
__device__ __forceinline__ uint32_t row_of(const uint32_t* offsets, uint32_t n_rows, uint64_t c) {
  uint32_t lo = 0, hi = n_rows;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if ((uint64_t)offsets[mid] <= c) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ bool merge_verify(const uint32_t* r, uint32_t lr, const uint32_t* s, uint32_t ls,
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

__global__ __launch_bounds__(256) void k_verify_lane(VerifyArgs A) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t n = A.n_candidates;
  // uniform trip count for the whole wave so emit_pairs sees every lane
  const uint64_t wave_base = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) & ~(uint64_t)(WAVE - 1);
  for (uint64_t wb = wave_base; wb < n; wb += stride) {
    const uint64_t c = wb + lane_id();
    bool pass = false;
    uint32_t rid = 0, sid = 0, ov = 0;
    if (c < n) {
      const uint32_t row = row_of(A.offsets, A.n_probes, c);
      rid = A.probes[row];
      sid = A.candidates[c];
      const uint32_t lr = A.probe.sizes[rid];
      const uint32_t ls = A.indexed.sizes[sid];
      const uint32_t mo = dev_minoverlap(A.bounds, lr, ls);
      pass = merge_verify(A.probe.tokens + A.probe.starts[rid], lr, A.indexed.tokens + A.indexed.starts[sid], ls,
                          mo, &ov);
    }
    emit_pairs(pass, rid, sid, ov, A.out.r, A.out.s, A.out.ov, &A.out.counters[CNT_PAIRS], A.out.capacity);
  }
}

// Elements of a among the first `diag` of the merge of a and b, ties taking
// b first (include/hyset/helpers.hpp:96-108).
__device__ __forceinline__ uint32_t merge_path_split(const uint32_t* a, uint32_t na, const uint32_t* b, uint32_t nb,
                                                     uint32_t diag) {
  uint32_t lo = diag > nb ? diag - nb : 0;
  uint32_t hi = min(diag, na);
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < b[diag - 1 - mid]) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_verify_wave(VerifyArgs A) {
  const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x / WAVE);
  const uint64_t w0 = (uint64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  const uint32_t lane = lane_id();
  for (uint64_t c = w0; c < A.n_candidates; c += nwaves) {
    const uint32_t row = row_of(A.offsets, A.n_probes, c);
    const uint32_t rid = A.probes[row];
    const uint32_t sid = A.candidates[c];
    const uint32_t lr = A.probe.sizes[rid];
    const uint32_t ls = A.indexed.sizes[sid];
    const uint32_t* r = A.probe.tokens + A.probe.starts[rid];
    const uint32_t* s = A.indexed.tokens + A.indexed.starts[sid];
    const uint32_t total = lr + ls;
    const uint32_t vt = (total + WAVE - 1) / WAVE;
    const uint32_t d0 = min(lane * vt, total);
    const uint32_t d1 = min(d0 + vt, total);
    uint32_t i = merge_path_split(r, lr, s, ls, d0);
    uint32_t k = d0 - i;
    uint32_t cnt = 0;
    for (uint32_t d = d0; d < d1; ++d) {
      if (i < lr && (k >= ls || r[i] < s[k])) {
        ++i;
      } else {
        cnt += (i < lr && r[i] == s[k]);
        ++k;
      }
    }
    const uint32_t ov = wave_sum(cnt);
    const uint32_t mo = dev_minoverlap(A.bounds, lr, ls);
    if (lane == 0 && ov >= mo) {
      unsigned long long idx = atomicAdd(&A.out.counters[CNT_PAIRS], 1ull);
      if (idx < A.out.capacity) {
        A.out.r[idx] = rid;
        A.out.s[idx] = sid;
        A.out.ov[idx] = ov;
      }
    }
  }
}

void launch_verify(hipStream_t st, const VerifyArgs& a, int scenario) {
  if (a.n_candidates == 0 || a.n_probes == 0) return;
  if (scenario == 2) {
    uint64_t blocks = (a.n_candidates + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_verify_wave, dim3((unsigned)blocks), dim3(256), 0, st, a);
  } else {
    uint64_t blocks = (a.n_candidates + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_verify_lane, dim3((unsigned)blocks), dim3(256), 0, st, a);
  }
  HS_CHECK_HIP(hipGetLastError());
}

}  // namespace hs
