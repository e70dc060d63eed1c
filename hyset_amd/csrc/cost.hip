// Cost models and the CPU/GPU split of a join workload.
//
// Reference: include/hyset/statistics.hpp:120-283 (partition features:
// size, average and standard deviation of set sizes, standard deviation of
// the prefix-token frequency histogram, universe and index size; linear cost
// |sum coeff * feature + intercept| per technique), include/hyset/cost.hpp:
// 28-58 (argmin technique per partition pair) and src/cost_model.cu:296-323
// (coefficients per technique and threshold).  The reference ships no
// coefficients (its cost_model.cu is not built), so the analytic model here
// is fitted by ridge least squares to timings this library measures; the
// learned model (mlp.hip) is trained on the same timings.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "ctx.h"
#include "device_common.h"

using namespace hs;

namespace {

// acc[0] += sum x, acc[1] += sum x^2 over sizes[first .. last]
__global__ void k_size_moments(const uint32_t* __restrict__ sizes, uint32_t first, uint32_t last, double* acc) {
  __shared__ double s0[256], s1[256];
  double a = 0, b = 0;
  for (uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x; i <= last; i += gridDim.x * blockDim.x) {
    const double x = sizes[i];
    a += x;
    b += x * x;
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  __syncthreads();
  for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      s0[threadIdx.x] += s0[threadIdx.x + s];
      s1[threadIdx.x] += s1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&acc[0], s0[0]);
    atomicAdd(&acc[1], s1[0]);
  }
}

// acc[2] += sum c, acc[3] += sum c^2 with c = offsets[t+1] - offsets[t]
__global__ void k_list_moments(const uint32_t* __restrict__ offsets, uint32_t universe, double* acc) {
  __shared__ double s0[256], s1[256];
  double a = 0, b = 0;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < universe; t += gridDim.x * blockDim.x) {
    const double c = offsets[t + 1] - offsets[t];
    a += c;
    b += c * c;
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  __syncthreads();
  for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      s0[threadIdx.x] += s0[threadIdx.x + s];
      s1[threadIdx.x] += s1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&acc[2], s0[0]);
    atomicAdd(&acc[3], s1[0]);
  }
}

__global__ void k_zero_f64(double* p, uint32_t n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0.0;
}

// hist[t] += 1 for every midprefix token t of the sets [a0, a0 + na) and
// [b0, b0 + nb), one thread per set (the histogram kernel of
// include/hyset/index.hpp:217-231 over the prefixes of two partitions)
__global__ void k_prefix_hist(DevColl c, const uint32_t* __restrict__ midprefix, uint32_t L, uint32_t a0,
                              uint32_t na, uint32_t b0, uint32_t nb, uint32_t universe, uint32_t* hist) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < na + nb; k += gridDim.x * blockDim.x) {
    const uint32_t i = k < na ? a0 + k : b0 + (k - na);
    const uint32_t l = c.sizes[i];
    const uint32_t p = midprefix[l < L ? l : L];
    const uint32_t* t = c.tokens + c.starts[i];
    for (uint32_t j = 0; j < p; ++j) {
      const uint32_t tok = t[j];
      if (tok < universe) atomicAdd(&hist[tok], 1u);
    }
  }
}

// acc[2] += sum c, acc[3] += sum c^2 over the counts hist[0 .. universe)
__global__ void k_count_moments(const uint32_t* __restrict__ hist, uint32_t universe, double* acc) {
  __shared__ double s0[256], s1[256];
  double a = 0, b = 0;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < universe; t += gridDim.x * blockDim.x) {
    const double c = hist[t];
    a += c;
    b += c * c;
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  __syncthreads();
  for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      s0[threadIdx.x] += s0[threadIdx.x + s];
      s1[threadIdx.x] += s1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&acc[2], s0[0]);
    atomicAdd(&acc[3], s1[0]);
  }
}

// Midprefix entries of the sets of range r (sets are sorted by size).
uint64_t midprefix_entries(const hs_ctx* ctx, const hs_coll* coll, const hs_range& r) {
  const auto& sf = coll->h_size_first;
  uint64_t acc = 0;
  for (uint32_t l = 0; l + 1 < (uint32_t)sf.size(); ++l) {
    const uint32_t lo = std::max(r.first, sf[l]), hi = std::min(r.last + 1, sf[l + 1]);
    if (hi > lo) acc += (uint64_t)(hi - lo) * ctx->tables.midprefix[std::min(l, ctx->tables.L)];
  }
  return acc;
}

// Cholesky solve of the (k x k) symmetric positive-definite system A x = b.
bool cholesky_solve(std::vector<double> A, std::vector<double>& b, uint32_t k) {
  for (uint32_t j = 0; j < k; ++j) {
    const double diag = A[j * k + j];
    double d = diag;
    for (uint32_t p = 0; p < j; ++p) d -= A[j * k + p] * A[j * k + p];
    if (!(d > 1e-12 * diag)) return false;  // (numerically) singular
    d = std::sqrt(d);
    A[j * k + j] = d;
    for (uint32_t i = j + 1; i < k; ++i) {
      double s = A[i * k + j];
      for (uint32_t p = 0; p < j; ++p) s -= A[i * k + p] * A[j * k + p];
      A[i * k + j] = s / d;
    }
  }
  for (uint32_t i = 0; i < k; ++i) {
    double s = b[i];
    for (uint32_t p = 0; p < i; ++p) s -= A[i * k + p] * b[p];
    b[i] = s / A[i * k + i];
  }
  for (int i = (int)k - 1; i >= 0; --i) {
    double s = b[i];
    for (uint32_t p = i + 1; p < k; ++p) s -= A[p * k + i] * b[p];
    b[i] = s / A[i * k + i];
  }
  return true;
}

}  // namespace

extern "C" {

int hs_features(hs_ctx* ctx, hs_coll* coll, const hs_range* range, float out[HS_NFEATURES]) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    HS_REQUIRE(coll && out, HS_E_INVALID, "NULL argument");
    HS_REQUIRE(ctx->device >= 0 && coll->n_joins == 0, HS_E_INVALID, "features need a device collection");
    const hs_range r = clamp_range(range, coll->n);
    HS_REQUIRE(!range_empty(r), HS_E_INVALID, "empty range");
    ensure_bounds(ctx, coll->maxlen);
    hipStream_t st = ctx->stream;
    // prefix-token histogram = list lengths of the range's self-join index
    const auto& ptab = ctx->tables.midprefix;
    const auto& sf = coll->h_size_first;
    const uint32_t nl = (uint32_t)sf.size();
    ctx->h_cum.ensure((size_t)nl * 4);
    ctx->d_cum.ensure((size_t)nl * 4);
    uint32_t* cum = ctx->h_cum.as<uint32_t>();
    uint64_t acc = 0;
    for (uint32_t l = 0; l + 1 < nl; ++l) {
      cum[l] = (uint32_t)acc;
      const uint32_t lo = std::max(r.first, sf[l]), hi = std::min(r.last + 1, sf[l + 1]);
      if (hi > lo) acc += (uint64_t)(hi - lo) * ptab[l];
    }
    cum[nl - 1] = (uint32_t)acc;
    HS_CHECK_HIP(hipMemcpyAsync(ctx->d_cum.ptr, cum, (size_t)nl * 4, hipMemcpyHostToDevice, st));
    build_index(st, coll->dev(), r.first, r.last, ctx->dev_bounds.midprefix, coll->size_first.as<uint32_t>(),
                ctx->d_cum.as<uint32_t>(), coll->universe, acc, ctx->ib, nullptr);
    ctx->feat.ensure(4 * sizeof(double));
    double* dacc = ctx->feat.as<double>();
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st, dacc, 4u);
    // statistics.hpp:173-184 reduces sizes over [startID, endID) but counts
    // endID - startID + 1 sets; the moments here follow it (a one-set range
    // keeps its set)
    const uint32_t n = r.last - r.first + 1;
    const uint32_t m_last = n > 1 ? r.last - 1 : r.last, m = m_last - r.first + 1;
    hipLaunchKernelGGL(k_size_moments, dim3(std::min<uint32_t>((m + 255) / 256, 1024)), dim3(256), 0, st,
                       coll->sizes.as<uint32_t>(), r.first, m_last, dacc);
    hipLaunchKernelGGL(k_list_moments, dim3(std::min<uint32_t>((coll->universe + 255) / 256, 1024)), dim3(256), 0, st,
                       ctx->ib.offsets.as<uint32_t>(), coll->universe, dacc);
    HS_CHECK_HIP(hipGetLastError());
    double h[4];
    HS_CHECK_HIP(hipMemcpyAsync(h, dacc, sizeof(h), hipMemcpyDeviceToHost, st));
    HS_CHECK_HIP(hipStreamSynchronize(st));
    const double mean = h[0] / m, var = std::max(0.0, h[1] / m - mean * mean);
    const double U = coll->universe, fmean = h[2] / U, fvar = std::max(0.0, h[3] / U - fmean * fmean);
    out[0] = (float)n;                // size
    out[1] = (float)mean;             // averageSetSize
    out[2] = (float)std::sqrt(var);   // setSizeStd
    out[3] = (float)std::sqrt(fvar);  // tokenFreqDistribution
    out[4] = (float)U;                // universeSize
    out[5] = (float)acc;              // indexSize
    out[6] = (float)ctx->t;           // threshold
  });
}

int hs_features_pair(hs_ctx* ctx, hs_coll* coll, const hs_range* left, const hs_range* right,
                     float out[HS_NFEATURES]) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    HS_REQUIRE(coll && left && right && out, HS_E_INVALID, "NULL argument");
    HS_REQUIRE(ctx->device >= 0 && coll->n_joins == 0, HS_E_INVALID, "features need a device collection");
    const hs_range a = clamp_range(left, coll->n), b = clamp_range(right, coll->n);
    HS_REQUIRE(!range_empty(a) && !range_empty(b) && left->last < coll->n && right->last < coll->n, HS_E_INVALID,
               "ranges must be non-empty and inside the collection");
    ensure_bounds(ctx, coll->maxlen);
    hipStream_t st = ctx->stream;
    const uint32_t U = coll->universe;
    ctx->feat.ensure(4 * sizeof(double));
    ctx->feat_hist.ensure((size_t)std::max<uint32_t>(U, 1) * 4);
    double* dacc = ctx->feat.as<double>();
    uint32_t* hist = ctx->feat_hist.as<uint32_t>();
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st, dacc, 4u);
    launch_fill_u32(st, hist, U, 0u);
    // sizes of [first, last) of each range (statistics.hpp:236-237); the
    // slots of the two last sets stay zero
    const uint32_t na = a.last - a.first + 1, nb = b.last - b.first + 1;
    if (na > 1)
      hipLaunchKernelGGL(k_size_moments, dim3(std::min<uint32_t>((na + 254) / 256, 1024)), dim3(256), 0, st,
                         coll->sizes.as<uint32_t>(), a.first, a.last - 1, dacc);
    if (nb > 1)
      hipLaunchKernelGGL(k_size_moments, dim3(std::min<uint32_t>((nb + 254) / 256, 1024)), dim3(256), 0, st,
                         coll->sizes.as<uint32_t>(), b.first, b.last - 1, dacc);
    const uint32_t nsets = na + nb;
    hipLaunchKernelGGL(k_prefix_hist, dim3(std::min<uint32_t>((nsets + 255) / 256, 4096)), dim3(256), 0, st,
                       coll->dev(), ctx->dev_bounds.midprefix, ctx->dev_bounds.L, a.first, na, b.first, nb, U, hist);
    if (U)
      hipLaunchKernelGGL(k_count_moments, dim3(std::min<uint32_t>((U + 255) / 256, 1024)), dim3(256), 0, st, hist,
                         U, dacc);
    HS_CHECK_HIP(hipGetLastError());
    double h[4];
    HS_CHECK_HIP(hipMemcpyAsync(h, dacc, sizeof(h), hipMemcpyDeviceToHost, st));
    HS_CHECK_HIP(hipStreamSynchronize(st));
    const double m = (double)na + nb;
    const double mean = h[0] / m, var = std::max(0.0, h[1] / m - mean * mean);
    const double Ud = std::max<uint32_t>(U, 1), fmean = h[2] / Ud, fvar = std::max(0.0, h[3] / Ud - fmean * fmean);
    out[0] = (float)m;                               // size
    out[1] = (float)mean;                            // averageSetSize
    out[2] = (float)std::sqrt(var);                  // setSizeStd
    out[3] = (float)std::sqrt(fvar);                 // tokenFreqDistribution
    out[4] = (float)U;                               // universeSize
    out[5] = (float)midprefix_entries(ctx, coll, a);  // indexSize: the left partition's entries
    out[6] = (float)ctx->t;                          // threshold
  });
}

int hs_linear_fit(const double* X, const double* Y, uint32_t n, uint32_t d, uint32_t n_out, double ridge,
                  double* coef) {
  if (!X || !Y || !coef || n == 0 || d == 0 || n_out == 0 || d > 64) return HS_E_INVALID;
  const uint32_t k = d + 1;  // features + intercept
  std::vector<double> A((size_t)k * k, 0.0);
  for (uint32_t r = 0; r < n; ++r)
    for (uint32_t i = 0; i < k; ++i) {
      const double xi = i < d ? X[(size_t)r * d + i] : 1.0;
      for (uint32_t j = 0; j < k; ++j) A[i * k + j] += xi * (j < d ? X[(size_t)r * d + j] : 1.0);
    }
  for (uint32_t i = 0; i < d; ++i) A[i * k + i] += ridge;
  for (uint32_t o = 0; o < n_out; ++o) {
    std::vector<double> b(k, 0.0);
    for (uint32_t r = 0; r < n; ++r)
      for (uint32_t i = 0; i < k; ++i) b[i] += (i < d ? X[(size_t)r * d + i] : 1.0) * Y[(size_t)r * n_out + o];
    if (!cholesky_solve(A, b, k)) return HS_E_INVALID;
    for (uint32_t i = 0; i < k; ++i) coef[(size_t)i * n_out + o] = b[i];
  }
  return HS_OK;
}

int hs_linear_predict(const double* coef, const double* X, uint32_t n, uint32_t d, uint32_t n_out, double* Y) {
  if (!coef || !X || !Y) return HS_E_INVALID;
  for (uint32_t r = 0; r < n; ++r)
    for (uint32_t o = 0; o < n_out; ++o) {
      double y = coef[(size_t)d * n_out + o];
      for (uint32_t i = 0; i < d; ++i) y += coef[(size_t)i * n_out + o] * X[(size_t)r * d + i];
      Y[(size_t)r * n_out + o] = std::fabs(y);  // statistics.hpp:150-160 takes the absolute value
    }
  return HS_OK;
}

int hs_split_run(hs_ctx* ctx, uint32_t n_units, hs_coll* const* colls, const int32_t* technique, uint32_t cpu_threads,
                 uint64_t* counts, double* unit_ms, double* wall_ms) {
  if (!ctx) return HS_E_INVALID;
  return guarded(ctx, [&] {
    HS_REQUIRE(colls && technique && counts, HS_E_INVALID, "NULL argument");
    uint32_t L = 0;
    for (uint32_t u = 0; u < n_units; ++u) {
      HS_REQUIRE(colls[u] && colls[u]->n_joins == 0 && colls[u]->device == ctx->device, HS_E_INVALID, "bad unit");
      HS_REQUIRE(technique[u] >= HS_TECH_FGSS && technique[u] <= HS_TECH_CPU, HS_E_INVALID, "bad technique");
      L = std::max(L, colls[u]->maxlen);
    }
    HS_REQUIRE(ctx->device >= 0, HS_E_INVALID, "the split needs a device context");
    ensure_bounds(ctx, L);  // no device join rebuilds the tables the CPU threads read
    const BoundTables T = ctx->tables;
    std::vector<uint32_t> cpu_units, gpu_units;
    for (uint32_t u = 0; u < n_units; ++u) (technique[u] == HS_TECH_CPU ? cpu_units : gpu_units).push_back(u);
    // host threads claim the largest units first (by token count), the
    // longest-first order of the planner's list schedule
    std::stable_sort(cpu_units.begin(), cpu_units.end(),
                     [&](uint32_t a, uint32_t b) { return colls[a]->n_tokens > colls[b]->n_tokens; });
    auto t0 = std::chrono::steady_clock::now();
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> pool;
    const uint32_t nt = std::max<uint32_t>(1, cpu_threads);
    for (uint32_t k = 0; k < nt && !cpu_units.empty(); ++k)
      pool.emplace_back([&] {
        for (;;) {
          const uint32_t i = next.fetch_add(1);
          if (i >= cpu_units.size()) break;
          const uint32_t u = cpu_units[i];
          const hs_coll& c = *colls[u];
          auto a = std::chrono::steady_clock::now();
          counts[u] = host_join(c, c, true, hs_range{0, c.n - 1}, hs_range{0, c.n - 1}, T, 1, nullptr);
          if (unit_ms) unit_ms[u] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
        }
      });
    // the device units run on this thread, one join each, while the CPU works
    int rc = HS_OK;
    std::string err;
    for (uint32_t u : gpu_units) {
      hs_result res{};
      res.count_only = 1;
      hs_tune tune{};
      tune.bitmap_bits = 64;  // the width the split's models are measured and fitted at
      const int32_t algo = technique[u] == HS_TECH_FGSS ? HS_ALGO_FGSS
                           : technique[u] == HS_TECH_BITMAP ? HS_ALGO_BITMAP : HS_ALGO_HYBRID;
      auto a = std::chrono::steady_clock::now();
      int r = hs_join(ctx, colls[u], colls[u], nullptr, nullptr, algo, &tune, &res);
      if (r != HS_OK && rc == HS_OK) {
        rc = r;
        err = ctx->err;
      }
      counts[u] = res.count;
      if (unit_ms) unit_ms[u] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
    }
    for (auto& th : pool) th.join();
    if (wall_ms) *wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc != HS_OK) throw Error{rc, err};
  });
}

}  // extern "C"
