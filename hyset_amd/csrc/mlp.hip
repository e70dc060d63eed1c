// A small multilayer perceptron trained and evaluated on the device: the
// learned cost model of the CPU/GPU split.
//
// Reference: the analytic split of include/hyset/statistics.hpp:132-161 and
// include/hyset/cost.hpp:28-58 is a linear cost per technique; this network
// (in -> h1 -> h2 -> out, ReLU hidden layers, linear output) predicts the
// log-time of every technique from the same partition features instead.
// Training is full-batch Adam on the mean squared error.  Every kernel is
// deterministic: each output element is summed by one thread in a fixed
// order (no atomics), so a run reproduces exactly.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "ctx.h"

using namespace hs;

namespace {

// Z[r, o] = b[o] + sum_i W[o, i] X[r, i];  A = relu(Z) when relu
__global__ void k_dense_fwd(const float* __restrict__ X, const float* __restrict__ W, const float* __restrict__ b,
                            float* __restrict__ Z, float* __restrict__ A, uint32_t n, uint32_t in, uint32_t out,
                            int relu) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * out) return;
  const uint32_t r = idx / out, o = idx % out;
  float z = b[o];
  const float* x = X + (size_t)r * in;
  const float* w = W + (size_t)o * in;
  for (uint32_t i = 0; i < in; ++i) z = fmaf(w[i], x[i], z);
  Z[idx] = z;
  A[idx] = relu ? fmaxf(z, 0.f) : z;
}

// dP = 2 (P - Y) / (n out); se[idx] = (P - Y)^2
__global__ void k_mse_grad(const float* __restrict__ P, const float* __restrict__ Y, float* __restrict__ dP,
                           float* __restrict__ se, uint32_t total) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const float d = P[idx] - Y[idx];
  dP[idx] = 2.f * d / (float)total;
  se[idx] = d * d;
}

// loss = sum(se) / total, by one workgroup in a fixed order
__global__ void k_mean(const float* __restrict__ se, uint32_t total, float* __restrict__ loss) {
  __shared__ float part[256];
  float acc = 0.f;
  for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) acc += se[i];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = part[0] / (float)total;
}

// dZ = dA * (Z > 0)
__global__ void k_relu_back(const float* __restrict__ dA, const float* __restrict__ Z, float* __restrict__ dZ,
                            uint32_t total) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < total) dZ[idx] = Z[idx] > 0.f ? dA[idx] : 0.f;
}

// dW[o, i] = sum_r dZ[r, o] X[r, i];  db[o] = sum_r dZ[r, o]  (thread per (o, i), i == in for db)
__global__ void k_dense_back_w(const float* __restrict__ dZ, const float* __restrict__ X, float* __restrict__ dW,
                               float* __restrict__ db, uint32_t n, uint32_t in, uint32_t out) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= out * (in + 1)) return;
  const uint32_t o = idx / (in + 1), i = idx % (in + 1);
  float acc = 0.f;
  if (i < in) {
    for (uint32_t r = 0; r < n; ++r) acc = fmaf(dZ[(size_t)r * out + o], X[(size_t)r * in + i], acc);
    dW[(size_t)o * in + i] = acc;
  } else {
    for (uint32_t r = 0; r < n; ++r) acc += dZ[(size_t)r * out + o];
    db[o] = acc;
  }
}

// dX[r, i] = sum_o dZ[r, o] W[o, i]
__global__ void k_dense_back_x(const float* __restrict__ dZ, const float* __restrict__ W, float* __restrict__ dX,
                               uint32_t n, uint32_t in, uint32_t out) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * in) return;
  const uint32_t r = idx / in, i = idx % in;
  float acc = 0.f;
  for (uint32_t o = 0; o < out; ++o) acc = fmaf(dZ[(size_t)r * out + o], W[(size_t)o * in + i], acc);
  dX[idx] = acc;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint32_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// (X - mean) / std per column
__global__ void k_standardize(const float* __restrict__ X, const float* __restrict__ mean,
                              const float* __restrict__ stdv, float* __restrict__ out, uint32_t n, uint32_t in) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * in) return;
  const uint32_t c = idx % in;
  out[idx] = (X[idx] - mean[c]) / stdv[c];
}

unsigned blocks(uint64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct hs_mlp {
  hs_ctx* ctx = nullptr;
  uint32_t in = 0, h1 = 0, h2 = 0, out = 0;
  uint32_t nparams = 0;
  DevBuf params, grads, m, v, mean, stdv;
  // activations for up to `cap` rows
  uint32_t cap = 0;
  DevBuf X, Xs, Z1, A1, Z2, A2, Z3, P, Y, dP, se, dA2, dZ2, dA1, dZ1, dX, loss;
  uint64_t step = 0;

  float* W1() { return params.as<float>(); }
  float* b1() { return W1() + (size_t)h1 * in; }
  float* W2() { return b1() + h1; }
  float* b2() { return W2() + (size_t)h2 * h1; }
  float* W3() { return b2() + h2; }
  float* b3() { return W3() + (size_t)out * h2; }
  float* g(size_t off) { return grads.as<float>() + off; }

  void ensure_rows(uint32_t n) {
    if (n <= cap) return;
    const size_t r = n;
    X.ensure(r * in * 4);
    Xs.ensure(r * in * 4);
    Z1.ensure(r * h1 * 4);
    A1.ensure(r * h1 * 4);
    Z2.ensure(r * h2 * 4);
    A2.ensure(r * h2 * 4);
    Z3.ensure(r * out * 4);
    P.ensure(r * out * 4);
    Y.ensure(r * out * 4);
    dP.ensure(r * out * 4);
    se.ensure(r * out * 4);
    dA2.ensure(r * h2 * 4);
    dZ2.ensure(r * h2 * 4);
    dA1.ensure(r * h1 * 4);
    dZ1.ensure(r * h1 * 4);
    dX.ensure(r * in * 4);
    loss.ensure(4);
    cap = n;
  }

  void forward(hipStream_t st, uint32_t n) {
    hipLaunchKernelGGL(k_standardize, dim3(blocks((uint64_t)n * in)), dim3(256), 0, st, X.as<float>(), mean.as<float>(),
                       stdv.as<float>(), Xs.as<float>(), n, in);
    hipLaunchKernelGGL(k_dense_fwd, dim3(blocks((uint64_t)n * h1)), dim3(256), 0, st, Xs.as<float>(), W1(), b1(),
                       Z1.as<float>(), A1.as<float>(), n, in, h1, 1);
    hipLaunchKernelGGL(k_dense_fwd, dim3(blocks((uint64_t)n * h2)), dim3(256), 0, st, A1.as<float>(), W2(), b2(),
                       Z2.as<float>(), A2.as<float>(), n, h1, h2, 1);
    hipLaunchKernelGGL(k_dense_fwd, dim3(blocks((uint64_t)n * out)), dim3(256), 0, st, A2.as<float>(), W3(), b3(),
                       Z3.as<float>(), P.as<float>(), n, h2, out, 0);
  }

  // loss and gradients of the rows in X / Y (forward included)
  void backward(hipStream_t st, uint32_t n) {
    forward(st, n);
    const uint32_t tot = n * out;
    hipLaunchKernelGGL(k_mse_grad, dim3(blocks(tot)), dim3(256), 0, st, P.as<float>(), Y.as<float>(), dP.as<float>(),
                       se.as<float>(), tot);
    hipLaunchKernelGGL(k_mean, dim3(1), dim3(256), 0, st, se.as<float>(), tot, loss.as<float>());
    const size_t oW1 = 0, ob1 = oW1 + (size_t)h1 * in, oW2 = ob1 + h1, ob2 = oW2 + (size_t)h2 * h1, oW3 = ob2 + h2,
                 ob3 = oW3 + (size_t)out * h2;
    // layer 3 (linear output)
    hipLaunchKernelGGL(k_dense_back_w, dim3(blocks((uint64_t)out * (h2 + 1))), dim3(256), 0, st, dP.as<float>(),
                       A2.as<float>(), g(oW3), g(ob3), n, h2, out);
    hipLaunchKernelGGL(k_dense_back_x, dim3(blocks((uint64_t)n * h2)), dim3(256), 0, st, dP.as<float>(), W3(),
                       dA2.as<float>(), n, h2, out);
    hipLaunchKernelGGL(k_relu_back, dim3(blocks((uint64_t)n * h2)), dim3(256), 0, st, dA2.as<float>(), Z2.as<float>(),
                       dZ2.as<float>(), n * h2);
    // layer 2
    hipLaunchKernelGGL(k_dense_back_w, dim3(blocks((uint64_t)h2 * (h1 + 1))), dim3(256), 0, st, dZ2.as<float>(),
                       A1.as<float>(), g(oW2), g(ob2), n, h1, h2);
    hipLaunchKernelGGL(k_dense_back_x, dim3(blocks((uint64_t)n * h1)), dim3(256), 0, st, dZ2.as<float>(), W2(),
                       dA1.as<float>(), n, h1, h2);
    hipLaunchKernelGGL(k_relu_back, dim3(blocks((uint64_t)n * h1)), dim3(256), 0, st, dA1.as<float>(), Z1.as<float>(),
                       dZ1.as<float>(), n * h1);
    // layer 1
    hipLaunchKernelGGL(k_dense_back_w, dim3(blocks((uint64_t)h1 * (in + 1))), dim3(256), 0, st, dZ1.as<float>(),
                       Xs.as<float>(), g(oW1), g(ob1), n, in, h1);
    HS_CHECK_HIP(hipGetLastError());
  }
};

namespace {
void upload_rows(hs_mlp* M, hipStream_t st, const float* X, const float* Y, uint32_t n) {
  M->ensure_rows(n);
  HS_CHECK_HIP(hipMemcpyAsync(M->X.ptr, X, (size_t)n * M->in * 4, hipMemcpyHostToDevice, st));
  if (Y) HS_CHECK_HIP(hipMemcpyAsync(M->Y.ptr, Y, (size_t)n * M->out * 4, hipMemcpyHostToDevice, st));
}
}  // namespace

extern "C" {

int hs_mlp_create(hs_ctx* ctx, uint32_t n_in, uint32_t h1, uint32_t h2, uint32_t n_out, uint64_t seed, hs_mlp** out) {
  if (!ctx || !out) return HS_E_INVALID;
  *out = nullptr;
  auto M = std::make_unique<hs_mlp>();
  int rc = guarded(ctx, [&] {
    HS_REQUIRE(ctx->device >= 0, HS_E_INVALID, "the network trains on the device");
    HS_REQUIRE(n_in && h1 && h2 && n_out && n_in <= 256 && h1 <= 1024 && h2 <= 1024 && n_out <= 64, HS_E_INVALID,
               "layer sizes out of range");
    M->ctx = ctx;
    M->in = n_in;
    M->h1 = h1;
    M->h2 = h2;
    M->out = n_out;
    M->nparams = h1 * n_in + h1 + h2 * h1 + h2 + n_out * h2 + n_out;
    // Glorot-uniform weights, zero biases, from a seeded xorshift stream
    std::vector<float> p(M->nparams, 0.f);
    uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&] {
      x ^= x << 13;
      x ^= x >> 7;
      x ^= x << 17;
      return (double)(x >> 11) * (1.0 / 9007199254740992.0);
    };
    size_t o = 0;
    auto layer = [&](uint32_t fan_in, uint32_t fan_out) {
      const double lim = std::sqrt(6.0 / (fan_in + fan_out));
      for (uint32_t k = 0; k < fan_in * fan_out; ++k) p[o++] = (float)((2.0 * rnd() - 1.0) * lim);
      o += fan_out;  // biases stay zero
    };
    layer(n_in, h1);
    layer(h1, h2);
    layer(h2, n_out);
    M->params.ensure(M->nparams * 4);
    M->grads.ensure(M->nparams * 4);
    M->m.ensure(M->nparams * 4);
    M->v.ensure(M->nparams * 4);
    HS_CHECK_HIP(hipMemcpy(M->params.ptr, p.data(), M->nparams * 4, hipMemcpyHostToDevice));
    std::vector<float> zeros(M->nparams, 0.f), ones(n_in, 1.f), zin(n_in, 0.f);
    HS_CHECK_HIP(hipMemcpy(M->m.ptr, zeros.data(), M->nparams * 4, hipMemcpyHostToDevice));
    HS_CHECK_HIP(hipMemcpy(M->v.ptr, zeros.data(), M->nparams * 4, hipMemcpyHostToDevice));
    M->mean.ensure(n_in * 4);
    M->stdv.ensure(n_in * 4);
    HS_CHECK_HIP(hipMemcpy(M->mean.ptr, zin.data(), n_in * 4, hipMemcpyHostToDevice));
    HS_CHECK_HIP(hipMemcpy(M->stdv.ptr, ones.data(), n_in * 4, hipMemcpyHostToDevice));
  });
  if (rc != HS_OK) return rc;
  *out = M.release();
  return HS_OK;
}

void hs_mlp_destroy(hs_mlp* M) { delete M; }

uint32_t hs_mlp_num_params(const hs_mlp* M) { return M ? M->nparams : 0; }

int hs_mlp_set_normalization(hs_mlp* M, const float* mean, const float* stdv) {
  if (!M || !mean || !stdv) return HS_E_INVALID;
  return guarded(M->ctx, [&] {
    std::vector<float> s(stdv, stdv + M->in);
    for (float& x : s)
      if (!(x > 0.f)) x = 1.f;
    HS_CHECK_HIP(hipMemcpy(M->mean.ptr, mean, M->in * 4, hipMemcpyHostToDevice));
    HS_CHECK_HIP(hipMemcpy(M->stdv.ptr, s.data(), M->in * 4, hipMemcpyHostToDevice));
  });
}

int hs_mlp_get_params(const hs_mlp* M, float* out) {
  if (!M || !out) return HS_E_INVALID;
  return guarded(M->ctx, [&] { HS_CHECK_HIP(hipMemcpy(out, M->params.ptr, M->nparams * 4, hipMemcpyDeviceToHost)); });
}

int hs_mlp_set_params(hs_mlp* M, const float* in) {
  if (!M || !in) return HS_E_INVALID;
  return guarded(M->ctx, [&] { HS_CHECK_HIP(hipMemcpy(M->params.ptr, in, M->nparams * 4, hipMemcpyHostToDevice)); });
}

int hs_mlp_predict(hs_mlp* M, const float* X, uint32_t n, float* Y) {
  if (!M || !X || !Y) return HS_E_INVALID;
  return guarded(M->ctx, [&] {
    if (n == 0) return;
    hipStream_t st = M->ctx->stream;
    upload_rows(M, st, X, nullptr, n);
    M->forward(st, n);
    HS_CHECK_HIP(hipMemcpyAsync(Y, M->P.ptr, (size_t)n * M->out * 4, hipMemcpyDeviceToHost, st));
    HS_CHECK_HIP(hipStreamSynchronize(st));
  });
}

int hs_mlp_loss_grad(hs_mlp* M, const float* X, const float* Y, uint32_t n, float* loss, float* grad) {
  if (!M || !X || !Y || !loss) return HS_E_INVALID;
  return guarded(M->ctx, [&] {
    HS_REQUIRE(n > 0, HS_E_INVALID, "no rows");
    hipStream_t st = M->ctx->stream;
    upload_rows(M, st, X, Y, n);
    M->backward(st, n);
    HS_CHECK_HIP(hipMemcpyAsync(loss, M->loss.ptr, 4, hipMemcpyDeviceToHost, st));
    if (grad) HS_CHECK_HIP(hipMemcpyAsync(grad, M->grads.ptr, M->nparams * 4, hipMemcpyDeviceToHost, st));
    HS_CHECK_HIP(hipStreamSynchronize(st));
  });
}

int hs_mlp_train(hs_mlp* M, const float* X, const float* Y, uint32_t n, uint32_t epochs, float lr, float* final_loss) {
  if (!M || !X || !Y) return HS_E_INVALID;
  return guarded(M->ctx, [&] {
    HS_REQUIRE(n > 0 && epochs > 0 && lr > 0.f, HS_E_INVALID, "bad training arguments");
    hipStream_t st = M->ctx->stream;
    upload_rows(M, st, X, Y, n);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (uint32_t e = 0; e < epochs; ++e) {
      M->backward(st, n);
      ++M->step;
      const float c1 = 1.f - std::pow(b1, (float)M->step), c2 = 1.f - std::pow(b2, (float)M->step);
      hipLaunchKernelGGL(k_adam, dim3(blocks(M->nparams)), dim3(256), 0, st, M->params.as<float>(),
                         M->grads.as<float>(), M->m.as<float>(), M->v.as<float>(), M->nparams, lr, b1, b2, eps, c1,
                         c2);
    }
    HS_CHECK_HIP(hipGetLastError());
    if (final_loss) {
      M->backward(st, n);
      HS_CHECK_HIP(hipMemcpyAsync(final_loss, M->loss.ptr, 4, hipMemcpyDeviceToHost, st));
    }
    HS_CHECK_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"
