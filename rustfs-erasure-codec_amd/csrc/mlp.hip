// mlp.hip -- the fp32 network of the learned kernel-variant dispatcher:
// x (n_in <= 16 features) -> 64 ReLU -> 64 ReLU -> y (n_out <= 32 predicted
// log-times, one per kernel variant), with hand-written forward, backward and
// Adam kernels.  Training minimises the weighted squared error
//   L = sum_s sum_o w[s][o] (y[s][o] - t[s][o])^2 / sum w
// over a whole training set held in HBM (full-batch steps).
//
// Device layout: parameters padded to kF = 16 inputs and kO = 32 outputs
// (padding weights start at zero and receive zero gradient, so they stay zero):
//   W1[kH][kF] b1[kH] W2[kH][kH] b2[kH] W3[kO][kH] b3[kO].
// One step is three launches:
//   1. mlp_sample_kernel: a lane per sample runs the forward pass with its
//      activations in VGPRs (weights broadcast from LDS), then the backward pass
//      to the per-sample deltas dy, dh2, dh1, which it writes with h1, h2;
//   2. mlp_wgrad_kernel: a lane per (parameter, chunk of 256 samples) sums that
//      parameter's gradient over the chunk;
//   3. mlp_reduce_kernel: a lane per parameter sums the chunk partials in a fixed
//      order (bit-reproducible) and applies Adam, or stores the gradient.
#include <cmath>
#include <cstring>

#include "internal.h"

namespace rsgpu {

constexpr int kF = 16, kH = 64, kO = 32;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kF;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kO * kH;
constexpr int kP = kOffB3 + kO;  // 7328 padded parameters
constexpr int kChunk = 256;      // samples per weight-gradient partial

struct SampleArgs {
  const float* params;
  const float* x;  // [B][n_in]
  const float* t;  // [B][n_out] targets (training)
  const float* w;  // [B][n_out] weights (training; 0 = no target)
  float* y;        // [B][n_out] outputs (inference) or nullptr
  float* h1;       // [B][kH]
  float* h2;       // [B][kH]
  float* d1;       // [B][kH]
  float* d2;       // [B][kH]
  float* d3;       // [B][kO]
  float* loss_part;  // [gridDim.x]
  float inv_wsum;
  int32_t B, n_in, n_out, train;
};

__global__ __launch_bounds__(256) void mlp_sample_kernel(const SampleArgs a) {
  __shared__ float P[kP];
  __shared__ float red[256];
  for (int i = threadIdx.x; i < kP; i += blockDim.x) P[i] = a.params[i];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  float lsum = 0.f;
  if (s < a.B) {
    float x[kF];
#pragma unroll
    for (int f = 0; f < kF; ++f) x[f] = f < a.n_in ? a.x[(size_t)s * a.n_in + f] : 0.f;
    float h1[kH];
#pragma unroll
    for (int i = 0; i < kH; ++i) {
      float z = P[kOffB1 + i];
#pragma unroll
      for (int f = 0; f < kF; ++f) z = fmaf(P[kOffW1 + i * kF + f], x[f], z);
      h1[i] = fmaxf(z, 0.f);
    }
    float h2[kH];
#pragma unroll
    for (int j = 0; j < kH; ++j) {
      float z = P[kOffB2 + j];
#pragma unroll
      for (int i = 0; i < kH; ++i) z = fmaf(P[kOffW2 + j * kH + i], h1[i], z);
      h2[j] = fmaxf(z, 0.f);
    }
    float dy[kO];
#pragma unroll
    for (int o = 0; o < kO; ++o) {
      float z = P[kOffB3 + o];
#pragma unroll
      for (int j = 0; j < kH; ++j) z = fmaf(P[kOffW3 + o * kH + j], h2[j], z);
      if (o < a.n_out && a.y) a.y[(size_t)s * a.n_out + o] = z;
      dy[o] = 0.f;
      if (a.train && o < a.n_out) {
        const float wt = a.w[(size_t)s * a.n_out + o];
        const float r = z - a.t[(size_t)s * a.n_out + o];
        lsum = fmaf(wt * r, r * a.inv_wsum, lsum);
        dy[o] = 2.f * wt * r * a.inv_wsum;
      }
    }
    if (a.train) {
      float dh2[kH];
#pragma unroll
      for (int j = 0; j < kH; ++j) {
        float g = 0.f;
#pragma unroll
        for (int o = 0; o < kO; ++o) g = fmaf(P[kOffW3 + o * kH + j], dy[o], g);
        dh2[j] = h2[j] > 0.f ? g : 0.f;
      }
#pragma unroll
      for (int i = 0; i < kH; ++i) {
        float g = 0.f;
#pragma unroll
        for (int j = 0; j < kH; ++j) g = fmaf(P[kOffW2 + j * kH + i], dh2[j], g);
        a.d1[(size_t)s * kH + i] = h1[i] > 0.f ? g : 0.f;
        a.h1[(size_t)s * kH + i] = h1[i];
      }
#pragma unroll
      for (int j = 0; j < kH; ++j) {
        a.d2[(size_t)s * kH + j] = dh2[j];
        a.h2[(size_t)s * kH + j] = h2[j];
      }
#pragma unroll
      for (int o = 0; o < kO; ++o) a.d3[(size_t)s * kO + o] = dy[o];
    }
  }
  if (!a.train) return;
  red[threadIdx.x] = lsum;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.loss_part[blockIdx.x] = red[0];
}

struct GradArgs {
  const float* x;
  const float* h1;
  const float* h2;
  const float* d1;
  const float* d2;
  const float* d3;
  float* partial;  // [n_chunks][kP]
  int32_t B, n_in;
};

__global__ __launch_bounds__(256) void mlp_wgrad_kernel(const GradArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kP) return;
  const int s0 = blockIdx.y * kChunk;
  const int s1 = min(a.B, s0 + kChunk);
  // gradient of parameter p = sum_s delta[s][row] * input[s][col]
  const float* delta;
  const float* in = nullptr;
  int dstride, row, istride = 0, col = 0;
  bool x_in = false;
  if (p < kOffB1) {
    row = (p - kOffW1) / kF, col = (p - kOffW1) % kF;
    delta = a.d1, dstride = kH, x_in = true;
  } else if (p < kOffW2) {
    row = p - kOffB1, delta = a.d1, dstride = kH;
  } else if (p < kOffB2) {
    row = (p - kOffW2) / kH, col = (p - kOffW2) % kH;
    delta = a.d2, dstride = kH, in = a.h1, istride = kH;
  } else if (p < kOffW3) {
    row = p - kOffB2, delta = a.d2, dstride = kH;
  } else if (p < kOffB3) {
    row = (p - kOffW3) / kH, col = (p - kOffW3) % kH;
    delta = a.d3, dstride = kO, in = a.h2, istride = kH;
  } else {
    row = p - kOffB3, delta = a.d3, dstride = kO;
  }
  float g = 0.f;
  if (x_in) {
    if (col < a.n_in)
      for (int s = s0; s < s1; ++s) g = fmaf(delta[(size_t)s * dstride + row], a.x[(size_t)s * a.n_in + col], g);
  } else if (in) {
    for (int s = s0; s < s1; ++s) g = fmaf(delta[(size_t)s * dstride + row], in[(size_t)s * istride + col], g);
  } else {
    for (int s = s0; s < s1; ++s) g += delta[(size_t)s * dstride + row];
  }
  a.partial[(size_t)blockIdx.y * kP + p] = g;
}

struct ReduceArgs {
  float* params;
  float* m1;
  float* m2;
  float* grad;  // grad-only mode output, else nullptr
  const float* partial;
  const float* loss_part;
  float* loss;  // [1]
  int32_t n_chunks, n_loss_parts;
  float lr, c1, c2;  // Adam step size and bias corrections 1 - beta^t
};

__global__ __launch_bounds__(256) void mlp_reduce_kernel(const ReduceArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) {
    float l = 0.f;
    for (int i = 0; i < a.n_loss_parts; ++i) l += a.loss_part[i];
    a.loss[0] = l;
  }
  if (p >= kP) return;
  float g = 0.f;
  for (int c = 0; c < a.n_chunks; ++c) g += a.partial[(size_t)c * kP + p];
  if (a.grad) {
    a.grad[p] = g;
    return;
  }
  constexpr float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float m = b1 * a.m1[p] + (1.f - b1) * g;
  const float v = b2 * a.m2[p] + (1.f - b2) * g * g;
  a.m1[p] = m;
  a.m2[p] = v;
  const float mh = m / a.c1, vh = v / a.c2;
  a.params[p] -= a.lr * mh / (sqrtf(vh) + eps);
}

}  // namespace rsgpu

using namespace rsgpu;

struct rsgpu_mlp {
  int n_in = 0, n_out = 0, device = 0;
  DeviceBuffer params, m1, m2, grad, loss, partial, loss_part, h1, h2, d1, d2, d3;
  size_t cap = 0;  // samples the activation buffers hold
  int64_t t = 0;   // Adam steps taken
};

namespace {

size_t n_user_params(int n_in, int n_out) {
  return (size_t)kH * n_in + kH + kH * kH + kH + (size_t)n_out * kH + n_out;
}

// user layout (W1[64][n_in] b1 W2[64][64] b2 W3[n_out][64] b3) <-> padded device layout
void pad_params(int n_in, int n_out, const float* u, std::vector<float>& d) {
  d.assign(kP, 0.f);
  size_t q = 0;
  for (int i = 0; i < kH; ++i)
    for (int f = 0; f < n_in; ++f) d[kOffW1 + i * kF + f] = u[q++];
  for (int i = 0; i < kH; ++i) d[kOffB1 + i] = u[q++];
  for (int i = 0; i < kH * kH; ++i) d[kOffW2 + i] = u[q++];
  for (int i = 0; i < kH; ++i) d[kOffB2 + i] = u[q++];
  for (int o = 0; o < n_out; ++o)
    for (int j = 0; j < kH; ++j) d[kOffW3 + o * kH + j] = u[q++];
  for (int o = 0; o < n_out; ++o) d[kOffB3 + o] = u[q++];
}

void unpad_params(int n_in, int n_out, const float* d, float* u) {
  size_t q = 0;
  for (int i = 0; i < kH; ++i)
    for (int f = 0; f < n_in; ++f) u[q++] = d[kOffW1 + i * kF + f];
  for (int i = 0; i < kH; ++i) u[q++] = d[kOffB1 + i];
  for (int i = 0; i < kH * kH; ++i) u[q++] = d[kOffW2 + i];
  for (int i = 0; i < kH; ++i) u[q++] = d[kOffB2 + i];
  for (int o = 0; o < n_out; ++o)
    for (int j = 0; j < kH; ++j) u[q++] = d[kOffW3 + o * kH + j];
  for (int o = 0; o < n_out; ++o) u[q++] = d[kOffB3 + o];
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int ensure_cap(rsgpu_mlp* h, size_t B) {
  if (B <= h->cap) return RSGPU_OK;
  const size_t chunks = (B + kChunk - 1) / kChunk, blocks = (B + 255) / 256;
  if (h->h1.alloc(B * kH * 4) != hipSuccess || h->h2.alloc(B * kH * 4) != hipSuccess ||
      h->d1.alloc(B * kH * 4) != hipSuccess || h->d2.alloc(B * kH * 4) != hipSuccess ||
      h->d3.alloc(B * kO * 4) != hipSuccess || h->partial.alloc(chunks * kP * 4) != hipSuccess ||
      h->loss_part.alloc(blocks * 4) != hipSuccess)
    return RSGPU_ERR_OUT_OF_MEMORY;
  h->cap = B;
  return RSGPU_OK;
}

// Forward + backward + gradient partials + reduce; adam = false stores the gradient.
int step(rsgpu_mlp* h, const float* x, const float* t, const float* w, size_t B, float inv_wsum,
         bool adam, float lr, hipStream_t s) {
  int rc = ensure_cap(h, B);
  if (rc) return rc;
  const unsigned blocks = (unsigned)((B + 255) / 256), chunks = (unsigned)((B + kChunk - 1) / kChunk);
  SampleArgs sa{};
  sa.params = (const float*)h->params.p;
  sa.x = x;
  sa.t = t;
  sa.w = w;
  sa.h1 = (float*)h->h1.p;
  sa.h2 = (float*)h->h2.p;
  sa.d1 = (float*)h->d1.p;
  sa.d2 = (float*)h->d2.p;
  sa.d3 = (float*)h->d3.p;
  sa.loss_part = (float*)h->loss_part.p;
  sa.inv_wsum = inv_wsum;
  sa.B = (int32_t)B;
  sa.n_in = h->n_in;
  sa.n_out = h->n_out;
  sa.train = 1;
  hipLaunchKernelGGL(mlp_sample_kernel, dim3(blocks), dim3(256), 0, s, sa);
  GradArgs ga{x, sa.h1, sa.h2, sa.d1, sa.d2, sa.d3, (float*)h->partial.p, (int32_t)B, h->n_in};
  hipLaunchKernelGGL(mlp_wgrad_kernel, dim3((kP + 255) / 256, chunks), dim3(256), 0, s, ga);
  ReduceArgs ra{};
  ra.params = (float*)h->params.p;
  ra.m1 = (float*)h->m1.p;
  ra.m2 = (float*)h->m2.p;
  ra.grad = adam ? nullptr : (float*)h->grad.p;
  ra.partial = (const float*)h->partial.p;
  ra.loss_part = sa.loss_part;
  ra.loss = (float*)h->loss.p;
  ra.n_chunks = (int32_t)chunks;
  ra.n_loss_parts = (int32_t)blocks;
  if (adam) {
    ++h->t;
    ra.lr = lr;
    ra.c1 = 1.f - std::pow(0.9f, (float)h->t);
    ra.c2 = 1.f - std::pow(0.999f, (float)h->t);
  }
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kP + 255) / 256), dim3(256), 0, s, ra);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_hip_error(e, "mlp step");
    return RSGPU_ERR_HIP;
  }
  return RSGPU_OK;
}

// sum of the weights (device) -> host, for the loss normalisation
int weight_sum(const float* w, size_t n, hipStream_t s, float* out) {
  std::vector<float> hw(n);
  if (hipMemcpyAsync(hw.data(), w, n * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return RSGPU_ERR_HIP;
  double sum = 0;
  for (float v : hw) sum += v;
  if (!(sum > 0)) return RSGPU_ERR_INVALID_ARGUMENT;
  *out = (float)(1.0 / sum);
  return RSGPU_OK;
}

}  // namespace

extern "C" {

int rsgpu_mlp_new(int32_t n_in, int32_t n_out, uint64_t seed, rsgpu_mlp** out) {
  if (!out || n_in < 1 || n_in > kF || n_out < 1 || n_out > kO) return RSGPU_ERR_INVALID_ARGUMENT;
  auto* h = new rsgpu_mlp();
  h->n_in = n_in;
  h->n_out = n_out;
  (void)hipGetDevice(&h->device);
  if (h->params.alloc(kP * 4) != hipSuccess || h->m1.alloc(kP * 4) != hipSuccess ||
      h->m2.alloc(kP * 4) != hipSuccess || h->grad.alloc(kP * 4) != hipSuccess ||
      h->loss.alloc(64) != hipSuccess) {
    delete h;
    return RSGPU_ERR_OUT_OF_MEMORY;
  }
  // He-uniform weights, zero biases; padding stays zero.
  std::vector<float> u(n_user_params(n_in, n_out), 0.f);
  uint64_t st = seed;
  auto uni = [&](float bound) {
    return (float)((splitmix(st) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * bound;
  };
  size_t q = 0;
  const float a1 = std::sqrt(6.f / n_in), a2 = std::sqrt(6.f / kH);
  for (int i = 0; i < kH * n_in; ++i) u[q++] = uni(a1);
  q += kH;
  for (int i = 0; i < kH * kH; ++i) u[q++] = uni(a2);
  q += kH;
  for (int i = 0; i < n_out * kH; ++i) u[q++] = uni(a2);
  std::vector<float> d;
  pad_params(n_in, n_out, u.data(), d);
  std::vector<float> z(kP, 0.f);
  if (hipMemcpy(h->params.p, d.data(), kP * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->m1.p, z.data(), kP * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->m2.p, z.data(), kP * 4, hipMemcpyHostToDevice) != hipSuccess) {
    delete h;
    return RSGPU_ERR_HIP;
  }
  *out = h;
  return RSGPU_OK;
}

void rsgpu_mlp_free(rsgpu_mlp* h) { delete h; }

int rsgpu_mlp_dims(const rsgpu_mlp* h, int32_t* n_in, int32_t* n_hidden, int32_t* n_out,
                   size_t* n_params) {
  if (!h || !n_in || !n_hidden || !n_out || !n_params) return RSGPU_ERR_INVALID_ARGUMENT;
  *n_in = h->n_in;
  *n_hidden = kH;
  *n_out = h->n_out;
  *n_params = n_user_params(h->n_in, h->n_out);
  return RSGPU_OK;
}

int rsgpu_mlp_get_params(const rsgpu_mlp* h, float* out, size_t n) {
  if (!h || !out || n != n_user_params(h->n_in, h->n_out)) return RSGPU_ERR_INVALID_ARGUMENT;
  std::vector<float> d(kP);
  if (hipMemcpy(d.data(), h->params.p, kP * 4, hipMemcpyDeviceToHost) != hipSuccess) return RSGPU_ERR_HIP;
  unpad_params(h->n_in, h->n_out, d.data(), out);
  return RSGPU_OK;
}

int rsgpu_mlp_set_params(rsgpu_mlp* h, const float* in, size_t n) {
  if (!h || !in || n != n_user_params(h->n_in, h->n_out)) return RSGPU_ERR_INVALID_ARGUMENT;
  std::vector<float> d, z(kP, 0.f);
  pad_params(h->n_in, h->n_out, in, d);
  if (hipMemcpy(h->params.p, d.data(), kP * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->m1.p, z.data(), kP * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->m2.p, z.data(), kP * 4, hipMemcpyHostToDevice) != hipSuccess)
    return RSGPU_ERR_HIP;
  h->t = 0;
  return RSGPU_OK;
}

int rsgpu_mlp_forward(rsgpu_mlp* h, const float* x, size_t batch, float* y, void* stream) {
  if (!h || !x || !y || batch == 0 || batch > (1u << 30)) return RSGPU_ERR_INVALID_ARGUMENT;
  hipStream_t s = (hipStream_t)stream;
  SampleArgs sa{};
  sa.params = (const float*)h->params.p;
  sa.x = x;
  sa.y = y;
  sa.B = (int32_t)batch;
  sa.n_in = h->n_in;
  sa.n_out = h->n_out;
  sa.train = 0;
  hipLaunchKernelGGL(mlp_sample_kernel, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, s, sa);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_hip_error(e, "mlp_sample_kernel");
    return RSGPU_ERR_HIP;
  }
  return RSGPU_OK;
}

int rsgpu_mlp_loss_grad(rsgpu_mlp* h, const float* x, const float* t, const float* w, size_t batch,
                        float* loss, float* grad, size_t n, void* stream) {
  if (!h || !x || !t || !w || !loss || !grad || batch == 0 || batch > (1u << 30) ||
      n != n_user_params(h->n_in, h->n_out))
    return RSGPU_ERR_INVALID_ARGUMENT;
  hipStream_t s = (hipStream_t)stream;
  float inv = 0;
  int rc = weight_sum(w, batch * h->n_out, s, &inv);
  if (rc) return rc;
  if ((rc = step(h, x, t, w, batch, inv, false, 0.f, s))) return rc;
  std::vector<float> d(kP);
  if (hipMemcpyAsync(d.data(), h->grad.p, kP * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(loss, h->loss.p, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return RSGPU_ERR_HIP;
  unpad_params(h->n_in, h->n_out, d.data(), grad);
  return RSGPU_OK;
}

int rsgpu_mlp_train(rsgpu_mlp* h, const float* x, const float* t, const float* w, size_t batch,
                    int32_t steps, float lr, float* loss, void* stream) {
  if (!h || !x || !t || !w || batch == 0 || batch > (1u << 30) || steps < 0 || !(lr > 0))
    return RSGPU_ERR_INVALID_ARGUMENT;
  hipStream_t s = (hipStream_t)stream;
  float inv = 0;
  int rc = weight_sum(w, batch * h->n_out, s, &inv);
  if (rc) return rc;
  for (int32_t i = 0; i < steps; ++i)
    if ((rc = step(h, x, t, w, batch, inv, true, lr, s))) return rc;
  if (loss) {
    // loss of the parameters before the last update
    if (hipMemcpyAsync(loss, h->loss.p, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return RSGPU_ERR_HIP;
  }
  return RSGPU_OK;
}

}  // extern "C"
