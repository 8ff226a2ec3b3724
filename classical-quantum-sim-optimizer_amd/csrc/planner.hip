// Fusion-planner cost model: a 2-hidden-layer, 64-unit FP32 MLP trained on
// the GPU by these kernels.
//
// Parameters are one packed FP32 vector (kMlpParams entries):
//   W1[F][H], b1[H], W2[H][H], b2[H], w3[H], b3      (F = kFeatures = 16 inputs, H = 64)
// Model: h1 = relu(x W1 + b1), h2 = relu(h1 W2 + b2), y = h2 . w3 + b3.
// Loss: mean squared error over the batch.
// Every sum runs in a fixed sequential order (one thread per output), so a
// training run is deterministic; the CPU oracle restates the same math in FP64
// for the gradient checks.
#include <hip/hip_runtime.h>

#include "planner.hpp"

namespace dq {
namespace mlp {
namespace {

constexpr int F = kFeatures;
constexpr int H = kHidden;
constexpr int OW1 = 0, OB1 = OW1 + F * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H;

// One workgroup of H threads per sample: forward pass, activations kept.
__global__ __launch_bounds__(H) void k_forward(const float* __restrict__ w, const float* __restrict__ x, int nb,
                                               float* __restrict__ h1, float* __restrict__ h2,
                                               float* __restrict__ y) {
  __shared__ float s1[H], s2[H];
  const int b = blockIdx.x, j = threadIdx.x;
  if (b >= nb) return;
  float a = w[OB1 + j];
  for (int f = 0; f < F; ++f) a = __builtin_fmaf(x[b * F + f], w[OW1 + f * H + j], a);
  a = a > 0.f ? a : 0.f;
  s1[j] = a;
  if (h1) h1[b * H + j] = a;
  __syncthreads();
  float c = w[OB2 + j];
  for (int i = 0; i < H; ++i) c = __builtin_fmaf(s1[i], w[OW2 + i * H + j], c);
  c = c > 0.f ? c : 0.f;
  s2[j] = c;
  if (h2) h2[b * H + j] = c;
  __syncthreads();
  if (j == 0) {
    float o = w[OB3];
    for (int i = 0; i < H; ++i) o = __builtin_fmaf(s2[i], w[OW3 + i], o);
    y[b] = o;
  }
}

// Per-sample output and hidden-layer deltas.
__global__ __launch_bounds__(H) void k_backward_act(const float* __restrict__ w, const float* __restrict__ h1,
                                                    const float* __restrict__ h2, const float* __restrict__ y,
                                                    const float* __restrict__ t, int nb, float* __restrict__ dy,
                                                    float* __restrict__ d1, float* __restrict__ d2) {
  __shared__ float s2[H];
  const int b = blockIdx.x, j = threadIdx.x;
  if (b >= nb) return;
  const float g = 2.f * (y[b] - t[b]) / (float)nb;
  if (j == 0) dy[b] = g;
  const float e2 = h2[b * H + j] > 0.f ? g * w[OW3 + j] : 0.f;
  s2[j] = e2;
  d2[b * H + j] = e2;
  __syncthreads();
  float e1 = 0.f;
  for (int k = 0; k < H; ++k) e1 = __builtin_fmaf(s2[k], w[OW2 + j * H + k], e1);
  d1[b * H + j] = h1[b * H + j] > 0.f ? e1 : 0.f;
}

// One thread per parameter: gradient summed over the batch in sample order.
__global__ __launch_bounds__(256) void k_param_grad(const float* __restrict__ x, const float* __restrict__ h1,
                                                    const float* __restrict__ h2, const float* __restrict__ dy,
                                                    const float* __restrict__ d1, const float* __restrict__ d2,
                                                    int nb, float* __restrict__ g) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  float acc = 0.f;
  if (p < OB1) {
    const int f = p / H, j = p % H;
    for (int b = 0; b < nb; ++b) acc = __builtin_fmaf(x[b * F + f], d1[b * H + j], acc);
  } else if (p < OW2) {
    const int j = p - OB1;
    for (int b = 0; b < nb; ++b) acc += d1[b * H + j];
  } else if (p < OB2) {
    const int i = (p - OW2) / H, j = (p - OW2) % H;
    for (int b = 0; b < nb; ++b) acc = __builtin_fmaf(h1[b * H + i], d2[b * H + j], acc);
  } else if (p < OW3) {
    const int j = p - OB2;
    for (int b = 0; b < nb; ++b) acc += d2[b * H + j];
  } else if (p < OB3) {
    const int j = p - OW3;
    for (int b = 0; b < nb; ++b) acc = __builtin_fmaf(h2[b * H + j], dy[b], acc);
  } else {
    for (int b = 0; b < nb; ++b) acc += dy[b];
  }
  g[p] = acc;
}

__global__ void k_loss(const float* __restrict__ y, const float* __restrict__ t, int nb, float* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  float acc = 0.f;
  for (int b = 0; b < nb; ++b) {
    const float d = y[b] - t[b];
    acc = __builtin_fmaf(d, d, acc);
  }
  *out = acc / (float)nb;
}

// Adam (Kingma & Ba, 2015) with bias correction; step is 1-based.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                              float b2, float eps, int step) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  const float gi = g[p];
  const float mi = b1 * m[p] + (1.f - b1) * gi;
  const float vi = b2 * v[p] + (1.f - b2) * gi * gi;
  m[p] = mi;
  v[p] = vi;
  const float mh = mi / (1.f - powf(b1, (float)step));
  const float vh = vi / (1.f - powf(b2, (float)step));
  w[p] -= lr * mh / (sqrtf(vh) + eps);
}

// Gather a mini-batch by index (device-side shuffle order).
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ X, const float* __restrict__ T,
                                                const uint32_t* __restrict__ idx, int nb, float* __restrict__ x,
                                                float* __restrict__ t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb * (F + 1)) return;
  const int b = i / (F + 1), f = i % (F + 1);
  const uint32_t src = idx[b];
  if (f < F) x[b * F + f] = X[(uint64_t)src * F + f];
  else t[b] = T[src];
}

}  // namespace

void forward(const float* w, const float* x, int nb, float* h1, float* h2, float* y, hipStream_t st) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_forward, dim3(nb), dim3(H), 0, st, w, x, nb, h1, h2, y);
}

void grad(const float* w, const float* x, const float* t, int nb, Scratch& s, float* g, float* loss,
          hipStream_t st) {
  hipLaunchKernelGGL(k_forward, dim3(nb), dim3(H), 0, st, w, x, nb, s.h1, s.h2, s.y);
  hipLaunchKernelGGL(k_backward_act, dim3(nb), dim3(H), 0, st, w, s.h1, s.h2, s.y, t, nb, s.dy, s.d1, s.d2);
  hipLaunchKernelGGL(k_param_grad, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, x, s.h1, s.h2, s.dy, s.d1,
                     s.d2, nb, g);
  if (loss) hipLaunchKernelGGL(k_loss, dim3(1), dim3(1), 0, st, s.y, t, nb, loss);
}

void adam(float* w, const float* g, float* m, float* v, float lr, int step, hipStream_t st) {
  hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, w, g, m, v, lr, 0.9f, 0.999f, 1e-8f,
                     step);
}

void gather(const float* X, const float* T, const uint32_t* idx, int nb, float* x, float* t, hipStream_t st) {
  const int total = nb * (F + 1);
  hipLaunchKernelGGL(k_gather, dim3((total + 255) / 256), dim3(256), 0, st, X, T, idx, nb, x, t);
}

}  // namespace mlp
}  // namespace dq
