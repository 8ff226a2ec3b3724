// Learned fusion planner: MLP kernels (planner.hip) and block features.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace dq {
namespace mlp {

constexpr int kFeatures = 16;
constexpr int kHidden = 64;
constexpr int kMlpParams = kFeatures * kHidden + kHidden + kHidden * kHidden + kHidden + kHidden + 1;

struct Scratch {
  float *h1, *h2, *y, *dy, *d1, *d2;
};

void forward(const float* w, const float* x, int nb, float* h1, float* h2, float* y, hipStream_t st);
// Forward + backward over one batch; g receives kMlpParams gradients of the
// mean squared error, *loss the loss (if non-null).
void grad(const float* w, const float* x, const float* t, int nb, Scratch& s, float* g, float* loss,
          hipStream_t st);
void adam(float* w, const float* g, float* m, float* v, float lr, int step, hipStream_t st);
void gather(const float* X, const float* T, const uint32_t* idx, int nb, float* x, float* t, hipStream_t st);

}  // namespace mlp
}  // namespace dq
