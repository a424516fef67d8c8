// Refinement-indicator MLP: declarations shared by mlp.hip (kernels) and
// mlp_api.cpp (host + C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace drf {

constexpr int kMlpIn = 128, kMlpHidden = 64;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden +
                           kMlpHidden + kMlpHidden + 1; // 12,481
constexpr int kMlpQuant = kMlpHidden * kMlpIn + kMlpHidden * kMlpHidden + kMlpHidden;
constexpr int kMlpMaxParts = 512; // partial gradients (workgroups) per training step

// One Adam update (lr = 0: gradient only).  c1 = 1 / (1 - beta1^t),
// c2 = 1 / (1 - beta2^t).
struct AdamStep {
    float lr, beta1, beta2, eps, c1, c2;
};

int mlp_grad_parts(int64_t n);
hipError_t launch_mlp_forward(int64_t n, const uint16_t *X, const float *w, const uint16_t *wq,
                              float *pred, hipStream_t s);
hipError_t launch_mlp_grad(int64_t n, const uint16_t *X, const float *y, float *w, uint16_t *wq,
                           float *m, float *v, float *partial, float *lpart, float *grad,
                           float *loss, const AdamStep &adam, hipStream_t s);
hipError_t launch_mlp_classify(int64_t n, const float *pred, float refine_level,
                               float coarsen_level, uint8_t *action, hipStream_t s);

} // namespace drf
