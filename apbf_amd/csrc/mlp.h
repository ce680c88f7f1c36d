// Learned lambda warm start: network shape, parameter layout and the
// inference launcher used by the scene.
#pragma once
#include "apbf_internal.h"

namespace apbf {

constexpr int MLP_IN = 16;  // per-particle features
constexpr int MLP_H = 64;   // hidden width (two hidden layers)
// Flat FP32 parameter vector: W1[n][k] (n < 64 hidden, k < 16 inputs), b1,
// W2[n][k] (n, k < 64), b2, w3[k], b3.
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr int MLP_PARAMS = MLP_OFF_B3 + 1;  // 5313
// Normalisation vector: feature means [16], feature scales [16], output scale.
constexpr int MLP_NORM = 2 * MLP_IN + 1;

void launch_mlp_infer(const float* params, const float* norm, const float* x, const uint32_t* d_len, uint32_t cap,
                      float* out, hipStream_t s);

}  // namespace apbf

// Device state of one network (opaque apbf_mlp of the C interface).
struct apbf_mlp {
    float* params = nullptr;  // MLP_PARAMS FP32 master
    float* norm = nullptr;    // 2 * MLP_IN + 1
    float* m1 = nullptr;
    float* m2 = nullptr;
    float* grad = nullptr;
    float* partial = nullptr;  // max blocks x (MLP_PARAMS + 1)
    float* xs = nullptr;       // gathered batch rows
    float* loss = nullptr;     // per step
    uint32_t* perm = nullptr;       // the call's sample order (a copy the captured steps read)
    uint32_t* step_base = nullptr;  // Adam steps done before the current call (device word)
    uint32_t max_batch = 0, max_blocks = 0, max_steps = 0, max_n = 0;
    // a training call's steps captured as one graph, replayed while the call's
    // arguments stay the same (one call per epoch)
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    const float* g_features = nullptr;
    const float* g_targets = nullptr;
    uint32_t g_n = 0, g_batch = 0, g_steps = 0;
    float g_lr = 0.0f;
};
