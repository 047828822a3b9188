// Rank predictor network (8 -> 64 -> 64 -> 1, GELU, FP32): layout of the
// parameter vector and the device launchers.
#pragma once

#include <hip/hip_runtime.h>

namespace htg {

constexpr int MLP_IN      = 8;
constexpr int MLP_H       = 64;
constexpr int MLP_W1      = 0;
constexpr int MLP_B1      = MLP_W1 + MLP_H * MLP_IN;
constexpr int MLP_W2      = MLP_B1 + MLP_H;
constexpr int MLP_B2      = MLP_W2 + MLP_H * MLP_H;
constexpr int MLP_W3      = MLP_B2 + MLP_H;
constexpr int MLP_B3      = MLP_W3 + MLP_H;
constexpr int MLP_NPARAMS = MLP_B3 + 1; // 4801

void launch_mlp_predict(const float *params, const float *mean, const float *stdv, const float *feat, int n,
                        float *out, hipStream_t st);
// Tiles of a batch of B samples (partial-gradient rows needed).
int mlp_tiles(int B);
// Gradient of the batch loss at params into grad; samples idx[0..B) of feat /
// target (or 0..B-1 when idx is null). part holds mlp_tiles(B) x MLP_NPARAMS
// floats, loss B per-sample losses.
void launch_mlp_grad(const float *params, const float *mean, const float *stdv, const float *feat,
                     const float *target, const int *idx, int B, float *part, double *loss, float *grad,
                     hipStream_t st);
void launch_adam(float *params, float *m, float *v, const float *grad, float lr, int t, hipStream_t st);

} // namespace htg
