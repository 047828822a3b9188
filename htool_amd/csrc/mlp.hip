// Rank predictor network on the GPU: an 8 -> 64 -> 64 -> 1 perceptron with
// GELU activations, FP32. Forward, backward, a deterministic weight-gradient
// reduction and Adam, all hand-written.
//
// Parameter vector (MLP_NPARAMS floats): W1 [64][8] (row j: weights of hidden
// unit j), b1 [64], W2 [64][64] (row j: weights of unit j of layer 2), b2 [64],
// w3 [64], b3 [1]. Inputs are normalised features (f - mean) / std.
// Loss over a batch of B samples: (1/B) sum 0.5 (out_b - y_b)^2. Work is done
// in tiles of 64 samples per workgroup with the activations in LDS.
#include "mlp.hpp"

#include <cmath>

namespace htg {
namespace {

constexpr int TS  = 64;  // samples per tile (one workgroup)
constexpr int TPB = 256; // threads per workgroup
constexpr int LD  = 65;  // padded row stride of the 64-wide LDS matrices

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
// d/dx [x Phi(x)] = Phi(x) + x phi(x)
__device__ __forceinline__ float gelu_grad(float x) {
    const float Phi = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float phi = 0.39894228040143268f * expf(-0.5f * x * x);
    return Phi + x * phi;
}

// Parameters in LDS: W1 [64][8], b1, W2 [64][LD] (padded), b2, w3, b3.
struct SParams {
    float W1[MLP_H * MLP_IN], b1[MLP_H], W2[MLP_H * LD], b2[MLP_H], w3[MLP_H], b3;
};

__device__ void stage_params(const float *params, SParams &s) {
    for (int i = threadIdx.x; i < MLP_H * MLP_IN; i += TPB)
        s.W1[i] = params[MLP_W1 + i];
    for (int i = threadIdx.x; i < MLP_H * MLP_H; i += TPB)
        s.W2[(i / MLP_H) * LD + i % MLP_H] = params[MLP_W2 + i];
    for (int i = threadIdx.x; i < MLP_H; i += TPB) {
        s.b1[i] = params[MLP_B1 + i];
        s.b2[i] = params[MLP_B2 + i];
        s.w3[i] = params[MLP_W3 + i];
    }
    if (threadIdx.x == 0)
        s.b3 = params[MLP_B3];
}

// Forward pass of a tile: X [TS][8] -> H1, H2 [TS][LD] (and G1, G2 = GELU'
// of the pre-activations when G1 is not null), out[TS].
__device__ void tile_forward(const SParams &p, const float (*X)[MLP_IN], float *H1, float *H2, float *G1, float *G2,
                             float *out) {
    const int r0 = 4 * (threadIdx.x / 16), c0 = 4 * (threadIdx.x % 16);
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int b = r0 + u, j = c0 + v;
            float a     = p.b1[j];
#pragma unroll
            for (int i = 0; i < MLP_IN; i++)
                a += p.W1[j * MLP_IN + i] * X[b][i];
            H1[b * LD + j] = gelu(a);
            if (G1)
                G1[b * LD + j] = gelu_grad(a);
        }
    __syncthreads();
    float acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++)
            acc[u][v] = p.b2[c0 + v];
    for (int i = 0; i < MLP_H; i++) {
        float h[4], w[4];
#pragma unroll
        for (int u = 0; u < 4; u++)
            h[u] = H1[(r0 + u) * LD + i];
#pragma unroll
        for (int v = 0; v < 4; v++)
            w[v] = p.W2[(c0 + v) * LD + i];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                acc[u][v] += h[u] * w[v];
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            H2[(r0 + u) * LD + c0 + v] = gelu(acc[u][v]);
            if (G2)
                G2[(r0 + u) * LD + c0 + v] = gelu_grad(acc[u][v]);
        }
    __syncthreads();
    if (threadIdx.x < TS) {
        const int b = threadIdx.x;
        float o     = p.b3;
        for (int j = 0; j < MLP_H; j++)
            o += p.w3[j] * H2[b * LD + j];
        out[b] = o;
    }
    __syncthreads();
}

__device__ void load_tile(const float *feat, const float *mean, const float *stdv, const int *idx, int b0, int n,
                          float (*X)[MLP_IN]) {
    for (int e = threadIdx.x; e < TS * MLP_IN; e += TPB) {
        const int b = e / MLP_IN, i = e % MLP_IN;
        float v     = 0;
        if (b0 + b < n) {
            const int src = idx ? idx[b0 + b] : b0 + b;
            v             = (feat[(size_t)src * MLP_IN + i] - mean[i]) / stdv[i];
        }
        X[b][i] = v;
    }
}

__global__ void __launch_bounds__(TPB) k_mlp_predict(const float *params, const float *mean, const float *stdv,
                                                     const float *feat, int n, float *out) {
    __shared__ SParams p;
    __shared__ float X[TS][MLP_IN];
    __shared__ float H1[TS * LD], H2[TS * LD], o[TS];
    stage_params(params, p);
    const int b0 = blockIdx.x * TS;
    load_tile(feat, mean, stdv, nullptr, b0, n, X);
    __syncthreads();
    tile_forward(p, X, H1, H2, nullptr, nullptr, o);
    if (threadIdx.x < TS && b0 + threadIdx.x < n)
        out[b0 + threadIdx.x] = o[threadIdx.x];
}

// Forward and backward pass of one tile of the batch (samples idx[b0..], or
// b0.. when idx is null), then the tile's partial gradient of every
// parameter, written to part[tile][MLP_NPARAMS]. Samples past B contribute 0.
__global__ void __launch_bounds__(TPB) k_mlp_tile_grad(const float *params, const float *mean, const float *stdv,
                                                       const float *feat, const float *target, const int *idx, int B,
                                                       float *part, double *loss) {
    __shared__ SParams p;
    __shared__ float X[TS][MLP_IN];
    __shared__ float H1[TS * LD], H2[TS * LD], G1[TS * LD], G2[TS * LD], o[TS], dO[TS];
    stage_params(params, p);
    const int b0 = blockIdx.x * TS;
    load_tile(feat, mean, stdv, idx, b0, B, X);
    __syncthreads();
    tile_forward(p, X, H1, H2, G1, G2, o);
    if (threadIdx.x < TS) {
        const int b = threadIdx.x;
        float g     = 0;
        if (b0 + b < B) {
            const int src = idx ? idx[b0 + b] : b0 + b;
            const float r = o[b] - target[src];
            loss[b0 + b]  = 0.5 * (double)r * (double)r;
            g             = r / (float)B;
        }
        dO[b] = g;
    }
    __syncthreads();
    const int r0 = 4 * (threadIdx.x / 16), c0 = 4 * (threadIdx.x % 16);
    // dZ2 = dOut w3 * GELU'(Z2), in place of G2
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++)
            G2[(r0 + u) * LD + c0 + v] *= dO[r0 + u] * p.w3[c0 + v];
    __syncthreads();
    // dZ1 = (dZ2 W2) * GELU'(Z1), in place of G1
    {
        float acc[4][4] = {};
        for (int j = 0; j < MLP_H; j++) {
            float d[4], w[4];
#pragma unroll
            for (int u = 0; u < 4; u++)
                d[u] = G2[(r0 + u) * LD + j];
#pragma unroll
            for (int v = 0; v < 4; v++)
                w[v] = p.W2[j * LD + c0 + v];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    acc[u][v] += d[u] * w[v];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                G1[(r0 + u) * LD + c0 + v] *= acc[u][v];
    }
    __syncthreads();
    float *pg = part + (size_t)blockIdx.x * MLP_NPARAMS;
    // dW2[j][i] = sum_b dZ2[b][j] H1[b][i]: thread owns j in r0.., i in c0..
    {
        float acc[4][4] = {};
        for (int b = 0; b < TS; b++) {
            float d[4], h[4];
#pragma unroll
            for (int u = 0; u < 4; u++)
                d[u] = G2[b * LD + r0 + u];
#pragma unroll
            for (int v = 0; v < 4; v++)
                h[v] = H1[b * LD + c0 + v];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    acc[u][v] += d[u] * h[v];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                pg[MLP_W2 + (r0 + u) * MLP_H + c0 + v] = acc[u][v];
    }
    // dW1[j][i] = sum_b dZ1[b][j] X[b][i]
    for (int e = threadIdx.x; e < MLP_H * MLP_IN; e += TPB) {
        const int j = e / MLP_IN, i = e % MLP_IN;
        float a     = 0;
        for (int b = 0; b < TS; b++)
            a += G1[b * LD + j] * X[b][i];
        pg[MLP_W1 + e] = a;
    }
    if (threadIdx.x < MLP_H) {
        const int j = threadIdx.x;
        float a1 = 0, a2 = 0, a3 = 0;
        for (int b = 0; b < TS; b++) {
            a1 += G1[b * LD + j];
            a2 += G2[b * LD + j];
            a3 += dO[b] * H2[b * LD + j];
        }
        pg[MLP_B1 + j] = a1;
        pg[MLP_B2 + j] = a2;
        pg[MLP_W3 + j] = a3;
    }
    if (threadIdx.x == 0) {
        float a = 0;
        for (int b = 0; b < TS; b++)
            a += dO[b];
        pg[MLP_B3] = a;
    }
}

// Sums the tiles' partial gradients in tile order (FP64 accumulation).
__global__ void k_mlp_reduce(const float *part, int ntiles, float *grad) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= MLP_NPARAMS)
        return;
    double a = 0;
    for (int t = 0; t < ntiles; t++)
        a += part[(size_t)t * MLP_NPARAMS + q];
    grad[q] = (float)a;
}

// Adam (Kingma and Ba, 2015) with bias correction; t is the 1-based step.
__global__ void k_adam(float *params, float *m, float *v, const float *grad, float lr, float beta1, float beta2,
                       float eps, int t) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAMS)
        return;
    const float g = grad[p];
    m[p]          = beta1 * m[p] + (1.0f - beta1) * g;
    v[p]          = beta2 * v[p] + (1.0f - beta2) * g * g;
    const float mh = m[p] / (1.0f - powf(beta1, (float)t));
    const float vh = v[p] / (1.0f - powf(beta2, (float)t));
    params[p] -= lr * mh / (sqrtf(vh) + eps);
}

} // namespace

int mlp_tiles(int B) { return (B + TS - 1) / TS; }

void launch_mlp_predict(const float *params, const float *mean, const float *stdv, const float *feat, int n,
                        float *out, hipStream_t st) {
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_mlp_predict, dim3(mlp_tiles(n)), dim3(TPB), 0, st, params, mean, stdv, feat, n, out);
}

void launch_mlp_grad(const float *params, const float *mean, const float *stdv, const float *feat,
                     const float *target, const int *idx, int B, float *part, double *loss, float *grad,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_tile_grad, dim3(mlp_tiles(B)), dim3(TPB), 0, st, params, mean, stdv, feat, target, idx,
                       B, part, loss);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAMS + 255) / 256), dim3(256), 0, st, part, mlp_tiles(B), grad);
}

void launch_adam(float *params, float *m, float *v, const float *grad, float lr, int t, hipStream_t st) {
    hipLaunchKernelGGL(k_adam, dim3((MLP_NPARAMS + 255) / 256), dim3(256), 0, st, params, m, v, grad, lr, 0.9f,
                       0.999f, 1e-8f, t);
}

} // namespace htg
