// mlp_kernels.hpp -- the learned cost model: an 8 -> 32 -> 32 -> 3 MLP
// (ReLU hidden layers, linear outputs) that predicts, for one vertex, the
// log2 of the lane-cycles each of three kernel families would spend on it
// (0: the wave-level kernels, 1: the workgroup kernel with an LDS table,
// 2: the chunked heavy-vertex kernels).  It has no reference counterpart:
// the reference bins by fixed degree thresholds (utils.h:44-85,
// nova_propogator.cuh:210-311).
//
// Kernels (FP32, one lane per sample; the layers are far too small for
// matrix cores: 2*(8*32 + 32*32 + 32*3) = 2752 FLOP per sample):
//   k_mlp_forward   y = f(x) for n samples
//   k_mlp_backward  mean-squared error over the measured outputs (mask),
//                   per-sample gradients reduced per workgroup in LDS and
//                   added to the gradient buffer with atomics
//   k_mlp_adam      Adam update of the 1,443 parameters
// Parameter layout (floats): W1[32][8] b1[32] W2[32][32] b2[32] W3[3][32] b3[3].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lpa {

constexpr int MLP_IN = 8, MLP_H = 32, MLP_OUT = 3;
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_H * MLP_IN, MLP_W2 = MLP_B1 + MLP_H,
              MLP_B2 = MLP_W2 + MLP_H * MLP_H, MLP_W3 = MLP_B2 + MLP_H, MLP_B3 = MLP_W3 + MLP_OUT * MLP_H,
              MLP_NPARAM = MLP_B3 + MLP_OUT;

struct MlpAct {
    float h1[MLP_H], h2[MLP_H], y[MLP_OUT];
};

__device__ __forceinline__ void mlp_eval(const float *__restrict__ P, const float *x, MlpAct &a) {
#pragma unroll
    for (int j = 0; j < MLP_H; j++) {
        float s = P[MLP_B1 + j];
#pragma unroll
        for (int i = 0; i < MLP_IN; i++) s += P[MLP_W1 + j * MLP_IN + i] * x[i];
        a.h1[j] = s > 0.f ? s : 0.f;
    }
#pragma unroll
    for (int j = 0; j < MLP_H; j++) {
        float s = P[MLP_B2 + j];
#pragma unroll
        for (int i = 0; i < MLP_H; i++) s += P[MLP_W2 + j * MLP_H + i] * a.h1[i];
        a.h2[j] = s > 0.f ? s : 0.f;
    }
#pragma unroll
    for (int j = 0; j < MLP_OUT; j++) {
        float s = P[MLP_B3 + j];
#pragma unroll
        for (int i = 0; i < MLP_H; i++) s += P[MLP_W3 + j * MLP_H + i] * a.h2[i];
        a.y[j] = s;
    }
}

__global__ void __launch_bounds__(256) k_mlp_forward(const float *__restrict__ params, const float *__restrict__ x,
                                                     float *__restrict__ y, uint64_t n) {
    __shared__ float P[MLP_NPARAM];
    for (int i = threadIdx.x; i < MLP_NPARAM; i += blockDim.x) P[i] = params[i];
    __syncthreads();
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float xi[MLP_IN];
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) xi[i] = x[t * MLP_IN + i];
    MlpAct a;
    mlp_eval(P, xi, a);
#pragma unroll
    for (int j = 0; j < MLP_OUT; j++) y[t * MLP_OUT + j] = a.y[j];
}

// loss = sum over samples and masked outputs of (y - target)^2 / count;
// grad += d loss / d params; *loss_acc += per-sample squared errors.
// Samples are idx[0..n) when idx is given, else 0..n-1.
__global__ void __launch_bounds__(256) k_mlp_backward(const float *__restrict__ params, const float *__restrict__ x,
                                                      const float *__restrict__ target,
                                                      const uint8_t *__restrict__ mask,
                                                      const uint32_t *__restrict__ idx, uint64_t n, float scale,
                                                      float *__restrict__ grad, float *__restrict__ loss_acc) {
    __shared__ float P[MLP_NPARAM];
    __shared__ float G[MLP_NPARAM];
    __shared__ float L;
    for (int i = threadIdx.x; i < MLP_NPARAM; i += blockDim.x) {
        P[i] = params[i];
        G[i] = 0.f;
    }
    if (threadIdx.x == 0) L = 0.f;
    __syncthreads();
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) {
        const uint64_t t = idx ? idx[q] : q;
        float xi[MLP_IN];
#pragma unroll
        for (int i = 0; i < MLP_IN; i++) xi[i] = x[t * MLP_IN + i];
        MlpAct a;
        mlp_eval(P, xi, a);
        float dy[MLP_OUT];
        float l = 0.f;
#pragma unroll
        for (int j = 0; j < MLP_OUT; j++) {
            const bool m = mask[t * MLP_OUT + j] != 0;
            const float e = m ? a.y[j] - target[t * MLP_OUT + j] : 0.f;
            l += e * e;
            dy[j] = 2.f * e * scale;
        }
        atomicAdd(&L, l * scale);
        float dh2[MLP_H];
#pragma unroll
        for (int i = 0; i < MLP_H; i++) dh2[i] = 0.f;
#pragma unroll
        for (int j = 0; j < MLP_OUT; j++) {
            if (dy[j] == 0.f) continue;
            atomicAdd(&G[MLP_B3 + j], dy[j]);
#pragma unroll
            for (int i = 0; i < MLP_H; i++) {
                atomicAdd(&G[MLP_W3 + j * MLP_H + i], dy[j] * a.h2[i]);
                dh2[i] += dy[j] * P[MLP_W3 + j * MLP_H + i];
            }
        }
        float dh1[MLP_H];
#pragma unroll
        for (int i = 0; i < MLP_H; i++) dh1[i] = 0.f;
#pragma unroll
        for (int j = 0; j < MLP_H; j++) {
            const float g = a.h2[j] > 0.f ? dh2[j] : 0.f;
            if (g == 0.f) continue;
            atomicAdd(&G[MLP_B2 + j], g);
#pragma unroll
            for (int i = 0; i < MLP_H; i++) {
                atomicAdd(&G[MLP_W2 + j * MLP_H + i], g * a.h1[i]);
                dh1[i] += g * P[MLP_W2 + j * MLP_H + i];
            }
        }
#pragma unroll
        for (int j = 0; j < MLP_H; j++) {
            const float g = a.h1[j] > 0.f ? dh1[j] : 0.f;
            if (g == 0.f) continue;
            atomicAdd(&G[MLP_B1 + j], g);
#pragma unroll
            for (int i = 0; i < MLP_IN; i++) atomicAdd(&G[MLP_W1 + j * MLP_IN + i], g * xi[i]);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < MLP_NPARAM; i += blockDim.x)
        if (G[i] != 0.f) atomicAdd(&grad[i], G[i]);
    if (threadIdx.x == 0) atomicAdd(loss_acc, L);
}

// Adam; c1 = 1/(1-b1^t), c2 = 1/(1-b2^t) are the bias corrections of step t.
__global__ void k_mlp_adam(float *__restrict__ params, const float *__restrict__ grad, float *__restrict__ m1,
                           float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_NPARAM) return;
    const float g = grad[i];
    const float a = b1 * m1[i] + (1.f - b1) * g;
    const float b = b2 * m2[i] + (1.f - b2) * g * g;
    m1[i] = a;
    m2[i] = b;
    const float ah = a * c1;
    const float bh = b * c2;
    params[i] -= lr * ah / (sqrtf(bh) + eps);
}

__global__ void k_zero_f32(float *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

}  // namespace lpa
