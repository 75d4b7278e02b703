// mlp.hip -- the fusion cost model's network: F inputs -> 64 -> 64 -> 1, ReLU,
// FP32, trained with mean squared error and Adam, all on the GPU.
//
// Parameter vector (kMlpParams floats, row-major):
//   W1[64][F], b1[64], W2[64][64], b2[64], w3[64], b3
// One wave (64 lanes) owns one sample at a time; lane t is hidden unit t of
// both layers.  The backward kernel recomputes the forward pass per sample
// and accumulates lane t's gradient rows (W2 row t, W1 row t, biases, w3[t])
// in registers across its share of the batch, then adds them to the global
// gradient with one atomic per element.
#include <hip/hip_runtime.h>

#include "mlp.hpp"

namespace dmx {

namespace {

constexpr int H = kMlpHidden;
constexpr int F = kMlpFeatures;

struct View {
    const float* W1;
    const float* b1;
    const float* W2;
    const float* b2;
    const float* w3;
    const float* b3;
};

__device__ __forceinline__ View view(const float* p)
{
    View v;
    v.W1 = p;
    v.b1 = v.W1 + H * F;
    v.W2 = v.b1 + H;
    v.b2 = v.W2 + H * H;
    v.w3 = v.b2 + H;
    v.b3 = v.w3 + H;
    return v;
}

__device__ __forceinline__ float wave_sum(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

}  // namespace

// Y[s] = f(X[s]) for s < count; blocks of 64 threads (one wave), grid-stride
// over samples.
__global__ void __launch_bounds__(64) k_mlp_forward(const float* __restrict__ params,
                                                    const float* __restrict__ X, int count,
                                                    float* __restrict__ Y)
{
    __shared__ float s_x[F];
    __shared__ float s_h1[H];
    const View v = view(params);
    const int t = threadIdx.x;
    for (int s = blockIdx.x; s < count; s += gridDim.x) {
        if (t < F) s_x[t] = X[(size_t)s * F + t];
        __syncthreads();
        float a1 = v.b1[t];
        for (int f = 0; f < F; f++) a1 += v.W1[t * F + f] * s_x[f];
        s_h1[t] = a1 > 0.f ? a1 : 0.f;
        __syncthreads();
        float a2 = v.b2[t];
        for (int j = 0; j < H; j++) a2 += v.W2[t * H + j] * s_h1[j];
        const float h2 = a2 > 0.f ? a2 : 0.f;
        const float y = wave_sum(v.w3[t] * h2);
        if (t == 0) Y[s] = y + v.b3[0];
        __syncthreads();
    }
}

// G += d/dparams of (1/norm) * sum_s (f(X[s]) - T[s])^2 over s < count.
// G must be zeroed by the caller.  Blocks of 64 threads.
__global__ void __launch_bounds__(64) k_mlp_backward(const float* __restrict__ params,
                                                     const float* __restrict__ X,
                                                     const float* __restrict__ T, int count,
                                                     float norm, float* __restrict__ G)
{
    __shared__ float s_x[F];
    __shared__ float s_h1[H];
    __shared__ float s_dh2[H];
    __shared__ float s_W2[H * H];
    const View v = view(params);
    const int t = threadIdx.x;
    for (int i = t; i < H * H; i += H) s_W2[i] = v.W2[i];

    float gW2[H], gW1[F];
    for (int j = 0; j < H; j++) gW2[j] = 0.f;
    for (int f = 0; f < F; f++) gW1[f] = 0.f;
    float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f;
    __syncthreads();

    for (int s = blockIdx.x; s < count; s += gridDim.x) {
        if (t < F) s_x[t] = X[(size_t)s * F + t];
        __syncthreads();
        float a1 = v.b1[t];
        for (int f = 0; f < F; f++) a1 += v.W1[t * F + f] * s_x[f];
        const float h1 = a1 > 0.f ? a1 : 0.f;
        s_h1[t] = h1;
        __syncthreads();
        float a2 = v.b2[t];
        for (int j = 0; j < H; j++) a2 += s_W2[t * H + j] * s_h1[j];
        const float h2 = a2 > 0.f ? a2 : 0.f;
        const float y = wave_sum(v.w3[t] * h2) + v.b3[0];
        const float dy = 2.f * (y - T[s]) / norm;
        const float dh2 = a2 > 0.f ? dy * v.w3[t] : 0.f;
        gw3 += dy * h2;
        gb3 += dy;
        gb2 += dh2;
        for (int j = 0; j < H; j++) gW2[j] += dh2 * s_h1[j];
        s_dh2[t] = dh2;
        __syncthreads();
        float dh1 = 0.f;
        for (int j = 0; j < H; j++) dh1 += s_W2[j * H + t] * s_dh2[j];
        dh1 = a1 > 0.f ? dh1 : 0.f;
        gb1 += dh1;
        for (int f = 0; f < F; f++) gW1[f] += dh1 * s_x[f];
        __syncthreads();
    }
    float* GW1 = G;
    float* Gb1 = GW1 + H * F;
    float* GW2 = Gb1 + H;
    float* Gb2 = GW2 + H * H;
    float* Gw3 = Gb2 + H;
    float* Gb3 = Gw3 + H;
    for (int f = 0; f < F; f++) atomicAdd(&GW1[t * F + f], gW1[f]);
    atomicAdd(&Gb1[t], gb1);
    for (int j = 0; j < H; j++) atomicAdd(&GW2[t * H + j], gW2[j]);
    atomicAdd(&Gb2[t], gb2);
    atomicAdd(&Gw3[t], gw3);
    if (t == 0) atomicAdd(&Gb3[0], gb3);
}

// Adam step on every parameter, then G <- 0.
__global__ void k_mlp_adam(float* __restrict__ P, float* __restrict__ G, float* __restrict__ M,
                           float* __restrict__ V, int n, float lr, float b1, float b2, float eps,
                           float c1, float c2)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float g = G[i];
        const float m = b1 * M[i] + (1.f - b1) * g;
        const float v2 = b2 * V[i] + (1.f - b2) * g * g;
        M[i] = m;
        V[i] = v2;
        P[i] -= lr * (m / c1) / (sqrtf(v2 / c2) + eps);
        G[i] = 0.f;
    }
}

__global__ void k_mlp_zero(float* __restrict__ G, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        G[i] = 0.f;
}

hipError_t mlp_forward(const float* d_params, const float* d_X, int count, float* d_Y,
                       hipStream_t stream)
{
    const int grid = count < 4096 ? (count > 0 ? count : 1) : 4096;
    k_mlp_forward<<<grid, 64, 0, stream>>>(d_params, d_X, count, d_Y);
    return hipGetLastError();
}

hipError_t mlp_backward(const float* d_params, const float* d_X, const float* d_T, int count,
                        float norm, float* d_G, hipStream_t stream)
{
    const int grid = count < 512 ? (count > 0 ? count : 1) : 512;
    k_mlp_backward<<<grid, 64, 0, stream>>>(d_params, d_X, d_T, count, norm, d_G);
    return hipGetLastError();
}

hipError_t mlp_adam(float* d_P, float* d_G, float* d_M, float* d_V, float lr, int step,
                    hipStream_t stream)
{
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.f, c2 = 1.f;
    for (int i = 0; i < step; i++) {
        c1 *= b1;
        c2 *= b2;
    }
    k_mlp_adam<<<(kMlpParams + 255) / 256, 256, 0, stream>>>(d_P, d_G, d_M, d_V, kMlpParams, lr,
                                                              b1, b2, 1e-8f, 1.f - c1, 1.f - c2);
    return hipGetLastError();
}

hipError_t mlp_zero(float* d_G, hipStream_t stream)
{
    k_mlp_zero<<<(kMlpParams + 255) / 256, 256, 0, stream>>>(d_G, kMlpParams);
    return hipGetLastError();
}

}  // namespace dmx
