// Learned troubled-cell indicator (gfx950): a 4 -> 64 -> 64 -> 64 -> 1
// perceptron with GELU hidden layers and a sigmoid output that predicts the
// DG/FV blending factor alpha of an element from the log modal energy
// fractions of rho and p in the top one and top two mode shells.
//
// The reference has no learned indicator (its blending traces are in
// setups/mhd/3d/_common_/setup_checkpoint_mod.f90:395-450); this is the
// library's own network, trained on (features, alpha) pairs that the
// Persson-Peraire indicator produces during the method's own runs.
//
// Parameter vector (FP32, canonical layout, NPARAMS = 8705):
//   W1[64][4], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], w4[64], b4
// with W[j][k] the weight from input k to unit j.  The inference layout
// used by the stage kernel stores W2 and W3 in blocks of four inputs,
// W[j][4b..4b+3] at 4 (64 b + j): lane j reads its unit's four weights as one
// 16-byte load and a wavefront reads 1 KB contiguous per block (16 loads per
// layer instead of 64).
//
// Kernels map one hidden unit to one lane of a 64-lane wavefront; an
// activation of another unit is read with v_readlane (uniform index).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nemo {

constexpr int MLP_IN = 4, MLP_H = 64;
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_H * MLP_IN, MLP_W2 = MLP_B1 + MLP_H, MLP_B2 = MLP_W2 + MLP_H * MLP_H,
              MLP_W3 = MLP_B2 + MLP_H, MLP_B3 = MLP_W3 + MLP_H * MLP_H, MLP_W4 = MLP_B3 + MLP_H,
              MLP_B4 = MLP_W4 + MLP_H, MLP_NP = MLP_B4 + 1;
static_assert(MLP_NP == 8705, "parameter count");

// Feature of a modal energy fraction f in [0, 1]: 0.1 ln(f + 1e-12).
__device__ __forceinline__ float mlp_feature(float f) { return 0.1f * __logf(f + 1.0e-12f); }

// GELU in its tanh form, 0.5 x (1 + tanh(c (x + 0.044715 x^3))),
// c = sqrt(2/pi) (Hendrycks & Gimpel 2016), and its derivative.
// tanh(y) = 1 - 2/(exp(2y) + 1).
__device__ __forceinline__ float gelu_t(float x) {
    const float y = 0.7978845608028654f * (x + 0.044715f * x * x * x);
    return 1.0f - 2.0f / (__expf(2.0f * y) + 1.0f);
}
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + gelu_t(x)); }
__device__ __forceinline__ float gelu_grad(float x) {
    const float t = gelu_t(x);
    const float dy = 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x * x);
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * dy;
}
__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }

__device__ __forceinline__ float lane_bcast(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

// Wave-wide sum (64 lanes) by butterfly shuffles.
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Forward pass for NB samples held uniformly by the wavefront (feat[s][k] is
// the same in every lane), with the inference-layout parameters P.  Lane j
// computes unit j of every hidden layer; y[s] (uniform) is the sigmoid output.
template <int NB>
__device__ __forceinline__ void mlp_forward_wave(const float* __restrict__ P, const float (&feat)[NB][MLP_IN],
                                                 float (&y)[NB]) {
    const int j = threadIdx.x & 63;
    float a1[NB], a2[NB], a3[NB];
    {
        const float w0 = P[MLP_W1 + j * 4 + 0], w1 = P[MLP_W1 + j * 4 + 1], w2 = P[MLP_W1 + j * 4 + 2],
                    w3 = P[MLP_W1 + j * 4 + 3], b = P[MLP_B1 + j];
#pragma unroll
        for (int s = 0; s < NB; ++s)
            a1[s] = gelu(b + w0 * feat[s][0] + w1 * feat[s][1] + w2 * feat[s][2] + w3 * feat[s][3]);
    }
    {
        float z[NB];
        const float b = P[MLP_B2 + j];
#pragma unroll
        for (int s = 0; s < NB; ++s) z[s] = b;
        for (int kb = 0; kb < MLP_H / 4; ++kb) {
            const float4 w = *reinterpret_cast<const float4*>(P + MLP_W2 + kb * 4 * MLP_H + j * 4);
#pragma unroll
            for (int s = 0; s < NB; ++s) {
                z[s] += w.x * lane_bcast(a1[s], 4 * kb);
                z[s] += w.y * lane_bcast(a1[s], 4 * kb + 1);
                z[s] += w.z * lane_bcast(a1[s], 4 * kb + 2);
                z[s] += w.w * lane_bcast(a1[s], 4 * kb + 3);
            }
        }
#pragma unroll
        for (int s = 0; s < NB; ++s) a2[s] = gelu(z[s]);
    }
    {
        float z[NB];
        const float b = P[MLP_B3 + j];
#pragma unroll
        for (int s = 0; s < NB; ++s) z[s] = b;
        for (int kb = 0; kb < MLP_H / 4; ++kb) {
            const float4 w = *reinterpret_cast<const float4*>(P + MLP_W3 + kb * 4 * MLP_H + j * 4);
#pragma unroll
            for (int s = 0; s < NB; ++s) {
                z[s] += w.x * lane_bcast(a2[s], 4 * kb);
                z[s] += w.y * lane_bcast(a2[s], 4 * kb + 1);
                z[s] += w.z * lane_bcast(a2[s], 4 * kb + 2);
                z[s] += w.w * lane_bcast(a2[s], 4 * kb + 3);
            }
        }
#pragma unroll
        for (int s = 0; s < NB; ++s) a3[s] = gelu(z[s]);
    }
    const float w4 = P[MLP_W4 + j], b4 = P[MLP_B4];
#pragma unroll
    for (int s = 0; s < NB; ++s) y[s] = sigmoid(b4 + wave_sum(w4 * a3[s]));
}

// ---------------------------------------------------------------------------
// Training kernels (canonical layout)
// ---------------------------------------------------------------------------

// Forward of n samples for the host API (inference layout): each wavefront
// takes groups of 4 samples.
static __global__ __launch_bounds__(64) void mlp_forward_kernel(const float* __restrict__ P, long long n,
                                                         const float* __restrict__ feat, float* __restrict__ out) {
    const long long groups = (n + 3) / 4;
    for (long long g = blockIdx.x; g < groups; g += gridDim.x) {
        float f[4][MLP_IN], y[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const long long i = g * 4 + s < n ? g * 4 + s : n - 1;
#pragma unroll
            for (int k = 0; k < MLP_IN; ++k) f[s][k] = feat[i * MLP_IN + k];
        }
        mlp_forward_wave<4>(P, f, y);
        if (threadIdx.x < 4 && g * 4 + threadIdx.x < n) {
            const int s = threadIdx.x;
            out[g * 4 + s] = s == 0 ? y[0] : (s == 1 ? y[1] : (s == 2 ? y[2] : y[3]));
        }
    }
}

// Loss and gradient of L = 1/(2B) sum_s (y_s - t_s)^2 over one minibatch
// (sample ids idx[off .. off+B), or off .. off+B when idx is null).  Block b
// takes samples b, b + gridDim.x, ... and writes its partial gradient
// (canonical layout) to part[b][NP] and its partial loss to part_loss[b];
// mlp_reduce_kernel sums the partials in a fixed order (deterministic).
// The weights are staged in LDS (rows padded to 65 floats so that both
// row and column reads are conflict-free).
static __global__ __launch_bounds__(64) void mlp_grad_kernel(const float* __restrict__ P, const float* __restrict__ feat,
                                                      const float* __restrict__ target,
                                                      const long long* __restrict__ idx, long long off, int B,
                                                      float* __restrict__ part, float* __restrict__ part_loss) {
    constexpr int LD = MLP_H + 1;
    __shared__ float sW2[MLP_H * LD], sW3[MLP_H * LD];
    const int j = threadIdx.x;
    for (int r = 0; r < MLP_H; ++r) {
        sW2[r * LD + j] = P[MLP_W2 + r * MLP_H + j];
        sW3[r * LD + j] = P[MLP_W3 + r * MLP_H + j];
    }
    __syncthreads();
    const float w10 = P[MLP_W1 + j * 4 + 0], w11 = P[MLP_W1 + j * 4 + 1], w12 = P[MLP_W1 + j * 4 + 2],
                w13 = P[MLP_W1 + j * 4 + 3];
    const float b1 = P[MLP_B1 + j], b2 = P[MLP_B2 + j], b3 = P[MLP_B3 + j], w4 = P[MLP_W4 + j], b4 = P[MLP_B4];
    float gW1[4] = {0, 0, 0, 0}, gb1 = 0, gb2 = 0, gb3 = 0, gw4 = 0, gb4 = 0, loss = 0;
    float gW2[MLP_H], gW3[MLP_H];
#pragma unroll
    for (int k = 0; k < MLP_H; ++k) gW2[k] = gW3[k] = 0.0f;
    const float invB = 1.0f / float(B);
    for (int s = blockIdx.x; s < B; s += gridDim.x) {
        const long long i = idx ? idx[off + s] : off + s;
        const float x0 = feat[i * 4 + 0], x1 = feat[i * 4 + 1], x2 = feat[i * 4 + 2], x3 = feat[i * 4 + 3];
        const float t = target[i];
        // forward
        const float z1 = b1 + w10 * x0 + w11 * x1 + w12 * x2 + w13 * x3;
        const float a1 = gelu(z1);
        float z2 = b2;
        for (int k = 0; k < MLP_H; ++k) z2 += sW2[j * LD + k] * lane_bcast(a1, k);
        const float a2 = gelu(z2);
        float z3 = b3;
        for (int k = 0; k < MLP_H; ++k) z3 += sW3[j * LD + k] * lane_bcast(a2, k);
        const float a3 = gelu(z3);
        const float z4 = b4 + wave_sum(w4 * a3);
        const float y = sigmoid(z4);
        const float r = y - t;
        loss += 0.5f * r * r * invB;
        // backward
        const float d4 = r * y * (1.0f - y) * invB;
        gb4 += d4;
        gw4 += d4 * a3;
        const float d3 = d4 * w4 * gelu_grad(z3);
        gb3 += d3;
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) gW3[k] += d3 * lane_bcast(a2, k);
        float da2 = 0.0f;  // sum_r W3[r][j] d3_r
        for (int rr = 0; rr < MLP_H; ++rr) da2 += sW3[rr * LD + j] * lane_bcast(d3, rr);
        const float d2 = da2 * gelu_grad(z2);
        gb2 += d2;
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) gW2[k] += d2 * lane_bcast(a1, k);
        float da1 = 0.0f;
        for (int rr = 0; rr < MLP_H; ++rr) da1 += sW2[rr * LD + j] * lane_bcast(d2, rr);
        const float d1 = da1 * gelu_grad(z1);
        gb1 += d1;
        gW1[0] += d1 * x0;
        gW1[1] += d1 * x1;
        gW1[2] += d1 * x2;
        gW1[3] += d1 * x3;
    }
    float* g = part + (long long)blockIdx.x * MLP_NP;
#pragma unroll
    for (int k = 0; k < 4; ++k) g[MLP_W1 + j * 4 + k] = gW1[k];
    g[MLP_B1 + j] = gb1;
#pragma unroll
    for (int k = 0; k < MLP_H; ++k) {
        g[MLP_W2 + j * MLP_H + k] = gW2[k];
        g[MLP_W3 + j * MLP_H + k] = gW3[k];
    }
    g[MLP_B2 + j] = gb2;
    g[MLP_B3 + j] = gb3;
    g[MLP_W4 + j] = gw4;
    if (j == 0) {
        g[MLP_B4] = gb4;  // uniform across lanes
        part_loss[blockIdx.x] = loss;
    }
}

// Sum of the per-block partial gradients in block order (deterministic);
// with adam != 0 also one Adam step (Kingma & Ba 2015) on P (canonical) with
// moments m, v and bias corrections c1 = 1 - beta1^t, c2 = 1 - beta2^t.
static __global__ void mlp_reduce_kernel(const float* __restrict__ part, int nparts, const float* __restrict__ part_loss,
                                  float* __restrict__ grad, float* __restrict__ loss, float* __restrict__ P,
                                  float* __restrict__ m, float* __restrict__ v, int adam, float lr, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && loss) {
        float l = 0.0f;
        for (int b = 0; b < nparts; ++b) l += part_loss[b];
        *loss += l;
    }
    if (p >= MLP_NP) return;
    float s = 0.0f;
    for (int b = 0; b < nparts; ++b) s += part[(long long)b * MLP_NP + p];
    if (grad) grad[p] = s;
    if (adam) {
        const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
        const float mm = beta1 * m[p] + (1.0f - beta1) * s;
        const float vv = beta2 * v[p] + (1.0f - beta2) * s * s;
        m[p] = mm;
        v[p] = vv;
        P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    }
}

// Canonical -> inference layout (W2 and W3 in blocks of four inputs).
static __global__ void mlp_inference_layout_kernel(const float* __restrict__ P, float* __restrict__ Q) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NP) return;
    if (p >= MLP_W2 && p < MLP_B2) {
        const int r = (p - MLP_W2) / MLP_H, c = (p - MLP_W2) % MLP_H;
        Q[MLP_W2 + (c >> 2) * 4 * MLP_H + r * 4 + (c & 3)] = P[p];
    } else if (p >= MLP_W3 && p < MLP_B3) {
        const int r = (p - MLP_W3) / MLP_H, c = (p - MLP_W3) % MLP_H;
        Q[MLP_W3 + (c >> 2) * 4 * MLP_H + r * 4 + (c & 3)] = P[p];
    } else {
        Q[p] = P[p];
    }
}

}  // namespace nemo
