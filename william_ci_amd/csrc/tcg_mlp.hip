// The per-quartet network: 6 -> 64 -> 64 -> 2, ReLU, BF16 operands on
// v_mfma_f32_16x16x16bf16_1k with FP32 accumulation (gfx950).
//
// Parameters (FP32 master copy, converted to BF16 at load), in this order:
//   W1[6][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2]       (kMlpParams floats)
// Outputs: y0 = log10 of the quartet's largest integral, y1 = the index of
// the quadrature order it needs (both regressed, mean-squared error).
//
//   k_mlp_fwd    inference: one wavefront per 16 samples, the hidden layers
//                in LDS between the matrix-core steps.
//   k_mlp_step   forward + backward of a mini-batch: four wavefronts x 16
//                samples per workgroup; weight gradients as MFMA products
//                over the sample dimension, reduced in LDS, then one global
//                FP32 atomic per parameter and workgroup.
//   k_mlp_adam   Adam update of the FP32 master parameters, gradients zeroed.
//
// MFMA lane layout (16x16x16, BF16): lane l supplies A[l%16][4(l/16)+i] and
// B[4(l/16)+i][l%16], i = 0..3; D[4(l/16)+v][l%16] = acc[v].
#include <hip/hip_runtime.h>

#include <cstdint>

#include "tcg_device.hpp"

namespace tcg {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 6, kHid = 64, kOut = 2;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kIn * kHid, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid * kOut, kMlpParamsDev = kOffB3 + kOut;

// round to nearest even
__device__ __forceinline__ short bf16(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ __forceinline__ f32x4 mma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// forward of the 16 samples of one wavefront; X: [16][kIn] rows in LDS-free
// registers (xs: this lane's sample row l%16), H1/H2: [16][64] in LDS.
// Returns the output tile D3 (cols 0..1 valid).
__device__ __forceinline__ f32x4 wave_forward(const float *__restrict__ P, const float *xrow, bool valid, int lane,
                                              float *H1, float *H2) {
    const int r = lane & 15, q = lane >> 4;
    // layer 1: A = X (16 x 16, features 0..5), B = W1 (16 x 64)
    bf16x4 a;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int k = 4 * q + i;
        a[i] = (valid && k < kIn) ? bf16(xrow[k]) : (short)0;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        bf16x4 b;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = 4 * q + i;
            b[i] = k < kIn ? bf16(P[kOffW1 + k * kHid + 16 * t + r]) : (short)0;
        }
        f32x4 d = mma(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float h = d[v] + P[kOffB1 + 16 * t + r];
            H1[(4 * q + v) * kHid + 16 * t + r] = h > 0.f ? h : 0.f;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // layer 2: A = H1 (16 x 64), B = W2 (64 x 64)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            bf16x4 aa, b;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = 16 * kt + 4 * q + i;
                aa[i] = bf16(H1[r * kHid + k]);
                b[i] = bf16(P[kOffW2 + k * kHid + 16 * t + r]);
            }
            d = mma(aa, b, d);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float h = d[v] + P[kOffB2 + 16 * t + r];
            H2[(4 * q + v) * kHid + 16 * t + r] = h > 0.f ? h : 0.f;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // layer 3: A = H2 (16 x 64), B = W3 (64 x 16, cols 0..1)
    f32x4 d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
        bf16x4 aa, b;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = 16 * kt + 4 * q + i;
            aa[i] = bf16(H2[r * kHid + k]);
            b[i] = r < kOut ? bf16(P[kOffW3 + k * kOut + r]) : (short)0;
        }
        d = mma(aa, b, d);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) d[v] += r < kOut ? P[kOffB3 + r] : 0.f;
    return d;
}

__global__ __launch_bounds__(256) void k_mlp_fwd(const float *__restrict__ P, const float *__restrict__ X, int n,
                                                 float *__restrict__ Y) {
    __shared__ float H1s[4][16 * kHid], H2s[4][16 * kHid];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    for (long long base = ((long long)blockIdx.x * 4 + wv) * 16; base < n; base += (long long)gridDim.x * 64) {
        const long long s = base + r;
        const bool valid = s < n;
        float xrow[kIn];
#pragma unroll
        for (int k = 0; k < kIn; ++k) xrow[k] = valid ? X[s * kIn + k] : 0.f;
        const f32x4 d = wave_forward(P, xrow, valid, lane, H1s[wv], H2s[wv]);
        // lane (r, q) holds D[row = 4q + v][col = r]
        if (r < kOut) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long so = base + 4 * q + v;
                if (so < n) Y[so * kOut + r] = d[v];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// forward + backward; loss = (1/n_total) sum_s sum_o wo (y - t)^2, grads added into G
__global__ __launch_bounds__(256) void k_mlp_step(const float *__restrict__ P, const float *__restrict__ X,
                                                  const float *__restrict__ T, int n, float inv_total, float w0,
                                                  float w1, float *__restrict__ G, float *__restrict__ loss) {
    __shared__ float H1s[4][16 * kHid], H2s[4][16 * kHid], dHs[4][16 * kHid], Gl[kMlpParamsDev];
    __shared__ float dYs[4][16 * 16], Xs[4][16 * 16];
    __shared__ float lsum[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    for (int i = threadIdx.x; i < kMlpParamsDev; i += 256) Gl[i] = 0.f;
    if (lane == 0) lsum[wv] = 0.f;
    __syncthreads();
    float *H1 = H1s[wv], *H2 = H2s[wv], *dH = dHs[wv], *dY = dYs[wv], *Xw = Xs[wv];
    for (long long base = ((long long)blockIdx.x * 4 + wv) * 16; base < n; base += (long long)gridDim.x * 64) {
        const long long s = base + r;
        const bool valid = s < n;
        float xrow[kIn];
#pragma unroll
        for (int k = 0; k < kIn; ++k) xrow[k] = valid ? X[s * kIn + k] : 0.f;
        // X tile (16 x 16, features padded) for the W1 gradient
        if (q == 0) {
#pragma unroll
            for (int k = 0; k < 16; ++k) Xw[r * 16 + k] = k < kIn ? xrow[k] : 0.f;
        }
        const f32x4 d = wave_forward(P, xrow, valid, lane, H1, H2);
        // dY (16 x 16, cols 0..1), loss
        float lw = 0.f;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const long long so = base + 4 * q + v;
            float g = 0.f;
            if (r < kOut && so < n) {
                const float e = d[v] - T[so * kOut + r];
                const float w = r == 0 ? w0 : w1;
                lw += w * e * e;
                g = 2.f * w * e * inv_total;
            }
            dY[(4 * q + v) * 16 + r] = g;
        }
        for (int off = 32; off > 0; off >>= 1) lw += __shfl_xor(lw, off, 64);
        if (lane == 0) lsum[wv] += lw;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // dH2 = dY W3^T (.) [H2 > 0]: A = dY (16 x 16), B[k = o][n = j] = W3[j][o]
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            bf16x4 a, b;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = 4 * q + i;
                a[i] = bf16(dY[r * 16 + k]);
                b[i] = k < kOut ? bf16(P[kOffW3 + (16 * t + r) * kOut + k]) : (short)0;
            }
            const f32x4 g = mma(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 4 * q + v, col = 16 * t + r;
                dH[row * kHid + col] = H2[row * kHid + col] > 0.f ? g[v] : 0.f;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // dW3[j][o] = sum_s H2[s][j] dY[s][o]: A[row = j][k = s] = H2[s][j], B[k = s][n = o] = dY[s][o]
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            bf16x4 a, b;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = 4 * q + i;
                a[i] = bf16(H2[k * kHid + 16 * t + r]);
                b[i] = bf16(dY[k * 16 + r]);
            }
            const f32x4 g = mma(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
            if (r < kOut) {
#pragma unroll
                for (int v = 0; v < 4; ++v) atomicAdd(&Gl[kOffW3 + (16 * t + 4 * q + v) * kOut + r], g[v]);
            }
        }
        // db3, db2
        if (lane < kOut) {
            float sacc = 0.f;
            for (int k = 0; k < 16; ++k) sacc += dY[k * 16 + lane];
            atomicAdd(&Gl[kOffB3 + lane], sacc);
        }
        {
            float sacc = 0.f;
            for (int k = 0; k < 16; ++k) sacc += dH[k * kHid + lane];
            atomicAdd(&Gl[kOffB2 + lane], sacc);
        }
        // dW2[i][j] = sum_s H1[s][i] dH2[s][j]: A[row = i][k = s] = H1[s][i], B[k = s][n = j] = dH2[s][j]
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            bf16x4 a;
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = bf16(H1[(4 * q + i) * kHid + 16 * mt + r]);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                bf16x4 b;
#pragma unroll
                for (int i = 0; i < 4; ++i) b[i] = bf16(dH[(4 * q + i) * kHid + 16 * t + r]);
                const f32x4 g = mma(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
                for (int v = 0; v < 4; ++v) atomicAdd(&Gl[kOffW2 + (16 * mt + 4 * q + v) * kHid + 16 * t + r], g[v]);
            }
        }
        // dH1 = dH2 W2^T (.) [H1 > 0]: A = dH2 (16 x 64), B[k = j][n = i] = W2[i][j]; kept in registers
        f32x4 g1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 g = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                bf16x4 a, b;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int k = 16 * kt + 4 * q + i;
                    a[i] = bf16(dH[r * kHid + k]);
                    b[i] = bf16(P[kOffW2 + (16 * t + r) * kHid + k]);
                }
                g = mma(a, b, g);
            }
            g1[t] = g;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 4 * q + v, col = 16 * t + r;
                dH[row * kHid + col] = H1[row * kHid + col] > 0.f ? g1[t][v] : 0.f;
            }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // db1
        {
            float sacc = 0.f;
            for (int k = 0; k < 16; ++k) sacc += dH[k * kHid + lane];
            atomicAdd(&Gl[kOffB1 + lane], sacc);
        }
        // dW1[f][j] = sum_s X[s][f] dH1[s][j]: A[row = f][k = s] = X[s][f], B[k = s][n = j] = dH1[s][j]
        {
            bf16x4 a;
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = bf16(Xw[(4 * q + i) * 16 + r]);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                bf16x4 b;
#pragma unroll
                for (int i = 0; i < 4; ++i) b[i] = bf16(dH[(4 * q + i) * kHid + 16 * t + r]);
                const f32x4 g = mma(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int f = 4 * q + v;
                    if (f < kIn) atomicAdd(&Gl[kOffW1 + f * kHid + 16 * t + r], g[v]);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kMlpParamsDev; i += 256)
        if (Gl[i] != 0.f) atomicAdd(&G[i], Gl[i]);
    if (threadIdx.x == 0) atomicAdd(loss, (lsum[0] + lsum[1] + lsum[2] + lsum[3]) * inv_total);
}

__global__ void k_mlp_adam(float *__restrict__ P, float *__restrict__ G, float *__restrict__ M, float *__restrict__ V,
                           float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParamsDev) return;
    const float g = G[i];
    const float m = b1 * M[i] + (1.f - b1) * g;
    const float v = b2 * V[i] + (1.f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
    G[i] = 0.f;
}

int mlp_param_count() { return kMlpParamsDev; }

hipError_t launch_mlp_fwd(const float *P, const float *X, int n, float *Y, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int grid = (n + 63) / 64 < 4096 ? (n + 63) / 64 : 4096;
    hipLaunchKernelGGL(k_mlp_fwd, dim3(grid), dim3(256), 0, st, P, X, n, Y);
    return hipGetLastError();
}

hipError_t launch_mlp_step(const float *P, const float *X, const float *T, int n, float inv_total, float w0, float w1,
                           float *G, float *loss, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int grid = (n + 63) / 64 < 256 ? (n + 63) / 64 : 256;
    hipLaunchKernelGGL(k_mlp_step, dim3(grid), dim3(256), 0, st, P, X, T, n, inv_total, w0, w1, G, loss);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float *P, float *G, float *M, float *V, float lr, float b1, float b2, float eps, float c1,
                           float c2, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParamsDev + 255) / 256), dim3(256), 0, st, P, G, M, V, lr, b1, b2, eps,
                       c1, c2);
    return hipGetLastError();
}

}  // namespace tcg
