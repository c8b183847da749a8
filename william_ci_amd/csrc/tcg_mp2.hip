// Gamma-supercell MP2 on the GPU (gfx950): the full (non-SAO) ERI layout,
// the (ia|jb) transform as FP64 matrix-core GEMMs, and the MP2 sums.
//
//   k_translate_rows  rows i > nfuc of ERI(nBas,nBas,nBas,nBas) from the rows of
//                     the unit cell: (i j|k l) = (i0 j'|k' l'), all indices
//                     shifted by the lattice vector taking i to the unit cell
//                     (src/Tools/symmetric_matrix.f90:62-131).
//   k_dgemm_nn        strided-batched C_b = A_b B_b, column-major, FP64 on
//                     v_mfma_f64_16x16x4f64: 64x64 tiles per workgroup of four
//                     wavefronts (32x32 each, 2x2 MFMA tiles), K in steps of 16
//                     staged through LDS.
//   k_mp2_sums        E2d = -sum (ia|jb)^2 / D, E2x = -sum (ia|jb)(ib|ja) / D,
//                     D = e_a + e_b - e_i - e_j (src/Energies/MP2/MP2.f90:39-68),
//                     one partial pair per workgroup (summed on the host in a
//                     fixed order).
#include <hip/hip_runtime.h>

#include "tcg_device.hpp"

namespace tcg {

typedef double dbl4 __attribute__((ext_vector_type(4)));

__global__ void k_translate_rows(double *__restrict__ eri, int nb, int nf, int ncell, const int *__restrict__ relc) {
    const size_t n4 = (size_t)nb * nb * nb * nb;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n4; e += (size_t)gridDim.x * blockDim.x) {
        size_t r = e;
        const int i = (int)(r % nb); r /= nb;
        if (i < nf) continue;
        const int j = (int)(r % nb); r /= nb;
        const int k = (int)(r % nb);
        const int l = (int)(r / nb);
        const int ci = i / nf;
        const int* rel = relc + (size_t)ci * ncell;
        const int j2 = rel[j / nf] * nf + j % nf;
        const int k2 = rel[k / nf] * nf + k % nf;
        const int l2 = rel[l / nf] * nf + l % nf;
        eri[e] = eri[(size_t)(i % nf) + nb * ((size_t)j2 + nb * ((size_t)k2 + (size_t)nb * l2))];
    }
}

hipError_t launch_translate_rows(double *eri, int nb, int nf, int ncell, const int *relc, hipStream_t st) {
    if (nb <= nf) return hipSuccess;
    hipLaunchKernelGGL(k_translate_rows, dim3(8192), dim3(256), 0, st, eri, nb, nf, ncell, relc);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// GEMM.  MFMA operand layout (16x16x4, FP64): lane l supplies A[l%16][l/16]
// and B[l/16][l%16]; the result D[4v + l/16][l%16] sits in acc[v].
// ---------------------------------------------------------------------------
constexpr int kTM = 64, kTN = 64, kTK = 16, kPad = 2;

__global__ __launch_bounds__(256) void k_dgemm_nn(int M, int N, int K, const double *__restrict__ A, long lda,
                                                  long long sA, const double *__restrict__ B, long ldb, long long sB,
                                                  double *__restrict__ C, long ldc, long long sC, int tilesM,
                                                  int tilesN) {
    __shared__ double As[kTK][kTM + kPad];
    __shared__ double Bs[kTK][kTN + kPad];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    long long bid = blockIdx.x;
    const int tm = (int)(bid % tilesM); bid /= tilesM;
    const int tn = (int)(bid % tilesN);
    const long long b = bid / tilesN;
    A += b * sA;
    B += b * sB;
    C += b * sC;
    const int m0 = tm * kTM, n0 = tn * kTN;
    const int wm = (wv & 1) * 32, wn = (wv >> 1) * 32;
    dbl4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = dbl4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += kTK) {
        // A tile: 64 (m) x 16 (k); thread: m = t % 64, k = t / 64 + 4 i
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = t & 63, k = (t >> 6) + 4 * i;
            const int gm = m0 + m, gk = k0 + k;
            As[k][m] = (gm < M && gk < K) ? A[gm + lda * gk] : 0.0;
        }
        // B tile: 16 (k) x 64 (n); thread: k = t % 16, n = t / 16 + 16 i
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = t & 15, n = (t >> 4) + 16 * i;
            const int gk = k0 + k, gn = n0 + n;
            Bs[k][n] = (gk < K && gn < N) ? B[gk + ldb * gn] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTK; kk += 4) {
            double a[2], bb[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) a[x] = As[kk + (lane >> 4)][wm + 16 * x + (lane & 15)];
#pragma unroll
            for (int y = 0; y < 2; ++y) bb[y] = Bs[kk + (lane >> 4)][wn + 16 * y + (lane & 15)];
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], bb[y], acc[x][y], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gm = m0 + wm + 16 * x + 4 * v + (lane >> 4);
                const int gn = n0 + wn + 16 * y + (lane & 15);
                if (gm < M && gn < N) C[gm + ldc * gn] = acc[x][y][v];
            }
}

hipError_t launch_dgemm_nn(int M, int N, int K, const double *A, long lda, long long sA, const double *B, long ldb,
                           long long sB, double *C, long ldc, long long sC, long long batch, hipStream_t st) {
    if (M <= 0 || N <= 0 || batch <= 0) return hipSuccess;
    const int tilesM = (M + kTM - 1) / kTM, tilesN = (N + kTN - 1) / kTN;
    const long long blocks = (long long)tilesM * tilesN * batch;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_dgemm_nn, dim3((unsigned)blocks), dim3(256), 0, st, M, N, K, A, lda, sA, B, ldb, sB, C, ldc,
                       sC, tilesM, tilesN);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// MP2 sums over X = the transformed (ia|jb), stored with strides
// (si, sa, sj, sb) for (i, a, j, b); e = orbital energies, occupied first.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mp2_sums(const double *__restrict__ X, const double *__restrict__ e, int o,
                                                  int v, long long si, long long sa, long long sj, long long sb,
                                                  double *__restrict__ part) {
    __shared__ double sd[256], sx[256];
    const size_t n = (size_t)o * v * o * v;
    double d = 0.0, x = 0.0;
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (size_t)gridDim.x * 256) {
        size_t r = idx;
        const int i = (int)(r % o); r /= o;
        const int a = (int)(r % v); r /= v;
        const int j = (int)(r % o);
        const int b = (int)(r / o);
        const double g = X[i * si + a * sa + j * sj + b * sb];       // (ia|jb)
        const double gx = X[i * si + b * sa + j * sj + a * sb];      // (ib|ja)
        const double D = e[o + a] + e[o + b] - e[i] - e[j];
        d -= g * g / D;
        x -= g * gx / D;
    }
    sd[threadIdx.x] = d;
    sx[threadIdx.x] = x;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sd[threadIdx.x] += sd[threadIdx.x + s];
            sx[threadIdx.x] += sx[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sd[0];
        part[2 * blockIdx.x + 1] = sx[0];
    }
}

hipError_t launch_mp2_sums(const double *X, const double *e, int o, int v, long long si, long long sa, long long sj,
                           long long sb, double *part, int nblocks, hipStream_t st) {
    hipLaunchKernelGGL(k_mp2_sums, dim3(nblocks), dim3(256), 0, st, X, e, o, v, si, sa, sj, sb, part);
    return hipGetLastError();
}

}  // namespace tcg
