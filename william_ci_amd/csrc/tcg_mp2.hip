// Gamma-supercell MP2 on the GPU (gfx950): the full (non-SAO) ERI layout,
// the (ia|jb) transform as FP64 matrix-core GEMMs, and the MP2 sums.
//
//   k_translate_rows  rows i > nfuc of ERI(nBas,nBas,nBas,nBas) from the rows of
//                     the unit cell: (i j|k l) = (i0 j'|k' l'), all indices
//                     shifted by the lattice vector taking i to the unit cell
//                     (src/Tools/symmetric_matrix.f90:62-131).
//   k_dgemm_nn        strided-batched C_b = alpha A_b B_b + beta C_b, column-major, FP64 on
//                     v_mfma_f64_16x16x4f64: 64x64 tiles per workgroup of four
//                     wavefronts (32x32 each, 2x2 MFMA tiles), K in steps of 16
//                     staged through LDS.
//   k_mp2_sums        E2d = -sum (ia|jb)^2 / D, E2x = -sum (ia|jb)(ib|ja) / D,
//                     D = e_a + e_b - e_i - e_j (src/Energies/MP2/MP2.f90:39-68),
//                     one partial pair per workgroup (summed on the host in a
//                     fixed order).
#include <hip/hip_runtime.h>

#include <algorithm>

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

__global__ __launch_bounds__(256) void k_dgemm_nn(int M, int N, int K, double alpha, const double *__restrict__ A,
                                                  long lda, long long sA, const double *__restrict__ B, long ldb,
                                                  long long sB, double beta, double *__restrict__ C, long ldc,
                                                  long long sC, int tilesM, int tilesN) {
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
                if (gm < M && gn < N) {
                    double *cp = C + gm + ldc * gn;
                    *cp = beta == 0.0 ? alpha * acc[x][y][v] : alpha * acc[x][y][v] + beta * *cp;
                }
            }
}

hipError_t launch_dgemm_nn(int M, int N, int K, double alpha, const double *A, long lda, long long sA, const double *B,
                           long ldb, long long sB, double beta, double *C, long ldc, long long sC, long long batch,
                           hipStream_t st) {
    if (M <= 0 || N <= 0 || batch <= 0) return hipSuccess;
    const int tilesM = (M + kTM - 1) / kTM, tilesN = (N + kTN - 1) / kTN;
    const long long blocks = (long long)tilesM * tilesN * batch;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_dgemm_nn, dim3((unsigned)blocks), dim3(256), 0, st, M, N, K, alpha, A, lda, sA, B, ldb, sB,
                       beta, C, ldc, sC, tilesM, tilesN);
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

// ---------------------------------------------------------------------------
// k-point MP2 (src/Energies/SAO/MP2/AO_to_MO_SAO.f90:34-283, MP2_SAO.f90:6-151)
// after the lattice Fourier transforms (GEMMs on the host side):
//   S(a,b,k2,c,k3,d,k4) = 1/N sum ERI(a, b R2, c R3, d R4) e^{+i k2 R2 - i k3 R3 + i k4 R4}
// k_kmo: ERI_kMO(p,q,r,s,k1,k2,k3) = sum conj C_k1(a,p) C_k2(b,q) conj C_k3(c,r)
//        C_k4(d,s) S, k4 = k1 - k2 + k3; one thread per element of the (ov|ov)
//        block (p, r < o_hi; q, s >= v_lo).
// k_kmp2_sums: kb = ki - ka + kj, |D| < 1e-8 left out and counted.
// Cells and k points are numbered x fastest.
// ---------------------------------------------------------------------------
struct KGrid {
    int nx, ny, nz;
    __device__ int add(int k1, int k2, int k3) const {    // k1 - k2 + k3
        const int x = ((k1 % nx) - (k2 % nx) + (k3 % nx) + 2 * nx) % nx;
        const int y = (((k1 / nx) % ny) - ((k2 / nx) % ny) + ((k3 / nx) % ny) + 2 * ny) % ny;
        const int z = ((k1 / (nx * ny)) - (k2 / (nx * ny)) + (k3 / (nx * ny)) + 2 * nz) % nz;
        return x + nx * (y + ny * z);
    }
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {   // conj(a) b
    return double2{a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}

__global__ __launch_bounds__(256) void k_kmo(const double *__restrict__ Sre, const double *__restrict__ Sim,
                                             const double2 *__restrict__ Ck, int nf, int N, KGrid g, int o_hi,
                                             int v_lo, double2 *__restrict__ out) {
    const int nv = nf - v_lo;
    const long long per = (long long)o_hi * nv * o_hi * nv;
    const long long total = per * N * N * N;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        long long r_ = t;
        const int p = (int)(r_ % o_hi); r_ /= o_hi;
        const int q = v_lo + (int)(r_ % nv); r_ /= nv;
        const int r = (int)(r_ % o_hi); r_ /= o_hi;
        const int s = v_lo + (int)(r_ % nv); r_ /= nv;
        const int k1 = (int)(r_ % N); r_ /= N;
        const int k2 = (int)(r_ % N);
        const int k3 = (int)(r_ / N);
        const int k4 = g.add(k1, k2, k3);
        const double2 *C1 = Ck + (size_t)k1 * nf * nf, *C2 = Ck + (size_t)k2 * nf * nf;
        const double2 *C3 = Ck + (size_t)k3 * nf * nf, *C4 = Ck + (size_t)k4 * nf * nf;
        // S(a, b, k2, c, k3, d, k4): strides 1, nf, nf^2, nf^2 N, nf^3 N, nf^3 N^2, nf^4 N^2
        const size_t sb = nf, sc = (size_t)nf * nf * N, sd = (size_t)nf * nf * nf * N * N;
        const size_t base = (size_t)nf * nf * k2 + (size_t)nf * nf * nf * N * k3 + (size_t)nf * nf * nf * nf * N * N * k4;
        double2 acc = {0.0, 0.0};
        for (int d = 0; d < nf; ++d) {
            double2 accd = {0.0, 0.0};
            for (int c = 0; c < nf; ++c) {
                double2 accc = {0.0, 0.0};
                for (int b = 0; b < nf; ++b) {
                    double2 accb = {0.0, 0.0};
                    const size_t o0 = base + b * sb + c * sc + d * sd;
                    for (int a = 0; a < nf; ++a) {
                        const double2 sv = {Sre[o0 + a], Sim[o0 + a]};
                        const double2 m = cmulc(C1[a + (size_t)nf * p], sv);
                        accb.x += m.x;
                        accb.y += m.y;
                    }
                    const double2 m = cmul(C2[b + (size_t)nf * q], accb);
                    accc.x += m.x;
                    accc.y += m.y;
                }
                const double2 m = cmulc(C3[c + (size_t)nf * r], accc);
                accd.x += m.x;
                accd.y += m.y;
            }
            const double2 m = cmul(C4[d + (size_t)nf * s], accd);
            acc.x += m.x;
            acc.y += m.y;
        }
        out[t] = acc;
    }
}

// ---------------------------------------------------------------------------
// k_dft_lattice: the lattice Fourier transform along one cell index,
//   Y[o][k][i] = scale sum_R X[o][R][i] (cos + i sgn sin)(k R),   R, k < N <= 64,
// X, Y viewed as [outer][N][inner] (inner fastest); the rows (i, o) are
// flattened, 16 per wavefront.  On the matrix cores the transposed product
// Y^T = W^T X^T is formed (16 k x 4 R times 4 R x 16 rows), so that a lane
// loads and stores 16 consecutive rows; W (cos, sgn sin) sits in LDS for the
// workgroup's lifetime.  kCplx: X has an imaginary part.
// ---------------------------------------------------------------------------
template <bool kCplx>
__global__ __launch_bounds__(256) void k_dft_lattice(const double *__restrict__ Xr, const double *__restrict__ Xi,
                                                     long long I, long long outer, int N,
                                                     const double *__restrict__ cosm, const double *__restrict__ sinm,
                                                     double sgn, double scale, double *__restrict__ Yr,
                                                     double *__restrict__ Yi) {
    __shared__ double Wc[64][65];
    __shared__ double Ws[64][65];
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int R = e >> 6, k = e & 63;
        const bool in = R < N && k < N;
        Wc[R][k] = in ? cosm[R + (size_t)N * k] : 0.0;
        Ws[R][k] = in ? sgn * sinm[R + (size_t)N * k] : 0.0;
    }
    __syncthreads();
    const long long rows = I * outer, ntile = (rows + 15) / 16;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, lr = lane & 15, lq = lane >> 4;
    const int nk = (N + 15) >> 4;
    for (long long tile = (long long)blockIdx.x * 4 + wv; tile < ntile; tile += (long long)gridDim.x * 4) {
        const long long r = tile * 16 + lr;
        const bool rv = r < rows;
        const long long rr = rv ? r : rows - 1;
        const long long inner = rr % I, ob = rr / I;
        const size_t base = (size_t)inner + (size_t)I * N * ob;          // (row, R) at base + I R
        dbl4 re[4], im[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            re[t] = dbl4{0.0, 0.0, 0.0, 0.0};
            im[t] = dbl4{0.0, 0.0, 0.0, 0.0};
        }
        for (int R0 = 0; R0 < N; R0 += 4) {
            const int R = R0 + lq;
            const double ar = R < N ? Xr[base + (size_t)I * R] : 0.0;
            const double ai = (kCplx && R < N) ? Xi[base + (size_t)I * R] : 0.0;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                if (t < nk) {
                    const double wc = Wc[R][16 * t + lr], ws = Ws[R][16 * t + lr];
                    re[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(wc, ar, re[t], 0, 0, 0);
                    im[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ws, ar, im[t], 0, 0, 0);
                    if (kCplx) {
                        re[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ws, ai, re[t], 0, 0, 0);
                        im[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(wc, ai, im[t], 0, 0, 0);
                    }
                }
            }
        }
        if (rv) {
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int k = 16 * t + 4 * v + lq;
                    if (t < nk && k < N) {
                        Yr[base + (size_t)I * k] = scale * re[t][v];
                        Yi[base + (size_t)I * k] = scale * im[t][v];
                    }
                }
        }
    }
}

hipError_t launch_dft_lattice(const double *Xr, const double *Xi, long long I, long long outer, int N, const double *cosm,
                              const double *sinm, double sgn, double scale, double *Yr, double *Yi, hipStream_t st) {
    if (N > 64 || I <= 0 || outer <= 0) return hipErrorInvalidValue;
    const long long ntile = (I * outer + 15) / 16;
    const int grid = (int)std::min<long long>((ntile + 3) / 4, 2048);
    if (Xi)
        hipLaunchKernelGGL(k_dft_lattice<true>, dim3(grid), dim3(256), 0, st, Xr, Xi, I, outer, N, cosm, sinm, sgn,
                           scale, Yr, Yi);
    else
        hipLaunchKernelGGL(k_dft_lattice<false>, dim3(grid), dim3(256), 0, st, Xr, Xi, I, outer, N, cosm, sinm, sgn,
                           scale, Yr, Yi);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// k_kmo_block: the four quarter transforms of one S block (k2, k3, k4) in LDS
// (nf <= 8), one workgroup per block: S is read once, coalesced by 25-element
// runs; the (ov|ov) block of k1 = k4 + k2 - k3 is written out.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_kmo_block(const double *__restrict__ Sre, const double *__restrict__ Sim,
                                                   const double2 *__restrict__ Ck, int nf, int N, KGrid g, int o_hi,
                                                   int v_lo, double2 *__restrict__ out) {
    extern __shared__ double2 sm[];
    const int nf2 = nf * nf, nf4 = nf2 * nf2, nv = nf - v_lo, per = o_hi * nv;
    double2 *T0 = sm, *T1 = sm + nf4;
    const int kb = blockIdx.x;
    const int k2 = kb % N, k3 = (kb / N) % N, k4 = kb / (N * N);
    const int k1 = g.add(k4, k3, k2);                    // k4 - k3 + k2
    const size_t sc = (size_t)nf2 * N, sd = (size_t)nf2 * nf * N * N;
    const size_t base = (size_t)nf2 * k2 + (size_t)nf2 * nf * N * k3 + (size_t)nf4 * N * N * k4;
    for (int e = threadIdx.x; e < nf4; e += blockDim.x) {          // T0[a + nf (b + nf (c + nf d))]
        const int ab = e % nf2, c = (e / nf2) % nf, d = e / (nf2 * nf);
        const size_t o = base + ab + sc * c + sd * d;
        T0[e] = double2{Sre[o], Sim[o]};
    }
    __syncthreads();
    const double2 *C1 = Ck + (size_t)k1 * nf2, *C2 = Ck + (size_t)k2 * nf2;
    const double2 *C3 = Ck + (size_t)k3 * nf2, *C4 = Ck + (size_t)k4 * nf2;
    // T1(p, b, c, d) = sum_a conj C1(a, p) T0(a, b, c, d), p < o_hi
    for (int e = threadIdx.x; e < o_hi * nf * nf2; e += blockDim.x) {
        const int p = e % o_hi, bcd = e / o_hi;
        double2 acc = {0.0, 0.0};
        for (int a = 0; a < nf; ++a) {
            const double2 m = cmulc(C1[a + nf * p], T0[a + nf * bcd]);
            acc.x += m.x;
            acc.y += m.y;
        }
        T1[e] = acc;                                      // T1[p + o_hi (b + nf (c + nf d))]
    }
    __syncthreads();
    // T0(p, q, c, d) = sum_b C2(b, q) T1(p, b, c, d), q >= v_lo
    for (int e = threadIdx.x; e < per * nf2; e += blockDim.x) {
        const int p = e % o_hi, q = (e / o_hi) % nv, cd = e / per;
        double2 acc = {0.0, 0.0};
        for (int b = 0; b < nf; ++b) {
            const double2 m = cmul(C2[b + nf * (v_lo + q)], T1[p + o_hi * (b + nf * cd)]);
            acc.x += m.x;
            acc.y += m.y;
        }
        T0[e] = acc;                                      // T0[pq + per (c + nf d)]
    }
    __syncthreads();
    // T1(pq, r, d) = sum_c conj C3(c, r) T0(pq, c, d), r < o_hi
    for (int e = threadIdx.x; e < per * o_hi * nf; e += blockDim.x) {
        const int pq = e % per, r = (e / per) % o_hi, d = e / (per * o_hi);
        double2 acc = {0.0, 0.0};
        for (int c = 0; c < nf; ++c) {
            const double2 m = cmulc(C3[c + nf * r], T0[pq + per * (c + nf * d)]);
            acc.x += m.x;
            acc.y += m.y;
        }
        T1[e] = acc;                                      // T1[pq + per (r + o_hi d)]
    }
    __syncthreads();
    // out(pq, rs) = sum_d C4(d, s) T1(pq, r, d), s >= v_lo
    const size_t ob = (size_t)per * per * (k1 + (size_t)N * (k2 + (size_t)N * k3));
    for (int e = threadIdx.x; e < per * per; e += blockDim.x) {
        const int pq = e % per, rs = e / per, r = rs % o_hi, s = rs / o_hi;
        double2 acc = {0.0, 0.0};
        for (int d = 0; d < nf; ++d) {
            const double2 m = cmul(C4[d + nf * (v_lo + s)], T1[pq + per * (r + o_hi * d)]);
            acc.x += m.x;
            acc.y += m.y;
        }
        out[ob + e] = acc;
    }
}

hipError_t launch_kmo(const double *Sre, const double *Sim, const double2 *Ck, int nf, int N, int nx, int ny, int nz,
                      int o_hi, int v_lo, double2 *out, hipStream_t st) {
    KGrid g{nx, ny, nz};
    if (nf <= 8) {
        const size_t lds = 2 * sizeof(double2) * (size_t)nf * nf * nf * nf;
        hipLaunchKernelGGL(k_kmo_block, dim3((unsigned)((size_t)N * N * N)), dim3(256), lds, st, Sre, Sim, Ck, nf, N, g,
                           o_hi, v_lo, out);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_kmo, dim3(4096), dim3(256), 0, st, Sre, Sim, Ck, nf, N, g, o_hi, v_lo, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// k_dft_mo: the lattice Fourier transform along one (orbital, cell) index pair
// j = c + nf R of X fused with the orbital contraction of that pair:
//   Y[i + I (m + M k) + I M N o] =
//       scale sum_c T[(k nf + c) M + m] sum_R (cos + i sgn sin)(k.cell_R) X[x(i) + sJ (c + nf R) + sO o]
// (Y += with A.accum: the sum over R split into slabs of cells, one call each)
// R over the NR input cells (cos/sin tables [R + NR k]), k over N points, x(i) =
// imap[i] (a gather of the inner rows) or i, for the workgroup's chunk m in
// [MR blockIdx.y, MR (blockIdx.y + 1)) of the M outputs.  The orbital index is
// contracted as soon as its transform is in the accumulators, so the transformed
// tensor with all nf orbitals is never stored.  X real (kCplx false) or split
// complex; Y split complex.  The rows (i, o) are flattened, 16 per tile, and each
// DFT is a matrix-core product as in k_dft_lattice (W in LDS).  KT = 4: a wave
// does every k tile of its own 16 rows; KT = 1: the four waves share 16 rows,
// wave w doing k tile w (room for more accumulators per k).
// ---------------------------------------------------------------------------

template <bool kCplx, int KT, int MR>
__global__ __launch_bounds__(256) void k_dft_mo(DftMoArgs A) {
    __shared__ double Wc[64][65];
    __shared__ double Ws[64][65];
    const int N = A.N, NR = A.NR, nf = A.nf, M = A.M;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int R = e >> 6, k = e & 63;
        const bool in = R < NR && k < N;
        Wc[R][k] = in ? A.cosm[R + (size_t)NR * k] : 0.0;
        Ws[R][k] = in ? A.sgn * A.sinm[R + (size_t)NR * k] : 0.0;
    }
    __syncthreads();
    const long long I = A.I, rows = I * A.O, ntile = (rows + 15) / 16;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, lr = lane & 15, lq = lane >> 4;
    const int nk = (N + 15) >> 4;
    const int m0 = MR * blockIdx.y;
    const int t_lo = KT == 4 ? 0 : wv;
    const long long tile0 = KT == 4 ? (long long)blockIdx.x * 4 + wv : (long long)blockIdx.x;
    const long long tstep = KT == 4 ? (long long)gridDim.x * 4 : (long long)gridDim.x;
    const size_t sR = (size_t)A.sJ * nf;
    for (long long tile = tile0; tile < ntile; tile += tstep) {
        const long long r = tile * 16 + lr;
        const bool rv = r < rows;
        const long long rr = rv ? r : rows - 1;
        const long long inner = rr % I, ob = rr / I;
        const size_t base = (size_t)(A.imap ? A.imap[inner] : inner) + (size_t)A.sO * ob;
        dbl4 yr[MR][KT], yi[MR][KT];
#pragma unroll
        for (int mm = 0; mm < MR; ++mm)
#pragma unroll
            for (int tt = 0; tt < KT; ++tt) {
                yr[mm][tt] = dbl4{0.0, 0.0, 0.0, 0.0};
                yi[mm][tt] = dbl4{0.0, 0.0, 0.0, 0.0};
            }
        for (int c = 0; c < nf; ++c) {
            dbl4 re[KT], im[KT];
#pragma unroll
            for (int tt = 0; tt < KT; ++tt) {
                re[tt] = dbl4{0.0, 0.0, 0.0, 0.0};
                im[tt] = dbl4{0.0, 0.0, 0.0, 0.0};
            }
            const size_t bc = base + (size_t)A.sJ * c;
            for (int R0 = 0; R0 < NR; R0 += 4) {
                const int R = R0 + lq;
                const double ar = R < NR ? A.Xr[bc + sR * R] : 0.0;
                const double ai = (kCplx && R < NR) ? A.Xi[bc + sR * R] : 0.0;
#pragma unroll
                for (int tt = 0; tt < KT; ++tt) {
                    const int t = t_lo + tt;
                    if (t < nk) {
                        const double wc = Wc[R][16 * t + lr], ws = Ws[R][16 * t + lr];
                        re[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(wc, ar, re[tt], 0, 0, 0);
                        im[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ws, ar, im[tt], 0, 0, 0);
                        if (kCplx) {
                            re[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ws, ai, re[tt], 0, 0, 0);
                            im[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(wc, ai, im[tt], 0, 0, 0);
                        }
                    }
                }
            }
            // fold orbital c into the outputs: lane holds k = 16 t + 4 v + lq of row lr
#pragma unroll
            for (int tt = 0; tt < KT; ++tt) {
                const int t = t_lo + tt;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int k = 16 * t + 4 * v + lq;
                    if (t >= nk || k >= N) continue;
                    const double2 *Tk = A.T + ((size_t)k * nf + c) * M;
#pragma unroll
                    for (int mm = 0; mm < MR; ++mm) {
                        const int m = m0 + mm;
                        if (m >= M) continue;
                        const double2 w = Tk[m];
                        yr[mm][tt][v] += w.x * re[tt][v] - w.y * im[tt][v];
                        yi[mm][tt][v] += w.x * im[tt][v] + w.y * re[tt][v];
                    }
                }
            }
        }
        if (rv) {
            const size_t ob_off = (size_t)inner + (size_t)I * M * N * ob;
#pragma unroll
            for (int tt = 0; tt < KT; ++tt) {
                const int t = t_lo + tt;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int k = 16 * t + 4 * v + lq;
                    if (t >= nk || k >= N) continue;
#pragma unroll
                    for (int mm = 0; mm < MR; ++mm) {
                        const int m = m0 + mm;
                        if (m >= M) continue;
                        const size_t o = ob_off + (size_t)I * (m + (size_t)M * k);
                        if (A.accum) {
                            A.Yr[o] += A.scale * yr[mm][tt][v];
                            A.Yi[o] += A.scale * yi[mm][tt][v];
                        } else {
                            A.Yr[o] = A.scale * yr[mm][tt][v];
                            A.Yi[o] = A.scale * yi[mm][tt][v];
                        }
                    }
                }
            }
        }
    }
}

template <bool kCplx, int KT, int MR>
static hipError_t dft_mo_launch(const DftMoArgs &A, hipStream_t st) {
    const long long ntile = (A.I * A.O + 15) / 16;
    const long long g = KT == 4 ? (ntile + 3) / 4 : ntile;
    const dim3 grid((unsigned)std::min<long long>(g, 4096), (unsigned)((A.M + MR - 1) / MR));
    hipLaunchKernelGGL((k_dft_mo<kCplx, KT, MR>), grid, dim3(256), 0, st, A);
    return hipGetLastError();
}

hipError_t launch_dft_mo(const DftMoArgs &A, hipStream_t st) {
    if (A.N > 64 || A.N < 1 || A.NR > 64 || A.NR < 1 || A.I <= 0 || A.O <= 0 || A.M <= 0 || A.nf <= 0)
        return hipErrorInvalidValue;
    const int nk = (A.N + 15) / 16;
    const bool cplx = A.Xi != nullptr;
    if (nk < 3 || A.M <= 2) {                      // every k tile in one wave
        if (cplx) return A.M == 1 ? dft_mo_launch<true, 4, 1>(A, st) : dft_mo_launch<true, 4, 2>(A, st);
        return A.M == 1 ? dft_mo_launch<false, 4, 1>(A, st) : dft_mo_launch<false, 4, 2>(A, st);
    }
    if (cplx) return A.M <= 4 ? dft_mo_launch<true, 1, 4>(A, st) : dft_mo_launch<true, 1, 8>(A, st);
    return A.M <= 4 ? dft_mo_launch<false, 1, 4>(A, st) : dft_mo_launch<false, 1, 8>(A, st);
}

// flags[j2] = 1 for every bra function j2 = g mod nbas of a listed non-zero column g
__global__ void k_list_j2_flags(const long long *__restrict__ list, long long n, int nbas,
                                unsigned char *__restrict__ flags) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        flags[list[i] % nbas] = 1;
}

hipError_t launch_list_j2_flags(const long long *list, long long n, int nbas, unsigned char *flags, hipStream_t st) {
    const int grid = (int)std::min<long long>((n + 255) / 256, 4096);
    if (n > 0) hipLaunchKernelGGL(k_list_j2_flags, dim3(grid > 0 ? grid : 1), dim3(256), 0, st, list, n, nbas, flags);
    return hipGetLastError();
}

// Last quarter transform of the (ov|ov) blocks: the unit-cell index a with
// conj C_k1(a, p), k1 = k4 - k3 + k2, from
//   Y[a + nf (q + nv (k2 + N (r + o_hi (k3 + N (s + nv k4)))))]    (split complex)
// into out[pq + per (rs + per (k1 + N (k2 + N k3)))], pq = p + o_hi q, rs = r + o_hi s,
// per = o_hi nv (the layout k_kmp2_sums reads).
__global__ __launch_bounds__(256) void k_kmo_last(const double *__restrict__ Yr, const double *__restrict__ Yi,
                                                  const double2 *__restrict__ Ck, int nf, int N, KGrid g, int o_hi,
                                                  int v_lo, double2 *__restrict__ out) {
    const int nv = nf - v_lo, per = o_hi * nv;
    const long long total = (long long)per * per * N * N * N;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        long long r_ = t;
        const int p = (int)(r_ % o_hi); r_ /= o_hi;
        const int q = (int)(r_ % nv); r_ /= nv;
        const int r = (int)(r_ % o_hi); r_ /= o_hi;
        const int s = (int)(r_ % nv); r_ /= nv;
        const int k1 = (int)(r_ % N); r_ /= N;
        const int k2 = (int)(r_ % N);
        const int k3 = (int)(r_ / N);
        const int k4 = g.add(k1, k2, k3);
        const size_t yb = (size_t)nf * (q + (size_t)nv * (k2 + (size_t)N * (r + (size_t)o_hi * (k3 + (size_t)N * (s + (size_t)nv * k4)))));
        const double2 *C1 = Ck + (size_t)k1 * nf * nf + (size_t)nf * p;
        double2 acc{0.0, 0.0};
        for (int a = 0; a < nf; ++a) {
            const double2 y{Yr[yb + a], Yi[yb + a]};
            const double2 m = cmulc(C1[a], y);
            acc.x += m.x;
            acc.y += m.y;
        }
        out[t] = acc;
    }
}

hipError_t launch_kmo_last(const double *Yr, const double *Yi, const double2 *Ck, int nf, int N, int nx, int ny,
                           int nz, int o_hi, int v_lo, double2 *out, hipStream_t st) {
    KGrid g{nx, ny, nz};
    const long long per = (long long)o_hi * (nf - v_lo);
    const long long total = per * per * N * N * N;
    const int grid = (int)std::min<long long>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_kmo_last, dim3(grid), dim3(256), 0, st, Yr, Yi, Ck, nf, N, g, o_hi, v_lo, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_kmp2_sums(const double2 *__restrict__ ekmo, const double *__restrict__ ek,
                                                   const int *__restrict__ occ, int nf, int N, KGrid g, int o_hi,
                                                   int v_lo, double *__restrict__ part,
                                                   unsigned long long *__restrict__ nzero) {
    __shared__ double sd[256], sx[256];
    const int nv = nf - v_lo;
    const long long total = (long long)N * N * N * o_hi * o_hi * nv * nv;
    double d = 0.0, x = 0.0;
    unsigned long long nz = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        long long r_ = t;
        const int i = (int)(r_ % o_hi); r_ /= o_hi;
        const int a = v_lo + (int)(r_ % nv); r_ /= nv;
        const int j = (int)(r_ % o_hi); r_ /= o_hi;
        const int b = v_lo + (int)(r_ % nv); r_ /= nv;
        const int ki = (int)(r_ % N); r_ /= N;
        const int ka = (int)(r_ % N);
        const int kj = (int)(r_ / N);
        const int kb = g.add(ki, ka, kj);
        if (occ[ki * nf + i] != 2 || occ[kj * nf + j] != 2 || occ[ka * nf + a] != 0 || occ[kb * nf + b] != 0) continue;
        const double D = ek[ka * nf + a] + ek[kb * nf + b] - ek[ki * nf + i] - ek[kj * nf + j];
        // ERI_kMO(p,q,r,s,k1,k2,k3): dims (o_hi, nv, o_hi, nv, N, N, N)
        const size_t blk = (size_t)o_hi * nv * o_hi * nv;
        const double2 gd = ekmo[(size_t)i + o_hi * ((size_t)(a - v_lo) + nv * ((size_t)j + (size_t)o_hi * (b - v_lo))) +
                                blk * ((size_t)ki + (size_t)N * ((size_t)ka + (size_t)N * kj))];
        if (fabs(D) < 1e-8) {
            ++nz;
            continue;
        }
        const double2 gx = ekmo[(size_t)i + o_hi * ((size_t)(b - v_lo) + nv * ((size_t)j + (size_t)o_hi * (a - v_lo))) +
                                blk * ((size_t)ki + (size_t)N * ((size_t)kb + (size_t)N * kj))];
        d -= (gd.x * gd.x + gd.y * gd.y) / D;
        x -= (gd.x * gx.x + gd.y * gx.y) / D;
    }
    if (nz) atomicAdd(nzero, nz);
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

hipError_t launch_kmp2_sums(const double2 *ekmo, const double *ek, const int *occ, int nf, int N, int nx, int ny,
                            int nz, int o_hi, int v_lo, double *part, int nblocks, unsigned long long *nzero,
                            hipStream_t st) {
    KGrid g{nx, ny, nz};
    hipLaunchKernelGGL(k_kmp2_sums, dim3(nblocks), dim3(256), 0, st, ekmo, ek, occ, nf, N, g, o_hi, v_lo, part, nzero);
    return hipGetLastError();
}

}  // namespace tcg
