// Density fitting on the GPU (gfx950): the kernels of tcg_df.cpp.
//
//   k_metric_factor  per k point: pivoted Cholesky of the Hermitian metric
//                    V(k) (nau <= 64 in LDS; k_metric_factor_g up to 256) with the stopping rule of LAPACK's
//                    ?pstrf (stop when the largest remaining diagonal <= tol),
//                    then f = P (L^-1)^H on the kept columns
//                    (src/module/df_toroidal.f90:1174-1205); only for the
//                    first k of each pair (k, -k).
//   k_metric_mirror  f(-k) = conj f(k) for the second k of each pair
//                    (:589-592), so that B = 1/N Re sum_k W(k) e^{-ikR} is a
//                    factor of the fitted integrals (W(-k) = conj W(k)).
//   k_df_gamma       G(Q, R) = sum_x B(x, Q) Ps(x, R)  (x = (mu0, nu) pairs)
//   k_df_jvec        J(x) = sum_Q B(x, Q) gamma(Q) (partial sums over chunks of Q, then their sum)
//   k_df_kt_y/_k     the exchange in k space (df_toroidal.f90:893-954):
//                    Kt(kk) = -1/(2N) sum_{k1,Q} M P(k1) M^H,  M = Bh_Q(k1, kk - k1)
//   k_df_l           L(p, q, Q, k1, k2) = sum_ab conj C_k1(a, p) C_k2(b, q) F(a, b, k2, Q, k2 - k1)
//   k_df_eri         ERI(p,q,r,s,k1,k2,k3) = 1/N sum_Q L(p,q,Q,k1,k2) L(r,s,Q,k3,k4)
//   k_df_l_f32, k_df_eri_mfma32, k_kmp2_sums_f32: the same in FP32 -- L with FP32
//                    arithmetic, the (ov|ov) assembly as FP32 matrix-core GEMMs with
//                    FP32 accumulation, the energy sums in FP32 with a compensated
//                    (hi, lo) FP32 pair per thread and per block
// Cells and k points are numbered x fastest; cdiff[c1 * N + c2] = c1 - c2.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "tcg_device.hpp"

namespace tcg {

__device__ __forceinline__ double2 zmul(double2 a, double2 b) { return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ double2 zmulc(double2 a, double2 b) {   // a conj(b)
    return double2{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
__device__ __forceinline__ double2 zconj(double2 a) { return double2{a.x, -a.y}; }

// V: [k][nau*nau] (re, im separate, column-major), f out likewise.
__global__ __launch_bounds__(64) void k_metric_factor(const double *__restrict__ Vre, const double *__restrict__ Vim,
                                                      int n, const int *__restrict__ mate, double tol,
                                                      double *__restrict__ fre, double *__restrict__ fim,
                                                      int *__restrict__ rank, double *__restrict__ dmin) {
    __shared__ double2 A[64][65];
    __shared__ double2 Li[64][65];
    __shared__ int perm[64];
    __shared__ double red_v[64];
    __shared__ int red_i[64];
    const int k = blockIdx.x, t = threadIdx.x;
    if (mate[k] < k) return;                     // the conjugate of an earlier k: k_metric_mirror
    const size_t off = (size_t)k * n * n;
    // Hermitian part of V(k)
    for (int j = 0; j < n; ++j)
        if (t < n) {
            const double2 a = {Vre[off + t + (size_t)n * j], Vim[off + t + (size_t)n * j]};
            const double2 b = {Vre[off + j + (size_t)n * t], Vim[off + j + (size_t)n * t]};
            A[t][j] = double2{0.5 * (a.x + b.x), 0.5 * (a.y - b.y)};
        }
    if (t < n) perm[t] = t;
    __syncthreads();
    int m = n;
    for (int j = 0; j < n; ++j) {
        // pivot: largest remaining diagonal (first one on ties)
        red_v[t] = (t >= j && t < n) ? A[t][t].x : -1e300;
        red_i[t] = t;
        __syncthreads();
        for (int s = 32; s > 0; s >>= 1) {
            if (t < s) {
                const double v2 = red_v[t + s];
                const int i2 = red_i[t + s];
                if (v2 > red_v[t] || (v2 == red_v[t] && i2 < red_i[t])) {
                    red_v[t] = v2;
                    red_i[t] = i2;
                }
            }
            __syncthreads();
        }
        const double dmax = red_v[0];
        const int p = red_i[0];
        __syncthreads();
        if (!(dmax > tol)) {
            m = j;
            break;
        }
        // symmetric swap of rows/columns j and p
        if (p != j) {
            if (t < n) {
                const double2 x = A[t][j];
                A[t][j] = A[t][p];
                A[t][p] = x;
            }
            __syncthreads();
            if (t < n) {
                const double2 x = A[j][t];
                A[j][t] = A[p][t];
                A[p][t] = x;
            }
            if (t == 0) {
                const int x = perm[j];
                perm[j] = perm[p];
                perm[p] = x;
            }
            __syncthreads();
        }
        const double ljj = sqrt(A[j][j].x);
        __syncthreads();
        if (t == j) A[j][j] = double2{ljj, 0.0};
        if (t > j && t < n) A[t][j] = double2{A[t][j].x / ljj, A[t][j].y / ljj};
        __syncthreads();
        // trailing update A(i, c) -= L(i, j) conj(L(c, j)), i, c > j
        if (t > j && t < n) {
            const double2 lij = A[t][j];
            for (int c = j + 1; c < n; ++c) {
                const double2 u = zmulc(lij, A[c][j]);
                A[t][c] = double2{A[t][c].x - u.x, A[t][c].y - u.y};
            }
        }
        __syncthreads();
    }
    // L^-1 of the leading m x m block, one column per lane (forward substitution)
    if (t < m) {
        for (int i = 0; i < m; ++i) {
            double2 s = {i == t ? 1.0 : 0.0, 0.0};
            for (int c = t; c < i; ++c) {
                const double2 u = zmul(A[i][c], Li[c][t]);
                s.x -= u.x;
                s.y -= u.y;
            }
            const double d = A[i][i].x;
            Li[i][t] = i < t ? double2{0.0, 0.0} : double2{s.x / d, s.y / d};
        }
    }
    __syncthreads();
    // f(perm(i), j) = conj(Linv(j, i)), i, j < m; zero elsewhere
    for (int j = 0; j < n; ++j)
        if (t < n) {
            fre[off + t + (size_t)n * j] = 0.0;
            fim[off + t + (size_t)n * j] = 0.0;
        }
    __syncthreads();
    for (int j = 0; j < m; ++j)
        if (t < m) {
            const double2 v = zconj(Li[j][t]);
            fre[off + perm[t] + (size_t)n * j] = v.x;
            fim[off + perm[t] + (size_t)n * j] = v.y;
        }
    if (t == 0) {
        rank[k] = m;
        dmin[k] = m > 0 ? A[m - 1][m - 1].x * A[m - 1][m - 1].x : 0.0;
    }
}

// The same factorisation for 64 < n <= 256: one workgroup of 256 threads per k, the
// working matrix and L^-1 in global memory (Aw, Lw: [k][n][n], row-major per k).
__global__ __launch_bounds__(256) void k_metric_factor_g(const double *__restrict__ Vre, const double *__restrict__ Vim,
                                                         int n, const int *__restrict__ mate, double tol,
                                                         double2 *__restrict__ Aw, double2 *__restrict__ Lw,
                                                         double *__restrict__ fre, double *__restrict__ fim,
                                                         int *__restrict__ rank, double *__restrict__ dmin) {
    __shared__ int perm[256];
    __shared__ double red_v[256];
    __shared__ int red_i[256];
    const int k = blockIdx.x, t = threadIdx.x;
    if (mate[k] < k) return;
    const size_t off = (size_t)k * n * n;
    double2 *A = Aw + off, *Li = Lw + off;              // A[r * n + c]
    for (int j = 0; j < n; ++j)
        if (t < n) {
            const double2 a = {Vre[off + t + (size_t)n * j], Vim[off + t + (size_t)n * j]};
            const double2 b = {Vre[off + j + (size_t)n * t], Vim[off + j + (size_t)n * t]};
            A[(size_t)t * n + j] = double2{0.5 * (a.x + b.x), 0.5 * (a.y - b.y)};
        }
    if (t < n) perm[t] = t;
    __syncthreads();
    int m = n;
    for (int j = 0; j < n; ++j) {
        red_v[t] = (t >= j && t < n) ? A[(size_t)t * n + t].x : -1e300;
        red_i[t] = t;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (t < s) {
                const double v2 = red_v[t + s];
                const int i2 = red_i[t + s];
                if (v2 > red_v[t] || (v2 == red_v[t] && i2 < red_i[t])) {
                    red_v[t] = v2;
                    red_i[t] = i2;
                }
            }
            __syncthreads();
        }
        const double dmax = red_v[0];
        const int p = red_i[0];
        __syncthreads();
        if (!(dmax > tol)) {
            m = j;
            break;
        }
        if (p != j) {
            if (t < n) {
                const double2 x = A[(size_t)t * n + j];
                A[(size_t)t * n + j] = A[(size_t)t * n + p];
                A[(size_t)t * n + p] = x;
            }
            __syncthreads();
            if (t < n) {
                const double2 x = A[(size_t)j * n + t];
                A[(size_t)j * n + t] = A[(size_t)p * n + t];
                A[(size_t)p * n + t] = x;
            }
            if (t == 0) {
                const int x = perm[j];
                perm[j] = perm[p];
                perm[p] = x;
            }
            __syncthreads();
        }
        const double ljj = sqrt(A[(size_t)j * n + j].x);
        __syncthreads();
        if (t == j) A[(size_t)j * n + j] = double2{ljj, 0.0};
        if (t > j && t < n) {
            const double2 v = A[(size_t)t * n + j];
            A[(size_t)t * n + j] = double2{v.x / ljj, v.y / ljj};
        }
        __syncthreads();
        if (t > j && t < n) {
            const double2 lij = A[(size_t)t * n + j];
            for (int c = j + 1; c < n; ++c) {
                const double2 u = zmulc(lij, A[(size_t)c * n + j]);
                const double2 x = A[(size_t)t * n + c];
                A[(size_t)t * n + c] = double2{x.x - u.x, x.y - u.y};
            }
        }
        __syncthreads();
    }
    if (t < m) {
        for (int i = 0; i < m; ++i) {
            double2 sacc = {i == t ? 1.0 : 0.0, 0.0};
            for (int c = t; c < i; ++c) {
                const double2 u = zmul(A[(size_t)i * n + c], Li[(size_t)c * n + t]);
                sacc.x -= u.x;
                sacc.y -= u.y;
            }
            const double d = A[(size_t)i * n + i].x;
            Li[(size_t)i * n + t] = i < t ? double2{0.0, 0.0} : double2{sacc.x / d, sacc.y / d};
        }
    }
    __syncthreads();
    for (int j = 0; j < n; ++j)
        if (t < n) {
            fre[off + t + (size_t)n * j] = 0.0;
            fim[off + t + (size_t)n * j] = 0.0;
        }
    __syncthreads();
    for (int j = 0; j < m; ++j)
        if (t < m) {
            const double2 v = zconj(Li[(size_t)j * n + t]);
            fre[off + perm[t] + (size_t)n * j] = v.x;
            fim[off + perm[t] + (size_t)n * j] = v.y;
        }
    if (t == 0) {
        rank[k] = m;
        const double lm = m > 0 ? A[(size_t)(m - 1) * n + (m - 1)].x : 0.0;
        dmin[k] = lm * lm;
    }
}

__global__ __launch_bounds__(256) void k_metric_mirror(int n, int nk, const int *__restrict__ mate,
                                                       double *__restrict__ fre, double *__restrict__ fim,
                                                       int *__restrict__ rank, double *__restrict__ dmin) {
    const size_t nn = (size_t)n * n;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < nn * nk; g += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(g / nn);
        const int mk = mate[k];
        if (mk >= k) continue;
        const size_t e = g % nn;
        fre[g] = fre[(size_t)mk * nn + e];
        fim[g] = -fim[(size_t)mk * nn + e];
        if (e == 0) {
            rank[k] = rank[mk];
            dmin[k] = dmin[mk];
        }
    }
}

hipError_t launch_metric_factor(const double *Vre, const double *Vim, int n, int nk, const int *mate, double tol,
                                double2 *work, double *fre, double *fim, int *rank, double *dmin, hipStream_t st) {
    if (n <= 64) {
        hipLaunchKernelGGL(k_metric_factor, dim3(nk), dim3(64), 0, st, Vre, Vim, n, mate, tol, fre, fim, rank, dmin);
    } else if (n <= 256 && work) {
        double2 *Aw = work, *Lw = work + (size_t)nk * n * n;
        hipLaunchKernelGGL(k_metric_factor_g, dim3(nk), dim3(256), 0, st, Vre, Vim, n, mate, tol, Aw, Lw, fre, fim,
                           rank, dmin);
    } else {
        return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t tot = (size_t)n * n * nk;
    const int grid = (int)std::min<size_t>((tot + 255) / 256, 4096);
    hipLaunchKernelGGL(k_metric_mirror, dim3(grid), dim3(256), 0, st, n, nk, mate, fre, fim, rank, dmin);
    return hipGetLastError();
}

// G(Q, R) = sum_x B(x, Q) Ps(x, R); one block per Q, threads over x
__global__ __launch_bounds__(256) void k_df_gamma(const double *__restrict__ B, const double *__restrict__ Ps, long nx,
                                                  int naux, int N, double *__restrict__ G) {
    __shared__ double red[256];
    const int Q = blockIdx.x;
    for (int R = 0; R < N; ++R) {
        double s = 0.0;
        for (long x = threadIdx.x; x < nx; x += 256) s += B[x + nx * (size_t)Q] * Ps[x + nx * (size_t)R];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) G[Q + (size_t)naux * R] = red[0];
        __syncthreads();
    }
}

// J(x) = sum_Q B(x, Q) gamma(Q): nx rows are too few threads to stream B (nx naux
// doubles) at HBM speed, so the Q range is split over kJChunks workgroup rows, each
// writing its partial sums; a second pass adds the chunks in order (the same J on
// every run)
constexpr int kJChunks = kDfJChunks;
__global__ __launch_bounds__(256) void k_df_jvec_part(const double *__restrict__ B, const double *__restrict__ gam,
                                                      long nx, int naux, double *__restrict__ part) {
    const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nx) return;
    const int per = (naux + kJChunks - 1) / kJChunks, q0 = blockIdx.y * per, q1 = min(naux, q0 + per);
    double s = 0.0;
    for (int Q = q0; Q < q1; ++Q) s += B[x + nx * (size_t)Q] * gam[Q];
    part[x + nx * (size_t)blockIdx.y] = s;
}

__global__ __launch_bounds__(256) void k_df_jvec(const double *__restrict__ part, long nx, double *__restrict__ J) {
    const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nx) return;
    double s = 0.0;
    for (int c = 0; c < kJChunks; ++c) s += part[x + nx * (size_t)c];
    J[x] = s;
}

hipError_t launch_df_gamma(const double *B, const double *Ps, int nf, int nb, int nau, int N, double *G,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_df_gamma, dim3(nau * N), dim3(256), 0, st, B, Ps, (long)nf * nb, nau * N, N, G);
    return hipGetLastError();
}

// part: kJChunks nf nb doubles
hipError_t launch_df_jvec(const double *B, const double *gam, int nf, int nb, int naux, double *part, double *J,
                          hipStream_t st) {
    const long nx = (long)nf * nb;
    const unsigned gx = (unsigned)((nx + 255) / 256);
    hipLaunchKernelGGL(k_df_jvec_part, dim3(gx, kJChunks), dim3(256), 0, st, B, gam, nx, naux, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_df_jvec, dim3(gx), dim3(256), 0, st, part, nx, J);
    return hipGetLastError();
}

// Y(mu, nu', Q, k1, kk) = sum_nu M(mu, nu) P(k1)(nu, nu'), M(mu, nu) = F(mu, nu, k1, Q, kap = k1 - kk)
__global__ __launch_bounds__(256) void k_df_kt_y(const double *__restrict__ Fr, const double *__restrict__ Fi,
                                                 const double2 *__restrict__ Pk, int nf, int nau, int N,
                                                 const int *__restrict__ cdiff, double2 *__restrict__ Y) {
    const size_t nff = (size_t)nf * nf;
    const size_t total = nff * nau * N * N;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        size_t r = t;
        const int mu = (int)(r % nf); r /= nf;
        const int nu2 = (int)(r % nf); r /= nf;
        const int Q = (int)(r % nau); r /= nau;
        const int k1 = (int)(r % N);
        const int kk = (int)(r / N);
        const int kap = cdiff[k1 * N + kk];
        const size_t fb = (size_t)mu + nff * (k1 + (size_t)N * (Q + (size_t)nau * kap));
        const double2 *P1 = Pk + nff * k1 + (size_t)nf * nu2;
        double2 acc = {0.0, 0.0};
        for (int nu = 0; nu < nf; ++nu) {
            const double2 u = zmul(double2{Fr[fb + (size_t)nf * nu], Fi[fb + (size_t)nf * nu]}, P1[nu]);
            acc.x += u.x;
            acc.y += u.y;
        }
        Y[t] = acc;
    }
}

// Kt(kk)(mu, mu') = scale sum_{k1, Q, nu'} Y(mu, nu', Q, k1, kk) conj M(mu', nu'); one block per (mu, mu', kk)
__global__ __launch_bounds__(256) void k_df_kt_k(const double *__restrict__ Fr, const double *__restrict__ Fi,
                                                 const double2 *__restrict__ Y, int nf, int nau, int N,
                                                 const int *__restrict__ cdiff, double scale,
                                                 double2 *__restrict__ Kt) {
    __shared__ double2 red[256];
    const size_t nff = (size_t)nf * nf;
    const int mu = blockIdx.x % nf, mu2 = (blockIdx.x / nf) % nf, kk = blockIdx.x / (nf * nf);
    const int terms = N * nau * nf;
    double2 acc = {0.0, 0.0};
    for (int w = threadIdx.x; w < terms; w += blockDim.x) {
        const int nu2 = w % nf, Q = (w / nf) % nau, k1 = w / (nf * nau);
        const int kap = cdiff[k1 * N + kk];
        const size_t fi = (size_t)mu2 + nf * (nu2 + nf * (k1 + (size_t)N * (Q + (size_t)nau * kap)));
        const double2 y = Y[(size_t)mu + nf * (nu2 + nf * (Q + (size_t)nau * (k1 + (size_t)N * kk)))];
        const double2 u = zmulc(y, double2{Fr[fi], Fi[fi]});
        acc.x += u.x;
        acc.y += u.y;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[threadIdx.x].x += red[threadIdx.x + w].x;
            red[threadIdx.x].y += red[threadIdx.x + w].y;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) Kt[nff * kk + mu + (size_t)nf * mu2] = double2{scale * red[0].x, scale * red[0].y};
}

hipError_t launch_df_kt(const double *Fr, const double *Fi, const double2 *Pk, int nf, int nau, int N, const int *cdiff,
                        double2 *Y, double2 *Kt, hipStream_t st) {
    const size_t total = (size_t)nf * nf * nau * N * N;
    const int grid = (int)std::min<size_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_df_kt_y, dim3(grid), dim3(256), 0, st, Fr, Fi, Pk, nf, nau, N, cdiff, Y);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_df_kt_k, dim3(nf * nf * N), dim3(256), 0, st, Fr, Fi, Y, nf, nau, N, cdiff, -0.5 / N, Kt);
    return hipGetLastError();
}

// L(p, q, Q, k1, k2), p < o_hi, v_lo <= q < nf; F(a, b, k2, Q, kap) complex (Fr, Fi)
__global__ __launch_bounds__(256) void k_df_l(const double *__restrict__ Fr, const double *__restrict__ Fi,
                                              const double2 *__restrict__ Ck, int nf,
                                              int nau, int N, int o_hi, int v_lo, const int *__restrict__ cdiff,
                                              double2 *__restrict__ L) {
    const int nv = nf - v_lo;
    const size_t total = (size_t)o_hi * nv * nau * N * N;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        size_t r = t;
        const int p = (int)(r % o_hi); r /= o_hi;
        const int q = v_lo + (int)(r % nv); r /= nv;
        const int Q = (int)(r % nau); r /= nau;
        const int k1 = (int)(r % N);
        const int k2 = (int)(r / N);
        const int kap = cdiff[k2 * N + k1];
        const double2 *C1 = Ck + (size_t)k1 * nf * nf, *C2 = Ck + (size_t)k2 * nf * nf;
        const size_t fb = (size_t)nf * nf * (k2 + (size_t)N * (Q + (size_t)nau * kap));
        double2 acc = {0.0, 0.0};
        for (int b = 0; b < nf; ++b) {
            double2 s = {0.0, 0.0};
            for (int a = 0; a < nf; ++a) {
                const size_t fi = fb + a + (size_t)nf * b;
                const double2 u = zmul(zconj(C1[a + (size_t)nf * p]), double2{Fr[fi], Fi[fi]});
                s.x += u.x;
                s.y += u.y;
            }
            const double2 u = zmul(C2[b + (size_t)nf * q], s);
            acc.x += u.x;
            acc.y += u.y;
        }
        L[t] = acc;
    }
}

// ERI(p,q,r,s,k1,k2,k3) = 1/N sum_Q L(p,q,Q,k1,k2) L(r,s,Q,k3,k4), k4 = k1 - k2 + k3.
// One block per (k1, k2): L(:,:,:,k1,k2) staged in LDS, every k3 from it.
__global__ __launch_bounds__(256) void k_df_eri_lds(const double2 *__restrict__ L, int nau, int N, int per,
                                                    const int *__restrict__ cdiff, const int *__restrict__ cadd,
                                                    double2 *__restrict__ out) {
    extern __shared__ double2 L1s[];             // per * nau
    const int k1 = blockIdx.x % N, k2 = blockIdx.x / N;
    const double2 *L1 = L + (size_t)per * nau * (k1 + (size_t)N * k2);
    for (int i = threadIdx.x; i < per * nau; i += blockDim.x) L1s[i] = L1[i];
    __syncthreads();
    const int dk = cdiff[k1 * N + k2];
    const double invN = 1.0 / N;
    const int pp = per * per;
    for (int w = threadIdx.x; w < pp * N; w += blockDim.x) {
        const int k3 = w / pp, e = w - k3 * pp;
        const int pq = e % per, rs = e / per;
        const int k4 = cadd[dk * N + k3];
        const double2 *L2 = L + (size_t)per * nau * (k3 + (size_t)N * k4);
        double2 acc = {0.0, 0.0};
        for (int Q = 0; Q < nau; ++Q) {
            const double2 u = zmul(L1s[pq + per * Q], L2[rs + per * Q]);
            acc.x += u.x;
            acc.y += u.y;
        }
        out[pq + (size_t)per * (rs + (size_t)per * (k1 + (size_t)N * (k2 + (size_t)N * k3)))] =
            double2{acc.x * invN, acc.y * invN};
    }
}

__global__ __launch_bounds__(256) void k_df_eri(const double2 *__restrict__ L, int nau, int N, int o_hi, int nv,
                                                const int *__restrict__ cdiff, const int *__restrict__ cadd,
                                                double2 *__restrict__ out) {
    const size_t per = (size_t)o_hi * nv;
    const size_t total = per * per * N * N * N;
    const double invN = 1.0 / N;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        size_t r_ = t;
        const size_t pq = r_ % per; r_ /= per;
        const size_t rs = r_ % per; r_ /= per;
        const int k1 = (int)(r_ % N); r_ /= N;
        const int k2 = (int)(r_ % N);
        const int k3 = (int)(r_ / N);
        const int k4 = cadd[cdiff[k1 * N + k2] * N + k3];
        const double2 *L1 = L + pq + per * (size_t)nau * (k1 + (size_t)N * k2);
        const double2 *L2 = L + rs + per * (size_t)nau * (k3 + (size_t)N * k4);
        double2 acc = {0.0, 0.0};
        for (int Q = 0; Q < nau; ++Q) {
            const double2 u = zmul(L1[per * Q], L2[per * Q]);
            acc.x += u.x;
            acc.y += u.y;
        }
        out[t] = double2{acc.x * invN, acc.y * invN};
    }
}

// ---------------------------------------------------------------------------
// FP32 path of the DF k-MP2.
// ---------------------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));

// L(p, q, Q, k1, k2) as k_df_l, in FP32 arithmetic (FP32 output)
__global__ __launch_bounds__(256) void k_df_l_f32(const double *__restrict__ Fr, const double *__restrict__ Fi,
                                                  const double2 *__restrict__ Ck, int nf, int nau, int N, int o_hi,
                                                  int v_lo, const int *__restrict__ cdiff, float2 *__restrict__ L) {
    const int nv = nf - v_lo;
    const size_t total = (size_t)o_hi * nv * nau * N * N;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        size_t r = t;
        const int p = (int)(r % o_hi); r /= o_hi;
        const int q = v_lo + (int)(r % nv); r /= nv;
        const int Q = (int)(r % nau); r /= nau;
        const int k1 = (int)(r % N);
        const int k2 = (int)(r / N);
        const int kap = cdiff[k2 * N + k1];
        const double2 *C1 = Ck + (size_t)k1 * nf * nf, *C2 = Ck + (size_t)k2 * nf * nf;
        const size_t fb = (size_t)nf * nf * (k2 + (size_t)N * (Q + (size_t)nau * kap));
        float ax = 0.f, ay = 0.f;
        for (int b = 0; b < nf; ++b) {
            float sx = 0.f, sy = 0.f;
            for (int a = 0; a < nf; ++a) {
                const size_t fi = fb + a + (size_t)nf * b;
                const double2 c = C1[a + (size_t)nf * p];
                const float cx = (float)c.x, cy = -(float)c.y, fx = (float)Fr[fi], fy = (float)Fi[fi];
                sx += cx * fx - cy * fy;
                sy += cx * fy + cy * fx;
            }
            const double2 c2 = C2[b + (size_t)nf * q];
            const float cx = (float)c2.x, cy = (float)c2.y;
            ax += cx * sx - cy * sy;
            ay += cx * sy + cy * sx;
        }
        L[t] = float2{ax, ay};
    }
}

// (ov|ov) on the FP32 matrix cores.  For one momentum transfer dk = k1 - k2 =
// k4 - k3 (blockIdx.y) the blocks form one complex GEMM over Q:
//   rows i = (pq, k1), k2 = k1 - dk;  columns j = (rs, k3), k4 = k3 + dk
//   ERI(i, j) = 1/N sum_Q L(pq, Q, k1, k2) L(rs, Q, k3, k4)
// in 16 x 16 tiles (one per wavefront), v_mfma_f32_16x16x4f32 with FP32
// accumulators: Re = sum ArBr - AiBi, Im = sum ArBi + AiBr (four real
// products per step of 4 Q).  A operand: lane holds row lr of the tile at
// Q0 + lq; B operand: column lr at Q0 + lq; D: rows 4 lq + v, column lr.
__global__ __launch_bounds__(256) void k_df_eri_mfma32(const float2 *__restrict__ L, int nau, int N, int per,
                                                       const int *__restrict__ cdiff, const int *__restrict__ cadd,
                                                       float2 *__restrict__ out) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, lr = lane & 15, lq = lane >> 4;
    const int M = per * N, nt = (M + 15) / 16, dk = blockIdx.y;
    const float invN = 1.0f / (float)N;
    for (int tile = blockIdx.x * 4 + wv; tile < nt * nt; tile += gridDim.x * 4) {
        const int ti = tile % nt, tj = tile / nt;
        const int i = 16 * ti + lr, j = 16 * tj + lr;
        const bool iv = i < M, jv = j < M;
        const int pq = iv ? i % per : 0, k1 = iv ? i / per : 0;
        const int k2 = cdiff[k1 * N + dk];
        const int rs = jv ? j % per : 0, k3 = jv ? j / per : 0;
        const int k4 = cadd[dk * N + k3];
        const float2 *Ar = L + pq + (size_t)per * nau * (k1 + (size_t)N * k2);
        const float2 *Bc = L + rs + (size_t)per * nau * (k3 + (size_t)N * k4);
        floatx4 rr = {0.f, 0.f, 0.f, 0.f}, ii = rr, ri = rr, ir = rr;
        for (int Q0 = 0; Q0 < nau; Q0 += 4) {
            const int Q = Q0 + lq;
            const float2 a = (iv && Q < nau) ? Ar[(size_t)per * Q] : float2{0.f, 0.f};
            const float2 b = (jv && Q < nau) ? Bc[(size_t)per * Q] : float2{0.f, 0.f};
            rr = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, rr, 0, 0, 0);
            ii = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.y, ii, 0, 0, 0);
            ri = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.y, ri, 0, 0, 0);
            ir = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.x, ir, 0, 0, 0);
        }
        if (!jv) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i2 = 16 * ti + 4 * lq + v;
            if (i2 >= M) continue;
            const int pq2 = i2 % per, kk1 = i2 / per, kk2 = cdiff[kk1 * N + dk];
            out[pq2 + (size_t)per * (rs + (size_t)per * (kk1 + (size_t)N * (kk2 + (size_t)N * k3)))] =
                float2{(rr[v] - ii[v]) * invN, (ri[v] + ir[v]) * invN};
        }
    }
}

// s + t as an unevaluated FP32 pair (hi, lo): Knuth's two-sum of hi and t, the
// rounding error added to lo
__device__ __forceinline__ void acc2(float &hi, float &lo, float t) {
    const float s = hi + t;
    const float bb = s - hi;
    const float err = (hi - (s - bb)) + (t - bb);
    hi = s;
    lo += err;
}

// kMP2_calc in FP32 (the k_kmp2_sums loop): terms from the FP32 (ov|ov), each
// sum kept as a compensated (hi, lo) FP32 pair, per thread and over the block;
// part[4 b ..] = d_hi, d_lo, x_hi, x_lo of block b
__global__ __launch_bounds__(256) void k_kmp2_sums_f32(const float2 *__restrict__ ekmo, const double *__restrict__ ek,
                                                       const int *__restrict__ occ, int nf, int N,
                                                       const int *__restrict__ cadd, const int *__restrict__ cdiff,
                                                       int o_hi, int v_lo, float *__restrict__ part,
                                                       unsigned long long *__restrict__ nzero) {
    __shared__ float sdh[256], sdl[256], sxh[256], sxl[256];
    const int nv = nf - v_lo;
    const long long total = (long long)N * N * N * o_hi * o_hi * nv * nv;
    float dh = 0.f, dl = 0.f, xh = 0.f, xl = 0.f;
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
        const int kb = cadd[cdiff[ki * N + ka] * N + kj];                // ki - ka + kj
        if (occ[ki * nf + i] != 2 || occ[kj * nf + j] != 2 || occ[ka * nf + a] != 0 || occ[kb * nf + b] != 0) continue;
        const float D = (float)(ek[ka * nf + a] + ek[kb * nf + b] - ek[ki * nf + i] - ek[kj * nf + j]);
        const size_t blk = (size_t)o_hi * nv * o_hi * nv;
        const float2 gd = ekmo[(size_t)i + o_hi * ((size_t)(a - v_lo) + nv * ((size_t)j + (size_t)o_hi * (b - v_lo))) +
                               blk * ((size_t)ki + (size_t)N * ((size_t)ka + (size_t)N * kj))];
        if (fabsf(D) < 1e-8f) {
            ++nz;
            continue;
        }
        const float2 gx = ekmo[(size_t)i + o_hi * ((size_t)(b - v_lo) + nv * ((size_t)j + (size_t)o_hi * (a - v_lo))) +
                               blk * ((size_t)ki + (size_t)N * ((size_t)kb + (size_t)N * kj))];
        acc2(dh, dl, -(gd.x * gd.x + gd.y * gd.y) / D);
        acc2(xh, xl, -(gd.x * gx.x + gd.y * gx.y) / D);
    }
    if (nz) atomicAdd(nzero, nz);
    sdh[threadIdx.x] = dh;
    sdl[threadIdx.x] = dl;
    sxh[threadIdx.x] = xh;
    sxl[threadIdx.x] = xl;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            float h = sdh[threadIdx.x], l = sdl[threadIdx.x] + sdl[threadIdx.x + s];
            acc2(h, l, sdh[threadIdx.x + s]);
            sdh[threadIdx.x] = h;
            sdl[threadIdx.x] = l;
            h = sxh[threadIdx.x];
            l = sxl[threadIdx.x] + sxl[threadIdx.x + s];
            acc2(h, l, sxh[threadIdx.x + s]);
            sxh[threadIdx.x] = h;
            sxl[threadIdx.x] = l;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x] = sdh[0];
        part[4 * blockIdx.x + 1] = sdl[0];
        part[4 * blockIdx.x + 2] = sxh[0];
        part[4 * blockIdx.x + 3] = sxl[0];
    }
}

hipError_t launch_df_mp2_f32(const double *Fr, const double *Fi, const double2 *Ck, int nf, int nau, int N, int o_hi,
                             int v_lo, const int *cdiff, const int *cadd, const double *ek, const int *occ,
                             float2 *L, float2 *eri, float *part, int nblocks, unsigned long long *nzero,
                             hipStream_t st) {
    hipLaunchKernelGGL(k_df_l_f32, dim3(2048), dim3(256), 0, st, Fr, Fi, Ck, nf, nau, N, o_hi, v_lo, cdiff, L);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int per = o_hi * (nf - v_lo), M = per * N, nt = (M + 15) / 16;
    const int gx = std::max(1, std::min((nt * nt + 3) / 4, 1024));
    hipLaunchKernelGGL(k_df_eri_mfma32, dim3(gx, N), dim3(256), 0, st, L, nau, N, per, cdiff, cadd, eri);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_kmp2_sums_f32, dim3(nblocks), dim3(256), 0, st, eri, ek, occ, nf, N, cadd, cdiff, o_hi, v_lo,
                       part, nzero);
    return hipGetLastError();
}

hipError_t launch_df_l(const double *Fr, const double *Fi, const double2 *Ck, int nf, int nau, int N, int o_hi,
                       int v_lo, const int *cdiff, double2 *L, hipStream_t st) {
    hipLaunchKernelGGL(k_df_l, dim3(2048), dim3(256), 0, st, Fr, Fi, Ck, nf, nau, N, o_hi, v_lo, cdiff, L);
    return hipGetLastError();
}

hipError_t launch_df_eri(const double2 *L, int nau, int N, int o_hi, int nv, const int *cdiff, const int *cadd,
                         double2 *out, hipStream_t st) {
    const int per = o_hi * nv;
    const size_t lds = sizeof(double2) * (size_t)per * nau;
    if (lds <= 65536) {
        hipLaunchKernelGGL(k_df_eri_lds, dim3(N * N), dim3(256), lds, st, L, nau, N, per, cdiff, cadd, out);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_df_eri, dim3(4096), dim3(256), 0, st, L, nau, N, o_hi, nv, cdiff, cadd, out);
    return hipGetLastError();
}

}  // namespace tcg
