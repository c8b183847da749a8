// gfx950 kernels of the engine: zero fill, and J/K from the stored ERI.
// The quartet pipeline itself is in tcg_rows.hip.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "tcg_device.hpp"

namespace tcg {

__global__ void k_fill(double *p, size_t n, double v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u32(unsigned *p, size_t n, unsigned v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill_u32(unsigned *p, size_t n, unsigned v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill(double *p, size_t n, double v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// J and K from the stored chemist-order ERI(n1,nbas,nbas,nbas)
// (src/Energies/hartree_exchange_potential.f90:30-202):
//   J(a,b) += E(a,b,c,d) P(c,d),   K(a,d) += -1/2 E(a,b,c,d) P(b,c)
// The array is a matrix E[f][g] with f = a + n1*b (contiguous plane) and
// g = c + nbas*d.  Pass 1: block j streams the columns g in [j*G, j*G+G)
// once, coalesced over f, and writes its partial J plane and its partial K
// rows (K needs the sum over b of the threads sharing a).  Pass 2 sums the
// partials.  HBM-bound: every ERI element is read exactly once.
// ---------------------------------------------------------------------------
constexpr int kJKGmax = 32;
static inline int jk_g(int nbas) { return nbas < kJKGmax ? nbas : kJKGmax; }   // columns per block (<= nbas)

__global__ __launch_bounds__(256) void k_jk_partial(const double *__restrict__ eri, const double *__restrict__ P,
                                                   double *__restrict__ jpart, double *__restrict__ kpart,
                                                   int n1, int nbas, int G) {
    extern __shared__ double kacc[];        // [2][n1]
    const size_t plane = (size_t)n1 * nbas;
    const size_t ncol = (size_t)nbas * nbas;
    const size_t g0 = (size_t)blockIdx.x * G;
    const size_t g1 = (g0 + G < ncol) ? g0 + G : ncol;
    const int d0 = (int)(g0 / nbas);
    for (int i = threadIdx.x; i < 2 * n1; i += blockDim.x) kacc[i] = 0.0;
    __syncthreads();
    for (size_t f = threadIdx.x; f < plane; f += blockDim.x) {
        const int a = (int)(f % n1), b = (int)(f / n1);
        double jv = 0.0, k0 = 0.0, k1 = 0.0;
        for (size_t g = g0; g < g1; ++g) {
            const double e = eri[f + plane * g];
            const int c = (int)(g % nbas), d = (int)(g / nbas);
            jv += e * P[g];
            const double t = e * P[b + (size_t)nbas * c];
            if (d == d0) k0 += t; else k1 += t;
        }
        jpart[(size_t)blockIdx.x * plane + f] = jv;
        atomicAdd(&kacc[a], k0);
        atomicAdd(&kacc[n1 + a], k1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * n1; i += blockDim.x) kpart[(size_t)blockIdx.x * 2 * n1 + i] = kacc[i];
}

__global__ void k_jk_reduce(const double *__restrict__ jpart, const double *__restrict__ kpart, double *__restrict__ J,
                            double *__restrict__ K, int n1, int nbas, int nblk, int G) {
    const size_t plane = (size_t)n1 * nbas;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < plane) {
        double s = 0.0;
        for (int bk = 0; bk < nblk; ++bk) s += jpart[(size_t)bk * plane + idx];
        const int a = (int)(idx % n1), b = (int)(idx / n1);
        J[a + (size_t)nbas * b] = s;
        // K(a,d) with d = b here: blocks whose column range touches d
        const int d = b;
        double k = 0.0;
        const size_t gfirst = (size_t)d * nbas, glast = gfirst + nbas - 1;
        const int bk0 = (int)(gfirst / G), bk1 = (int)(glast / G);
        for (int bk = bk0; bk <= bk1; ++bk) {
            const int dd0 = (int)(((size_t)bk * G) / nbas);
            k += kpart[(size_t)bk * 2 * n1 + (d == dd0 ? 0 : n1) + a];
        }
        K[a + (size_t)nbas * d] = -0.5 * k;
    }
}

size_t jk_work_doubles(int n1, int nbas) {
    const size_t ncol = (size_t)nbas * nbas;
    const size_t nblk = (ncol + jk_g(nbas) - 1) / jk_g(nbas);
    return nblk * ((size_t)n1 * nbas + 2 * (size_t)n1);
}

hipError_t launch_jk_stored(const double *eri, const double *P, double *J, double *K, int n1, int nbas,
                            double *jwork, hipStream_t st) {
    const size_t ncol = (size_t)nbas * nbas;
    const int G = jk_g(nbas);
    const int nblk = (int)((ncol + G - 1) / G);
    const size_t plane = (size_t)n1 * nbas;
    double *jpart = jwork;
    double *kpart = jwork + (size_t)nblk * plane;
    hipLaunchKernelGGL(k_jk_partial, dim3(nblk), dim3(256), sizeof(double) * 2 * n1, st, eri, P, jpart, kpart, n1,
                       nbas, G);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_jk_reduce, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, st, jpart, kpart, J, K, n1,
                       nbas, nblk, G);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Column-sparse J/K (the reference's sparse path,
// src/Energies/hartree_exchange_potential.f90:54-155): the columns
// ERI(:,b,c,d) with a non-zero element are listed once per tensor, ordered by
// d (for K) and by b (for J); each SCF iteration then reads only them.
//   K(a,d) = -1/2 sum_{(b,c) in list_d} ERI(a,b,c,d) P(b,c)   (both orders; P symmetric)
//   J(a,b) =      sum_{(c,d) in list_b} ERI(a,b,c,d) P(c,d)   (chemist)
//   J(a,c) =      sum_{(b,d) in list_c} ERI(a,b,c,d) P(b,d)   (physicist)
// One block per d (K) or per J column: no global atomics.
// ---------------------------------------------------------------------------
// flags and keys per column; the per-d and per-target counts are gathered in
// LDS histograms by persistent workgroups (one global atomic per bin and
// workgroup instead of one per non-zero column).
__global__ __launch_bounds__(256) void k_col_flags(const double *__restrict__ eri, int n1, int nbas, int physicist,
                                                   unsigned char *__restrict__ flags, int *__restrict__ cnt_d,
                                                   int *__restrict__ cnt_t, unsigned *__restrict__ key_t) {
    extern __shared__ int hist[];                 // [nbas] per d, [nbas] per target
    for (int i = threadIdx.x; i < 2 * nbas; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const size_t ncol = (size_t)nbas * nbas * nbas;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < ncol; g += (size_t)gridDim.x * blockDim.x) {
        bool nz = false;
        for (int a = 0; a < n1; ++a) nz |= eri[(size_t)a + (size_t)n1 * g] != 0.0;
        flags[g] = nz ? 1 : 0;
        const int b = (int)(g % nbas), c = (int)((g / nbas) % nbas), d = (int)(g / ((size_t)nbas * nbas));
        const int t = physicist ? c : b;
        key_t[g] = (unsigned)t;
        if (nz) {
            atomicAdd(&hist[d], 1);
            atomicAdd(&hist[nbas + t], 1);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbas; i += blockDim.x) {
        if (hist[i]) atomicAdd(&cnt_d[i], hist[i]);
        if (hist[nbas + i]) atomicAdd(&cnt_t[i], hist[nbas + i]);
    }
}

// Column list entries g = b + nbas (c + nbas d); K(a,d) for the block's d, J(a,t)
// for the block's target t (t = b chemist, t = c physicist: J(a,c) += ERI(a,b,c,d)
// P(b,d), src/Energies/hartree_exchange_potential.f90:136-146).  Each output
// column's list is split over gridDim.y workgroups (enough workgroups to fill
// the chip when the columns are few); their sums are added to the zeroed output.
// Few rows (n1 <= NA): threads over the list entries, register accumulators
// per row, one LDS reduction per block.
template <int NA>
__global__ __launch_bounds__(256) void k_jk_sparse_small(const double *__restrict__ eri, const double *__restrict__ P,
                                                         const long long *__restrict__ list,
                                                         const long long *__restrict__ off, int n1, int nbas, int is_j,
                                                         int physicist, double *__restrict__ out) {
    __shared__ double acc[NA];
    const int col = blockIdx.x;
    for (int a = threadIdx.x; a < NA; a += blockDim.x) acc[a] = 0.0;
    __syncthreads();
    double loc[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) loc[a] = 0.0;
    const long long o0 = off[col], o1 = off[col + 1];
    const long long len = o1 - o0, chunk = (len + gridDim.y - 1) / gridDim.y;
    const long long i0 = o0 + chunk * blockIdx.y, i1 = i0 + chunk < o1 ? i0 + chunk : o1;
    for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const long long g = list[i];
        const int b = (int)(g % nbas), c = (int)((g / nbas) % nbas), d = (int)(g / ((long long)nbas * nbas));
        double pw;
        if (!is_j) pw = P[b + (size_t)nbas * c];
        else pw = physicist ? P[b + (size_t)nbas * d] : P[c + (size_t)nbas * d];
        const double *e = eri + (size_t)n1 * g;
#pragma unroll
        for (int a = 0; a < NA; ++a)
            if (a < n1) loc[a] += e[a] * pw;
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        double v = loc[a];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (a < n1 && (threadIdx.x & 63) == 0) atomicAdd(&acc[a], v);
    }
    __syncthreads();
    if (i1 <= i0) return;
    const double f = is_j ? 1.0 : -0.5;
    for (int a = threadIdx.x; a < n1; a += blockDim.x) atomicAdd(&out[a + (size_t)nbas * col], f * acc[a]);
}

// Many rows (n1 > 32, the full layout): threads over the rows a (coalesced
// column reads), the block walks its part of the list.
__global__ __launch_bounds__(256) void k_jk_sparse_rows(const double *__restrict__ eri, const double *__restrict__ P,
                                                        const long long *__restrict__ list,
                                                        const long long *__restrict__ off, int n1, int nbas, int is_j,
                                                        int physicist, double *__restrict__ out) {
    const int col = blockIdx.x;
    const double f = is_j ? 1.0 : -0.5;
    const long long o0 = off[col], o1 = off[col + 1];
    const long long len = o1 - o0, chunk = (len + gridDim.y - 1) / gridDim.y;
    const long long i0 = o0 + chunk * blockIdx.y, i1 = i0 + chunk < o1 ? i0 + chunk : o1;
    if (i1 <= i0) return;
    for (int a0 = 0; a0 < n1; a0 += blockDim.x) {
        const int a = a0 + threadIdx.x;
        double acc = 0.0;
        for (long long i = i0; i < i1; ++i) {
            const long long g = list[i];
            const int b = (int)(g % nbas), c = (int)((g / nbas) % nbas), d = (int)(g / ((long long)nbas * nbas));
            double pw;
            if (!is_j) pw = P[b + (size_t)nbas * c];
            else pw = physicist ? P[b + (size_t)nbas * d] : P[c + (size_t)nbas * d];
            if (a < n1) acc += eri[(size_t)a + (size_t)n1 * g] * pw;
        }
        if (a < n1) atomicAdd(&out[a + (size_t)nbas * col], f * acc);
    }
}

__global__ void k_zero_rows(double *__restrict__ out, int n1, int nbas) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < (long long)n1 * nbas) out[(idx % n1) + (size_t)nbas * (idx / n1)] = 0.0;
}

static hipError_t launch_one(const double *eri, const double *P, const long long *list, const long long *off, int n1,
                             int nbas, int is_j, int physicist, int splits, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_rows, dim3((unsigned)(((long long)n1 * nbas + 255) / 256)), dim3(256), 0, st, out, n1,
                       nbas);
    const dim3 grid(nbas, splits);
    if (n1 <= 8)
        hipLaunchKernelGGL(k_jk_sparse_small<8>, grid, dim3(256), 0, st, eri, P, list, off, n1, nbas, is_j,
                           physicist, out);
    else if (n1 <= 16)
        hipLaunchKernelGGL(k_jk_sparse_small<16>, grid, dim3(256), 0, st, eri, P, list, off, n1, nbas, is_j,
                           physicist, out);
    else if (n1 <= 32)
        hipLaunchKernelGGL(k_jk_sparse_small<32>, grid, dim3(256), 0, st, eri, P, list, off, n1, nbas, is_j,
                           physicist, out);
    else
        hipLaunchKernelGGL(k_jk_sparse_rows, grid, dim3(256), 0, st, eri, P, list, off, n1, nbas, is_j,
                           physicist, out);
    return hipGetLastError();
}

hipError_t launch_col_flags(const double *eri, int n1, int nbas, int physicist, unsigned char *flags, int *cnt_d,
                            int *cnt_t, unsigned *key_t, hipStream_t st) {
    const size_t ncol = (size_t)nbas * nbas * nbas;
    if ((size_t)nbas * 2 * sizeof(int) > 65536) return hipErrorInvalidValue;     // nbas <= 8192
    const int grid = (int)std::min<size_t>((ncol + 255) / 256, 4096);
    hipLaunchKernelGGL(k_col_flags, dim3(grid), dim3(256), sizeof(int) * 2 * (size_t)nbas, st, eri, n1, nbas,
                       physicist, flags, cnt_d, cnt_t, key_t);
    return hipGetLastError();
}

// nz = entries of each list (the non-zero columns).  Rows a < n1 of J and K
// are written; the others are left as they are.
hipError_t launch_jk_sparse(const double *eri, const double *P, const long long *list_d, const long long *off_d,
                            const long long *list_t, const long long *off_t, int n1, int nbas, int physicist,
                            long long nz, double *J, double *K, hipStream_t st) {
    // about 2048 workgroups per pass, each at least 512 list entries on average
    long long splits = (2048 + nbas - 1) / nbas;
    const long long per_col = nz / (nbas > 0 ? nbas : 1);
    splits = std::min<long long>(splits, std::max<long long>(1, per_col / 512));
    splits = std::max<long long>(1, std::min<long long>(splits, 64));
    hipError_t e = launch_one(eri, P, list_d, off_d, n1, nbas, 0, physicist, (int)splits, K, st);
    if (e != hipSuccess) return e;
    return launch_one(eri, P, list_t, off_t, n1, nbas, 1, physicist, (int)splits, J, st);
}

__global__ void k_gather_keys(const unsigned *__restrict__ keys, const long long *__restrict__ idx, long long n,
                              unsigned *__restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = keys[idx[i]];
}

hipError_t launch_gather_keys(const unsigned *keys, const long long *idx, long long n, unsigned *out, hipStream_t st) {
    hipLaunchKernelGGL(k_gather_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, keys, idx, n, out);
    return hipGetLastError();
}

}  // namespace tcg
