// Density screening of the direct J/K (gfx950): the shell quartets whose
// integral bound times the largest density element they multiply is below a
// threshold leave the list of one J/K build.  With a density difference
// P_n - P_(n-1) this is the incremental Fock build of integral-direct SCF: the
// change of J and K is built from the change of P, and late iterations, where
// the change is small, evaluate few quartets.
//
//   k_pair_dmax      Dsh(X, Y) = max |P(i, j)| over i in shell X of the unit
//                    cell, j in shell Y (any cell): [loc(X) + nshuc Y]
//   k_density_keep   keep[q] = bound(q) max_{pairs XY of q} Dsh(XY) >= tau,
//                    bound = the Schwarz product Qsh(AB) Qsh(CD), or the
//                    smaller of it and the network's bound (10^lbound)
//   k_gather_list    the kept quartets (with their rules and bounds)
//
// The density is lattice periodic, P(i + R, j + R) = P(i, j), so the block of a
// shell pair depends only on the two shells' relative cell: the pair (X, Y)
// reads Dsh(loc X, loc Y + nshuc relc(cell X, cell Y)).  The six pairs of a
// quartet (AB, CD, AC, AD, BC, BD) are the density blocks its eight
// permutations contract with, in J and in K.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "tcg_device.hpp"

namespace tcg {

__global__ __launch_bounds__(256) void k_pair_dmax(const double *__restrict__ P, int nb, const Shell *__restrict__ sh,
                                                   int nshuc, int nsh, double *__restrict__ Dsh) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nshuc * nsh) return;
    const int X = t % nshuc, Y = t / nshuc;
    const Shell &sx = sh[X], &sy = sh[Y];
    double m = 0.0;
    for (int j = 0; j < sy.ncomp; ++j)
        for (int i = 0; i < sx.ncomp; ++i) m = fmax(m, fabs(P[(size_t)(sx.first_ao + i) + (size_t)nb * (sy.first_ao + j)]));
    Dsh[t] = m;
}

__device__ __forceinline__ double pair_d(const Shell *__restrict__ sh, const int *__restrict__ relc, int ncell,
                                         int nshuc, const double *__restrict__ Dsh, int X, int Y) {
    const Shell &sx = sh[X], &sy = sh[Y];
    const int yr = sy.loc + nshuc * relc[sx.cell * ncell + sy.cell];
    return Dsh[sx.loc + (size_t)nshuc * yr];
}

__global__ __launch_bounds__(256) void k_density_keep(const Quartet *__restrict__ q4, int nq4,
                                                      const Shell *__restrict__ sh, const double *__restrict__ Qsh,
                                                      int nsh, int nshuc, const int *__restrict__ relc, int ncell,
                                                      const double *__restrict__ Dsh, const float *__restrict__ lbound,
                                                      double tau, unsigned char *__restrict__ keep) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq4; q += gridDim.x * blockDim.x) {
        const Quartet Q = q4[q];
        const int A = Q.s[0], B = Q.s[1], C = Q.s[2], D = Q.s[3];
        double d = pair_d(sh, relc, ncell, nshuc, Dsh, A, B);
        d = fmax(d, pair_d(sh, relc, ncell, nshuc, Dsh, C, D));
        d = fmax(d, pair_d(sh, relc, ncell, nshuc, Dsh, A, C));
        d = fmax(d, pair_d(sh, relc, ncell, nshuc, Dsh, A, D));
        d = fmax(d, pair_d(sh, relc, ncell, nshuc, Dsh, B, C));
        d = fmax(d, pair_d(sh, relc, ncell, nshuc, Dsh, B, D));
        double bound = Qsh[(size_t)A + (size_t)nsh * B] * Qsh[(size_t)C + (size_t)nsh * D];
        if (lbound) bound = fmin(bound, exp10((double)lbound[q]));
        keep[q] = bound * d >= tau ? 1 : 0;
    }
}

__global__ __launch_bounds__(256) void k_gather_list(const Quartet *__restrict__ q4, const int *__restrict__ qrule,
                                                     const float *__restrict__ lb, const long long *__restrict__ idx,
                                                     const unsigned long long *__restrict__ n,
                                                     Quartet *__restrict__ q4_out, int *__restrict__ rule_out,
                                                     float *__restrict__ lb_out) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < (long long)*n) {
        const long long i = idx[j];
        q4_out[j] = q4[i];
        if (qrule) rule_out[j] = qrule[i];
        if (lb) lb_out[j] = lb[i];
    }
}

// out = alpha a + beta b (out may be a or b)
__global__ __launch_bounds__(256) void k_axpby(const double *a, const double *b, double *out, size_t n, double alpha,
                                               double beta) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = alpha * a[i] + beta * b[i];
}

hipError_t launch_axpby(const double *a, const double *b, double *out, size_t n, double alpha, double beta,
                        hipStream_t st) {
    if (n == 0) return hipSuccess;
    const size_t g = std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_axpby, dim3((unsigned)g), dim3(256), 0, st, a, b, out, n, alpha, beta);
    return hipGetLastError();
}

hipError_t launch_pair_dmax(const double *P, int nb, const Shell *sh, int nshuc, int nsh, double *Dsh,
                            hipStream_t st) {
    const int n = nshuc * nsh;
    hipLaunchKernelGGL(k_pair_dmax, dim3((n + 255) / 256), dim3(256), 0, st, P, nb, sh, nshuc, nsh, Dsh);
    return hipGetLastError();
}

hipError_t launch_density_keep(const Quartet *q4, int nq4, const Shell *sh, const double *Qsh, int nsh, int nshuc,
                               const int *relc, int ncell, const double *Dsh, const float *lbound, double tau,
                               unsigned char *keep, hipStream_t st) {
    if (nq4 <= 0) return hipSuccess;
    const int grid = (int)std::min<long long>(((long long)nq4 + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(k_density_keep, dim3(grid), dim3(256), 0, st, q4, nq4, sh, Qsh, nsh, nshuc, relc, ncell, Dsh,
                       lbound, tau, keep);
    return hipGetLastError();
}

hipError_t launch_gather_list(const Quartet *q4, const int *qrule, const float *lb, const long long *idx,
                              const unsigned long long *n, long long nmax, Quartet *q4_out, int *rule_out,
                              float *lb_out, hipStream_t st) {
    const long long g = (nmax + 255) / 256;
    hipLaunchKernelGGL(k_gather_list, dim3((unsigned)(g > 0 ? g : 1)), dim3(256), 0, st, q4, qrule, lb, idx, n, q4_out,
                       rule_out, lb_out);
    return hipGetLastError();
}

}  // namespace tcg
