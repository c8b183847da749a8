// Setup kernels of the ERI build: pair Fourier vectors and the list of
// unique shell quartets, both generated on the GPU.
#include <hip/hip_runtime.h>

#include "tcg_device.hpp"

namespace tcg {

__constant__ int c_perm8_setup[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                        {2, 3, 0, 1}, {3, 2, 0, 1}, {2, 3, 1, 0}, {3, 2, 1, 0}};

// exp(-x) I_0(x): power series below 20, Hankel asymptotic series above,
// summed to the smallest term (the host's i0_scaled).
__device__ double d_i0_scaled(double x) {
    if (x <= 0.0) return 1.0;
    if (x <= 20.0) {
        const double q = 0.25 * x * x;
        double term = 1.0, sum = 1.0;
        for (int k = 1; k < 200; ++k) {
            term *= q / ((double)k * (double)k);
            sum += term;
            if (term < 1e-17 * sum) break;
        }
        return sum * exp(-x);
    }
    double term = 1.0, sum = 1.0, prev = 1e300;
    for (int k = 1; k < 200; ++k) {
        const double t = term * (2.0 * k - 1) * (2.0 * k - 1) / (8.0 * k * x);
        if (t >= prev || t < 1e-18 * sum) break;
        prev = t;
        term = t;
        sum += term;
    }
    return sum / sqrt(2.0 * 3.14159265358979323846 * x);
}

// One thread per (primitive pair, periodic direction): exp(-A) I_n(A) by
// Miller's recurrence started at M = N + lA + lB (the reference's
// bessel_scaled_vector_fast), then for every power pair (ia, ib) the vector
//   V_m = sum_k d_k exp(-A) I_|m-k|(A),  m = 0..N,
// d_k the Fourier coefficients of sin^ia(u+phA) sin^ib(u+phB).
__global__ void k_pair_vectors(const VecJob *__restrict__ jobs, int njobs, double *__restrict__ ivbuf,
                               double2 *__restrict__ vec) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= njobs) return;
    const VecJob J = jobs[j];
    const int M = J.N + J.lA + J.lB;
    double *iv = ivbuf + J.ivoff;
    if (J.A < 1e-10) {
        for (int n = 0; n <= M; ++n) iv[n] = 0.0;
        iv[0] = 1.0;
    } else {
        const double r = 2.0 / J.A;
        double up = 0.0, cur = 1e-280;     // I_{k+1}, I_k
        iv[M] = cur;
        for (int k = M; k >= 1; --k) {
            const double dn = up + (double)k * r * cur;
            iv[k - 1] = dn;
            up = cur;
            cur = dn;
            if (dn > 1e250) {
                for (int m = k - 1; m <= M; ++m) iv[m] *= 1e-250;
                up *= 1e-250;
                cur *= 1e-250;
            }
        }
        const double s = d_i0_scaled(J.A) / iv[0];
        for (int n = 0; n <= M; ++n) iv[n] *= s;
    }
    const int nb = J.lB + 1;
    for (int ia = 0; ia <= J.lA; ++ia)
        for (int ib = 0; ib <= J.lB; ++ib) {
            double2 *V = vec + J.voff + (size_t)(ia * nb + ib) * (J.N + 1);
            const int K = ia + ib;
            if (K == 0) {
                for (int m = 0; m <= J.N; ++m) V[m] = double2{iv[m], 0.0};
                continue;
            }
            // d_k by repeated multiplication with e^{i ph} e^{iu}/(2i) - e^{-i ph} e^{-iu}/(2i)
            double dr[2 * 2 * kMaxL + 3], di[2 * 2 * kMaxL + 3];
            for (int k = 0; k < 2 * K + 1; ++k) { dr[k] = 0.0; di[k] = 0.0; }
            dr[K] = 1.0;
            for (int m = 1; m <= K; ++m) {
                const double ph = (m <= ia) ? J.phA : J.phB;
                double sp, cp;
                sincos(ph, &sp, &cp);
                // cp_ = (sin ph - i cos ph)/2 multiplies e^{+iu}; cm_ = (sin ph + i cos ph)/2 multiplies e^{-iu}
                double nr[2 * 2 * kMaxL + 3], ni[2 * 2 * kMaxL + 3];
                for (int k = 0; k < 2 * K + 1; ++k) { nr[k] = 0.0; ni[k] = 0.0; }
                for (int k = -(m - 1); k <= m - 1; ++k) {
                    const double ar = dr[k + K], ai = di[k + K];
                    nr[k + 1 + K] += 0.5 * (sp * ar + cp * ai);
                    ni[k + 1 + K] += 0.5 * (sp * ai - cp * ar);
                    nr[k - 1 + K] += 0.5 * (sp * ar - cp * ai);
                    ni[k - 1 + K] += 0.5 * (sp * ai + cp * ar);
                }
                for (int k = 0; k < 2 * K + 1; ++k) { dr[k] = nr[k]; di[k] = ni[k]; }
            }
            for (int m = 0; m <= J.N; ++m) {
                double sr = 0.0, si = 0.0;
                for (int k = -K; k <= K; ++k) {
                    const double b = iv[abs(m - k)];
                    sr += dr[k + K] * b;
                    si += di[k + K] * b;
                }
                V[m] = double2{sr, si};
            }
        }
}

hipError_t launch_pair_vectors(const VecJob *jobs, int njobs, double *ivbuf, double2 *vec, hipStream_t st) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_vectors, dim3((njobs + 63) / 64), dim3(64), 0, st, jobs, njobs, ivbuf, vec);
    return hipGetLastError();
}

// Unique shell quartets (src/Toroidal/ERI_toroidal.f90:342-401): SI in the
// unit cell, the local indices of SJ, SK, SL not below SI, both Schwarz
// screens, and (SI,SJ,SK,SL) the smallest member of its orbit under the 8
// index permutations x the lattice translations.  One thread per
// (SI, SJ, SK); pass 0 counts, pass 1 appends.
__global__ void k_enum_quartets(const Shell *__restrict__ sh, const int *__restrict__ relc,
                                const double *__restrict__ Qsh, double Qmax, int nshuc, int nsh, int ncell,
                                unsigned long long *__restrict__ count, Quartet *__restrict__ out,
                                unsigned long long cap) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nshuc * nsh * nsh) return;
    const int SI = (int)(t / ((long)nsh * nsh));
    const int SJ = (int)((t / nsh) % nsh);
    const int SK = (int)(t % nsh);
    if (sh[SJ].loc < SI || sh[SK].loc < SI) return;
    const double qij = Qsh[(size_t)SI + (size_t)nsh * SJ];
    if (qij * Qmax < kSchwarz) return;
    const int cK = sh[SK].cell, lK = sh[SK].loc;
    for (int SL = 0; SL < nsh; ++SL) {
        if (sh[SL].loc < SI) continue;
        if (qij * Qsh[(size_t)SK + (size_t)nsh * SL] < kSchwarz) continue;
        const int q[4] = {SI, SJ, SK, SL};
        const int qc[4] = {0, sh[SJ].cell, cK, sh[SL].cell};
        const int ql[4] = {SI, sh[SJ].loc, lK, sh[SL].loc};
        bool smaller = false;
        for (int ip = 1; ip < 8 && !smaller; ++ip) {
            int t4[4];
            for (int x = 0; x < 4; ++x) t4[x] = c_perm8_setup[ip][x];
            const int c0 = qc[t4[0]];
            for (int x = 0; x < 4; ++x) {
                const int rr = relc[(size_t)c0 * ncell + qc[t4[x]]] * nshuc + ql[t4[x]];
                if (rr < q[x]) { smaller = true; break; }
                if (rr > q[x]) break;
            }
        }
        if (smaller) continue;
        const unsigned long long pos = atomicAdd(count, 1ull);
        if (out && pos < cap) {
            Quartet Q;
            Q.s[0] = SI; Q.s[1] = SJ; Q.s[2] = SK; Q.s[3] = SL;
            Q.bra = SI * nsh + SJ;
            Q.ket = lK * nsh + relc[(size_t)cK * ncell + sh[SL].cell] * nshuc + sh[SL].loc;
            for (int d = 0; d < 3; ++d) Q.shift[d] = sh[SK].xyz[d] - sh[lK].xyz[d];
            Q.out = -1;
            Q.prim = -1;
            out[pos] = Q;
        }
    }
}

hipError_t launch_enum_quartets(const Shell *sh, const int *relc, const double *Qsh, double Qmax, int nshuc,
                                int nsh, int ncell, unsigned long long *count, Quartet *out,
                                unsigned long long cap, hipStream_t st) {
    const long n = (long)nshuc * nsh * nsh;
    hipLaunchKernelGGL(k_enum_quartets, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sh, relc, Qsh, Qmax,
                       nshuc, nsh, ncell, count, out, cap);
    return hipGetLastError();
}

}  // namespace tcg
