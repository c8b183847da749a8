// Learned per-quartet quadrature orders (gfx950).
//
//   k_features   six features of every surviving primitive quartet (one
//                thread each), normalised; a quartet outside the training
//                range is flagged.
//   k_rules      the rule of each shell quartet: the largest order index the
//                network asks for over its primitive quartets, one step of
//                margin added; any flagged primitive quartet -> the plain rule.
//                Rule numbers of the tables: 0 = the plain rule, r + 1 =
//                learned order r.
//   k_rule_hist  shell quartets per rule.
//
// Features of the primitive quartet (e1 A, e2 B | e3 C, e4 D) on a torus with
// ndp periodic directions d < ndp of lengths L_d (a_d = 2 pi / L_d, chord
// distances 2 (1 - cos a_d x) / a_d^2) and real directions d >= ndp, L = the
// shortest periodic length:
//   f0 = log10 rho,  rho = P Q / (P + Q),  P = e1 + e2,  Q = e3 + e4
//   f1 = log10 (rho L^2)
//   f2 = log10 (1 + rho R^2), R = distance of the pair centres (periodic
//        directions: circular weighted mean, chord; real ones: weighted mean)
//   f3 = (lA + lB + lC + lD) / 4
//   f4 = log10 (max(e) L^2)
//   f5 = max(log10 (K_AB K_CD), -30) / 10,  K_AB = exp(-e1 e2 / P d_AB^2)
#include <hip/hip_runtime.h>

#include "tcg_device.hpp"

namespace tcg {

__device__ __forceinline__ double chord2(double a, double dx) { return 2.0 * (1.0 - cos(a * dx)) / (a * a); }

__device__ void pair_centre(int ndp, const double *a, double e1, const double *A, double e2, const double *B,
                            double *P) {
    for (int d = 0; d < 3; ++d) {
        if (d < ndp) {
            const double s = e1 * sin(a[d] * A[d]) + e2 * sin(a[d] * B[d]);
            const double cc = e1 * cos(a[d] * A[d]) + e2 * cos(a[d] * B[d]);
            P[d] = atan2(s, cc) / a[d];
        } else {
            P[d] = (e1 * A[d] + e2 * B[d]) / (e1 + e2);
        }
    }
}

__device__ double pair_d2(int ndp, const double *a, const double *A, const double *B) {
    double r = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double dx = A[d] - B[d];
        r += d < ndp ? chord2(a[d], dx) : dx * dx;
    }
    return r;
}

__global__ __launch_bounds__(256) void k_features(EriArgs A, const int *__restrict__ owner, long long ns,
                                                  const Surv *__restrict__ surv, const double *__restrict__ expo,
                                                  const float *__restrict__ norm, float *__restrict__ feat,
                                                  unsigned char *__restrict__ ood) {
    const int ndp = A.ndp;
    double L = A.L[0];
    for (int d = 1; d < ndp; ++d) L = fmin(L, A.L[d]);
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < ns; g += (long long)gridDim.x * blockDim.x) {
        const Quartet &Q = A.q4[owner[g]];
        const Shell &sA = A.sh[Q.s[0]], &sB = A.sh[Q.s[1]], &sC = A.sh[Q.s[2]], &sD = A.sh[Q.s[3]];
        const int lt = sA.l + sB.l + sC.l + sD.l;
        {
            const Surv S = surv[g];
            const PrimPair &pb = A.pp[S.pb], &pk = A.pp[S.pk];
            const double e1 = expo[A.ao_poff[sA.first_ao] + pb.pi], e2 = expo[A.ao_poff[sB.first_ao] + pb.pj];
            const double e3 = expo[A.ao_poff[sC.first_ao] + pk.pi], e4 = expo[A.ao_poff[sD.first_ao] + pk.pj];
            const double P = e1 + e2, Qe = e3 + e4, rho = P * Qe / (P + Qe);
            double Pc[3], Qc[3];
            pair_centre(ndp, A.a, e1, sA.xyz, e2, sB.xyz, Pc);
            pair_centre(ndp, A.a, e3, sC.xyz, e4, sD.xyz, Qc);
            const double R2 = pair_d2(ndp, A.a, Pc, Qc);
            const double emax = fmax(fmax(e1, e2), fmax(e3, e4));
            const double lk = -(e1 * e2 / P * pair_d2(ndp, A.a, sA.xyz, sB.xyz) +
                                e3 * e4 / Qe * pair_d2(ndp, A.a, sC.xyz, sD.xyz)) / 2.302585092994046;
            double f[6];
            f[0] = log10(rho);
            f[1] = log10(rho * L * L);
            f[2] = log10(1.0 + rho * R2);
            f[3] = 0.25 * lt;
            f[4] = log10(emax * L * L);
            f[5] = fmax(lk, -30.0) / 10.0;
            // norm: mean[6], std[6], lo[6], hi[6]
            bool out = false;
            for (int k = 0; k < 6; ++k) {
                out = out || f[k] < norm[12 + k] || f[k] > norm[18 + k];
                feat[g * 6 + k] = (float)((f[k] - norm[k]) / norm[6 + k]);
            }
            ood[g] = out ? 1 : 0;
        }
    }
}

// Skipping (log_skip > -1e30): a shell quartet whose primitive quartets all lie in
// the training range is left out (rule kRuleSkip) when the bound head's largest
// log10 |integral| over them, plus log10 of their number and margin_b, is below
// log_skip, and its Schwarz product Qsh(bra) Qsh(ket) is below cap (cap > 0): the
// learned estimate decides inside the band [1e-13, cap) of the Schwarz screen,
// whose inequality keeps every skipped integral below cap.
__global__ __launch_bounds__(64) void k_rules(int nq4, const long long *__restrict__ soff, const float *__restrict__ pred,
                                              const unsigned char *__restrict__ ood, float margin, int plain,
                                              float log_skip, float margin_b, const Quartet *__restrict__ q4,
                                              const double *__restrict__ Qsh, int nsh, double cap,
                                              int *__restrict__ qrule, float *__restrict__ lbound) {
    for (int qi = blockIdx.x; qi < nq4; qi += gridDim.x) {
        int r = 0;
        float bmax = -1e30f;
        for (long long g = soff[qi] + threadIdx.x; g < soff[qi + 1]; g += 64) {
            int rg = plain;
            if (!ood[g]) {
                const float y = pred[g * 2 + 1] + margin;
                rg = y <= 0.f ? 0 : (int)ceilf(y);
                rg = rg > plain ? plain : rg;
                bmax = fmaxf(bmax, pred[g * 2]);
            } else {
                bmax = 1e30f;                        // outside the training range: never skipped
            }
            r = r > rg ? r : rg;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const int o = __shfl_xor(r, off, 64);
            r = r > o ? r : o;
            bmax = fmaxf(bmax, __shfl_xor(bmax, off, 64));
        }
        if (threadIdx.x == 0) {
            const long long n = soff[qi + 1] - soff[qi];
            bool skip = n > 0 && log_skip > -1e30f && bmax + log10f((float)n) + margin_b < log_skip;
            if (skip && cap > 0.0) {
                const Quartet &Q = q4[qi];
                skip = Qsh[(size_t)Q.s[0] + (size_t)nsh * Q.s[1]] * Qsh[(size_t)Q.s[2] + (size_t)nsh * Q.s[3]] < cap;
            }
            qrule[qi] = skip ? kRuleSkip : (r >= plain ? 0 : r + 1);
            // log10 of the learned bound of the whole shell quartet (largest predicted
            // integral, times the number of primitive quartets, times 10^margin_b)
            if (lbound) lbound[qi] = n > 0 ? bmax + log10f((float)n) + margin_b : -1e30f;
        }
    }
}

// keep[i] = 1 when quartet i is evaluated (its rule is not kRuleSkip)
__global__ void k_rule_keep(int nq4, const int *__restrict__ qrule, unsigned char *__restrict__ keep) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nq4) keep[i] = qrule[i] == kRuleSkip ? 0 : 1;
}

// q4_out[j] = q4[idx[j]], rule_out[j] = qrule[idx[j]]
__global__ void k_gather_quartets(const Quartet *__restrict__ q4, const int *__restrict__ qrule,
                                  const long long *__restrict__ idx, const unsigned long long *__restrict__ n,
                                  Quartet *__restrict__ q4_out, int *__restrict__ rule_out) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < (long long)*n) {
        q4_out[j] = q4[idx[j]];
        rule_out[j] = qrule[idx[j]];
    }
}

__global__ __launch_bounds__(256) void k_rule_hist(int nq4, const int *__restrict__ qrule,
                                                   unsigned long long *__restrict__ hist) {
    __shared__ unsigned h[16];
    if (threadIdx.x < 16) h[threadIdx.x] = 0;
    __syncthreads();
    for (int qi = blockIdx.x * blockDim.x + threadIdx.x; qi < nq4; qi += gridDim.x * blockDim.x)
        atomicAdd(&h[qrule[qi] & 15], 1u);
    __syncthreads();
    if (threadIdx.x < 16 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

hipError_t launch_features(const EriArgs &a, const int *owner, long long ns, const Surv *surv, const double *expo,
                           const float *norm, float *feat, unsigned char *ood, hipStream_t st) {
    if (ns <= 0) return hipSuccess;
    const long long blocks = (ns + 255) / 256;
    const int grid = blocks < (1 << 20) ? (int)blocks : (1 << 20);
    hipLaunchKernelGGL(k_features, dim3(grid), dim3(256), 0, st, a, owner, ns, surv, expo, norm, feat, ood);
    return hipGetLastError();
}

hipError_t launch_rules(int nq4, const long long *soff, const float *pred, const unsigned char *ood, float margin,
                        int plain, float log_skip, float margin_b, const Quartet *q4, const double *Qsh, int nsh,
                        double cap, int *qrule, float *lbound, hipStream_t st) {
    if (nq4 <= 0) return hipSuccess;
    const int grid = nq4 < (1 << 20) ? nq4 : (1 << 20);
    hipLaunchKernelGGL(k_rules, dim3(grid), dim3(64), 0, st, nq4, soff, pred, ood, margin, plain, log_skip, margin_b,
                       q4, Qsh, nsh, cap, qrule, lbound);
    return hipGetLastError();
}

hipError_t launch_rule_keep(int nq4, const int *qrule, unsigned char *keep, hipStream_t st) {
    hipLaunchKernelGGL(k_rule_keep, dim3((nq4 + 255) / 256 > 0 ? (nq4 + 255) / 256 : 1), dim3(256), 0, st, nq4, qrule,
                       keep);
    return hipGetLastError();
}

hipError_t launch_gather_quartets(const Quartet *q4, const int *qrule, const long long *idx,
                                  const unsigned long long *n, long long nmax, Quartet *q4_out, int *rule_out,
                                  hipStream_t st) {
    const long long g = (nmax + 255) / 256;
    hipLaunchKernelGGL(k_gather_quartets, dim3((unsigned)(g > 0 ? g : 1)), dim3(256), 0, st, q4, qrule, idx, n, q4_out,
                       rule_out);
    return hipGetLastError();
}

hipError_t launch_rule_hist(int nq4, const int *qrule, unsigned long long *hist, hipStream_t st) {
    if (nq4 <= 0) return hipSuccess;
    const int grid = (nq4 + 255) / 256 < 1024 ? (nq4 + 255) / 256 : 1024;
    hipLaunchKernelGGL(k_rule_hist, dim3(grid), dim3(256), 0, st, nq4, qrule, hist);
    return hipGetLastError();
}

}  // namespace tcg
