// Setup kernels of the ERI build: pair Fourier vectors and the list of
// unique shell quartets, both generated on the GPU.
#include <hip/hip_runtime.h>

#include "tcg_device.hpp"
#include "tcg_math.hpp"

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

// exp(-A) I_n(A), n = 0..M, M = N + lA + lB, one wavefront per (primitive
// pair, periodic direction), by Miller's recurrence
//   w_{k-1} = w_{k+1} + (2k/A) w_k,  w_{M+1} = 0,  w_M = 1
// (the reference's bessel_scaled_vector_fast start), normalised to
// exp(-A) I_0(A).  The M steps are split into 64 consecutive segments, one
// per lane, top segment at lane 0: each lane forms the 2x2 transfer matrix of
// its segment, (w_{lo}, w_{lo+1}) = T (w_{hi}, w_{hi+1}); an inclusive scan
// of these products over the lanes (scaled by powers of two, exponent kept
// apart) gives every segment its starting pair, and each lane then runs its
// own segment.  Inside one segment the values grow by less than
// (1 + 2M/A)^ceil(M/64), far inside the double range for the M = nmax(A) + l
// used here, so only the scan needs the exponents.
struct Mat2 {
    double a, b, c, d;
    int e;                                       // value = [[a b] [c d]] * 2^e
};

__device__ __forceinline__ Mat2 mat2_mul(const Mat2 &P, const Mat2 &Q) {
    Mat2 R{P.a * Q.a + P.b * Q.c, P.a * Q.b + P.b * Q.d, P.c * Q.a + P.d * Q.c, P.c * Q.b + P.d * Q.d, P.e + Q.e};
    const double m = fmax(fmax(fabs(R.a), fabs(R.b)), fmax(fabs(R.c), fabs(R.d)));
    if (m > 0.0) {
        int ex;
        frexp(m, &ex);
        R.a = ldexp(R.a, -ex);
        R.b = ldexp(R.b, -ex);
        R.c = ldexp(R.c, -ex);
        R.d = ldexp(R.d, -ex);
        R.e += ex;
    }
    return R;
}

__device__ __forceinline__ Mat2 mat2_shfl_up(const Mat2 &P, int delta) {
    return Mat2{__shfl_up(P.a, delta, 64), __shfl_up(P.b, delta, 64), __shfl_up(P.c, delta, 64),
                __shfl_up(P.d, delta, 64), __shfl_up(P.e, delta, 64)};
}

__global__ __launch_bounds__(64) void k_bessel_jobs(const VecJob *__restrict__ jobs, int njobs, double *__restrict__ ivbuf) {
    const int lane = threadIdx.x;
    for (int jb = blockIdx.x; jb < njobs; jb += gridDim.x) {
        const VecJob J = jobs[jb];
        const int M = J.N + J.lA + J.lB;
        double *iv = ivbuf + J.ivoff;
        if (J.A < 1e-10) {
            for (int n = lane; n <= M; n += 64) iv[n] = n == 0 ? 1.0 : 0.0;
            continue;
        }
        const double r = 2.0 / J.A;
        const int seg = (M + 63) / 64;
        const int khi = M - lane * seg, klo = khi - seg > 0 ? khi - seg : 0;
        // transfer matrix of this lane's segment
        Mat2 T{1.0, 0.0, 0.0, 1.0, 0};
        for (int k = khi; k > klo; --k) {
            const double ak = (double)k * r;
            const Mat2 n{ak * T.a + T.c, ak * T.b + T.d, T.a, T.b, 0};
            T = n;
        }
        T = mat2_mul(T, Mat2{1.0, 0.0, 0.0, 1.0, 0});          // scaled, exponent apart
        // inclusive scan, lane j: T_j T_{j-1} ... T_0
        Mat2 P = T;
#pragma unroll
        for (int dl = 1; dl < 64; dl <<= 1) {
            const Mat2 Q = mat2_shfl_up(P, dl);
            if (lane >= dl) P = mat2_mul(P, Q);
        }
        Mat2 S = mat2_shfl_up(P, 1);
        if (lane == 0) S = Mat2{1.0, 0.0, 0.0, 1.0, 0};
        // walk: (w_khi, w_khi+1) = S (1, 0); the lane covering k = 1 ends at w_0
        const int jbot = (M - 1) / seg;
        double u0 = S.a, u1 = S.c;
        for (int k = khi; k > klo; --k) {
            const double dn = u1 + (double)k * r * u0;
            u1 = u0;
            u0 = dn;
        }
        const double w0 = __shfl(u0, jbot, 64);
        const int e0 = __shfl(S.e, jbot, 64);
        const double sc = d_i0_scaled(J.A) / w0;
        const int de = S.e - e0;
        if (khi >= 1) {
            u0 = S.a;
            u1 = S.c;
            iv[khi] = ldexp(u0 * sc, de);
            for (int k = khi; k > klo; --k) {
                const double dn = u1 + (double)k * r * u0;
                u1 = u0;
                u0 = dn;
                if (k - 1 > klo || klo == 0) iv[k - 1] = ldexp(dn * sc, de);
            }
        }
    }
}

// Primitive-pair records of the representative shell pairs, on the GPU: one
// workgroup per shell pair, a thread per primitive pair (i major).  Per pair:
// prefactor prod_d exp(-2(e1+e2-p_d)/a_d^2) prod_real exp(-mu R^2),
// barycentres (toroidal per periodic direction, clifford_pair; the ordinary
// Gaussian centre in the real ones), exponent sum, coefficient bound, Bessel
// lengths N_d = min(nmax(2 p_d / a_d^2), ncap) and the Fourier-vector jobs with
// the coefficients d_k of sin^ia(u+phA) sin^ib(u+phB); Hermite E of the real
// directions (src/module/clifford_tools.f90:467-549,578-607,715-785;
// src/Toroidal/ERI_shell_block_toroidal.f90:133-209).  ps starts at 1 (the
// Schwarz pass sets it).
struct PairSetupArgs {
    const PairBlock *blk;
    const Shell *sh;
    const double *expo, *coef;
    const int *ao_poff;
    int ndp, ncap;
    double L[3], a[3];
    PrimPair *pp;
    VecJob *jobs;
    VecOut *vo;
    double2 *dco;
    double *herm;
};

__global__ __launch_bounds__(64) void k_pair_setup(PairSetupArgs S) {
    const PairBlock B = S.blk[blockIdx.x];
    const Shell sA = S.sh[B.S1], sB = S.sh[B.S2];
    const int nij = sA.nprim * sB.nprim, ndp = S.ndp, nreal = 3 - ndp;
    const int ncomb = (sA.l + 1) * (sB.l + 1), kE = 2 * kMaxL + 1, dstride = 2 * (sA.l + sB.l) + 1;
    for (int ij = threadIdx.x; ij < nij; ij += blockDim.x) {
        const int i = ij / sB.nprim, j = ij - i * sB.nprim;
        PrimPair P{};
        P.pi = i;
        P.pj = j;
        const double e1 = S.expo[sA.poff + i], e2 = S.expo[sB.poff + j];
        // esum = 0: two constant functions (the exponent-0 s function of the
        // density-fitting integrals, df_toroidal.f90:837-844)
        const double esum = e1 + e2, mu = esum > 0.0 ? e1 * e2 / esum : 0.0;
        P.q = esum;
        double fac = 1.0;
        for (int d = 0; d < 3; ++d) {
            if (d < ndp) {
                double pe, cen;
                clifford_pair(e1, e2, sA.xyz[d], sB.xyz[d], S.L[d], S.a[d], &pe, &cen);
                P.cen[d] = cen;
                fac *= exp(-2.0 * (esum - pe) / (S.a[d] * S.a[d]));
                const double Aarg = 2.0 * pe / (S.a[d] * S.a[d]);
                const int nm = nmax_est(Aarg);
                const int N = nm < S.ncap ? nm : S.ncap;
                const int jl = ij * ndp + d;
                P.N[d] = N;
                P.A[d] = (float)Aarg;
                P.voff[d] = B.vec0 + jl * ncomb * (S.ncap + 1);
                VecJob J;
                J.A = Aarg;
                J.phA = S.a[d] * (cen - sA.xyz[d]);
                J.phB = S.a[d] * (cen - sB.xyz[d]);
                J.N = N;
                J.lA = sA.l;
                J.lB = sB.l;
                J.voff = P.voff[d];
                J.ivoff = B.iv0 + (long long)jl * (S.ncap + sA.l + sB.l + 1);
                S.jobs[B.job0 + jl] = J;
                for (int ia = 0; ia <= sA.l; ++ia)
                    for (int ib = 0; ib <= sB.l; ++ib) {
                        const int cl = jl * ncomb + ia * (sB.l + 1) + ib;
                        VecOut o;
                        o.job = B.job0 + jl;
                        o.dco = B.dco0 + cl * dstride;
                        o.voff = J.voff + (ia * (sB.l + 1) + ib) * (N + 1);
                        o.K = ia + ib;
                        S.vo[B.vo0 + cl] = o;
                        sin_fourier_coefs(ia, ib, J.phA, J.phB, S.dco + o.dco);
                    }
            } else {
                const double cen = esum > 0.0 ? (e1 * sA.xyz[d] + e2 * sB.xyz[d]) / esum
                                              : 0.5 * (sA.xyz[d] + sB.xyz[d]);
                P.cen[d] = cen;
                fac *= exp(-mu * (sA.xyz[d] - sB.xyz[d]) * (sA.xyz[d] - sB.xyz[d]));
            }
        }
        P.fac = fac;
        P.eoff = B.herm0 + ij * nreal * ncomb * kE;
        double *E = S.herm + P.eoff;
        for (int x = 0; x < nreal * ncomb * kE; ++x) E[x] = 0.0;
        for (int dr = ndp; dr < 3; ++dr)
            for (int ia = 0; ia <= sA.l; ++ia)
                for (int ib = 0; ib <= sB.l; ++ib)
                    hermite_E(ia, ib, esum, P.cen[dr] - sA.xyz[dr], P.cen[dr] - sB.xyz[dr],
                              E + ((size_t)((dr - ndp) * (sA.l + 1) + ia) * (sB.l + 1) + ib) * kE);
        double ca = 0.0, cb = 0.0;
        for (int x = 0; x < sA.ncomp; ++x) ca = fmax(ca, fabs(S.coef[S.ao_poff[sA.first_ao + x] + i]));
        for (int x = 0; x < sB.ncomp; ++x) cb = fmax(cb, fabs(S.coef[S.ao_poff[sB.first_ao + x] + j]));
        P.cmax = ca * cb;
        P.ps = 1.0;
        S.pp[B.first + ij] = P;
    }
}

hipError_t launch_pair_setup(const PairBlock *blk, int nblk, const Shell *sh, const double *expo, const double *coef,
                             const int *ao_poff, int ndp, int ncap, const double L[3], const double a[3], PrimPair *pp,
                             VecJob *jobs, VecOut *vo, double2 *dco, double *herm, hipStream_t st) {
    if (nblk <= 0) return hipSuccess;
    PairSetupArgs S{blk, sh, expo, coef, ao_poff, ndp, ncap, {L[0], L[1], L[2]}, {a[0], a[1], a[2]}, pp, jobs, vo,
                    dco, herm};
    hipLaunchKernelGGL(k_pair_setup, dim3(nblk), dim3(64), 0, st, S);
    return hipGetLastError();
}

// V_m = sum_k d_k exp(-A) I_|m-k|(A), m = 0..N: one block per vector, the
// orders m across the threads.  d_k (Fourier coefficients of
// sin^ia(u+phA) sin^ib(u+phB), k = -K..K) from the pair set-up.
__global__ void k_pair_combine(const VecOut *__restrict__ vo, const VecJob *__restrict__ jobs,
                               const double *__restrict__ ivbuf, const double2 *__restrict__ dco,
                               double2 *__restrict__ vec) {
    const VecOut O = vo[blockIdx.x];
    const VecJob J = jobs[O.job];
    const double *iv = ivbuf + J.ivoff;
    const double2 *d = dco + O.dco;
    const int K = O.K;
    double2 *V = vec + O.voff;
    for (int m = threadIdx.x; m <= J.N; m += blockDim.x) {
        double sr = 0.0, si = 0.0;
        for (int k = -K; k <= K; ++k) {
            const double b = iv[abs(m - k)];
            sr += d[k + K].x * b;
            si += d[k + K].y * b;
        }
        V[m] = double2{sr, si};
    }
}

hipError_t launch_pair_vectors(const VecJob *jobs, int njobs, const VecOut *vo, int nvo, const double2 *dco,
                               double *ivbuf, double2 *vec, hipStream_t st) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bessel_jobs, dim3(njobs < 65536 ? njobs : 65536), dim3(64), 0, st, jobs, njobs, ivbuf);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nvo <= 0) return e;
    hipLaunchKernelGGL(k_pair_combine, dim3(nvo), dim3(128), 0, st, vo, jobs, ivbuf, dco, vec);
    return hipGetLastError();
}

// Schwarz tables from the diagonal blocks (ab|ab) of one pass
// (src/Toroidal/ERI_toroidal.f90:200-236 and :300-338).  Quartets [0, nshell)
// are shell blocks (SI in the unit cell, SJ anywhere): Q(i + R, j + R) =
// sqrt|(ij|ij)| for every lattice translation R, each entry written once from
// the pair whose first function lies in the unit cell.  The rest are one
// primitive pair each: ps of that pair = max over components.  One workgroup
// of 64 per quartet.
__global__ __launch_bounds__(64) void k_schwarz_q(const Quartet *__restrict__ q4, int nshell, const Shell *__restrict__ sh,
                                                  const ShellPair *__restrict__ sp, const double *__restrict__ blk,
                                                  const int *__restrict__ relc, int nf, int nb, int nc,
                                                  double *__restrict__ Q, PrimPair *__restrict__ pp) {
    const int x = blockIdx.x;
    const Quartet q = q4[x];
    const Shell A = sh[q.s[0]], B = sh[q.s[1]];
    const int nI = A.ncomp, nJ = B.ncomp;
    const double *b = blk + q.out;
    if (x >= nshell) {
        double m = 0.0;
        for (int t = threadIdx.x; t < nI * nJ; t += 64) {
            const int a = t % nI, bb = t / nI;
            m = fmax(m, sqrt(fabs(b[a + nI * (bb + nJ * (a + nI * bb))])));
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
        if (threadIdx.x == 0) pp[sp[q.bra].first + q.prim].ps = m;
        return;
    }
    for (int t = threadIdx.x; t < nI * nJ * nc; t += 64) {
        const int ab = t % (nI * nJ), ctr = t / (nI * nJ);
        const int a = ab % nI, bb = ab / nI;
        const double v = sqrt(fabs(b[a + nI * (bb + nJ * (a + nI * bb))]));
        const int i = A.first_ao + a, j = B.first_ao + bb;
        const int li = i % nf, cj = j / nf, lj = j % nf;
        const int r1 = ctr * nf + li;                                   // i moved to cell ctr
        const int r2 = relc[(size_t)relc[(size_t)ctr * nc] * nc + cj] * nf + lj;   // j moved with it
        Q[(size_t)r1 + (size_t)nb * r2] = v;
    }
}

// Qsh(S1, S2) = max of Q over the two shells' components; *qmax_bits = the
// largest Qsh as the bit pattern of a non-negative double (ordered as integers).
__global__ __launch_bounds__(64) void k_schwarz_sh(const Shell *__restrict__ sh, int nsh, int nb, const double *__restrict__ Q,
                                                   double *__restrict__ Qsh, unsigned long long *__restrict__ qmax_bits) {
    const long long t = (long long)blockIdx.x * 64 + threadIdx.x;
    double m = 0.0;
    if (t < (long long)nsh * nsh) {
        const int S1 = (int)(t % nsh), S2 = (int)(t / nsh);
        const Shell A = sh[S1], B = sh[S2];
        for (int b = 0; b < B.ncomp; ++b)
            for (int a = 0; a < A.ncomp; ++a) m = fmax(m, Q[(size_t)(A.first_ao + a) + (size_t)nb * (B.first_ao + b)]);
        Qsh[t] = m;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if (threadIdx.x == 0) atomicMax(qmax_bits, (unsigned long long)__double_as_longlong(m));
}

hipError_t launch_schwarz_tables(const Quartet *q4, int nquart, int nshell, const Shell *sh, const ShellPair *sp,
                                 const double *blk, const int *relc, int nf, int nb, int nc, int nsh, double *Q,
                                 double *Qsh, PrimPair *pp, unsigned long long *qmax_bits, hipStream_t st) {
    if (nquart <= 0 || nsh <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_schwarz_q, dim3(nquart), dim3(64), 0, st, q4, nshell, sh, sp, blk, relc, nf, nb, nc, Q, pp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const long long n = (long long)nsh * nsh;
    hipLaunchKernelGGL(k_schwarz_sh, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, sh, nsh, nb, Q, Qsh, qmax_bits);
    return hipGetLastError();
}

// Unique shell quartets (src/Toroidal/ERI_toroidal.f90:342-401): SI in the
// unit cell, the local indices of SJ, SK, SL not below SI, both Schwarz
// screens, and (SI,SJ,SK,SL) the smallest member of its orbit under the 8
// index permutations x the lattice translations.  One thread per candidate
// (SI, SJ, SK, SL); pass 0 counts (one atomic per wavefront), pass 1 appends.
// One thread per (SI, SJ, SK) and run of kSLRun consecutive SL: the unique
// quartets (SI in the unit cell, every member's local index >= SI, Schwarz
// Qsh(SI,SJ) Qsh(SK,SL) >= 1e-13, the smallest member of its orbit under the 8
// permutations with the first index translated to the unit cell), written in
// no particular order.  Positions: a wavefront-wide scan of the lanes' counts,
// then per workgroup the wavefronts' totals meet in LDS, one atomic per
// workgroup.
constexpr int kSLRun = 4;

__global__ __launch_bounds__(1024) void k_enum_quartets(const Shell *__restrict__ sh, const int *__restrict__ relc,
                                const double *__restrict__ Qsh, const unsigned long long *__restrict__ qmax_bits,
                                int nshuc, int nsh, int ncell,
                                unsigned long long *__restrict__ count, Quartet *__restrict__ out,
                                unsigned long long cap) {
    const int nrun = (nsh + kSLRun - 1) / kSLRun;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < (long long)nshuc * nsh * nsh * nrun;
    int SI = 0, SJ = 0, SK = 0, SL0 = 0;
    bool pre = live;
    if (pre) {
        SL0 = (int)(t % nrun) * kSLRun;
        SK = (int)((t / nrun) % nsh);
        SJ = (int)((t / ((long long)nrun * nsh)) % nsh);
        SI = (int)(t / ((long long)nrun * nsh * nsh));
        pre = sh[SJ].loc >= SI && sh[SK].loc >= SI;
    }
    double qij = 0.0;
    if (pre) {
        const double Qmax = __longlong_as_double((long long)*qmax_bits);
        qij = Qsh[(size_t)SI + (size_t)nsh * SJ];
        pre = qij * Qmax >= kSchwarz;
    }
    unsigned keepm = 0;
    if (pre) {
        const int cJ = sh[SJ].cell, cK = sh[SK].cell;
        const int lJ = sh[SJ].loc, lK = sh[SK].loc;
        for (int u = 0; u < kSLRun; ++u) {
            const int SL = SL0 + u;
            if (SL >= nsh) break;
            if (sh[SL].loc < SI || qij * Qsh[(size_t)SK + (size_t)nsh * SL] < kSchwarz) continue;
            const int q[4] = {SI, SJ, SK, SL};
            const int qc[4] = {0, cJ, cK, sh[SL].cell};
            const int ql[4] = {SI, lJ, lK, sh[SL].loc};
            bool keep = true;
            for (int ip = 1; ip < 8 && keep; ++ip) {
                const int c0 = qc[c_perm8_setup[ip][0]];
                for (int x = 0; x < 4; ++x) {
                    const int y = c_perm8_setup[ip][x];
                    const int rr = relc[(size_t)c0 * ncell + qc[y]] * nshuc + ql[y];
                    if (rr < q[x]) { keep = false; break; }
                    if (rr > q[x]) break;
                }
            }
            if (keep) keepm |= 1u << u;
        }
    }
    // this lane's count, its offset inside the wavefront (inclusive scan), the wavefront total
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const int mine = __popc(keepm);
    int incl = mine;
#pragma unroll
    for (int dl = 1; dl < 64; dl <<= 1) {
        const int v = __shfl_up(incl, dl, 64);
        if (lane >= dl) incl += v;
    }
    __shared__ unsigned wcnt[16];
    __shared__ unsigned long long wbase[16];
    if (lane == 63) wcnt[wv] = (unsigned)incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int w = 0; w < nwv; ++w) {
            wbase[w] = tot;
            tot += wcnt[w];
        }
        const unsigned long long b0 = tot ? atomicAdd(count, tot) : 0ull;
        for (int w = 0; w < nwv; ++w) wbase[w] += b0;
    }
    __syncthreads();
    if (!keepm || !out) return;
    unsigned long long pos = wbase[wv] + (unsigned)(incl - mine);
    const int lK = sh[SK].loc;
    for (int u = 0; u < kSLRun; ++u) {
        if (!(keepm & (1u << u))) continue;
        if (pos >= cap) return;
        const int SL = SL0 + u;
        Quartet Q;
        Q.s[0] = SI; Q.s[1] = SJ; Q.s[2] = SK; Q.s[3] = SL;
        Q.bra = SI * nsh + SJ;
        Q.ket = lK * nsh + relc[(size_t)sh[SK].cell * ncell + sh[SL].cell] * nshuc + sh[SL].loc;
        for (int d = 0; d < 3; ++d) Q.shift[d] = sh[SK].xyz[d] - sh[lK].xyz[d];
        Q.out = -1;
        Q.prim = -1;
        out[pos++] = Q;
    }
}

hipError_t launch_enum_quartets(const Shell *sh, const int *relc, const double *Qsh, const unsigned long long *qmax_bits,
                                int nshuc,
                                int nsh, int ncell, unsigned long long *count, Quartet *out,
                                unsigned long long cap, hipStream_t st) {
    const long long n = (long long)nshuc * nsh * nsh * ((nsh + kSLRun - 1) / kSLRun);
    const long long blocks = (n + 1023) / 1024;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_enum_quartets, dim3((unsigned)blocks), dim3(1024), 0, st, sh, relc, Qsh, qmax_bits, nshuc, nsh,
                       ncell, count, out, cap);
    return hipGetLastError();
}

}  // namespace tcg
