// gfx950 kernels of the toroidal ERI engine.
//
// k_eri_quartets: one wavefront per shell quartet (SI SJ|SK SL).  It follows
// the arithmetic of the reference kernel (src/Toroidal/ERI_shell_block_toroidal.f90:28-390)
// but organises it for CDNA4:
//
//   * Each surviving primitive quartet contributes, per periodic direction d
//     and per power combination (ia,ib,ic,id), one ROW
//         F_n = Re[ conj(V^bra_n) V^ket_n e^{i n phi} ],  n = 0..min(N_bra, N_ket)
//     where V are the pair Fourier vectors (sign +1 form; the bra vector of
//     the reference, built with sign -1, is its complex conjugate).
//   * The periodic factor at every quadrature node k is the contraction
//         I_d(t_k) = sum_n F_n T_d[n][k],  T_d[n][k] = c_n exp(-C_k) I_n(C_k)
//     (c_0 = 1, c_n = 2, zero beyond the node's truncation order).  Sixteen
//     rows x 16 nodes are one v_mfma_f64_16x16x4f64 tile per 4 orders n: the
//     lanes generate F on the VALU (complex products + phase rotation) while
//     the matrix core accumulates.  Node tiles whose truncation order is
//     exceeded stop early.
//   * Real directions use the McMurchie-Davidson form per node; the
//     component contraction sum_k w_k Ix Iy Iz and the coefficient products
//     are accumulated into the shell block in LDS.
//   * The block is then either written out (Schwarz / primitive-Schwarz
//     passes) or scattered with the 8 index permutations, each shifted so
//     that its first index lies in the unit cell
//     (src/Toroidal/ERI_toroidal.f90:416-454).
#include <hip/hip_runtime.h>

#include "tcg_device.hpp"

namespace tcg {

typedef double dbl4 __attribute__((ext_vector_type(4)));

__constant__ int c_perm8[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                  {2, 3, 0, 1}, {3, 2, 0, 1}, {2, 3, 1, 0}, {3, 2, 1, 0}};

constexpr int kChunk = 256;          // candidate primitive quartets per pass over the survivor list
constexpr int kMaxNodeTiles = kMaxNQ / kNodeTile;
constexpr int kMaxBlk = 6 * 6 * 6 * 6;  // (dd|dd)

__device__ __forceinline__ int wave_max_int(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

__global__ __launch_bounds__(64) void k_eri_quartets(EriArgs A) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const int r = lane & 15, kq = lane >> 4;
    const int cap = A.max_ncomb > 16 ? A.max_ncomb : 16;   // rows per direction in I
    const int nqp = A.nqp, nq = A.nq, ndp = A.ndp;
    const int nh = nqp / kNodeTile;
    const int tabrows = A.ncap + 5;

    for (int qi = blockIdx.x; qi < A.nq4; qi += gridDim.x) {
        const Quartet Q = A.q4[qi];
        const Shell sA = A.sh[Q.s[0]], sB = A.sh[Q.s[1]], sC = A.sh[Q.s[2]], sD = A.sh[Q.s[3]];
        const int lA = sA.l, lB = sB.l, lC = sC.l, lD = sD.l;
        const int nA = sA.ncomp, nB = sB.ncomp, nC = sC.ncomp, nD = sD.ncomp;
        const int ncomb = (lA + 1) * (lB + 1) * (lC + 1) * (lD + 1);
        const int nblk = nA * nB * nC * nD;
        const int NN = lA + lB + lC + lD;
        const ShellPair bp = A.sp[Q.bra], kp = A.sp[Q.ket];

        // LDS: I[3][cap][nqp] | blk[kMaxBlk] | sfac[kChunk] | sidx[kChunk][2] | msgn[16] | mdest[16]
        double *I = lds;
        double *blk = lds + (size_t)3 * cap * nqp;
        double *sfac = blk + kMaxBlk;
        int *sidx = (int *)(sfac + kChunk);
        double *msgn = sfac + 2 * kChunk;
        int *mdest = (int *)(msgn + 16);

        for (int i = lane; i < nblk; i += 64) blk[i] = 0.0;

        const long ncand = (Q.prim >= 0) ? 1 : (long)bp.count * kp.count;
        for (long c0 = 0; c0 < ncand; c0 += kChunk) {
            __syncthreads();
            // ---- survivors of this chunk (kernel prefactor + primitive Schwarz screens) ----
            int nsurv = 0;
            const long cend = (c0 + kChunk < ncand) ? c0 + kChunk : ncand;
            for (long base = c0; base < cend; base += 64) {
                long c = base + lane;
                bool ok = c < cend;
                int ij = 0, kl = 0;
                double fac = 0.0;
                if (ok) {
                    if (Q.prim >= 0) { ij = Q.prim; kl = Q.prim; }
                    else { ij = (int)(c / kp.count); kl = (int)(c % kp.count); }
                    const PrimPair &pb = A.pp[bp.first + ij];
                    const PrimPair &pk = A.pp[kp.first + kl];
                    fac = (A.fac_const * pb.fac) * pk.fac;
                    if (A.screen_prefactor && fabs(fac) * pb.cmax * pk.cmax <= kScreenTol) ok = false;
                    if (A.screen_prim && pb.ps * pk.ps < kPrimTol) ok = false;
                }
                unsigned long long m = __ballot(ok);
                int pos = nsurv + __popcll(m & ((1ull << lane) - 1ull));
                if (ok) {
                    sfac[pos] = fac;
                    sidx[2 * pos] = ij;
                    sidx[2 * pos + 1] = kl;
                }
                nsurv += __popcll(m);
            }
            __syncthreads();
            if (nsurv == 0) continue;
            if (lane == 0 && A.counters) {
                atomicAdd(&A.counters[0], (unsigned long long)nsurv);
                atomicAdd(&A.counters[1], (unsigned long long)nsurv * ncomb * ndp);
            }

            const int R = ncomb;                       // rows per group and direction
            const int gpt = (R <= 16) ? 16 / R : 1;    // groups per tile
            const int tpd = (R <= 16) ? 1 : (R + 15) / 16;
            const int ngs = (nsurv + gpt - 1) / gpt;
            for (int gs = 0; gs < ngs; ++gs) {
                const int g0 = gs * gpt;
                const int ng = (nsurv - g0 < gpt) ? nsurv - g0 : gpt;
                // ---------------- periodic directions: MFMA tiles ----------------
                for (int d = 0; d < ndp; ++d) {
                    for (int tt = 0; tt < tpd; ++tt) {
                        int gl, rr;
                        bool valid;
                        if (R <= 16) { gl = r / R; rr = r % R; valid = (gl < ng); }
                        else { gl = 0; rr = tt * 16 + r; valid = rr < R; }
                        const int combo = valid ? rr : 0;
                        int t = combo;
                        const int ia = t % (lA + 1); t /= (lA + 1);
                        const int ib = t % (lB + 1); t /= (lB + 1);
                        const int ic = t % (lC + 1);
                        const int id = t / (lC + 1);
                        int nm = -1;
                        const double2 *Gp = A.vec, *Hp = A.vec;
                        double phi = 0.0, sgn = 1.0;
                        if (valid) {
                            const int g = g0 + gl;
                            const PrimPair &pb = A.pp[bp.first + sidx[2 * g]];
                            const PrimPair &pk = A.pp[kp.first + sidx[2 * g + 1]];
                            const int bN = pb.N[d], kN = pk.N[d];
                            nm = bN < kN ? bN : kN;
                            Gp = A.vec + pb.voff[d] + (size_t)(ia * (lB + 1) + ib) * (bN + 1);
                            Hp = A.vec + pk.voff[d] + (size_t)(ic * (lD + 1) + id) * (kN + 1);
                            phi = A.a[d] * (pb.cen[d] - (pk.cen[d] + Q.shift[d]));
                            sgn = pow(A.a[d], -(double)(ia + ib + ic + id));
                        }
                        if (kq == 0) {
                            msgn[r] = sgn;
                            mdest[r] = valid ? gl * ncomb + combo : -1;
                        }
                        if (A.counters) {
                            // algorithmic flops of this row: the truncated node sums + the F products
                            unsigned long long fl = 0;
                            if (kq == 0 && valid) {
                                long s = 0;
                                for (int k = 0; k < nq; ++k) {
                                    const int tk = A.tnmax[d * nqp + k];
                                    s += (nm < tk ? nm : tk) + 1;
                                }
                                fl = (unsigned long long)(2 * s + 10 * (nm + 1));
                            }
                            for (int off = 32; off > 0; off >>= 1) fl += __shfl_xor(fl, off, 64);
                            if (lane == 0) atomicAdd(&A.counters[2], fl);
                        }
                        const int nmax_t = wave_max_int(nm);
                        int lim[kMaxNodeTiles];
                        int nend = -1;
                        for (int h = 0; h < nh; ++h) {
                            int lh = A.tab_lim[d * nh + h];
                            lim[h] = lh < nmax_t ? lh : nmax_t;
                            nend = lim[h] > nend ? lim[h] : nend;
                        }
                        dbl4 acc[kMaxNodeTiles];
                        for (int h = 0; h < kMaxNodeTiles; ++h) acc[h] = dbl4{0.0, 0.0, 0.0, 0.0};
                        double s4, c4, cn, sn;
                        sincos(4.0 * phi, &s4, &c4);
                        sincos((double)kq * phi, &sn, &cn);
                        const double *tab = A.tab + (size_t)d * tabrows * nqp;
                        for (int n0 = 0; n0 <= nend; n0 += 4) {
                            const int n = n0 + kq;
                            if (n0 > 0) {
                                if ((n0 & 63) == 0) {
                                    sincos((double)n * phi, &sn, &cn);
                                } else {
                                    double tc = cn * c4 - sn * s4;
                                    sn = sn * c4 + cn * s4;
                                    cn = tc;
                                }
                            }
                            double F = 0.0;
                            if (n <= nm) {
                                const double2 g = Gp[n], hv = Hp[n];
                                const double re = g.x * hv.x + g.y * hv.y;   // conj(g) * h
                                const double im = g.x * hv.y - g.y * hv.x;
                                F = re * cn - im * sn;
                            }
                            const double *trow = tab + (size_t)n * nqp + r;
#pragma unroll
                            for (int h = 0; h < kMaxNodeTiles; ++h) {
                                if (h < nh && n0 <= lim[h]) {
                                    const double b = trow[h * kNodeTile];
                                    acc[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(F, b, acc[h], 0, 0, 0);
                                }
                            }
                        }
                        __syncthreads();
                        // v_mfma_f64_16x16x4f64 result: lane (r, kq) holds D[row = 4*v + kq][col = r]
#pragma unroll
                        for (int h = 0; h < kMaxNodeTiles; ++h) {
                            if (h < nh) {
#pragma unroll
                                for (int v = 0; v < 4; ++v) {
                                    const int row = 4 * v + kq;
                                    const int dst = mdest[row];
                                    if (dst >= 0) I[((size_t)d * cap + dst) * nqp + h * kNodeTile + r] = acc[h][v] * msgn[row];
                                }
                            }
                        }
                        __syncthreads();
                    }
                }
                // ---------------- real directions: Hermite form per node ----------------
                for (int dr = ndp; dr < 3; ++dr) {
                    const int drr = dr - ndp;
                    for (int task = lane; task < ng * nq; task += 64) {
                        const int gl = task / nq, k = task % nq;
                        const int g = g0 + gl;
                        const PrimPair &pb = A.pp[bp.first + sidx[2 * g]];
                        const PrimPair &pk = A.pp[kp.first + sidx[2 * g + 1]];
                        const double p = pb.q, q = pk.q;
                        const double tk = A.tnode[k], t2 = tk * tk;
                        const double D2 = 1.0 / (p * q + (p + q) * t2);
                        const double w = p * q * t2 * D2;
                        const double rs = 3.14159265358979323846 * sqrt(D2);
                        const double Rr = pb.cen[dr] - (pk.cen[dr] + Q.shift[dr]);
                        double gt[4 * kMaxL + 1];
                        gt[0] = (Rr == 0.0) ? rs : rs * exp(-w * Rr * Rr);
                        double gm1 = 0.0, gc = gt[0];
                        for (int n = 0; n < NN; ++n) {
                            const double gp1 = 2.0 * w * Rr * gc - 2.0 * (double)n * w * gm1;
                            gt[n + 1] = gp1;
                            gm1 = gc;
                            gc = gp1;
                        }
                        const double *Eb = A.herm + pb.eoff + (size_t)drr * (lA + 1) * (lB + 1) * (2 * kMaxL + 1);
                        const double *Ek = A.herm + pk.eoff + (size_t)drr * (lC + 1) * (lD + 1) * (2 * kMaxL + 1);
                        for (int combo = 0; combo < ncomb; ++combo) {
                            int t = combo;
                            const int ia = t % (lA + 1); t /= (lA + 1);
                            const int ib = t % (lB + 1); t /= (lB + 1);
                            const int ic = t % (lC + 1);
                            const int id = t / (lC + 1);
                            const double *eb = Eb + (ia * (lB + 1) + ib) * (2 * kMaxL + 1);
                            const double *ek = Ek + (ic * (lD + 1) + id) * (2 * kMaxL + 1);
                            double SN[4 * kMaxL + 1];
                            for (int n = 0; n <= ia + ib + ic + id; ++n) SN[n] = 0.0;
                            for (int n = 0; n <= ia + ib; ++n)
                                for (int m = 0; m <= ic + id; ++m)
                                    SN[n + m] += ((n & 1) ? -1.0 : 1.0) * eb[n] * ek[m];
                            double v = 0.0;
                            for (int n = 0; n <= ia + ib + ic + id; ++n) v += SN[n] * gt[n];
                            I[((size_t)dr * cap + gl * ncomb + combo) * nqp + k] = v;
                        }
                    }
                }
                __syncthreads();
                if (A.counters && lane == 0)
                    atomicAdd(&A.counters[2], (unsigned long long)ng *
                                                  ((unsigned long long)ncomb * (3 - ndp) * 2 * nq * (NN + 1) +
                                                   3ull * nq * nblk));
                // ---------------- component contraction into the block ----------------
                for (int task = lane; task < ng * nblk; task += 64) {
                    const int gl = task / nblk, comp = task % nblk;
                    const int g = g0 + gl;
                    int t = comp;
                    const int ca = t % nA; t /= nA;
                    const int cb = t % nB; t /= nB;
                    const int cc = t % nC;
                    const int cd = t / nC;
                    const int *pa = A.ao_lxyz + 3 * (sA.first_ao + ca);
                    const int *pbw = A.ao_lxyz + 3 * (sB.first_ao + cb);
                    const int *pc = A.ao_lxyz + 3 * (sC.first_ao + cc);
                    const int *pd = A.ao_lxyz + 3 * (sD.first_ao + cd);
                    int cx[3];
                    for (int dd = 0; dd < 3; ++dd)
                        cx[dd] = pa[dd] + (lA + 1) * (pbw[dd] + (lB + 1) * (pc[dd] + (lC + 1) * pd[dd]));
                    const double *Ix = I + ((size_t)0 * cap + gl * ncomb + cx[0]) * nqp;
                    const double *Iy = I + ((size_t)1 * cap + gl * ncomb + cx[1]) * nqp;
                    const double *Iz = I + ((size_t)2 * cap + gl * ncomb + cx[2]) * nqp;
                    double val = 0.0;
                    for (int k = 0; k < nq; ++k) val += (Ix[k] * A.tweight[k]) * Iy[k] * Iz[k];
                    const PrimPair &pb = A.pp[bp.first + sidx[2 * g]];
                    const PrimPair &pk = A.pp[kp.first + sidx[2 * g + 1]];
                    const double cA = A.coef[A.ao_poff[sA.first_ao + ca] + pb.pi];
                    const double cB = A.coef[A.ao_poff[sB.first_ao + cb] + pb.pj];
                    const double cC = A.coef[A.ao_poff[sC.first_ao + cc] + pk.pi];
                    const double cD = A.coef[A.ao_poff[sD.first_ao + cd] + pk.pj];
                    atomicAdd(&blk[comp], sfac[g] * cA * cB * cC * cD * val);
                }
                __syncthreads();
            }
        }
        __syncthreads();
        // ---------------- output ----------------
        if (A.blk_out && Q.out >= 0) {
            for (int i = lane; i < nblk; i += 64) A.blk_out[Q.out + i] = blk[i];
        } else if (A.eri) {
            const size_t nb = A.nbas, n1 = A.n1;
            for (int task = lane; task < nblk * 8; task += 64) {
                const int comp = task >> 3, ip = task & 7;
                int t = comp;
                const int ca = t % nA; t /= nA;
                const int cb = t % nB; t /= nB;
                const int cc = t % nC;
                const int cd = t / nC;
                const int f4[4] = {sA.first_ao + ca, sB.first_ao + cb, sC.first_ao + cc, sD.first_ao + cd};
                if (A.Qao[(size_t)f4[0] + nb * f4[1]] * A.Qao[(size_t)f4[2] + nb * f4[3]] < kSchwarz) continue;
                int t4[4], pp[4];
                for (int x = 0; x < 4; ++x) t4[x] = f4[c_perm8[ip][x]];
                const int c0 = t4[0] / A.nfuc;
                for (int x = 0; x < 4; ++x) pp[x] = A.relc[c0 * A.ncell + t4[x] / A.nfuc] * A.nfuc + t4[x] % A.nfuc;
                const size_t idx = A.physicist
                    ? pp[0] + n1 * (pp[2] + nb * (pp[1] + nb * (size_t)pp[3]))
                    : pp[0] + n1 * (pp[1] + nb * (pp[2] + nb * (size_t)pp[3]));
                A.eri[idx] = blk[comp];
            }
        }
    }
}

size_t eri_lds_bytes(const EriArgs &a) {
    const int cap = a.max_ncomb > 16 ? a.max_ncomb : 16;
    return sizeof(double) * ((size_t)3 * cap * a.nqp + kMaxBlk + 2 * kChunk + 32);
}

hipError_t launch_eri_quartets(const EriArgs &a, hipStream_t st) {
    if (a.nq4 <= 0) return hipSuccess;
    const size_t lds = eri_lds_bytes(a);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_eri_quartets, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
    }
    const int grid = a.nq4 < (1 << 22) ? a.nq4 : (1 << 22);
    hipLaunchKernelGGL(k_eri_quartets, dim3(grid), dim3(64), lds, st, a);
    return hipGetLastError();
}

__global__ void k_fill(double *p, size_t n, double v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
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

}  // namespace tcg
