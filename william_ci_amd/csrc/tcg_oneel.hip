// Nuclear attraction matrix on the GPU (rows of the unit cell).
//
// Same t-quadrature as the reference's nuclear_attraction_integral_toroidal_nD
// (src/Toroidal/1D/1D_Nuclear_attraction_toroidal.f90:26-158):
//   V_ij = - sum_C Z_C  pref  sum_k w_k  prod_{d periodic} Ix_d(t_k)  prod_{d real} e_d s_d(t_k)
//   Ix_d(t_k) = sum_n F_n T_d[n][k],   F_n = Re[ e^{i n a_d (P_d - C_d)} G_n ]
// with G_n the bra pair Fourier vector (complex conjugate of the stored sign +1
// vector).  Sixteen nuclei are one MFMA row tile, so each F row costs one
// complex rotation per order and the node sums run on v_mfma_f64_16x16x4f64
// like the ERI kernel.
#include <hip/hip_runtime.h>

#include "tcg_device.hpp"

namespace tcg {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// v[d] of a by-value kernel argument array without dynamic indexing (which
// would copy the argument block to scratch memory)
__device__ __forceinline__ double sel3(const double v[3], int d) { return d == 0 ? v[0] : (d == 1 ? v[1] : v[2]); }

__device__ double overlap_real_1d(int j, int m, double p, double pa, double pb) {
    // sum_{k,q} C(j,k) C(m,q) pa^(j-k) pb^(m-q) <x^(k+q)>,  <x^n> = (n-1)!!/(2p)^(n/2) sqrt(pi/p)
    const double base = sqrt(3.14159265358979323846 / p);
    double acc = 0.0;
    double cj = 1.0;
    for (int k = 0; k <= j; ++k) {
        double cm = 1.0;
        for (int q = 0; q <= m; ++q) {
            const int n = k + q;
            if ((n & 1) == 0) {
                double mom = base;
                for (int t = 1; t < n; t += 2) mom *= (double)t / (2.0 * p);
                acc += cj * cm * pow(pa, (double)(j - k)) * pow(pb, (double)(m - q)) * mom;
            }
            cm = cm * (double)(m - q) / (double)(q + 1);
        }
        cj = cj * (double)(j - k) / (double)(k + 1);
    }
    return acc;
}

// One wavefront per (AO i of the unit cell, AO j, primitive pair p, group of 16
// nuclei); the blocks of one (i, j) add into V(i, j).  The order loop follows
// the row kernel: node tiles as slots sorted by their order limit, stages of
// unconditional MFMAs, two register sets loaded two steps ahead, the phase by
// rotation with a second rotation per 64 orders.
__global__ __launch_bounds__(64) void k_nuclear(EriArgs A, const double *__restrict__ nuc_xyz, const int *__restrict__ nuc_z,
                                                int natom, const int *__restrict__ ao_shell, const int *__restrict__ pair_rep,
                                                int ngroups, double *__restrict__ Vout) {
    extern __shared__ double Ix[];      // [3][16][nqp]
    __shared__ double part[64];
    const int lane = threadIdx.x, r = lane & 15, kq = lane >> 4;
    const int nf = A.nfuc, nb = A.nbas, nq = A.nq, nqp = A.nqp, ndp = A.ndp;
    const int nh = nqp / kNodeTile;
    const int tabrows = A.ncap + 5, tlast = tabrows - 1;
    const int npairs = nf * nb;
    const int pidx = blockIdx.x % npairs;        // i + nf * j
    const int chunk = blockIdx.x / npairs;
    const int p = chunk / ngroups, c0 = (chunk % ngroups) * 16;
    const int i = pidx % nf, j = pidx / nf;
    const ShellPair sp = A.sp[pair_rep[pidx]];
    if (p >= sp.count) return;
    const Shell &sA = A.sh[ao_shell[i]], &sB = A.sh[ao_shell[j]];
    const int lB = sB.l;
    const int *li = A.ao_lxyz + 3 * i, *lj = A.ao_lxyz + 3 * j;
    const PrimPair &pp = A.pp[sp.first + p];
    const double ci = A.coef[A.ao_poff[i] + pp.pi], cj = A.coef[A.ao_poff[j] + pp.pj];
    double pref = ci * cj * pp.fac * 2.0 / sqrt(3.14159265358979323846);
    for (int d = 0; d < ndp; ++d) pref *= pow(sel3(A.a, d), -(double)(li[d] + lj[d])) * sel3(A.L, d);
    const double esum = pp.q;
    const int c = c0 + r;
    const bool cvalid = c < natom;
    for (int d = 0; d < ndp; ++d) {
        const int N = pp.N[d];
        const double2 *Vp = A.vec + pp.voff[d] + (size_t)(li[d] * (lB + 1) + lj[d]) * (N + 1);
        const double phi = cvalid ? sel3(A.a, d) * (pp.cen[d] - nuc_xyz[3 * c + d]) : 0.0;
        double s4, c4, s64, c64, sb0, cb0;
        sincos(4.0 * phi, &s4, &c4);
        sincos(64.0 * phi, &s64, &c64);
        sincos((double)kq * phi, &sb0, &cb0);
        const double *tabd = A.tab + (size_t)d * tabrows * nqp;
        auto gidx = [&](int n) { return n <= N ? n : 0; };
        auto tidx = [&](int n) { return (size_t)(n < tlast ? n : tlast) * nqp; };
        for (int hg = 0; hg < nh; hg += 4) {
            int sl[4], spn[4];
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                int lh = -1;
                if (hg + h < nh) {
                    lh = A.tab_lim[d * nh + hg + h];
                    lh = lh < N ? lh : N;
                }
                sl[h] = lh;
                spn[h] = h;
            }
#pragma unroll
            for (int a = 1; a < 4; ++a)
#pragma unroll
                for (int b = a; b > 0; --b) {
                    const bool sw = sl[b - 1] < sl[b];
                    const int x = sl[b - 1], y = sl[b], px = spn[b - 1], py = spn[b];
                    sl[b - 1] = sw ? y : x;
                    sl[b] = sw ? x : y;
                    spn[b - 1] = sw ? py : px;
                    spn[b] = sw ? px : py;
                }
            const double *tab = tabd + hg * kNodeTile + r;
            int toff[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) toff[s] = sl[s] >= 0 ? spn[s] * kNodeTile : 0;
            dbl4 acc[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) acc[s] = dbl4{0.0, 0.0, 0.0, 0.0};
            double cn = cb0, sn = sb0, cb = cb0, sb = sb0;
            double2 vP = Vp[gidx(kq)], vQ = Vp[gidx(4 + kq)];
            double bP[4], bQ[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                bP[s] = tab[tidx(kq) + toff[s]];
                bQ[s] = tab[tidx(4 + kq) + toff[s]];
            }
            auto step = [&](int m, int cact, double2 &v, double *bb) {
                const int n = m + kq;
                if (m > 0) {
                    if ((m & 63) == 0) {
                        const double tb = cb * c64 - sb * s64;
                        sb = sb * c64 + cb * s64;
                        cb = tb;
                        cn = cb;
                        sn = sb;
                    } else {
                        const double tc = cn * c4 - sn * s4;
                        sn = sn * c4 + cn * s4;
                        cn = tc;
                    }
                }
                const double F = (cvalid && n <= N) ? v.x * cn + v.y * sn : 0.0;   // Re[e^{in phi} conj(v)]
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    if (s < cact) acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(F, bb[s], acc[s], 0, 0, 0);
                const int n2 = n + 8;
                v = Vp[gidx(n2)];
                const size_t t2 = tidx(n2);
#pragma unroll
                for (int s = 0; s < 4; ++s) bb[s] = s < cact ? tab[t2 + toff[s]] : 0.0;
            };
            int n0 = 0;
#pragma unroll
            for (int cact = 4; cact >= 1; --cact) {
                const int lend = sl[cact - 1];
                for (; n0 <= lend; n0 += 8) {
                    step(n0, cact, vP, bP);
                    step(n0 + 4, cact, vQ, bQ);
                }
            }
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (hg + spn[s] < nh)
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            Ix[((size_t)d * 16 + 4 * v + kq) * nqp + (hg + spn[s]) * kNodeTile + r] = acc[s][v];
        }
    }
    __syncthreads();
    // per (nucleus, node): real directions, weights, product; summed per lane
    double lsum = 0.0;
    for (int task = lane; task < 16 * nq; task += 64) {
        const int cr = task / nq, k = task % nq, cc = c0 + cr;
        if (cc >= natom) continue;
        double prod = A.tweight[k];
        for (int d = 0; d < ndp; ++d) prod *= Ix[((size_t)d * 16 + cr) * nqp + k];
        const double t2 = A.tnode[k] * A.tnode[k];
        const double ppt = esum + t2, eta = t2 / ppt;
        for (int d = ndp; d < 3; ++d) {
            const double pc = pp.cen[d] - nuc_xyz[3 * cc + d];
            const double pa = pp.cen[d] - sA.xyz[d], pb = pp.cen[d] - sB.xyz[d];
            prod *= exp(-esum * eta * pc * pc) * overlap_real_1d(li[d], lj[d], ppt, pa - eta * pc, pb - eta * pc);
        }
        lsum -= (double)nuc_z[cc] * prod;
    }
    part[lane] = pref * lsum;
    __syncthreads();
    if (lane == 0) {
        double s = 0.0;
        for (int x = 0; x < 64; ++x) s += part[x];
        atomicAdd(&Vout[i + (size_t)nb * j], s);
    }
}

hipError_t launch_nuclear(const EriArgs &a, const double *nuc_xyz, const int *nuc_z, int natom, const int *ao_shell,
                          const int *pair_rep, int max_prim_pairs, double *V, hipStream_t st) {
    const int npairs = a.nfuc * a.nbas;
    const int ngroups = (natom + 15) / 16;
    const long long blocks = (long long)npairs * max_prim_pairs * ngroups;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    const size_t lds = sizeof(double) * 3 * 16 * (size_t)a.nqp;
    hipLaunchKernelGGL(k_nuclear, dim3((unsigned)blocks), dim3(64), lds, st, a, nuc_xyz, nuc_z, natom, ao_shell,
                       pair_rep, ngroups, V);
    return hipGetLastError();
}

}  // namespace tcg
