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
#include "tcg_math.hpp"

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


// ---------------------------------------------------------------------------
// Overlap and kinetic rows and the AO normalisation (closed form per direction):
//   overlap_matrix_toroidal_nD / kinetic_matrix_toroidal_nD
//       e.g. src/Toroidal/1D/1D_Overlap_toroidal.f90:23, 1D_Kinetic_toroidal.f90:102
//   normalize_AO_toroidal_nD   src/basis/normalization/normalize_1D.f90:17-79
// A periodic direction of a primitive pair (e1 at A, e2 at B) is the product of
// exp(-2(e1+e2-pe)/a^2) and sum_k d_k exp(-Ar) I_|k|(Ar) L a^-(i+l), with
// Ar = 2 pe / a^2, pe and the barycentre from clifford_pair, d_k the Fourier
// coefficients of sin^i(u+phA) sin^l(u+phB); a real direction is the Hermite E_0
// of the Gaussian product times sqrt(pi/p).  The kinetic factor of a direction
// applies d^2/dx^2 to the ket: on the torus l(l-1) S(i,l-2) - l^2 a^2 S(i,l)
// - 2 e2 (2l+1) S_cos(i,l) + 4 e2^2 S(i,l+2) (S_cos: the extra cos(a(x-B))).
// One thread per element; the orders are small (lA + lB + 3), so every element
// runs Miller's recurrence for its primitive pairs itself.
// ---------------------------------------------------------------------------
namespace {

constexpr int kOneN = 2 * kMaxL + 3;        // highest Bessel order a kinetic factor reads
constexpr int kOneK = 2 * kMaxL + 2;        // highest sin-product degree (ket raised by 2)

// exp(-A) I_n(A), n = 0..N, Miller's backward recurrence from n0 with two running
// values, rescaled against overflow, normalised to exp(-A) I_0(A)
// (tcg_math.hpp::bessel_miller with the start of bessel_scaled_vector)
__device__ void bessel_dev(int N, double A, double *out) {
    for (int n = 0; n <= N; ++n) out[n] = 0.0;
    if (A < 1e-12) {
        out[0] = 1.0;
        return;
    }
    const int n0 = N + 30 + (int)sqrt(80.0 * A);
    double wk1 = 0.0, wk = 1e-280;              // w[k+1], w[k]
    if (n0 <= N) out[n0] = wk;
    const double r = 2.0 / A;
    for (int k = n0; k >= 1; --k) {
        double wm = wk1 + (double)k * r * wk;
        if (k - 1 <= N) out[k - 1] = wm;
        if (wm > 1e250) {
            wm *= 1e-250;
            wk *= 1e-250;
            for (int m = k - 1; m <= N; ++m) out[m] *= 1e-250;
        }
        wk1 = wk;
        wk = wm;
    }
    const double sc = i0_scaled(A) / out[0];
    for (int n = 0; n <= N; ++n) out[n] *= sc;
}

// d_k, k = -K..K, of sin^i(u+phA) sin^l(u+phB), K = i + l <= kOneK
__device__ void sin_fourier_dev(int i, int l, double phA, double phB, double2 *d) {
    const int K = i + l;
    double2 cur[2 * kOneK + 3], nxt[2 * kOneK + 3];
    for (int k = 0; k < 2 * K + 3; ++k) cur[k] = double2{0.0, 0.0};
    cur[K + 1] = double2{1.0, 0.0};
    const double sA = sin(phA), cA = cos(phA), sB = sin(phB), cB = cos(phB);
    for (int m = 1; m <= K; ++m) {
        const double sn = (m <= i) ? sA : sB, cs = (m <= i) ? cA : cB;
        const double pr = 0.5 * sn, pi = -0.5 * cs;        // e^{i ph}/(2i)
        const double mr = 0.5 * sn, mi = 0.5 * cs;         // -e^{-i ph}/(2i)
        for (int k = 0; k < 2 * K + 3; ++k) nxt[k] = double2{0.0, 0.0};
        for (int k = -(m - 1); k <= m - 1; ++k) {
            const double2 v = cur[k + K + 1];
            double2 &up = nxt[k + 1 + K + 1], &dn = nxt[k - 1 + K + 1];
            up.x += pr * v.x - pi * v.y;
            up.y += pr * v.y + pi * v.x;
            dn.x += mr * v.x - mi * v.y;
            dn.y += mr * v.y + mi * v.x;
        }
        for (int k = 0; k < 2 * K + 3; ++k) cur[k] = nxt[k];
    }
    for (int k = -K; k <= K; ++k) d[k + K] = cur[k + K + 1];
}

__device__ double torus_factor(int i, int l, double phA, double phB, double a, double L, const double *iv,
                               bool with_cos) {
    if (l < 0) return 0.0;
    const int K = i + l;
    double2 d[2 * kOneK + 1];
    sin_fourier_dev(i, l, phA, phB, d);
    double acc = 0.0;
    if (!with_cos) {
        for (int k = -K; k <= K; ++k) acc += d[k + K].x * iv[k < 0 ? -k : k];
    } else {
        // times cos(u + phB) = (e^{i phB} e^{iu} + e^{-i phB} e^{-iu}) / 2: real part
        const double cb = cos(phB), sb = sin(phB);
        for (int k = -K; k <= K; ++k) {
            const double2 v = d[k + K];
            const int kp = k + 1 < 0 ? -(k + 1) : k + 1, km = k - 1 < 0 ? -(k - 1) : k - 1;
            acc += 0.5 * (cb * v.x - sb * v.y) * iv[kp];
            acc += 0.5 * (cb * v.x + sb * v.y) * iv[km];
        }
    }
    return acc * L * pow(a, -(double)(i + l));
}

__device__ double real_factor(int j, int m, double p, double PA, double PB) {
    if (m < 0) return 0.0;
    double E[2 * 8 + 3];
    hermite_E(j, m, p, PA, PB, E);
    return E[0] * sqrt(kPi / p);
}

// overlap and kinetic of one primitive pair of AOs (powers li, lj; centres A, B)
__device__ void prim_st(const OneStArgs &g, const int *li, const int *lj, const double *A, const double *B, double e1,
                        double e2, bool kinetic, double *s_out, double *t_out) {
    const double esum = e1 + e2, mu = e1 * e2 / esum;
    double pref = 1.0, Sd[3], Dd[3];
    const int lt = li[0] + li[1] + li[2] + lj[0] + lj[1] + lj[2];
    for (int d = 0; d < 3; ++d) {
        const int l1 = li[d], l2 = lj[d];
        if (d < g.ndp) {
            const double ad = g.a[d], L = g.L[d];
            double pe, cen;
            clifford_pair(e1, e2, A[d], B[d], L, ad, &pe, &cen);
            pref *= exp(-2.0 * (esum - pe) / (ad * ad));
            const double Ar = 2.0 * pe / (ad * ad);
            double iv[kOneN + 1];
            bessel_dev(lt + 3 < kOneN ? lt + 3 : kOneN, Ar, iv);
            const double phA = ad * (cen - A[d]), phB = ad * (cen - B[d]);
            Sd[d] = torus_factor(l1, l2, phA, phB, ad, L, iv, false);
            if (kinetic) {
                double D = 0.0;
                if (l2 >= 2) D += (double)(l2 * (l2 - 1)) * torus_factor(l1, l2 - 2, phA, phB, ad, L, iv, false);
                if (l2 >= 1) D -= (double)(l2 * l2) * ad * ad * Sd[d];
                D -= 2.0 * e2 * (2 * l2 + 1) * torus_factor(l1, l2, phA, phB, ad, L, iv, true);
                D += 4.0 * e2 * e2 * torus_factor(l1, l2 + 2, phA, phB, ad, L, iv, false);
                Dd[d] = D;
            }
        } else {
            const double P = (e1 * A[d] + e2 * B[d]) / esum;
            pref *= exp(-mu * (A[d] - B[d]) * (A[d] - B[d]));
            const double PA = P - A[d], PB = P - B[d];
            Sd[d] = real_factor(l1, l2, esum, PA, PB);
            if (kinetic) {
                double D = -2.0 * e2 * (2 * l2 + 1) * Sd[d] + 4.0 * e2 * e2 * real_factor(l1, l2 + 2, esum, PA, PB);
                if (l2 >= 2) D += (double)(l2 * (l2 - 1)) * real_factor(l1, l2 - 2, esum, PA, PB);
                Dd[d] = D;
            }
        }
    }
    *s_out = pref * Sd[0] * Sd[1] * Sd[2];
    if (kinetic) *t_out = -0.5 * pref * (Dd[0] * Sd[1] * Sd[2] + Sd[0] * Dd[1] * Sd[2] + Sd[0] * Sd[1] * Dd[2]);
}

}  // namespace

// S(i, j), T(i, j) for i < nfuc, every j (column-major, leading dimension nbas), in
// two launches: one thread per primitive pair (p, q) of an AO pair (i, j) (every
// prim_st runs a Bessel recurrence of up to ~1000 terms on long tori, so a thread
// per AO pair left most of the GPU idle), then per AO pair the contraction sum in
// the order p, then q.  w: slot ((i + nf j) mp + p') mp + q' holds (S, T) of the
// pair's primitives p', q' (their positions in their AOs; mp = most primitives of
// an AO).
__global__ __launch_bounds__(256) void k_one_st_prim(OneStArgs g, int mp, double *__restrict__ w) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)g.nf * g.nb * mp * mp) return;
    const int qq = (int)(idx % mp), pp = (int)((idx / mp) % mp);
    const long long ij = idx / ((long long)mp * mp);
    const int i = (int)(ij % g.nf), j = (int)(ij / g.nf);
    const int p = g.poff[i] + pp, q = g.poff[j] + qq;
    if (p >= g.poff[i + 1] || q >= g.poff[j + 1]) return;
    double s, t;
    prim_st(g, g.lxyz + 3 * i, g.lxyz + 3 * j, g.center + 3 * i, g.center + 3 * j, g.expo[p], g.expo[q], true, &s,
            &t);
    w[2 * idx] = s;
    w[2 * idx + 1] = t;
}

__global__ __launch_bounds__(256) void k_one_st(OneStArgs g, int mp, const double *__restrict__ w) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)g.nf * g.nb) return;
    const int i = (int)(idx % g.nf), j = (int)(idx / g.nf);
    const double *wij = w + 2 * (size_t)idx * mp * mp;
    double S = 0.0, T = 0.0;
    for (int p = g.poff[i]; p < g.poff[i + 1]; ++p)
        for (int q = g.poff[j]; q < g.poff[j + 1]; ++q) {
            const size_t slot = (size_t)(p - g.poff[i]) * mp + (q - g.poff[j]);
            const double cc = g.coef[p] * g.coef[q];
            S += cc * wij[2 * slot];
            T += cc * wij[2 * slot + 1];
        }
    g.S[(size_t)i + (size_t)g.nb * j] = S;
    g.T[(size_t)i + (size_t)g.nb * j] = T;
}

// normalize_AO_toroidal_nD: every primitive of AO i to one, then the
// contraction: c_p = coef_p / sqrt(S_pp), out = c / sqrt(c^T S c), S the
// self-overlap of the AO's raw primitives (same centre).  The self-overlaps by one
// thread per primitive pair (w: slot (i mp + p') mp + q'), then one thread per AO.
__global__ __launch_bounds__(256) void k_norm_prim(OneStArgs g, int mp, double *__restrict__ w) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)g.nb * mp * mp) return;
    const int qq = (int)(idx % mp), pp = (int)((idx / mp) % mp), i = (int)(idx / ((long long)mp * mp));
    const int p0 = g.poff[i];
    if (p0 + pp >= g.poff[i + 1] || p0 + qq >= g.poff[i + 1]) return;
    double s, dummy;
    prim_st(g, g.lxyz + 3 * i, g.lxyz + 3 * i, g.center + 3 * i, g.center + 3 * i, g.expo[p0 + pp], g.expo[p0 + qq],
            false, &s, &dummy);
    w[idx] = s;
}

__global__ __launch_bounds__(64) void k_normalize(OneStArgs g, const double *__restrict__ raw, int mp,
                                                  const double *__restrict__ w, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nb) return;
    const int p0 = g.poff[i], np = g.poff[i + 1] - p0;
    const double *wi = w + (size_t)i * mp * mp;
    double norm2 = 0.0;
    for (int p = 0; p < np; ++p) {
        const double cp = raw[p0 + p] / sqrt(wi[p * mp + p]);
        for (int q = 0; q < np; ++q) norm2 += cp * (raw[p0 + q] / sqrt(wi[q * mp + q])) * wi[p * mp + q];
    }
    const double nrm = 1.0 / sqrt(norm2);
    for (int p = 0; p < np; ++p) out[p0 + p] = raw[p0 + p] / sqrt(wi[p * mp + p]) * nrm;
}

// rows of the other cells from the unit cell's (src/Tools/symmetric_matrix.f90:1-52):
// zero == 1 clears |M(i, j)| < 1e-14 in the unit-cell rows; zero == 0 copies
// M(a + nf ci, b + nf cj) = M(a, b + nf relc(ci, cj)) for the rows of the other cells
__global__ __launch_bounds__(256) void k_translate_fill(double *__restrict__ M, int nf, int nb, int nc,
                                                        const int *__restrict__ relc, int zero) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (zero) {
        if (idx >= (long long)nf * nb) return;
        const int i = (int)(idx % nf), j = (int)(idx / nf);
        double &m = M[(size_t)i + (size_t)nb * j];
        if (fabs(m) < 1e-14) m = 0.0;
        return;
    }
    if (idx >= (long long)(nb - nf) * nb) return;
    const int i = nf + (int)(idx % (nb - nf)), j = (int)(idx / (nb - nf));
    const int ci = i / nf, a = i % nf, cj = j / nf, b = j % nf;
    const int src = relc[(size_t)ci * nc + cj] * nf + b;
    M[(size_t)i + (size_t)nb * j] = M[(size_t)a + (size_t)nb * src];
}

hipError_t launch_one_st(const OneStArgs &g, int mp, double *work, hipStream_t st) {
    const long long n = (long long)g.nf * g.nb;
    hipLaunchKernelGGL(k_one_st_prim, dim3((unsigned)((n * mp * mp + 255) / 256)), dim3(256), 0, st, g, mp, work);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_one_st, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, mp, work);
    return hipGetLastError();
}

hipError_t launch_normalize(const OneStArgs &g, const double *raw, double *out, int mp, double *work,
                            hipStream_t st) {
    const long long n = (long long)g.nb * mp * mp;
    hipLaunchKernelGGL(k_norm_prim, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, mp, work);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_normalize, dim3((g.nb + 63) / 64), dim3(64), 0, st, g, raw, mp, work, out);
    return hipGetLastError();
}

hipError_t launch_translate_fill(double *M, int nf, int nb, int nc, const int *relc, hipStream_t st) {
    hipLaunchKernelGGL(k_translate_fill, dim3((unsigned)(((long long)nf * nb + 255) / 256)), dim3(256), 0, st, M, nf,
                       nb, nc, relc, 1);
    if (nb > nf)
        hipLaunchKernelGGL(k_translate_fill, dim3((unsigned)(((long long)(nb - nf) * nb + 255) / 256)), dim3(256), 0,
                           st, M, nf, nb, nc, relc, 0);
    return hipGetLastError();
}

}  // namespace tcg
