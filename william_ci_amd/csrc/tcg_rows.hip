// ERI quartet pipeline, row form (gfx950).
//
// The arithmetic is that of the reference kernel
// (src/Toroidal/ERI_shell_block_toroidal.f90:28-390); the work is regrouped so
// that the matrix cores see full, uniform tiles:
//
//   k_surv        one wavefront per shell quartet: the primitive quartets that
//                 survive the prefactor and primitive-Schwarz screens
//                 (ERI_shell_block_toroidal.f90:191,216-219), counted, then listed.
//   k_rows_emit   one ROW per (surviving primitive quartet, periodic direction d,
//                 power combination): offsets of its bra/ket Fourier vectors,
//                 length nm (row_length: the product's cut-off, at most
//                 min(N_bra, N_ket)), phase phi, power sum; sort key
//                 (rule, d, nm).
//   (radix sort of the rows by key, tcg_sort.hip)
//   k_rows_mfma   one wavefront per 16 consecutive sorted rows:
//                 I_d(t_k) = sum_n Re[conj(G_n) H_n e^{in phi}] T_d[n][k]
//                 with v_mfma_f64_16x16x4f64 (F on the VALU as the A operand,
//                 the t-table as B), loads two steps ahead of use.  Result:
//                 the node values of every row (nqp per row), in HBM.
//   k_real_rows   a few wavefronts per shell quartet, a lane per node, over its
//                 (primitive quartet, real direction) pairs: the Hermite-form
//                 factors of all power combinations.
//   k_contract    one wavefront per shell quartet: sum_k w_k Ix Iy Iz, coefficients, the shell
//                 block in LDS, then the 8-permutation scatter translated to
//                 the unit cell (src/Toroidal/ERI_toroidal.f90:416-454) or the
//                 block itself (Schwarz passes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "tcg_device.hpp"

namespace tcg {

typedef double dbl4 __attribute__((ext_vector_type(4)));

__constant__ int c_perm8_rows[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                       {2, 3, 0, 1}, {3, 2, 0, 1}, {2, 3, 1, 0}, {3, 2, 1, 0}};


__device__ __forceinline__ int wmax(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wmin(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

// rule of a quartet and its node count
__device__ __forceinline__ int quartet_rule(const EriArgs &A, int qi) { return A.qrule ? A.qrule[qi] : 0; }
__device__ __forceinline__ int rule_nodes(const EriArgs &A, int r) { return A.rule_nq ? A.rule_nq[r] : A.nq; }

__device__ __forceinline__ void quartet_shape(const EriArgs &A, const Quartet &Q, int l[4], int &ncomb) {
    for (int x = 0; x < 4; ++x) l[x] = A.sh[Q.s[x]].l;
    ncomb = (l[0] + 1) * (l[1] + 1) * (l[2] + 1) * (l[3] + 1);
}

// the s/p forms: (P, NBP) -> list 0..7 (P p shells, NBP of them in the bra; at
// most two per pair); list 8 holds the rest
__host__ __device__ constexpr int sp_list(int P, int NBP) {
    return P == 0 ? 0 : P == 1 ? 1 + NBP : P == 2 ? 3 + NBP : 5 + NBP;
}
constexpr int kSpLists = 9;

// the quartets the s/p form takes: s and p shells, at most three p shells
// (four: 81 accumulators per lane do not fit the registers), and rules of at
// most 64 nodes; returns the list, or -1
__device__ __forceinline__ int sp_form(const EriArgs &A, const Quartet &Q, int qi) {
    if (!A.sp_contract || rule_nodes(A, quartet_rule(A, qi)) > 64) return -1;
    const int lA = A.sh[Q.s[0]].l, lB = A.sh[Q.s[1]].l, lC = A.sh[Q.s[2]].l, lD = A.sh[Q.s[3]].l;
    if (lA > 1 || lB > 1 || lC > 1 || lD > 1) return -1;
    const int P = lA + lB + lC + lD;
    return P <= 3 ? sp_list(P, lA + lB) : -1;
}

// A slab build (slab_pos set) stores only the elements whose last index lies in
// one of the slab's cells.  Permutation (p1 p2|p3 p4) of the quartet (A B|C D)
// stores its block with the first index translated to the unit cell, so its last
// index lies in cell relc(cell p1, cell p4); the eight permutations pair
// (p1, p4) as (A,D) (B,D) (A,C) (B,C) (C,B) (D,B) (C,A) (D,A).  A quartet none of
// whose pairs reaches the slab's cells has nothing to store.
__device__ __forceinline__ bool slab_hit(const EriArgs &A, const Quartet &Q) {
    if (!A.slab_pos) return true;
    const int cA = A.sh[Q.s[0]].cell, cB = A.sh[Q.s[1]].cell, cC = A.sh[Q.s[2]].cell, cD = A.sh[Q.s[3]].cell;
    auto in = [&](int x, int y) { return A.slab_pos[A.relc[x * A.ncell + y]] >= 0; };
    return in(cA, cD) || in(cB, cD) || in(cA, cC) || in(cB, cC) || in(cC, cB) || in(cD, cB) || in(cC, cA) ||
           in(cD, cA);
}

// ---------------------------------------------------------------------------
// survivors: pass 0 counts (scount, rcount), pass 1 writes at soff; a quartet a
// slab build does not store has none
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_surv(EriArgs A, int pass, long long *__restrict__ scount,
                                             long long *__restrict__ rcount, const long long *__restrict__ soff,
                                             Surv *__restrict__ surv) {
    // 16 lanes per quartet, four quartets per wavefront (most quartets have few
    // candidate primitive quartets; a group's survivors keep their candidate order)
    const int lane = threadIdx.x & 63, gl = lane & 15, grp = lane >> 4;
    const int q0 = blockIdx.x * 16 + (threadIdx.x >> 4);
    // a list whose length is on the device: the entries past it count zero
    const int nlive = A.nq4_dev ? (int)min((unsigned long long)A.nq4, *A.nq4_dev) : A.nq4;
    for (int qi = q0; qi < A.nq4; qi += gridDim.x * 16) {
        if (qi >= nlive || !slab_hit(A, A.q4[qi])) {
            if (!pass && gl == 0) {
                scount[qi] = 0;
                rcount[qi] = 0;
            }
            continue;
        }
        const Quartet Q = A.q4[qi];
        const ShellPair bp = A.sp[Q.bra], kp = A.sp[Q.ket];
        const long ncand = (Q.prim >= 0) ? 1 : (long)bp.count * kp.count;
        long n = 0;
        const long base_out = pass ? soff[qi] : 0;
        for (long base = 0; base < ncand; base += 16) {
            const long c = base + gl;
            bool ok = c < ncand;
            int ij = 0, kl = 0;
            double fac = 0.0;
            if (ok) {
                if (Q.prim >= 0) { ij = Q.prim; kl = Q.prim; }
                else { ij = (int)(c / kp.count); kl = (int)(c % kp.count); }
                const PrimPair &pb = A.pp[bp.first + ij];
                const PrimPair &pk = A.pp[kp.first + kl];
                fac = (A.fac_const * pb.fac) * pk.fac;
                // prefactor screen: not in the primitive Schwarz table (quartets of one primitive pair)
                if (A.screen_prefactor && Q.prim < 0 && fabs(fac) * pb.cmax * pk.cmax <= kScreenTol) ok = false;
                if (A.screen_prim && pb.ps * pk.ps < kPrimTol) ok = false;
            }
            // this group's 16 bits of the wavefront's ballot (lanes of groups that left
            // the loop are inactive and contribute none)
            const unsigned m = (unsigned)((__ballot(ok) >> (16 * grp)) & 0xffffull);
            if (pass && ok) {
                const long pos = n + __popc(m & ((1u << gl) - 1u));
                Surv sv;
                sv.pb = bp.first + ij;
                sv.pk = kp.first + kl;
                sv.fac = fac;
                surv[base_out + pos] = sv;
            }
            n += __popc(m);
        }
        if (!pass && gl == 0) {
            int l[4], ncomb;
            quartet_shape(A, Q, l, ncomb);
            scount[qi] = n;
            rcount[qi] = n * A.ndp * ncomb;
        }
    }
}

hipError_t launch_surv(const EriArgs &a, int pass, long long *scount, long long *rcount, const long long *soff,
                       Surv *surv, hipStream_t st) {
    if (a.nq4 <= 0) return hipSuccess;
    const int nwg = (a.nq4 + 15) / 16;
    const int grid = nwg < (1 << 20) ? nwg : (1 << 20);
    hipLaunchKernelGGL(k_surv, dim3(grid), dim3(256), 0, st, a, pass, scount, rcount, soff, surv);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// rows
// ---------------------------------------------------------------------------
// Orders of one row F_n = Re[conj(G_n) H_n e^{in phi}]: each pair's Fourier vector
// is cut where its Bessel factor e^{-A} I_n(A) falls below ~e^{-40.5} of I_0 (N =
// 20 + 9 sqrt(A), the reference's truncation, clifford_tools.f90:566-572); the
// product of the two falls there already at n - L = 9 sqrt(A_b A_k / (A_b + A_k))
// (e^{-n^2/2A} per factor; L = the larger power sum, the shift of the sin-product
// Fourier coefficients), + 20 as the reference adds, so the row stops at the
// smaller of that and min(N_bra, N_ket).  The orders left out carry less than
// 1e-17 of the row's leading terms.
__device__ __forceinline__ int row_length(int bN, int kN, float Ab, float Ak, int Lb, int Lk) {
    int nm = bN < kN ? bN : kN;
    const float s = Ab + Ak;
    const float h = s > 0.f ? Ab * Ak / s : 0.f;
    const int cut = 20 + (Lb > Lk ? Lb : Lk) + (int)ceilf(9.0f * sqrtf(h));
    return cut < nm ? cut : nm;
}

// the last quartet whose first row is at most R: roff ascends, roff[0] = 0 <= R <
// roff[nq4]; quartets without rows share their offset with the next one, so the
// quartet found has rows
__device__ __forceinline__ int row_owner(const long long *__restrict__ roff, int nq4, long long R) {
    int lo = 0, hi = nq4;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (roff[mid] <= R)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// one thread per row (rows of all quartets, roff[nq4] of them): its quartet by
// binary search over the row offsets, then its record, key and value
__global__ __launch_bounds__(256) void k_rows_emit(EriArgs A, const long long *__restrict__ soff,
                                                  const long long *__restrict__ roff, const Surv *__restrict__ surv,
                                                  RowRec *__restrict__ rows, unsigned *__restrict__ keys,
                                                  int *__restrict__ vals) {
    const long long total = roff[A.nq4];
    for (long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x; row < total;
         row += (long long)gridDim.x * blockDim.x) {
        const int qi = row_owner(roff, A.nq4, row);
        const Quartet &Q = A.q4[qi];
        int l[4], ncomb;
        quartet_shape(A, Q, l, ncomb);
        const long s0 = soff[qi];
        const long r0 = roff[qi], rl = row - r0;
        {
            const int combo = (int)(rl % ncomb);
            const int d = (int)((rl / ncomb) % A.ndp);
            const long g = rl / ((long)ncomb * A.ndp);
            const Surv S = surv[s0 + g];
            const PrimPair &pb = A.pp[S.pb];
            const PrimPair &pk = A.pp[S.pk];
            int t = combo;
            const int ia = t % (l[0] + 1); t /= (l[0] + 1);
            const int ib = t % (l[1] + 1); t /= (l[1] + 1);
            const int ic = t % (l[2] + 1);
            const int id = t / (l[2] + 1);
            const int bN = pb.N[d], kN = pk.N[d];
            RowRec R;
            R.nm = row_length(bN, kN, pb.A[d], pk.A[d], l[0] + l[1], l[2] + l[3]);
            R.gofs = pb.voff[d] + (ia * (l[1] + 1) + ib) * (bN + 1);
            R.hofs = pk.voff[d] + (ic * (l[3] + 1) + id) * (kN + 1);
            const double ad = (d == 0) ? A.a[0] : (d == 1) ? A.a[1] : A.a[2];
            const double shd = Q.shift[d];
            // the phase factors of the row, once per row (the MFMA kernel needs no sincos)
            const double phi = ad * (pb.cen[d] - (pk.cen[d] + shd));
            sincos(phi, &R.s1, &R.c1);
            // e^{4 i phi} by two doublings of e^{i phi} (a few ulps); e^{64 i phi} directly
            const double c2 = (R.c1 - R.s1) * (R.c1 + R.s1), s2 = 2.0 * R.s1 * R.c1;
            R.c4 = (c2 - s2) * (c2 + s2);
            R.s4 = 2.0 * s2 * c2;
            sincos(64.0 * phi, &R.s64, &R.c64);
            R.sgn = pow(ad, -(double)(ia + ib + ic + id));
            const int rule = quartet_rule(A, qi);
            R.d = d | (rule << 8);
            rows[r0 + rl] = R;
            // key (rule, direction, order): the order field holds ncap + 1 (nm_bits), the
            // direction two bits above it, the rule above both
            keys[r0 + rl] = ((unsigned)rule << (A.nm_bits + 2)) | ((unsigned)d << A.nm_bits) | (unsigned)R.nm;
            vals[r0 + rl] = (int)(r0 + rl);
        }
    }
}

// the same with one wavefront per quartet, a lane per row (quartets of many rows:
// no search, the quartet's data read once per wavefront)
__global__ __launch_bounds__(256) void k_rows_emit_q(EriArgs A, const long long *__restrict__ soff,
                                                    const long long *__restrict__ roff, const Surv *__restrict__ surv,
                                                    RowRec *__restrict__ rows, unsigned *__restrict__ keys,
                                                    int *__restrict__ vals) {
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);     // four wavefronts, one quartet each
    for (int qi = blockIdx.x * 4 + wv; qi < A.nq4; qi += gridDim.x * 4) {
        const Quartet Q = A.q4[qi];
        int l[4], ncomb;
        quartet_shape(A, Q, l, ncomb);
        const long s0 = soff[qi], ns = soff[qi + 1] - s0;
        const long r0 = roff[qi], nr = ns * A.ndp * ncomb;
        for (long rl = lane; rl < nr; rl += 64) {
            const int combo = (int)(rl % ncomb);
            const int d = (int)((rl / ncomb) % A.ndp);
            const long g = rl / ((long)ncomb * A.ndp);
            const Surv S = surv[s0 + g];
            const PrimPair &pb = A.pp[S.pb];
            const PrimPair &pk = A.pp[S.pk];
            int t = combo;
            const int ia = t % (l[0] + 1); t /= (l[0] + 1);
            const int ib = t % (l[1] + 1); t /= (l[1] + 1);
            const int ic = t % (l[2] + 1);
            const int id = t / (l[2] + 1);
            const int bN = pb.N[d], kN = pk.N[d];
            RowRec R;
            R.nm = row_length(bN, kN, pb.A[d], pk.A[d], l[0] + l[1], l[2] + l[3]);
            R.gofs = pb.voff[d] + (ia * (l[1] + 1) + ib) * (bN + 1);
            R.hofs = pk.voff[d] + (ic * (l[3] + 1) + id) * (kN + 1);
            const double ad = (d == 0) ? A.a[0] : (d == 1) ? A.a[1] : A.a[2];
            const double shd = A.q4[qi].shift[d];
            // the phase factors of the row, once per row (the MFMA kernel needs no sincos)
            const double phi = ad * (pb.cen[d] - (pk.cen[d] + shd));
            sincos(phi, &R.s1, &R.c1);
            // e^{4 i phi} by two doublings of e^{i phi} (a few ulps); e^{64 i phi} directly
            const double c2 = (R.c1 - R.s1) * (R.c1 + R.s1), s2 = 2.0 * R.s1 * R.c1;
            R.c4 = (c2 - s2) * (c2 + s2);
            R.s4 = 2.0 * s2 * c2;
            sincos(64.0 * phi, &R.s64, &R.c64);
            R.sgn = pow(ad, -(double)(ia + ib + ic + id));
            const int rule = quartet_rule(A, qi);
            R.d = d | (rule << 8);
            rows[r0 + rl] = R;
            // key (rule, direction, order): the order field holds ncap + 1 (nm_bits), the
            // direction two bits above it, the rule above both
            keys[r0 + rl] = ((unsigned)rule << (A.nm_bits + 2)) | ((unsigned)d << A.nm_bits) | (unsigned)R.nm;
            vals[r0 + rl] = (int)(r0 + rl);
        }
    }
}

hipError_t launch_rows_emit(const EriArgs &a, const long long *soff, const long long *roff, const Surv *surv,
                            RowRec *rows, unsigned *keys, int *vals, long long nrows, hipStream_t st) {
    if (a.nq4 <= 0 || nrows <= 0) return hipSuccess;
    // nrows: the row count or an upper bound (the kernels stop at roff[nq4]).  Lists
    // of 32 or more rows per quartet on average keep a wavefront per quartet; shorter
    // ones (most quartets of s and p shells) a thread per row
    if (nrows >= 32ll * a.nq4) {
        const int nwg = (a.nq4 + 3) / 4;
        const int grid = nwg < (1 << 20) ? nwg : (1 << 20);
        hipLaunchKernelGGL(k_rows_emit_q, dim3(grid), dim3(256), 0, st, a, soff, roff, surv, rows, keys, vals);
        return hipGetLastError();
    }
    const long long nwg = (nrows + 255) / 256;
    const int grid = nwg < (1 << 20) ? (int)nwg : (1 << 20);
    hipLaunchKernelGGL(k_rows_emit, dim3(grid), dim3(256), 0, st, a, soff, roff, surv, rows, keys, vals);
    return hipGetLastError();
}

// One tile = 16 sorted rows; a workgroup takes four consecutive tiles (one
// per wavefront) at a time.  kNT = node tiles of 16 in this launch (4 for the
// 64-node rule, 2 for 32 nodes, 3 for 40).
//
// Lanes build F_n = Re[conj(G_n) H_n e^{in phi}] (the A operand of
// v_mfma_f64_16x16x4f64) for 16 rows x 4 orders; the t-table is the B
// operand, D accumulates 16 rows x 16 nodes over n in steps of 4.  The
// workgroup's four tiles run one (rule, direction) combination at a time (one
// in all but the groups where the sorted rows change combination): the table
// slice of S orders x the launch's nodes is staged in LDS by the whole
// workgroup, double-buffered (the next slice's global loads in flight while
// the current one feeds the matrix cores), so each table entry is read from
// L2 once per workgroup instead of once per wavefront.
//
// The node tiles run as slots sorted by their last order, descending: stage
// c has the c slots with the largest limits active, each stage a loop of
// unconditional MFMAs.  G and H are loaded two steps ahead; the phase
// advances by a rotation with a second rotation per 64 orders (no sincos).
template <int kNT>
__global__ __launch_bounds__(256) void k_rows_mfma(EriArgs A, const RowRec *__restrict__ rows,
                                                  const int *__restrict__ order, long long nrows, int hbase0,
                                                  double *__restrict__ rowval) {
    // node tiles hbase .. hbase + kNT - 1 (blockIdx.y: one group of kNT per grid row)
    const int hbase = hbase0 + kNT * (int)blockIdx.y;
    constexpr int S = 16;                            // orders per LDS slice
    constexpr int TW = kNT * kNodeTile;              // nodes of this launch
    constexpr int W = TW + 4;                        // LDS row stride (padded)
    constexpr int NE = S * TW / 256;                 // slice entries per thread
    __shared__ double tl[2][S * W];
    __shared__ int wdd[4][2], wnm[4];
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nqp = A.nqp, nht = A.nqp / kNodeTile;
    const int tabrows = A.ncap + 5, tlast = tabrows - 1;
    const long long ntiles = (nrows + 15) / 16, ngroups = (ntiles + 3) / 4;
    long long flops = 0;                              // this lane's rows' flop count (one atomic per wavefront)
    // the groups from the last (the rows sort by ascending order, so the costliest
    // groups start first and the cheap ones fill the tail)
    for (long long gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const long long grp = ngroups - 1 - gi;
        const long long tile = grp * 4 + wv;
        const long long pos = tile * 16 + r;
        // (a list sized by an upper bound ends in padding entries with index -1)
        int ridx = pos < nrows ? order[pos] : -1;
        const bool valid = ridx >= 0;
        int nm = -1, d = 0x7fffffff;                    // d: direction | rule << 8
        double sgn = 1.0, c1 = 1.0, s1 = 0.0, c4 = 1.0, s4 = 0.0, c64 = 1.0, s64 = 0.0;
        // byte offsets of the lane's G and H entries (order kq) in the pair vectors
        unsigned gob = 16u * (unsigned)kq, hob = 16u * (unsigned)kq;
        if (valid) {
            const RowRec R = rows[ridx];
            nm = R.nm;
            d = R.d;
            sgn = R.sgn;
            c1 = R.c1;
            s1 = R.s1;
            c4 = R.c4;
            s4 = R.s4;
            c64 = R.c64;
            s64 = R.s64;
            gob = 16u * (unsigned)(R.gofs + kq);
            hob = 16u * (unsigned)(R.hofs + kq);
        }
        if (A.counters && kq == 0 && valid && hbase == 0) {
            const int rl = d >> 8, dl = d & 0xff;
            const long long s = A.tsum[(size_t)(rl * A.ndp + dl) * (A.ncap + 1) + (nm < A.ncap ? nm : A.ncap)];
            flops += 2 * s + 10 * (nm + 1);
        }
        // e^{i n phi}: rotation by e^{4 i phi} inside each block of 64 orders; at a block start
        // the lane's base phase advances by e^{64 i phi} (the row's phase factors, formed once
        // per row by the emit kernel), so the rounding error stays within a few tens of ulps
        // for n up to ~1300; the lane's own e^{i kq phi} is a product of at most three e^{i phi}
        double cn0 = 1.0, sn0 = 0.0;
#pragma unroll
        for (int q = 1; q <= 3; ++q) {
            const double tc = cn0 * c1 - sn0 * s1, ts = sn0 * c1 + cn0 * s1;
            cn0 = q <= kq ? tc : cn0;
            sn0 = q <= kq ? ts : sn0;
        }
        // smallest combination above lo over the workgroup's four tiles (uniform)
        auto group_min = [&](int lo) {
            const int m = wmin(d > lo ? d : 0x7fffffff);
            if (lane == 0) wdd[wv][0] = m;
            __syncthreads();
            const int g = __builtin_amdgcn_readfirstlane(min(min(wdd[0][0], wdd[1][0]), min(wdd[2][0], wdd[3][0])));
            __syncthreads();
            return g;
        };
        // per combination dd: slots, accumulators, the order loop, the write-out
        auto run = [&](int dd) {
            const int nmd = (d == dd) ? nm : -1;      // rows of another combination contribute 0
            const int tr = (dd >> 8) * A.ndp + (dd & 0xff);   // table block of this rule and direction
            // (wavefront-uniform values in scalar registers: the order loops below are
            // scalar loops, not exec-masked ones)
            const int nmax_t = __builtin_amdgcn_readfirstlane(wmax(nmd));
            // node tiles as slots sorted by their last order, descending (uniform values);
            // padding tiles of a rule with fewer nodes have lim -1 and never run
            int sl[kNT], sp[kNT];
#pragma unroll
            for (int h = 0; h < kNT; ++h) {
                const int lh = A.tab_lim[tr * nht + hbase + h];
                sl[h] = lh < nmax_t ? lh : nmax_t;
                sp[h] = h;
            }
#pragma unroll
            for (int i = 1; i < kNT; ++i)
#pragma unroll
                for (int j = i; j > 0; --j) {
                    const bool sw = sl[j - 1] < sl[j];
                    const int a = sl[j - 1], b = sl[j], pa = sp[j - 1], pb = sp[j];
                    sl[j - 1] = sw ? b : a;
                    sl[j] = sw ? a : b;
                    sp[j - 1] = sw ? pb : pa;
                    sp[j] = sw ? pa : pb;
                }
            const double *tabg = A.tab + (size_t)tr * tabrows * nqp + hbase * kNodeTile;
            int toff[kNT];
#pragma unroll
            for (int j = 0; j < kNT; ++j) toff[j] = sp[j] * kNodeTile;
            dbl4 acc[kNT];
#pragma unroll
            for (int j = 0; j < kNT; ++j) acc[j] = dbl4{0.0, 0.0, 0.0, 0.0};
            double cn = cn0, sn = sn0, cb = cn0, sb = sn0;
            // Two register sets of G, H, each consumed by one step and refilled for the step
            // after next.  Orders past a row's nm read the entries that follow its vector
            // (the pair vectors end in a padding of ncap + 64 entries, so every such read
            // stays in the allocation), and orders past the table the zero rows at its end;
            // the selection of F or the zero table make their products vanish.
            const char *vb = reinterpret_cast<const char *>(A.vec);
            auto gl = [&](unsigned m) { return *reinterpret_cast<const double2 *>(vb + gob + 16u * m); };
            auto hl = [&](unsigned m) { return *reinterpret_cast<const double2 *>(vb + hob + 16u * m); };
            auto tidx = [&](int n) { return (size_t)(n < tlast ? n : tlast) * nqp; };
            double2 gP = gl(0), hP = hl(0);
            double2 gQ = gl(4), hQ = hl(4);
            int s0 = 0, cur = 0;
            auto fval = [&](int m, const double2 &g, const double2 &hh) {
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
                const double re = g.x * hh.x + g.y * hh.y;   // conj(g) * h
                const double im = g.x * hh.y - g.y * hh.x;
                return (n <= nmd) ? re * cn - im * sn : 0.0;
            };
            // two steps of 4 orders each, n0..n0+3 and n0+4..n0+7 (lane orders n0 + kq,
            // n0 + 4 + kq; both inside the LDS slice: n0 is a multiple of 8 below its
            // end), c slots active: the table operands of both steps are read from LDS
            // first, so that their latency runs under the F assembly; then the MFMAs of
            // the first step and of the second (the order of every accumulator's sum is
            // that of two single steps); then the G, H loads of the pair after next
            int n0 = 0;
            auto step2 = [&](int c) {
                const double *brow = &tl[cur][(n0 + kq - s0) * W + r];
                double b0[kNT], b1[kNT];
#pragma unroll
                for (int j = 0; j < kNT; ++j)
                    if (j < c) {
                        b0[j] = brow[toff[j]];
                        b1[j] = brow[4 * W + toff[j]];
                    }
                __builtin_amdgcn_sched_barrier(0);
                const double F0 = fval(n0, gP, hP);
                const double F1 = fval(n0 + 4, gQ, hQ);
#pragma unroll
                for (int j = 0; j < kNT; ++j)
                    if (j < c) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(F0, b0[j], acc[j], 0, 0, 0);
#pragma unroll
                for (int j = 0; j < kNT; ++j)
                    if (j < c) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(F1, b1[j], acc[j], 0, 0, 0);
                gP = gl((unsigned)(n0 + 8));
                hP = hl((unsigned)(n0 + 8));
                gQ = gl((unsigned)(n0 + 12));
                hQ = hl((unsigned)(n0 + 12));
            };
            // stage loop over [n0, s_end): stage c has the c slots with the largest limits
            // active, up to order sl[c-1] (steps in pairs: one extra step past a limit only
            // adds zero products)
            auto stages = [&](int s_end) {
#pragma unroll
                for (int c = kNT; c >= 1; --c) {
                    const int lend = sl[c - 1];
#pragma unroll 1
                    for (; n0 <= lend && n0 < s_end; n0 += 8) step2(c);
                }
            };
            {
                // the workgroup's largest order, then the slices
                if (lane == 0) wnm[wv] = nmax_t;
                __syncthreads();
                const int nmax_wg = __builtin_amdgcn_readfirstlane(max(max(wnm[0], wnm[1]), max(wnm[2], wnm[3])));
                double stg[NE];
                auto gload = [&](int sb0) {
#pragma unroll
                    for (int i = 0; i < NE; ++i) {
                        const int e = threadIdx.x + 256 * i;
                        stg[i] = tabg[tidx(sb0 + e / TW) + e % TW];
                    }
                };
                auto sstore = [&](int buf) {
#pragma unroll
                    for (int i = 0; i < NE; ++i) {
                        const int e = threadIdx.x + 256 * i;
                        tl[buf][(e / TW) * W + e % TW] = stg[i];
                    }
                };
                gload(0);
                sstore(0);
                __syncthreads();
                for (s0 = 0; s0 <= nmax_wg; s0 += S) {
                    const bool more = s0 + S <= nmax_wg;
                    if (more) gload(s0 + S);
                    stages(s0 + S);
                    if (more) sstore(cur ^ 1);
                    __syncthreads();
                    cur ^= 1;
                }
            }
            // lane (r, kq) holds D[row = 4v + kq][node = 16h + r]; the node tiles past the
            // rule's last node (limit -1) are never read and are not written
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = 4 * v + kq;
                const int rrow = __shfl(ridx, row, 64);
                const int rd = __shfl(d, row, 64);
                const double sg = __shfl(sgn, row, 64);
                if (rrow >= 0 && rd == dd) {
#pragma unroll
                    for (int j = 0; j < kNT; ++j)
                        if (sl[j] >= 0) rowval[(size_t)rrow * nqp + hbase * kNodeTile + toff[j] + r] = acc[j][v] * sg;
                }
            }
        };
        // the (rule, direction) combinations present in the workgroup's tiles, ascending
        // (one, except where the sorted rows change combination)
#pragma unroll 1
        for (int dd = group_min(-1); dd != 0x7fffffff; dd = group_min(dd)) run(dd);
    }
    if (A.counters) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) flops += __shfl_xor(flops, off, 64);
        if (lane == 0 && flops) atomicAdd(&A.counters[4], (unsigned long long)flops);
    }
}

hipError_t launch_rows_mfma(const EriArgs &a, const RowRec *rows, const int *order, long long nrows, double *rowval,
                            hipStream_t st) {
    if (nrows <= 0) return hipSuccess;
    const long long ntiles = (nrows + 15) / 16;
    const long long nwg = (ntiles + 3) / 4;                 // four wavefronts per workgroup, a tile each
    // at most as many workgroups as the chip holds at once (the LDS path's barriers
    // keep the four wavefronts of a workgroup together; groups of tiles are walked)
    // (an XCD-aware walk, each XCD a contiguous eighth of the sorted groups, was
    // slower: the eighths differ several-fold in cost, the orders growing along the sort)
    const int grid = nwg < 2048 ? (int)nwg : 2048;
    const int nht = a.nqp / kNodeTile;
    // short row lists (the Schwarz passes: few, long rows) give each node tile a grid row
    // of its own (four times the wavefronts, one MFMA per step each); long lists take up
    // to four node tiles per wavefront, one launch per group of four
    if (ntiles < 4 * 2048) {
        const int g1 = (int)nwg;
        hipLaunchKernelGGL(k_rows_mfma<1>, dim3(g1, nht), dim3(256), 0, st, a, rows, order, nrows, 0, rowval);
        return hipGetLastError();
    }
    for (int h0 = 0; h0 < nht; h0 += 4) {
        switch (nht - h0 < 4 ? nht - h0 : 4) {
        case 1: hipLaunchKernelGGL(k_rows_mfma<1>, dim3(grid), dim3(256), 0, st, a, rows, order, nrows, h0, rowval); break;
        case 2: hipLaunchKernelGGL(k_rows_mfma<2>, dim3(grid), dim3(256), 0, st, a, rows, order, nrows, h0, rowval); break;
        case 3: hipLaunchKernelGGL(k_rows_mfma<3>, dim3(grid), dim3(256), 0, st, a, rows, order, nrows, h0, rowval); break;
        default: hipLaunchKernelGGL(k_rows_mfma<4>, dim3(grid), dim3(256), 0, st, a, rows, order, nrows, h0, rowval); break;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// lane l's value of v (l wavefront-uniform)
__device__ __forceinline__ double readlane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// LDS writes of the wavefront visible to its other lanes (for LDS that one
// wavefront owns; no wait for outstanding global loads)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------
// real directions (ERI_shell_block_toroidal.f90:228-236,302-335): at node t,
// for a surviving primitive quartet with exponent sums p, q and distance R,
//   g_0 = pi sqrt(D) e^{-w R^2}, g_{N+1} = 2 w R g_N - 2 N w g_{N-1},
//   D = 1/(pq + (p+q) t^2), w = pq t^2 D,
// and the McMurchie-Davidson factor of every power combination
//   I = sum_{n,m} (-1)^n E_n E'_m g_{n+m},
// taken as sum_n (-1)^n E_n h_n with h_n = sum_m E'_m g_{n+m} (each h_n
// shared by the bra combinations of one ket combination).
// sqrt(D) by the hardware reciprocal square root estimate of pq + (p+q) t^2
// refined by two Newton steps y <- y (3 - x y^2) / 2, each of which squares the
// relative error (about 1 ulp after the second); D = y^2.  This replaces a
// division and a square root per node and direction.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

__device__ __forceinline__ void real_gt(double p, double q, double t2, double Rr, int NN, double gt[9]) {
    const double y = rsqrt_nr(p * q + (p + q) * t2);
    const double D2 = y * y;
    const double w = p * q * t2 * D2;
    const double rs = 3.14159265358979323846 * y;
    gt[0] = (Rr == 0.0) ? rs : rs * exp(-w * Rr * Rr);
    double gm1 = 0.0, gc = gt[0];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const double gp1 = 2.0 * w * Rr * gc - 2.0 * (double)n * w * gm1;
        gt[n + 1] = (n < NN) ? gp1 : 0.0;
        gm1 = gc;
        gc = gp1;
    }
}

// store(combo, value) for combo = ia + (lA+1)(ib + (lB+1)(ic + (lC+1) id));
// Eb, Ek: Hermite E of the bra / ket pair in this direction,
// [(ia*(lB+1)+ib)*kE + n], n <= 4 used (kE = 2 kMaxL + 1 as stored)
template <class F>
__device__ __forceinline__ void real_combos(int lA, int lB, int lC, int lD, const double *__restrict__ Eb,
                                            const double *__restrict__ Ek, const double gt[9], F &&store,
                                            int kE = 2 * kMaxL + 1) {
    const int nab = lA + lB;
    for (int id = 0; id <= lD; ++id)
        for (int ic = 0; ic <= lC; ++ic) {
            const double *ek = Ek + (ic * (lD + 1) + id) * kE;
            const int ncd = ic + id;
            double em[5];
#pragma unroll
            for (int m = 0; m < 5; ++m) em[m] = m <= ncd ? ek[m] : 0.0;
            double h[5];
#pragma unroll
            for (int n = 0; n < 5; ++n) {
                double sacc = 0.0;
                if (n <= nab) {
#pragma unroll
                    for (int m = 0; m < 5; ++m)
                        if (m <= ncd) sacc += em[m] * gt[n + m];
                }
                h[n] = sacc;
            }
            const int kc = ic + (lC + 1) * id;
            for (int ib = 0; ib <= lB; ++ib)
                for (int ia = 0; ia <= lA; ++ia) {
                    const double *eb = Eb + (ia * (lB + 1) + ib) * kE;
                    const int nb2 = ia + ib;
                    double v = 0.0;
#pragma unroll
                    for (int n = 0; n < 5; ++n)
                        if (n <= nb2) v += ((n & 1) ? -eb[n] : eb[n]) * h[n];
                    store(ia + (lA + 1) * (ib + (lB + 1) * kc), v);
                }
        }
}

// ---------------------------------------------------------------------------
// real-direction rows for the contraction: gridDim.y wavefronts per shell
// quartet (one for long quartet lists), lanes over the nodes, each taking every
// gridDim.y-th (surviving primitive quartet, real direction) step; written next
// to the periodic rows of the same quartet: with
// R = roff[q] / ndp (the quartet's first (group, combination) index) and g
// the group inside the quartet,
//   realval[(R * nreal + (g * nreal + dr) * ncomb + combo) * nqp + k].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_real_rows(EriArgs A, const long long *__restrict__ soff,
                                                  const long long *__restrict__ roff,
                                                  const Surv *__restrict__ surv, double *__restrict__ realval,
                                                  const int *__restrict__ list, const unsigned *__restrict__ cnt) {
    const int nreal = 3 - A.ndp, nqp = A.nqp, lane = threadIdx.x;
    const int nlist = list ? (int)*cnt : A.nq4;
    for (int e = blockIdx.x; e < nlist; e += gridDim.x) {
        const int qi = list ? list[e] : e;
        const Quartet *Q = A.q4 + qi;
        const int rule = quartet_rule(A, qi);
        const int nq = rule_nodes(A, rule);
        const int lA = A.sh[Q->s[0]].l, lB = A.sh[Q->s[1]].l, lC = A.sh[Q->s[2]].l, lD = A.sh[Q->s[3]].l;
        const int ncomb = (lA + 1) * (lB + 1) * (lC + 1) * (lD + 1);
        const int NN = lA + lB + lC + lD;
        const long long s0 = soff[qi], ns = soff[qi + 1] - s0;
        const long long R0 = roff[qi] / A.ndp;
        const double shift[3] = {Q->shift[0], Q->shift[1], Q->shift[2]};
        double t2v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = lane + 64 * h;
            const double tk = k < nq ? A.tnode[rule * nqp + k] : 0.0;
            t2v[h] = tk * tk;
        }
        // survivors in batches of 64: lane i reads survivor i of the batch (exponent
        // sums, distance per real direction, Hermite offsets); the steps take them from
        // that lane (readlane), so a step waits only on its Hermite coefficients
        for (long long gb = 0; gb < ns; gb += 64) {
            const int nb = ns - gb < 64 ? (int)(ns - gb) : 64;
            double bq, kq, Rl[2] = {0.0, 0.0};
            int ebo, eko;
            {
                const Surv S = surv[s0 + gb + (lane < nb ? lane : nb - 1)];
                const PrimPair &pb = A.pp[S.pb];
                const PrimPair &pk = A.pp[S.pk];
                bq = pb.q;
                kq = pk.q;
                ebo = pb.eoff;
                eko = pk.eoff;
#pragma unroll
                for (int drr = 0; drr < 2; ++drr)
                    if (drr < nreal) {
                        const int dr = A.ndp + drr;
                        Rl[drr] = pb.cen[dr] - (pk.cen[dr] + shift[dr]);
                    }
            }
            const int nit = nb * nreal;
            for (int it = blockIdx.y; it < nit; it += gridDim.y) {
                const int i = it / nreal, drr = it - i * nreal;
                const long long gl = gb + i;
                const double pq = readlane_d(bq, i), qq = readlane_d(kq, i);
                const double Rr = readlane_d(drr == 0 ? Rl[0] : Rl[1], i);
                const double *Eb = A.herm + __builtin_amdgcn_readlane(ebo, i) +
                                   (size_t)drr * (lA + 1) * (lB + 1) * (2 * kMaxL + 1);
                const double *Ek = A.herm + __builtin_amdgcn_readlane(eko, i) +
                                   (size_t)drr * (lC + 1) * (lD + 1) * (2 * kMaxL + 1);
                double *out = realval + ((size_t)(R0 * nreal + (gl * nreal + drr) * ncomb)) * nqp;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int k = lane + 64 * h;
                    if (h * 64 >= nq) break;
                    double gt[9];
                    real_gt(pq, qq, t2v[h], Rr, NN, gt);
                    real_combos(lA, lB, lC, lD, Eb, Ek, gt, [&](int combo, double v) {
                        if (k < nq) out[(size_t)combo * nqp + k] = v;
                    });
                }
            }
        }
    }
}

// every quartet of a (list == nullptr), or the *cnt quartets of list
static hipError_t real_rows(const EriArgs &a, const long long *soff, const long long *roff, const Surv *surv,
                            double *realval, const int *list, const unsigned *cnt, hipStream_t st) {
    if (a.ndp >= 3 || a.nq4 <= 0) return hipSuccess;
    // a list's length is on the device: a grid of at most 4096 wavefronts walks it
    const int cap = list ? 4096 : (1 << 20);
    const int grid = a.nq4 < cap ? a.nq4 : cap;
    // short quartet lists (the Schwarz passes: few quartets, up to nprim^4 survivors each)
    // spread the (survivor, direction) steps of a quartet over 16 wavefronts; long lists
    // keep one wavefront per quartet (measured faster there)
    const int gy = a.nq4 < 16384 ? 16 : 1;
    hipLaunchKernelGGL(k_real_rows, dim3(grid, gy), dim3(64), 0, st, a, soff, roff, surv, realval, list, cnt);
    return hipGetLastError();
}

hipError_t launch_real_rows(const EriArgs &a, const long long *soff, const long long *roff, const Surv *surv,
                            long long ns, double *realval, hipStream_t st) {
    if (ns <= 0) return hipSuccess;
    return real_rows(a, soff, roff, surv, realval, nullptr, nullptr, st);
}

// ---------------------------------------------------------------------------
// contraction, one wavefront per shell quartet: the block
//   blk(a,b,c,d) = sum_g fac_g cA cB cC cD sum_k w_k Ix Iy Iz.
// Each lane owns one component and a residue class of the surviving
// primitive quartets g: with fewer than 64 components, G = 64 / ncomp lanes
// share a component (g = lane / ncomp, + G, ...), their sums meet through
// part[] (the first 64 doubles of the per-wavefront staging) in a fixed order; with 64 or more,
// lanes loop over the components and every g.  No atomics: the block comes
// out the same on every run.  Nothing in a task waits on a chain of global
// loads: the Cartesian powers of the four shells are packed into one word per
// shell up front, the row addresses follow from (g, component), and the
// survivors' prefactors and primitive indices (from the pair index: pairs of a
// shell pair are ordered i major) and the four shells' contraction
// coefficients are staged in LDS, 64 survivors at a time.
// Node rows are 16-byte aligned (nqp a multiple of 16) and read two nodes per
// load.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long shell_powers(const EriArgs &A, const Shell &S) {
    unsigned long long w = 0;
    for (int c = 0; c < S.ncomp; ++c) {
        const int *p = A.ao_lxyz + 3 * (S.first_ao + c);
        w |= (unsigned long long)(p[0] | (p[1] << 2) | (p[2] << 4)) << (6 * c);
    }
    return w;
}

__device__ __forceinline__ void quartet_block(const EriArgs &A, const Quartet &Q, int rule, long s0, long ns,
                                              long r0, const Surv *__restrict__ surv,
                                              const double *__restrict__ rowval, const double *__restrict__ realval,
                                              double *blk, double *part, int lane) {
    const int nqp = A.nqp, nq = rule_nodes(A, rule), ndp = A.ndp, nreal = 3 - ndp;
    const double *tw = A.tweight + (size_t)rule * nqp;
    const Shell sA = A.sh[Q.s[0]], sB = A.sh[Q.s[1]], sC = A.sh[Q.s[2]], sD = A.sh[Q.s[3]];
    const int lA = sA.l, lB = sB.l, lC = sC.l, lD = sD.l;
    const int nA = sA.ncomp, nB = sB.ncomp, nC = sC.ncomp, nD = sD.ncomp;
    const int ncomb = (lA + 1) * (lB + 1) * (lC + 1) * (lD + 1);
    const int nblk = nA * nB * nC * nD;
    const int nq2 = nq >> 1;
    const double2 *tw2 = reinterpret_cast<const double2 *>(tw);
    const unsigned long long pwA = shell_powers(A, sA), pwB = shell_powers(A, sB), pwC = shell_powers(A, sC),
                             pwD = shell_powers(A, sD);
    const int pb0 = A.sp[Q.bra].first, pk0 = A.sp[Q.ket].first;   // first primitive pair of bra / ket
    const int npA = sA.nprim, npB = sB.nprim, npC = sC.nprim, npD = sD.nprim;
    const long R0 = r0 / ndp;
    // staged in LDS after part[64]: the prefactors and primitive indices of a chunk of
    // up to 64 survivors, and the contraction coefficients of the four shells
    // (cX[c np + p] = coefficient of component c, primitive p), so that a
    // (component, survivor) task reads LDS rather than a chain of global loads
    double *sfac = part + 64;
    int *sidx = reinterpret_cast<int *>(sfac + 64);            // 64 packed (pi, pj, qk, ql), 4 bits each
    double *cA = sfac + 96, *cB = cA + nA * npA, *cC = cB + nB * npB, *cD = cC + nC * npC;
    for (int e = lane; e < nA * npA; e += 64) cA[e] = A.coef[sA.poff + e];
    for (int e = lane; e < nB * npB; e += 64) cB[e] = A.coef[sB.poff + e];
    for (int e = lane; e < nC * npC; e += 64) cC[e] = A.coef[sC.poff + e];
    for (int e = lane; e < nD * npD; e += 64) cD[e] = A.coef[sD.poff + e];
    // one component, survivors g0, g0 + gstep, ... < gend of the staged chunk
    // starting at gbase: the coefficient-weighted node sums
    auto comp_sum = [&](int cp, long gbase, long g0, long gend, long gstep) {
        int t = cp;
        const int ca = t % nA; t /= nA;
        const int cb = t % nB; t /= nB;
        const int cc = t % nC;
        const int cd = t / nC;
        const unsigned pa = (unsigned)(pwA >> (6 * ca)), pbw = (unsigned)(pwB >> (6 * cb)),
                       pc = (unsigned)(pwC >> (6 * cc)), pd = (unsigned)(pwD >> (6 * cd));
        int cxyz[3];
#pragma unroll
        for (int dd = 0; dd < 3; ++dd) {
            const int sh = 2 * dd;
            cxyz[dd] = ((pa >> sh) & 3) + (lA + 1) * (((pbw >> sh) & 3) +
                       (lB + 1) * (((pc >> sh) & 3) + (lC + 1) * ((pd >> sh) & 3)));
        }
        const double *cfA = cA + ca * npA, *cfB = cB + cb * npB, *cfC = cC + cc * npC, *cfD = cD + cd * npD;
        double acc = 0.0;
        for (long gl = g0; gl < gend; gl += gstep) {
            const long g = gbase + gl;
            const double *I3[3];
#pragma unroll
            for (int dd = 0; dd < 3; ++dd) {
                if (dd < ndp)
                    I3[dd] = rowval + (size_t)(r0 + (g * ndp + dd) * ncomb + cxyz[dd]) * nqp;
                else
                    I3[dd] = realval + ((size_t)(R0 * nreal + (g * nreal + (dd - ndp)) * ncomb + cxyz[dd])) * nqp;
            }
            const double2 *X = reinterpret_cast<const double2 *>(I3[0]);
            const double2 *Y = reinterpret_cast<const double2 *>(I3[1]);
            const double2 *Z = reinterpret_cast<const double2 *>(I3[2]);
            // eight nodes per step with all of their loads issued before the products
            // (each lane walks its own rows, so the loads, not the arithmetic, set the pace)
            double v0 = 0.0, v1 = 0.0;
            int k2 = 0;
            for (; k2 + 4 <= nq2; k2 += 4) {
                double2 x[4], y[4], z[4], w[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    x[u] = X[k2 + u];
                    y[u] = Y[k2 + u];
                    z[u] = Z[k2 + u];
                    w[u] = tw2[k2 + u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    v0 += (x[u].x * w[u].x) * y[u].x * z[u].x;
                    v1 += (x[u].y * w[u].y) * y[u].y * z[u].y;
                }
            }
            for (; k2 < nq2; ++k2) {
                const double2 x = X[k2], y = Y[k2], z = Z[k2], w = tw2[k2];
                v0 += (x.x * w.x) * y.x * z.x;
                v1 += (x.y * w.y) * y.y * z.y;
            }
            double val = v0 + v1;
            if (nq & 1) val += (I3[0][nq - 1] * tw[nq - 1]) * I3[1][nq - 1] * I3[2][nq - 1];
            const int px = sidx[gl];
            acc += sfac[gl] * cfA[px & 15] * cfB[(px >> 4) & 15] * cfC[(px >> 8) & 15] * cfD[(px >> 12) & 15] * val;
        }
        return acc;
    };
    const bool wide = nblk >= 64;
    const int G = wide ? 1 : 64 / nblk, cpl = lane % nblk, gs = lane / nblk;
    double mine = 0.0;                                    // lane's running sum (fewer than 64 components)
    if (wide)
        for (int cp = lane; cp < nblk; cp += 64) blk[cp] = 0.0;
    for (long gb = 0; gb < ns; gb += 64) {
        const long nc = ns - gb < 64 ? ns - gb : 64;
        wave_sync();                                      // the previous chunk's readers are done
        if (lane < nc) {
            const Surv S = surv[s0 + gb + lane];
            const int ib = S.pb - pb0, ik = S.pk - pk0;
            const int pi = ib / npB, pj = ib - pi * npB, qk = ik / npD, ql = ik - qk * npD;
            sfac[lane] = S.fac;
            sidx[lane] = pi | (pj << 4) | (qk << 8) | (ql << 12);
        }
        wave_sync();
        if (wide) {
            for (int cp = lane; cp < nblk; cp += 64) blk[cp] += comp_sum(cp, gb, 0, nc, 1);
        } else if (gs < G) {
            mine += comp_sum(cpl, gb, gs, nc, G);
        }
    }
    if (!wide) {
        part[lane] = gs < G ? mine : 0.0;
        wave_sync();
        if (lane < nblk) {
            double sum = 0.0;
            for (int j = 0; j < G; ++j) sum += part[lane + j * nblk];
            blk[lane] = sum;
        }
    }
    if (A.counters && lane == 0)
        atomicAdd(&A.counters[2], (unsigned long long)ns * ((unsigned long long)ncomb * nreal * 2 * nq *
                                                                (lA + lB + lC + lD + 1) + 3ull * nq * nblk));
}

// ---------------------------------------------------------------------------
// contraction of a quartet of s and p shells by a group of 16 lanes over the
// nodes, four quartets per wavefront.
// With P p shells (rank j = 0..P-1 in the order A, B, C, D; NBP of them in
// the bra A, B) the power combination of a direction has one bit per p shell
// (set when the shell's component points along that direction): ncomb = 2^P,
// bits 0..NBP-1 the bra's, the rest the ket's; a component quartet is a digit
// string (dir_0 .. dir_{P-1}), dir in {x, y, z}, with
//   cx = sum_j [dir_j = x] 2^j, cy = sum_j [dir_j = y] 2^j, cz = sum_j [dir_j = z] 2^j,
// all known at compile time.  Per surviving primitive quartet the group loads
// the 2^P periodic-direction node rows it needs once, in runs of 16
// consecutive nodes (one 128-byte line per load); the real-direction factors
// (ERI_shell_block_toroidal.f90:302-335) are formed in registers at each node:
//   g_0 = pi sqrt(D) e^{-w R^2}, g_{N+1} = 2 w R g_N - 2 N w g_{N-1},
//   D = 1/(pq + (p+q) t^2), w = pq t^2 D,
//   I(bra bits, ket bits) = sum_n (-1)^n E_n sum_m E'_m g_{n+m}
// with the pairs' Hermite coefficients E, E' (in the order the separate
// real-row kernel sums them).  Ix Iy Iz then go into 3^P accumulators per
// lane; one butterfly sum per component at the end.  The contraction
// coefficients go into the factors: a p shell's coefficient for its component
// along direction d multiplies the direction-d factors whose combination has
// that shell's bit set, so every product Ix Iy Iz carries each p shell's
// coefficient exactly once; the s shells' coefficients, the prefactor and the
// node weight scale the x factors.  A quartet's block is a chain of dependent
// loads (quartet, shells, offsets, survivors, coefficients, rows, then the
// scatter's tables): four quartets per wavefront keep four such chains in
// flight.  The sums run in a fixed order: the block is the same on every run.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int pow3(int p) { return p == 0 ? 1 : 3 * pow3(p - 1); }
__host__ __device__ constexpr int popc(int v) { return v == 0 ? 0 : (v & 1) + popc(v >> 1); }

// Hermite coefficients of one pair and real direction for its p-bits (rank
// order): stored at [(ia (lB+1) + ib) kE + n]; with two p shells ia is bit 0
// and ib bit 1, with one the index is the bit itself.
template <int NB>
__device__ __forceinline__ constexpr int herm_index(int bits) {
    return NB == 2 ? (((bits & 1) << 1) | (bits >> 1)) : bits;
}

// the factors of one real direction at one node for every combination c < 2^P
// (bra bits c & (2^NBP - 1), ket bits c >> NBP)
template <int P, int NBP>
__device__ __forceinline__ void real_factors(double p, double q, double t2, double Rr, const double *__restrict__ Eb,
                                             const double *__restrict__ Ek, double out[1 << P]) {
    constexpr int NKP = P - NBP, kE = 2 * kMaxL + 1;
    const double y = rsqrt_nr(p * q + (p + q) * t2);
    const double D2 = y * y;
    const double w = p * q * t2 * D2;
    const double rs = 3.14159265358979323846 * y;
    double g[P + 1];
    g[0] = (Rr == 0.0) ? rs : rs * exp(-w * Rr * Rr);
    {
        double gm1 = 0.0, gc = g[0];
#pragma unroll
        for (int n = 0; n < P; ++n) {
            const double gp1 = 2.0 * w * Rr * gc - 2.0 * (double)n * w * gm1;
            g[n + 1] = gp1;
            gm1 = gc;
            gc = gp1;
        }
    }
#pragma unroll
    for (int kb = 0; kb < (1 << NKP); ++kb) {
        const double *ek = Ek + herm_index<NKP>(kb) * kE;
        double h[NBP + 1];
#pragma unroll
        for (int n = 0; n <= NBP; ++n) {
            double s = 0.0;
#pragma unroll
            for (int m = 0; m <= popc(kb); ++m) s += ek[m] * g[n + m];
            h[n] = s;
        }
#pragma unroll
        for (int bb = 0; bb < (1 << NBP); ++bb) {
            const double *eb = Eb + herm_index<NBP>(bb) * kE;
            double v = 0.0;
#pragma unroll
            for (int n = 0; n <= popc(bb); ++n) v += ((n & 1) ? -eb[n] : eb[n]) * h[n];
            out[bb | (kb << NBP)] = v;
        }
    }
}

// the block of one s/p quartet by the GL lanes of a group (gl = 0..GL-1): the
// rule's nodes (at most 64) in 64 / GL runs of GL; NDP periodic directions
template <int NDP, int P, int NBP, int GL>
__device__ __forceinline__ void quartet_block_sp(const EriArgs &A, int qi, const Quartet &Q, long s0, long ns,
                                                 long r0, const Surv *__restrict__ surv,
                                                 const double *__restrict__ rowval, double *blk, int gl) {
    constexpr int NC = 1 << P, NT = pow3(P), kE = 2 * kMaxL + 1, NREAL = 3 - NDP;
    const int rule = quartet_rule(A, qi);
    const int nqp = A.nqp, nq = rule_nodes(A, rule);
    const double *tw = A.tweight + (size_t)rule * nqp;
    const double *tn = A.tnode + (size_t)rule * nqp;
    const Shell sA = A.sh[Q.s[0]], sB = A.sh[Q.s[1]], sC = A.sh[Q.s[2]], sD = A.sh[Q.s[3]];
    const int pb0 = A.sp[Q.bra].first, pk0 = A.sp[Q.ket].first;
    const int npB = sB.nprim, npD = sD.nprim;
    // the p shells in rank order: position (0..3 = A..D), direction -> component map,
    // coefficient offset and primitive count
    int ppos[4] = {0, 0, 0, 0}, pof[4] = {0, 0, 0, 0}, pnp[4] = {1, 1, 1, 1};
    unsigned dm[4] = {0, 0, 0, 0};
    {
        const Shell *sh4[4] = {&sA, &sB, &sC, &sD};
        int j = 0;
#pragma unroll
        for (int x = 0; x < 4; ++x)
            if (sh4[x]->l == 1) {
                ppos[j & 3] = x;
                pof[j & 3] = sh4[x]->poff;
                pnp[j & 3] = sh4[x]->nprim;
                dm[j & 3] = sh4[x]->dirmap;
                ++j;
            }
    }
    // the node runs: node (lanes past the rule's last node: the last node, weight 0),
    // weight and t^2
    constexpr int NR = 64 / GL;
    int kr[NR];
    double w[NR], t2[NR];
    const int nrun = (nq + GL - 1) / GL;               // node runs holding the rule's nodes
#pragma unroll
    for (int h = 0; h < NR; ++h) {
        const int k = GL * h + gl;
        kr[h] = k < nq ? k : nq - 1;
        w[h] = k < nq ? tw[kr[h]] : 0.0;
        t2[h] = tn[kr[h]] * tn[kr[h]];
    }
    double acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = 0.0;
    // survivors in batches of GL: lane gl reads survivor gl of the batch and forms its
    // coefficients (s shells and prefactor; each p shell's along x, y, z), exponent
    // sums, real-direction distances and Hermite offsets; the steps take them from
    // that lane of the group (shuffle), U survivors and UH node runs per step with
    // their periodic row loads in flight together (a step past the batch's end
    // repeats its last survivor with coefficient 0)
    constexpr int U = P == 0 ? 2 : 1, UH = P <= 1 ? 4 : P == 2 ? 2 : 1;
    const int gbase = threadIdx.x & (64 - GL);
    constexpr int eb_stride = (1 << NBP) * kE, ek_stride = (1 << (P - NBP)) * kE;   // one real direction
    for (long gb = 0; gb < ns; gb += GL) {
        const int nb = ns - gb < GL ? (int)(ns - gb) : GL;
        double cs_l, cpd_l[3][P > 0 ? P : 1], bq_l = 0.0, kq_l = 0.0, R_l[NREAL > 0 ? NREAL : 1];
        int ebo_l = 0, eko_l = 0;
        {
            const Surv S = surv[s0 + gb + (gl < nb ? gl : nb - 1)];
            const int ib = S.pb - pb0, ik = S.pk - pk0;
            const int pi = ib / npB, pj = ib - pi * npB, qk = ik / npD, ql = ik - qk * npD;
            double cs = gl < nb ? S.fac : 0.0;
            cs *= sA.l == 0 ? A.coef[sA.poff + pi] : 1.0;
            cs *= sB.l == 0 ? A.coef[sB.poff + pj] : 1.0;
            cs *= sC.l == 0 ? A.coef[sC.poff + qk] : 1.0;
            cs *= sD.l == 0 ? A.coef[sD.poff + ql] : 1.0;
            cs_l = cs;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const int pr = ppos[j] == 0 ? pi : ppos[j] == 1 ? pj : ppos[j] == 2 ? qk : ql;
                const double *cj = A.coef + pof[j] + pr;
#pragma unroll
                for (int d = 0; d < 3; ++d) cpd_l[d][j] = cj[(int)((dm[j] >> (2 * d)) & 3) * pnp[j]];
            }
            if (NREAL > 0) {
                const PrimPair &ppb = A.pp[S.pb];
                const PrimPair &ppk = A.pp[S.pk];
                bq_l = ppb.q;
                kq_l = ppk.q;
                ebo_l = ppb.eoff;
                eko_l = ppk.eoff;
#pragma unroll
                for (int drr = 0; drr < NREAL; ++drr)
                    R_l[drr] = ppb.cen[NDP + drr] - (ppk.cen[NDP + drr] + Q.shift[NDP + drr]);
            }
        }
        for (int i0 = 0; i0 < nb; i0 += U) {
            // the U survivors' coefficients, periodic row addresses, exponent sums,
            // distances and Hermite coefficients
            double cs[U], cpd[U][3][P > 0 ? P : 1], pq[U], qq[U], Rr[U][NREAL > 0 ? NREAL : 1];
            const double *rp[U][NDP];
            const double *Eb[U], *Ek[U];
#pragma unroll
            for (int v = 0; v < U; ++v) {
                const int i = i0 + v;                          // < GL: U divides GL
                const int src = gbase + i;
                cs[v] = __shfl(cs_l, src, 64);
#pragma unroll
                for (int j = 0; j < P; ++j)
#pragma unroll
                    for (int d = 0; d < 3; ++d) cpd[v][d][j] = __shfl(cpd_l[d][j], src, 64);
                if (NREAL > 0) {
                    pq[v] = __shfl(bq_l, src, 64);
                    qq[v] = __shfl(kq_l, src, 64);
#pragma unroll
                    for (int drr = 0; drr < NREAL; ++drr) Rr[v][drr] = __shfl(R_l[drr], src, 64);
                    Eb[v] = A.herm + __shfl(ebo_l, src, 64);
                    Ek[v] = A.herm + __shfl(eko_l, src, 64);
                }
                const long g = gb + (i < nb ? i : nb - 1);
#pragma unroll
                for (int dd = 0; dd < NDP; ++dd) rp[v][dd] = rowval + (size_t)(r0 + (g * NDP + dd) * NC) * nqp;
            }
            // UH node runs at a time: every periodic row load first (runs past the rule's
            // last node load its last node, weight 0), then per run the real factors and
            // the products; the runs stop after the rule's last node
            for (int h0 = 0; h0 < NR; h0 += UH) {
                if (h0 >= nrun) break;
                double Xp[U][UH][NDP][NC];
#pragma unroll
                for (int v = 0; v < U; ++v)
#pragma unroll
                    for (int hh = 0; hh < UH; ++hh)
#pragma unroll
                        for (int dd = 0; dd < NDP; ++dd)
#pragma unroll
                            for (int c = 0; c < NC; ++c) Xp[v][hh][dd][c] = rp[v][dd][(size_t)c * nqp + kr[h0 + hh]];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int v = 0; v < U; ++v)
#pragma unroll
                    for (int hh = 0; hh < UH; ++hh) {
                        double F[3][NC];
#pragma unroll
                        for (int dd = 0; dd < NDP; ++dd)
#pragma unroll
                            for (int c = 0; c < NC; ++c) F[dd][c] = Xp[v][hh][dd][c];
#pragma unroll
                        for (int drr = 0; drr < NREAL; ++drr)
                            real_factors<P, NBP>(pq[v], qq[v], t2[h0 + hh], Rr[v][drr], Eb[v] + drr * eb_stride,
                                                 Ek[v] + drr * ek_stride, F[NDP + drr]);
                        const double cw = cs[v] * w[h0 + hh];
#pragma unroll
                        for (int c = 0; c < NC; ++c) F[0][c] *= cw;
                        // p shell j: its direction-d coefficient into the direction-d factors with bit j set
#pragma unroll
                        for (int j = 0; j < P; ++j)
#pragma unroll
                            for (int c = 0; c < NC; ++c)
                                if (c & (1 << j)) {
#pragma unroll
                                    for (int d = 0; d < 3; ++d) F[d][c] *= cpd[v][d][j];
                                }
#pragma unroll
                        for (int t = 0; t < NT; ++t) {
                            int cx = 0, cy = 0, cz = 0, u = t;
#pragma unroll
                            for (int j = 0; j < P; ++j) {
                                const int dir = u % 3;
                                u /= 3;
                                cx |= (dir == 0) << j;
                                cy |= (dir == 1) << j;
                                cz |= (dir == 2) << j;
                            }
                            acc[t] = fma(F[0][cx] * F[1][cy], F[2][cz], acc[t]);
                        }
                    }
            }
        }
    }
    // group sums and the block index of each digit string
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        double v = acc[t];
#pragma unroll
        for (int off = GL / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        int cp = 0, u = t, st = 1;
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int dir = u % 3;
            u /= 3;
            cp += (int)((dm[j] >> (2 * dir)) & 3) * st;
            st *= 3;
        }
        if (gl == 0) blk[cp] = v;
    }
    if (A.counters && gl == 0) {
        const int nreal = NREAL;
        atomicAdd(&A.counters[2], (unsigned long long)ns * ((unsigned long long)NC * nreal * 2 * nq *
                                                                (sA.l + sB.l + sC.l + sD.l + 1) + 3ull * nq * NT));
    }
}

// the stored position (ERI_toroidal.f90:433-449 index order, first index in
// the unit cell) of component quartet f4 under permutation ip
__device__ __forceinline__ void stored_position(const EriArgs &A, const int f4[4], int ip, int pp[4]) {
    int t4[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) t4[x] = f4[c_perm8_rows[ip][x]];
    const int c0 = t4[0] / A.nfuc;
#pragma unroll
    for (int x = 0; x < 4; ++x) pp[x] = A.relc[c0 * A.ncell + t4[x] / A.nfuc] * A.nfuc + t4[x] % A.nfuc;
}

// The components comp, comp + step, comp + 2 step, ... of a block, as digits
// comp = ca + nA (cb + nB (cc + nC cd)): the digits of the start and of the step
// are split once, each further component adds the step's digits with carries
// (no division per component)
struct CompWalk {
    int ca, cb, cc, cd;
    int sa, sb, sc, sd;
    int nA, nB, nC;
    __device__ __forceinline__ CompWalk(int comp, int step, int nA_, int nB_, int nC_) : nA(nA_), nB(nB_), nC(nC_) {
        split(comp, ca, cb, cc, cd);
        split(step, sa, sb, sc, sd);
    }
    __device__ __forceinline__ void split(int v, int &a, int &b, int &c, int &d) const {
        a = v % nA;
        v /= nA;
        b = v % nB;
        v /= nB;
        c = v % nC;
        d = v / nC;
    }
    // each digit stays below twice its radix before its carry, so one subtraction suffices
    __device__ __forceinline__ void next() {
        ca += sa;
        int k = ca >= nA;
        ca -= k ? nA : 0;
        cb += sb + k;
        k = cb >= nB;
        cb -= k ? nB : 0;
        cc += sc + k;
        k = cc >= nC;
        cc -= k ? nC : 0;
        cd += sd + k;
    }
};

// ---------------------------------------------------------------------------
// contraction and output, one wavefront per shell quartet: the block, then
// written out (Schwarz passes) or scattered to the 8 permutations translated
// to the unit cell (src/Toroidal/ERI_toroidal.f90:416-454).
// ---------------------------------------------------------------------------
// the finished block: written out (Schwarz passes) or scattered to the 8
// permutations translated to the unit cell
// (by nl lanes, lane = 0..nl-1)
__device__ __forceinline__ void write_block(const EriArgs &A, const Quartet &Q, const Shell &sA, const Shell &sB,
                                            const Shell &sC, const Shell &sD, const double *blk, int lane,
                                            int nl = 64) {
    const int nA = sA.ncomp, nB = sB.ncomp, nC = sC.ncomp, nD = sD.ncomp;
    const int nblk = nA * nB * nC * nD;
    if (A.blk_out && Q.out >= 0) {
        for (int i = lane; i < nblk; i += nl) A.blk_out[Q.out + i] = blk[i];
    } else if (A.eri) {
        const size_t nb = A.nbas, n1 = A.n1;
        // a lane takes the components lane, lane + nl, ...: the element screen once per
        // component, then its stored position under each of the 8 permutations, that
        // of component (0,0,0,0) plus the component's offsets (a shell lies inside one
        // cell, so the translation is the same for all its components)
        constexpr int perm[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                    {2, 3, 0, 1}, {3, 2, 0, 1}, {2, 3, 1, 0}, {3, 2, 1, 0}};
        const int f0[4] = {sA.first_ao, sB.first_ao, sC.first_ao, sD.first_ao};
        int pb[8][4];
        // the last index's base in the output: its cell's position in the slab (-1:
        // outside it; a shell lies inside one cell, so this holds for all its components)
        int p3b[8];
#pragma unroll
        for (int ip = 0; ip < 8; ++ip) {
            stored_position(A, f0, ip, pb[ip]);
            const int cl = pb[ip][3] / A.nfuc, off = pb[ip][3] - cl * A.nfuc;
            const int lc = A.slab_pos ? A.slab_pos[cl] : cl;
            p3b[ip] = lc >= 0 ? lc * A.nfuc + off : -1;
        }
        CompWalk cw(lane, nl, nA, nB, nC);
        for (int comp = lane; comp < nblk; comp += nl, cw.next()) {
            const int c4[4] = {cw.ca, cw.cb, cw.cc, cw.cd};
            if (A.Qao[(size_t)(f0[0] + c4[0]) + nb * (f0[1] + c4[1])] *
                    A.Qao[(size_t)(f0[2] + c4[2]) + nb * (f0[3] + c4[3])] <
                kSchwarz)
                continue;
            const double v = blk[comp];
#pragma unroll
            for (int ip = 0; ip < 8; ++ip) {
                int pp[4];
#pragma unroll
                for (int x = 0; x < 4; ++x) pp[x] = pb[ip][x] + c4[perm[ip][x]];
                if (p3b[ip] < 0) continue;                                   // outside the slab
                const size_t p3 = (size_t)(p3b[ip] + c4[perm[ip][3]]);
                const size_t idx = A.physicist ? pp[0] + n1 * (pp[2] + nb * (pp[1] + nb * p3))
                                               : pp[0] + n1 * (pp[1] + nb * (pp[2] + nb * p3));
                A.eri[idx] = v;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_contract(EriArgs A, const long long *__restrict__ soff,
                                                  const long long *__restrict__ roff, const Surv *__restrict__ surv,
                                                  const double *__restrict__ rowval,
                                                  const double *__restrict__ realval, const int *__restrict__ list,
                                                  const unsigned *__restrict__ cnt) {
    // four wavefronts per workgroup, one shell quartet each (every quartet, or the
    // cnt entries of list); per wavefront in LDS a block of max_blk and the staging (part_words)
    extern __shared__ double lds[];
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double *blk = lds + (size_t)wv * (A.max_blk + A.part_words), *part = blk + A.max_blk;
    const int n = list ? (int)*cnt : A.nq4;
    for (int e = blockIdx.x * 4 + wv; e < n; e += gridDim.x * 4) {
        const int qi = list ? list[e] : e;
        const Quartet Q = A.q4[qi];
        const Shell &sA = A.sh[Q.s[0]], &sB = A.sh[Q.s[1]], &sC = A.sh[Q.s[2]], &sD = A.sh[Q.s[3]];
        const long s0 = soff[qi], ns = soff[qi + 1] - s0, r0 = roff[qi];
        wave_sync();
        quartet_block(A, Q, quartet_rule(A, qi), s0, ns, r0, surv, rowval, realval, blk, part, lane);
        wave_sync();
        write_block(A, Q, sA, sB, sC, sD, blk, lane);
    }
}

// the quartets into one list per s/p form (sp_list, 0..7) and one (8) of the
// rest (lists[L * nq4 + i], cnt[L] entries): per workgroup of 1024 the
// wavefronts' counts meet in LDS, one atomic per workgroup and list
// (quartets without survivors have a zero block: the outputs start at zero, so
// they join no list)
__global__ __launch_bounds__(1024) void k_sp_classify(EriArgs A, const long long *__restrict__ soff,
                                                      int *__restrict__ lists, unsigned *__restrict__ cnt) {
    __shared__ unsigned wcnt[16][kSpLists], wbase[16][kSpLists];
    const int qi = blockIdx.x * 1024 + threadIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int L = -1;
    if (qi < A.nq4 && soff[qi + 1] > soff[qi]) {
        L = sp_form(A, A.q4[qi], qi);
        if (L < 0) L = kSpLists - 1;
    }
    unsigned long long mm = 0;
#pragma unroll
    for (int p = 0; p < kSpLists; ++p) {
        const unsigned long long m = __ballot(L == p);
        if (L == p) mm = m;
        if (lane == 0) wcnt[wv][p] = (unsigned)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < kSpLists) {
        const int p = threadIdx.x;
        unsigned tot = 0;
        for (int w = 0; w < 16; ++w) {
            wbase[w][p] = tot;
            tot += wcnt[w][p];
        }
        const unsigned b0 = tot ? atomicAdd(&cnt[p], tot) : 0u;
        for (int w = 0; w < 16; ++w) wbase[w][p] += b0;
    }
    __syncthreads();
    if (L >= 0) lists[(size_t)L * A.nq4 + wbase[wv][L] + __popcll(mm & ((1ull << lane) - 1))] = qi;
}

// the s/p quartets of one form (P p shells, NBP in the bra; list of cnt
// entries), GL lanes per quartet, 256 / GL quartets per workgroup; one launch
// per form keeps each kernel's registers to what its form needs
// ---------------------------------------------------------------------------
// Direct J/K (src/module/eri_direct.f90:262-374): every stored element of a
// block once, J(p1,p2) += E P(p3,p4), K(p1,p4) -= E P(p2,p3) / 2.
// "Once": a (component, permutation) pair counts only if no pair earlier in
// the reference's loop order (components a fastest, then permutations 1..8)
// reaches the same stored position -- the hash of the reference
// (:331-359) expressed as a test on the 8 inverse images of the position.
// Four wavefronts per workgroup, one shell quartet each; J and K of the
// workgroup accumulate in LDS (flushed once at the end with global atomics)
// when they fit, else directly in global memory.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool first_generator(const EriArgs &A, const Shell &sA, const Shell &sB, const Shell &sC,
                                                const Shell &sD, int comp, int ip, const int pp[4]) {
    constexpr int perm[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                {2, 3, 0, 1}, {3, 2, 0, 1}, {2, 3, 1, 0}, {3, 2, 1, 0}};
    const int key = comp * 8 + ip;
    const int first[4] = {sA.first_ao, sB.first_ao, sC.first_ao, sD.first_ao};
    const int ncmp[4] = {sA.ncomp, sB.ncomp, sC.ncomp, sD.ncomp};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        int q[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) q[perm[j][x]] = pp[x];         // permutation j maps q to pp
        const int c0 = q[0] / A.nfuc;
        bool inside = true;
        int off[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            const int qt = A.relc[c0 * A.ncell + q[x] / A.nfuc] * A.nfuc + q[x] % A.nfuc;
            off[x] = qt - first[x];
            inside = inside && off[x] >= 0 && off[x] < ncmp[x];
        }
        if (inside) {
            const int c2 = off[0] + ncmp[0] * (off[1] + ncmp[1] * (off[2] + ncmp[2] * off[3]));
            if (c2 * 8 + j < key) return false;
        }
    }
    return true;
}


// True when a permutation other than the identity, with the translation that
// brings its first index to cell 0, maps the shell quartet onto itself: only
// then can two (component, permutation) pairs of the quartet reach one stored
// position, and only then is first_generator needed.
__device__ __forceinline__ bool shell_stabilized(const EriArgs &A, const Shell &sA, const Shell &sB, const Shell &sC,
                                                 const Shell &sD) {
    const int f[4] = {sA.first_ao, sB.first_ao, sC.first_ao, sD.first_ao};
    bool stab = false;
#pragma unroll
    for (int g = 1; g < 8; ++g) {
        int pp[4];
        stored_position(A, f, g, pp);
        stab = stab || (pp[0] == f[0] && pp[1] == f[1] && pp[2] == f[2] && pp[3] == f[3]);
    }
    return stab;
}

// Zero the elements of a finished block that the element screen drops
// (Q(p1,p2) Q(p3,p4) < 1e-13, ERI_toroidal.f90:422), by nl lanes: the J/K
// epilogues then skip zeros instead of screening each (element, permutation).
__device__ __forceinline__ void screen_block(const EriArgs &A, const Shell &sA, const Shell &sB, const Shell &sC,
                                             const Shell &sD, double *blk, int lane, int nl) {
    const int nA = sA.ncomp, nB = sB.ncomp, nC = sC.ncomp, nD = sD.ncomp;
    const int nblk = nA * nB * nC * nD;
    const size_t nb = A.nbas;
    CompWalk cw(lane, nl, nA, nB, nC);
    for (int comp = lane; comp < nblk; comp += nl, cw.next()) {
        if (A.Qao[(size_t)(sA.first_ao + cw.ca) + nb * (sB.first_ao + cw.cb)] *
                A.Qao[(size_t)(sC.first_ao + cw.cc) + nb * (sD.first_ao + cw.cd)] < kSchwarz)
            blk[comp] = 0.0;
    }
}

// J and K of one finished, screened block (the direct path): every counted
// element adds J(p1,p2) += E P(p3,p4), K(p1,p4) -= E P(p2,p3) / 2 into Jl, Kl (LDS
// or global, leading dimension nfuc), by nl lanes (a multiple of 8: a lane keeps
// one permutation, its positions are those of component (0,0,0,0) plus offsets).
__device__ __forceinline__ void jk_block(const EriArgs &A, const Shell &sA, const Shell &sB, const Shell &sC,
                                         const Shell &sD, const double *blk, int lane, int nl, bool stab, double *Jl,
                                         double *Kl) {
    const int nA = sA.ncomp, nB = sB.ncomp, nC = sC.ncomp, nD = sD.ncomp;
    const int nblk = nA * nB * nC * nD;
    const int nf = A.nfuc, nb = A.nbas;
    const int ip = lane & 7;
    const int f0[4] = {sA.first_ao, sB.first_ao, sC.first_ao, sD.first_ao};
    int pb[4], px[4];
    stored_position(A, f0, ip, pb);
#pragma unroll
    for (int x = 0; x < 4; ++x) px[x] = c_perm8_rows[ip][x];
    for (int task = lane; task < nblk * 8; task += nl) {
        const int comp = task >> 3;
        const double val = blk[comp];
        if (val == 0.0) continue;
        int t = comp;
        const int ca = t % nA; t /= nA;
        const int cb = t % nB; t /= nB;
        const int cc = t % nC;
        const int cd = t / nC;
        const int c4[4] = {ca, cb, cc, cd};
        int pp[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) pp[x] = pb[x] + c4[px[x]];
        if (stab && !first_generator(A, sA, sB, sC, sD, comp, ip, pp)) continue;
        atomicAdd(&Jl[pp[0] + (size_t)nf * pp[1]], val * A.Pden[pp[2] + (size_t)nb * pp[3]]);
        atomicAdd(&Kl[pp[0] + (size_t)nf * pp[3]], -0.5 * val * A.Pden[pp[1] + (size_t)nb * pp[2]]);
    }
}

// The same for J and K in global memory, by nl lanes of one quartet: one
// permutation at a time, the counted elements add into the (n0 x n1) J and
// (n0 x n3) K sub-blocks the permutation maps them to (w[0..35] and w[36..71],
// LDS), reading the (n2 x n3) and (n1 x n2) density blocks staged in w[72..143]
// (and the transposed (n2 x n3) block in w[144..179]);
// each sub-block then leaves with one global atomic per entry instead of one per
// element.  Positions: those of component (0,0,0,0), pb, plus the offsets.
__device__ __forceinline__ void jk_block_sub(const EriArgs &A, const Shell &sA, const Shell &sB, const Shell &sC,
                                             const Shell &sD, const double *blk, int lane, int nl, bool stab,
                                             double *w, double *Jg, double *Kg) {
    const int nA = sA.ncomp, nB = sB.ncomp, nC = sC.ncomp, nD = sD.ncomp;
    const int nblk = nA * nB * nC * nD;
    const int nf = A.nfuc, nb = A.nbas;
    const int ncmp[4] = {nA, nB, nC, nD};
    const int f0[4] = {sA.first_ao, sB.first_ao, sC.first_ao, sD.first_ao};
    double *Jw = w, *Kw = w + 36, *Pj = w + 72, *Pk = w + 108, *Pjt = w + 144;
    for (int ip = 0; ip < 8; ++ip) {
        const int x0 = c_perm8_rows[ip][0], x1 = c_perm8_rows[ip][1], x2 = c_perm8_rows[ip][2],
                  x3 = c_perm8_rows[ip][3];
        const int n0 = ncmp[x0], n1 = ncmp[x1], n2 = ncmp[x2], n3 = ncmp[x3];
        // J terms of permutations ip and ip + 2 (ip & 2 == 0) meet in one J element (see
        // k_contract_jk): without stabilizers the first adds both, the second none
        const bool jpair = !stab && (ip & 2) == 0, jskip = !stab && (ip & 2) != 0;
        int pb[4];
        stored_position(A, f0, ip, pb);
        for (int e = lane; e < 36; e += nl) {
            Jw[e] = 0.0;
            Kw[e] = 0.0;
            if (e < n2 * n3) {
                const int a2 = pb[2] + e / n3, a3 = pb[3] + e % n3;
                Pj[e] = A.Pden[a2 + (size_t)nb * a3];
                Pjt[e] = A.Pden[a3 + (size_t)nb * a2];
            }
            if (e < n1 * n2) Pk[e] = A.Pden[(pb[1] + e / n2) + (size_t)nb * (pb[2] + e % n2)];
        }
        wave_sync();
        for (int comp = lane; comp < nblk; comp += nl) {
            const double val = blk[comp];
            if (val == 0.0) continue;
            int t = comp;
            const int ca = t % nA; t /= nA;
            const int cb = t % nB; t /= nB;
            const int cc = t % nC;
            const int cd = t / nC;
            const int c4[4] = {ca, cb, cc, cd};
            const int i0 = c4[x0], i1 = c4[x1], i2 = c4[x2], i3 = c4[x3];
            if (stab) {
                const int pp[4] = {pb[0] + i0, pb[1] + i1, pb[2] + i2, pb[3] + i3};
                if (!first_generator(A, sA, sB, sC, sD, comp, ip, pp)) continue;
            }
            if (jpair)
                atomicAdd(&Jw[i0 * n1 + i1], val * (Pj[i2 * n3 + i3] + Pjt[i2 * n3 + i3]));
            else if (!jskip)
                atomicAdd(&Jw[i0 * n1 + i1], val * Pj[i2 * n3 + i3]);
            atomicAdd(&Kw[i0 * n3 + i3], -0.5 * val * Pk[i1 * n2 + i2]);
        }
        wave_sync();
        if (!jskip)
            for (int e = lane; e < n0 * n1; e += nl) {
                const double v = Jw[e];
                if (v != 0.0) atomicAdd(&Jg[(pb[0] + e / n1) + (size_t)nf * (pb[1] + e % n1)], v);
            }
        for (int e = lane; e < n0 * n3; e += nl) {
            const double v = Kw[e];
            if (v != 0.0) atomicAdd(&Kg[(pb[0] + e / n3) + (size_t)nf * (pb[3] + e % n3)], v);
        }
        wave_sync();
    }
}

// JK = 0: the block goes to the ERI array (write_block); JK = 1: its J and K
// contributions (jk_block), into LDS copies of J and K flushed at the end when
// jk_lds, else straight into the global arrays.
template <int NDP, int P, int NBP, int GL, int JK>
__global__ __launch_bounds__(256) void k_contract_sp(EriArgs A, const int *__restrict__ list,
                                                     const unsigned *__restrict__ cnt,
                                                     const long long *__restrict__ soff,
                                                     const long long *__restrict__ roff,
                                                     const Surv *__restrict__ surv, const double *__restrict__ rowval,
                                                     int jk_lds) {
    constexpr int NG = 256 / GL;
    __shared__ double lds[NG * pow3(P)];
    __shared__ double jkw[JK ? NG : 1][JK ? 180 : 1];   // per-group J/K sub-blocks (global J/K)
    extern __shared__ double jkl[];
    const int grp = threadIdx.x / GL, gl = threadIdx.x & (GL - 1);
    double *blk = lds + grp * pow3(P);
    const size_t njk = (size_t)A.nfuc * A.nbas;
    double *Jl = (JK && jk_lds) ? jkl : A.Jacc, *Kl = (JK && jk_lds) ? jkl + njk : A.Kacc;
    if (JK && jk_lds) {
        for (size_t i = threadIdx.x; i < 2 * njk; i += 256) jkl[i] = 0.0;
        __syncthreads();
    }
    const long n = *cnt;
    for (long e0 = (long)blockIdx.x * NG; e0 < n; e0 += (long)gridDim.x * NG) {
        const long e = e0 + grp;
        const int qi = e < n ? list[e] : 0;
        const Quartet Q = A.q4[qi];
        if (e < n) {
            const long s0 = soff[qi], ns = soff[qi + 1] - s0, r0 = roff[qi];
            quartet_block_sp<NDP, P, NBP, GL>(A, qi, Q, s0, ns, r0, surv, rowval, blk, gl);
        }
        wave_sync();
        if (e < n) {
            const Shell &sA = A.sh[Q.s[0]], &sB = A.sh[Q.s[1]], &sC = A.sh[Q.s[2]], &sD = A.sh[Q.s[3]];
            if (JK) {
                screen_block(A, sA, sB, sC, sD, blk, gl, GL);
                wave_sync();
                const bool stab = shell_stabilized(A, sA, sB, sC, sD);
                if (jk_lds) jk_block(A, sA, sB, sC, sD, blk, gl, GL, stab, Jl, Kl);
                else jk_block_sub(A, sA, sB, sC, sD, blk, gl, GL, stab, jkw[JK ? grp : 0], Jl, Kl);
            } else {
                write_block(A, Q, sA, sB, sC, sD, blk, gl, GL);
            }
        }
        wave_sync();
    }
    if (JK && jk_lds) {
        __syncthreads();
        for (size_t i = threadIdx.x; i < njk; i += 256) {
            if (jkl[i] != 0.0) atomicAdd(&A.Jacc[i], jkl[i]);
            if (jkl[njk + i] != 0.0) atomicAdd(&A.Kacc[i], jkl[njk + i]);
        }
    }
}

// TCG_JK_GLOBAL=1 forces the global-memory J/K accumulation (the path of systems
// whose J/K rows exceed the LDS) so that it can be tested on small systems; read
// at every launch, so a test can switch it within one process
static bool jk_force_global() {
    const char *e = getenv("TCG_JK_GLOBAL");
    return e && e[0] == '1';
}

// J/K copies in LDS per workgroup up to this size (TCG_SP_JK_LDS_KB, default 150
// KiB: LDS atomics beat global ones even at one workgroup per CU); larger ones
// are accumulated in global memory
static size_t sp_jk_lds_bytes() {
    static const size_t v = [] {
        const char *e = getenv("TCG_SP_JK_LDS_KB");
        const long kb = e ? atol(e) : 150;
        return (size_t)(kb < 0 ? 0 : kb) * 1024;
    }();
    return v;
}

template <int NDP, int P, int NBP, int GL, int JK>
static void launch_sp1(const EriArgs &b, const int *lists, const unsigned *cnt, const long long *soff,
                       const long long *roff, const Surv *surv, const double *rowval, hipStream_t st) {
    constexpr int L = sp_list(P, NBP);
    const size_t jk_bytes = JK ? sizeof(double) * 2 * (size_t)b.nfuc * b.nbas : 0;
    const int jk_lds = JK && jk_bytes <= sp_jk_lds_bytes() && !jk_force_global() ? 1 : 0;
    int ng = (b.nq4 + 256 / GL - 1) / (256 / GL);
    ng = ng < (1 << 20) ? ng : (1 << 20);
    if (jk_lds) {
        // persistent: each workgroup flushes its copy once
        const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, (160 * 1024) / (jk_bytes + 4096)));
        ng = ng < 256 * per_cu ? ng : 256 * per_cu;
        if (jk_bytes > 64 * 1024)
            (void)hipFuncSetAttribute((const void *)k_contract_sp<NDP, P, NBP, GL, JK>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)jk_bytes);
    }
    hipLaunchKernelGGL((k_contract_sp<NDP, P, NBP, GL, JK>), dim3(ng), dim3(256), jk_lds ? jk_bytes : 0, st, b,
                       lists + (size_t)L * b.nq4, cnt + L, soff, roff, surv, rowval, jk_lds);
}

template <int NDP, int JK>
static void launch_sp_all(const EriArgs &b, const int *lists, const unsigned *cnt, const long long *soff,
                          const long long *roff, const Surv *surv, const double *rowval, hipStream_t st) {
    launch_sp1<NDP, 0, 0, 16, JK>(b, lists, cnt, soff, roff, surv, rowval, st);
    launch_sp1<NDP, 1, 0, 16, JK>(b, lists, cnt, soff, roff, surv, rowval, st);
    launch_sp1<NDP, 1, 1, 16, JK>(b, lists, cnt, soff, roff, surv, rowval, st);
    launch_sp1<NDP, 2, 0, 16, JK>(b, lists, cnt, soff, roff, surv, rowval, st);
    launch_sp1<NDP, 2, 1, 16, JK>(b, lists, cnt, soff, roff, surv, rowval, st);
    launch_sp1<NDP, 2, 2, 16, JK>(b, lists, cnt, soff, roff, surv, rowval, st);
    launch_sp1<NDP, 3, 1, 16, JK>(b, lists, cnt, soff, roff, surv, rowval, st);
    launch_sp1<NDP, 3, 2, 16, JK>(b, lists, cnt, soff, roff, surv, rowval, st);
}

size_t contract_lds_bytes(const EriArgs &a) { return 4 * sizeof(double) * ((size_t)a.max_blk + a.part_words); }

__global__ __launch_bounds__(256) void k_contract_jk(EriArgs A, const long long *__restrict__ soff,
                                                     const long long *__restrict__ roff,
                                                     const Surv *__restrict__ surv, const double *__restrict__ rowval,
                                                     const double *__restrict__ realval, int jk_lds,
                                                     const int *__restrict__ list, const unsigned *__restrict__ cnt) {
    // LDS: [4][max_blk] blocks, [4][5][sub_words] per-wave J/K sub-blocks and the P
    // sub-blocks they read, [4][part_words] the contraction's staging, then
    // J[nfuc*nbas], K[nfuc*nbas] when jk_lds
    extern __shared__ double lds[];
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nf = A.nfuc, nb = A.nbas;
    const size_t njk = (size_t)nf * nb;
    double *blk = lds + (size_t)wv * A.max_blk;
    const int sw = A.sub_words;
    double *Jw = lds + 4 * (size_t)A.max_blk + 5 * sw * wv, *Kw = Jw + sw, *Pj = Jw + 2 * sw, *Pk = Jw + 3 * sw,
           *Pjt = Jw + 4 * sw;
    double *part = lds + 4 * (size_t)A.max_blk + 20 * (size_t)sw + (size_t)A.part_words * wv;
    double *Jl = jk_lds ? lds + 4 * (size_t)A.max_blk + 20 * (size_t)sw + 4 * (size_t)A.part_words : A.Jacc;
    double *Kl = jk_lds ? Jl + njk : A.Kacc;
    if (jk_lds) {
        for (size_t i = threadIdx.x; i < 2 * njk; i += 256) Jl[i] = 0.0;
    }
    __syncthreads();
    const int nlist = list ? (int)*cnt : A.nq4;
    for (int e = blockIdx.x * 4 + wv; e < nlist; e += gridDim.x * 4) {
        const int qi = list ? list[e] : e;
        const Quartet Q = A.q4[qi];
        const Shell &sA = A.sh[Q.s[0]], &sB = A.sh[Q.s[1]], &sC = A.sh[Q.s[2]], &sD = A.sh[Q.s[3]];
        const int nA = sA.ncomp, nB = sB.ncomp, nC = sC.ncomp, nD = sD.ncomp;
        const int nblk = nA * nB * nC * nD;
        const long s0 = soff[qi], ns = soff[qi + 1] - s0, r0 = roff[qi];
        wave_sync();
        quartet_block(A, Q, quartet_rule(A, qi), s0, ns, r0, surv, rowval, realval, blk, part, lane);
        wave_sync();
        screen_block(A, sA, sB, sC, sD, blk, lane, 64);
        wave_sync();
        const bool stab = shell_stabilized(A, sA, sB, sC, sD);
        if (jk_lds) {
            // J/K in LDS: every counted element adds there directly
            jk_block(A, sA, sB, sC, sD, blk, lane, 64, stab, Jl, Kl);
            continue;
        }
        // J/K in global memory: one permutation at a time, the counted elements of the
        // block add into the (n0 x n1) J and (n0 x n3) K sub-blocks the permutation maps
        // them to (LDS), reading the (n2 x n3) and (n1 x n2) sub-blocks of P they need
        // (staged in LDS first); the sub-blocks then go out with one atomic per entry.
        // A component's stored position is that of component (0,0,0,0), pb, plus its
        // offsets (a shell lies inside one cell).
        // (the permutation loop unrolled: each permutation's index selections are fixed;
        // the lane's components walk by digits, split once per quartet)
        constexpr int perm[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                    {2, 3, 0, 1}, {3, 2, 0, 1}, {2, 3, 1, 0}, {3, 2, 1, 0}};
        const int ncmp[4] = {nA, nB, nC, nD};
        const int f0[4] = {sA.first_ao, sB.first_ao, sC.first_ao, sD.first_ao};
        const CompWalk cw0(lane, 64, nA, nB, nC);
#pragma unroll
        for (int ip = 0; ip < 8; ++ip) {
            const int x0 = perm[ip][0], x1 = perm[ip][1], x2 = perm[ip][2], x3 = perm[ip][3];
            const int n0 = ncmp[x0], n1 = ncmp[x1], n2 = ncmp[x2], n3 = ncmp[x3];
            // Permutations ip and ip + 2 (ip & 2 == 0: (ab|cd) and (ab|dc), (ba|cd) and
            // (ba|dc), ...) store at positions that differ only by swapping the last two
            // indices, so their J terms go to one J element: E P(p3,p4) + E P(p4,p3).
            // Without stabilizers every (component, permutation) counts, and the first
            // of the pair adds both terms (pjt: the transposed density block); with
            // stabilizers each permutation adds its own, as counted.
            const bool jpair = !stab && (ip & 2) == 0, jskip = !stab && (ip & 2) != 0;
            int pb[4];
            stored_position(A, f0, ip, pb);
            if (!jskip)
                for (int e2 = lane; e2 < n0 * n1; e2 += 64) Jw[e2] = 0.0;
            for (int e2 = lane; e2 < n0 * n3; e2 += 64) Kw[e2] = 0.0;
            if (!jskip)
                for (int e2 = lane; e2 < n2 * n3; e2 += 64) {
                    const int a2 = pb[2] + e2 / n3, a3 = pb[3] + e2 % n3;
                    Pj[e2] = A.Pden[a2 + (size_t)nb * a3];
                    if (jpair) Pjt[e2] = A.Pden[a3 + (size_t)nb * a2];
                }
            for (int e2 = lane; e2 < n1 * n2; e2 += 64)
                Pk[e2] = A.Pden[(pb[1] + e2 / n2) + (size_t)nb * (pb[2] + e2 % n2)];
            wave_sync();
            CompWalk cw = cw0;
            for (int comp = lane; comp < nblk; comp += 64, cw.next()) {
                const double val = blk[comp];
                if (val == 0.0) continue;
                const int c4[4] = {cw.ca, cw.cb, cw.cc, cw.cd};
                const int i0 = c4[x0], i1 = c4[x1], i2 = c4[x2], i3 = c4[x3];
                if (stab) {
                    const int pp[4] = {pb[0] + i0, pb[1] + i1, pb[2] + i2, pb[3] + i3};
                    if (!first_generator(A, sA, sB, sC, sD, comp, ip, pp)) continue;
                }
                if (jpair)
                    atomicAdd(&Jw[i0 * n1 + i1], val * (Pj[i2 * n3 + i3] + Pjt[i2 * n3 + i3]));
                else if (!jskip)
                    atomicAdd(&Jw[i0 * n1 + i1], val * Pj[i2 * n3 + i3]);
                atomicAdd(&Kw[i0 * n3 + i3], -0.5 * val * Pk[i1 * n2 + i2]);
            }
            wave_sync();
            if (!jskip)
                for (int e2 = lane; e2 < n0 * n1; e2 += 64) {
                    const double v = Jw[e2];
                    if (v != 0.0) atomicAdd(&Jl[(pb[0] + e2 / n1) + (size_t)nf * (pb[1] + e2 % n1)], v);
                }
            for (int e2 = lane; e2 < n0 * n3; e2 += 64) {
                const double v = Kw[e2];
                if (v != 0.0) atomicAdd(&Kl[(pb[0] + e2 / n3) + (size_t)nf * (pb[3] + e2 % n3)], v);
            }
            wave_sync();
        }
    }
    if (jk_lds) {
        __syncthreads();
        for (size_t i = threadIdx.x; i < njk; i += 256) {
            if (Jl[i] != 0.0) atomicAdd(&A.Jacc[i], Jl[i]);
            if (Kl[i] != 0.0) atomicAdd(&A.Kacc[i], Kl[i]);
        }
    }
}

hipError_t launch_contract_jk(const EriArgs &a, const long long *soff, const long long *roff, const Surv *surv,
                              const double *rowval, const double *realval, const int *list, const unsigned *cnt,
                              hipStream_t st) {
    if (a.nq4 <= 0) return hipSuccess;
    const size_t blk_bytes = sizeof(double) * (4 * (size_t)a.max_blk + 20 * (size_t)a.sub_words + 4 * (size_t)a.part_words);
    const size_t jk_bytes = sizeof(double) * 2 * (size_t)a.nfuc * a.nbas;
    const size_t lds_max = 160 * 1024;
    const bool force_global = jk_force_global();
    const int jk_lds = (blk_bytes + jk_bytes <= lds_max && !force_global) ? 1 : 0;
    const size_t lds = blk_bytes + (jk_lds ? jk_bytes : 0);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_contract_jk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    // persistent workgroups: as many as fit on the 256 CUs at this LDS size
    const int per_cu = jk_lds ? (int)std::max<size_t>(1, std::min<size_t>(4, lds_max / lds)) : 4;
    const int want = (a.nq4 + 3) / 4;
    const int grid = std::min(want, 256 * per_cu);
    hipLaunchKernelGGL(k_contract_jk, dim3(grid), dim3(256), lds, st, a, soff, roff, surv, rowval, realval, jk_lds,
                       list, cnt);
    return hipGetLastError();
}

hipError_t launch_contract(const EriArgs &a, const long long *soff, const long long *roff, const Surv *surv,
                           const double *rowval, double *realval, int *lists, unsigned *cnt, hipStream_t st) {
    if (a.nq4 <= 0) return hipSuccess;
    const int nwg = (a.nq4 + 3) / 4;
    const int grid = nwg < (1 << 20) ? nwg : (1 << 20);
    const size_t lds = contract_lds_bytes(a);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)k_contract, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    // short quartet lists (the Schwarz passes: few quartets, up to nprim^4 survivors
    // each) stay with the general kernel, one wavefront per quartet
    EriArgs b = a;
    if (b.nq4 < b.sp_min_quartets) b.sp_contract = 0;
    if (b.sp_contract) {
        // s/p quartets: lists by form, then one kernel per form (grids for the largest list)
        hipError_t e = hipMemsetAsync(cnt, 0, kSpLists * sizeof(unsigned), st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_sp_classify, dim3((a.nq4 + 1023) / 1024), dim3(1024), 0, st, b, soff, lists, cnt);
        // real-direction rows only for the quartets of the general kernel (the last list)
        const int *rest = lists + (size_t)(kSpLists - 1) * b.nq4;
        e = real_rows(b, soff, roff, surv, realval, rest, cnt + (kSpLists - 1), st);
        if (e != hipSuccess) return e;
        const bool jk = b.Jacc != nullptr;
        if (b.ndp == 1) jk ? launch_sp_all<1, 1>(b, lists, cnt, soff, roff, surv, rowval, st)
                           : launch_sp_all<1, 0>(b, lists, cnt, soff, roff, surv, rowval, st);
        else if (b.ndp == 2) jk ? launch_sp_all<2, 1>(b, lists, cnt, soff, roff, surv, rowval, st)
                                : launch_sp_all<2, 0>(b, lists, cnt, soff, roff, surv, rowval, st);
        else jk ? launch_sp_all<3, 1>(b, lists, cnt, soff, roff, surv, rowval, st)
                : launch_sp_all<3, 0>(b, lists, cnt, soff, roff, surv, rowval, st);
        // the rest (d shells, four p shells) from the last list (its length is on the
        // device: at most 4096 workgroups walk it)
        if (jk) return launch_contract_jk(b, soff, roff, surv, rowval, realval, rest, cnt + (kSpLists - 1), st);
        hipLaunchKernelGGL(k_contract, dim3(grid < 4096 ? grid : 4096), dim3(256), lds, st, b, soff, roff, surv,
                           rowval, realval, rest, cnt + (kSpLists - 1));
        return hipGetLastError();
    }
    {
        const hipError_t e = real_rows(b, soff, roff, surv, realval, nullptr, nullptr, st);
        if (e != hipSuccess) return e;
    }
    if (b.Jacc) return launch_contract_jk(b, soff, roff, surv, rowval, realval, nullptr, nullptr, st);
    hipLaunchKernelGGL(k_contract, dim3(grid), dim3(256), lds, st, b, soff, roff, surv, rowval, realval, nullptr,
                       nullptr);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// chunking of long quartet lists
// ---------------------------------------------------------------------------
// bounds[k] = first quartet q with roff[q] >= k * cap (k < nchunk), bounds[nchunk] = nq4
__global__ void k_chunk_bounds(const long long *__restrict__ roff, int nq4, long long cap, int nchunk,
                               int *__restrict__ bounds) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nchunk) return;
    if (k == nchunk) { bounds[k] = nq4; return; }
    const long long target = (long long)k * cap;
    int lo = 0, hi = nq4;                       // roff[0..nq4] is non-decreasing
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (roff[mid] < target) lo = mid + 1; else hi = mid;
    }
    bounds[k] = lo;
}

// dst[i] = src[i] - src[0], i = 0..n-1
__global__ void k_rebase(const long long *__restrict__ src, int n, long long *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] - src[0];
}

hipError_t launch_chunk_bounds(const long long *roff, int nq4, long long cap, int nchunk, int *bounds, hipStream_t st) {
    hipLaunchKernelGGL(k_chunk_bounds, dim3((nchunk + 1 + 63) / 64), dim3(64), 0, st, roff, nq4, cap, nchunk, bounds);
    return hipGetLastError();
}

hipError_t launch_rebase(const long long *src, int n, long long *dst, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rebase, dim3((n + 255) / 256), dim3(256), 0, st, src, n, dst);
    return hipGetLastError();
}

// owner[g] = the shell quartet of survivor g
__global__ __launch_bounds__(64) void k_surv_quartet(const long long *__restrict__ soff, int nq4, int *__restrict__ sq) {
    for (int qi = blockIdx.x; qi < nq4; qi += gridDim.x)
        for (long long g = soff[qi] + threadIdx.x; g < soff[qi + 1]; g += 64) sq[g] = qi;
}

hipError_t launch_surv_owner(const long long *soff, int nq4, int *owner, hipStream_t st) {
    if (nq4 <= 0) return hipSuccess;
    const int grid = nq4 < (1 << 20) ? nq4 : (1 << 20);
    hipLaunchKernelGGL(k_surv_quartet, dim3(grid), dim3(64), 0, st, soff, nq4, owner);
    return hipGetLastError();
}

}  // namespace tcg
