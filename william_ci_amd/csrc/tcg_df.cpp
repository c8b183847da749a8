// Density fitting (the reference's module df_toroidal, keyword
// density_fitting): df_init, df_JK, df_kMO, with the integrals from the
// engine and the linear algebra on the GPU.
//
//   (mu nu|la si) ~ sum_Q B^Q_{mu nu} B^Q_{la si},  B = (mu nu|P) [V^-1/2]_{PQ}
//
// * The 3- and 2-index integrals are quartets of the engine on a combined
//   basis (per cell: the AOs, the auxiliary functions, one exponent-0 s
//   function = the constant 1):  (mu nu|P) = (mu nu|P 1),  (P|Q) = (P 1|Q 1)
//   (df_toroidal.f90:555-637, dummy_at :837-844).
// * V is block-circulant over the cells: V(k) = sum_R (P0|Q R) e^{ikR}, a
//   pivoted Cholesky per k (tolerance 1e-10) gives f(k) = (L^-1)^H, and
//   B(R') = 1/N Re sum_k [sum_RP T3(RP) e^{ik RP}] f(k) e^{-ik R'}
//   (df_toroidal.f90:570-700); every product is an FP64 matrix-core GEMM.
// * df_JK (:850-957): J through gamma_Q; K in k space,
//   Kt(k) = -1/(2N) sum_{k1,Q} Bh_Q(k1, k-k1) P(k1) Bh_Q(k1, k-k1)^H, back to cells.
// * df_kMO (:981-1112): L^Q_pq(k1,k2) and (ov|ov) = 1/N sum_Q L L, then the
//   k-MP2 sums of tcg_kmp2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/tcg.h"
#include "tcg_devbuf.hpp"
#include "tcg_device.hpp"

using namespace tcg;

namespace tcg {
hipError_t launch_metric_factor(const double *Vre, const double *Vim, int n, int nk, const int *mate, double tol,
                                double2 *work, double *fre, double *fim, int *rank, double *dmin, hipStream_t st);
hipError_t launch_df_gamma(const double *B, const double *Ps, int nf, int nb, int nau, int N, double *G,
                           hipStream_t st);
hipError_t launch_df_jvec(const double *B, const double *gam, int nf, int nb, int naux, double *part, double *J,
                          hipStream_t st);
hipError_t launch_df_l(const double *Fr, const double *Fi, const double2 *Ck, int nf, int nau, int N, int o_hi,
                       int v_lo, const int *cdiff, double2 *L, hipStream_t st);
hipError_t launch_df_mp2_f32(const double *Fr, const double *Fi, const double2 *Ck, int nf, int nau, int N, int o_hi,
                             int v_lo, const int *cdiff, const int *cadd, const double *ek, const int *occ,
                             float2 *L, float2 *eri, float *part, int nblocks, unsigned long long *nzero,
                             hipStream_t st);
hipError_t launch_df_eri(const double2 *L, int nau, int N, int o_hi, int nv, const int *cdiff, const int *cadd,
                         double2 *out, hipStream_t st);
hipError_t launch_df_kt(const double *Fr, const double *Fi, const double2 *Pk, int nf, int nau, int N, const int *cdiff,
                        double2 *Y, double2 *Kt, hipStream_t st);
}  // namespace tcg

namespace {

constexpr double kMetricTol = 1e-10;     // pivoted Cholesky of (P|Q)(k), Gaussian aux basis (df_toroidal.f90:68)
constexpr double kDfSchwarz = 1e-13;     // df_schwarz (df_toroidal.f90:70)

}  // namespace

struct tcg_df {
    tcg_ctx *comb = nullptr;             // engine context on the combined basis
    int ndp = 1, nf = 0, nb = 0, nau = 0, N = 1, nfc = 0, ncl[3] = {1, 1, 1};
    std::vector<int> cdiff, cadd;
    DevBuf<int> d_cdiff, d_cadd;
    DevBuf<double> d_B;                    // df_B(nfuc, nBas, naux), column-major
    DevBuf<double> d_jpart;                // partial sums of the J pass
    // F(a, b0, k, Q, kap) = sum_{Rb, RQ} df_B(a, b0 + Rb, Q RQ) e^{ik Rb} e^{-i kap RQ} (real, imaginary)
    DevBuf<double> d_Fr, d_Fi;
    bool f_ready = false;
    DevBuf<double2> d_Pk, d_Y, d_Kt;       // exchange work arrays (kept between SCF iterations)
    std::vector<int> rank;
    double dmin = 0.0;
    double times[4] = {0, 0, 0, 0};
    std::string err;
    hipStream_t st = nullptr;
    int fail(int code, const std::string &m) {
        err = m;
        return code;
    }
};

#define DCHK(df, expr)                                                                              \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return (df)->fail(TCG_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static double now_ms_df() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// e^{+i k.R} (cells and k points x fastest): cos and sin as (R, k) column-major
static void phase_tables(const int ncl[3], int N, double scale, std::vector<double> &c, std::vector<double> &s) {
    c.assign((size_t)N * N, 0.0);
    s.assign((size_t)N * N, 0.0);
    for (int k = 0; k < N; ++k)
        for (int R = 0; R < N; ++R) {
            const double th = 2.0 * M_PI *
                              ((double)(k % ncl[0]) * (R % ncl[0]) / ncl[0] +
                               (double)((k / ncl[0]) % ncl[1]) * ((R / ncl[0]) % ncl[1]) / ncl[1] +
                               (double)(k / (ncl[0] * ncl[1])) * (R / (ncl[0] * ncl[1])) / ncl[2]);
            c[R + (size_t)N * k] = scale * std::cos(th);
            s[R + (size_t)N * k] = scale * std::sin(th);
        }
}

// complex product of real/imaginary parts: C = A B (+ C when acc), batched
static hipError_t zgemm_split(int M, int Nn, int K, const double *Ar, const double *Ai, long lda, long long sA,
                              const double *Br, const double *Bi, long ldb, long long sB, double *Cr, double *Ci,
                              long ldc, long long sC, long long batch, bool conjB, hipStream_t st) {
    const double sb = conjB ? -1.0 : 1.0;
    hipError_t e = launch_dgemm_nn(M, Nn, K, 1.0, Ar, lda, sA, Br, ldb, sB, 0.0, Cr, ldc, sC, batch, st);
    if (e == hipSuccess) e = launch_dgemm_nn(M, Nn, K, -sb, Ai, lda, sA, Bi, ldb, sB, 1.0, Cr, ldc, sC, batch, st);
    if (e == hipSuccess) e = launch_dgemm_nn(M, Nn, K, sb, Ar, lda, sA, Bi, ldb, sB, 0.0, Ci, ldc, sC, batch, st);
    if (e == hipSuccess) e = launch_dgemm_nn(M, Nn, K, 1.0, Ai, lda, sA, Br, ldb, sB, 1.0, Ci, ldc, sC, batch, st);
    return e;
}

// F of df_L_kap (df_toroidal.f90:1071-1090) for every kap at once, kept on the GPU: the
// exchange (Bh(k1,k2) = F(k1, kap = -k2), :898-953) and the MP2 L vectors read it.
static int build_F(tcg_df *df) {
    if (df->f_ready) return TCG_OK;
    const int nf = df->nf, N = df->N, nau = df->nau, naux = nau * N;
    std::vector<double> cphi, sphi;
    phase_tables(df->ncl, N, 1.0, cphi, sphi);
    DevBuf<double> dC, dS, Y1r, Y1i;
    DCHK(df, dC.upload(cphi));
    DCHK(df, dS.upload(sphi));
    // df_B viewed (a, b0, Rb, Q, RQ): first Rb -> k, batched over (Q, RQ)
    const size_t nff = (size_t)nf * nf, nY = nff * N * naux;
    DCHK(df, Y1r.alloc(nY));
    DCHK(df, Y1i.alloc(nY));
    DCHK(df, launch_dgemm_nn((int)nff, N, N, 1.0, df->d_B.p, (long)nff, (long long)(nff * N), dC.p, N, 0, 0.0, Y1r.p,
                             (long)nff, (long long)(nff * N), naux, df->st));
    DCHK(df, launch_dgemm_nn((int)nff, N, N, 1.0, df->d_B.p, (long)nff, (long long)(nff * N), dS.p, N, 0, 0.0, Y1i.p,
                             (long)nff, (long long)(nff * N), naux, df->st));
    // then RQ -> kap with e^{-i kap RQ}: rows (a, b0, k, Q), columns RQ
    DCHK(df, df->d_Fr.alloc(nY));
    DCHK(df, df->d_Fi.alloc(nY));
    const long rows = (long)(nff * N * nau);
    DCHK(df, zgemm_split((int)rows, N, N, Y1r.p, Y1i.p, rows, 0, dC.p, dS.p, N, 0, df->d_Fr.p, df->d_Fi.p, rows, 0, 1,
                         true, df->st));
    DCHK(df, hipStreamSynchronize(df->st));
    df->f_ready = true;
    return TCG_OK;
}

extern "C" {

int tcg_df_create(tcg_ctx *ctx, int nau_cell, const double *aux_center, const int *aux_lxyz,
                  const int *aux_prim_off, const double *aux_expo, const double *aux_coef, tcg_df **out) {
    if (!ctx || !out || nau_cell <= 0 || nau_cell > 256 || !aux_center || !aux_lxyz || !aux_prim_off || !aux_expo ||
        !aux_coef)
        return TCG_E_ARG;
    *out = nullptr;
    CtxView v{};
    if (ctx_view(ctx, &v) != TCG_OK || v.nbas == 0) return TCG_E_STATE;
    tcg_df *df = new tcg_df();
    df->ndp = v.ndp;
    df->nf = v.nfuc;
    df->nb = v.nbas;
    df->nau = nau_cell;
    for (int d = 0; d < 3; ++d) df->ncl[d] = v.ncell[d];
    const int N = v.ncell[0] * v.ncell[1] * v.ncell[2];
    df->N = N;
    const int nf = v.nfuc, nau = nau_cell, nfc = nf + nau + 1;
    df->nfc = nfc;
    if (hipStreamCreate(&df->st) != hipSuccess) {
        delete df;
        return TCG_E_DEVICE;
    }
    *out = df;
    const double t0 = now_ms_df();
    // ---- combined basis, cell by cell: AOs, auxiliary functions, the constant 1 ----
    std::vector<double> center, expo, coef;
    std::vector<int> lxyz, poff(1, 0);
    auto add = [&](const double *xyz, const int *l, const double *e, const double *cf, int np) {
        center.insert(center.end(), xyz, xyz + 3);
        lxyz.insert(lxyz.end(), l, l + 3);
        expo.insert(expo.end(), e, e + np);
        coef.insert(coef.end(), cf, cf + np);
        poff.push_back(poff.back() + np);
    };
    const int zero3[3] = {0, 0, 0};
    const double e0 = 0.0, c1 = 1.0;
    for (int c = 0; c < N; ++c) {
        for (int i = 0; i < nf; ++i) {
            const int g = c * nf + i;
            add(v.center + 3 * g, v.lxyz + 3 * g, v.expo + v.poff[g], v.coef + v.poff[g], v.poff[g + 1] - v.poff[g]);
        }
        for (int p = 0; p < nau; ++p) {
            const int g = c * nau + p;
            add(aux_center + 3 * g, aux_lxyz + 3 * g, aux_expo + aux_prim_off[g], aux_coef + aux_prim_off[g],
                aux_prim_off[g + 1] - aux_prim_off[g]);
        }
        add(aux_center + 3 * (c * nau), zero3, &e0, &c1, 1);          // dummy_at: centre of the cell's aux
    }
    int rc = tcg_create(v.ndp, v.L, v.ncell, v.fast ? v.nq : 32, &df->comb);
    if (rc) return df->fail(rc, "tcg_create (combined basis) failed");
    if (!v.fast && (rc = tcg_set_quadrature(df->comb, 0, v.nq))) return df->fail(rc, tcg_last_error(df->comb));
    rc = tcg_set_basis(df->comb, N * nfc, nfc, center.data(), lxyz.data(), poff.data(), expo.data(), coef.data());
    if (rc) return df->fail(rc, std::string("combined basis: ") + tcg_last_error(df->comb));
    // cell arithmetic
    df->cdiff.assign((size_t)N * N, 0);
    df->cadd.assign((size_t)N * N, 0);
    for (int a = 0; a < N; ++a)
        for (int b = 0; b < N; ++b) {
            const int ax = a % df->ncl[0], ay = (a / df->ncl[0]) % df->ncl[1], az = a / (df->ncl[0] * df->ncl[1]);
            const int bx = b % df->ncl[0], by = (b / df->ncl[0]) % df->ncl[1], bz = b / (df->ncl[0] * df->ncl[1]);
            auto idx = [&](int x, int y, int z) {
                return ((x % df->ncl[0] + df->ncl[0]) % df->ncl[0]) +
                       df->ncl[0] * (((y % df->ncl[1] + df->ncl[1]) % df->ncl[1]) +
                                     df->ncl[1] * ((z % df->ncl[2] + df->ncl[2]) % df->ncl[2]));
            };
            df->cdiff[(size_t)a * N + b] = idx(ax - bx, ay - by, az - bz);
            df->cadd[(size_t)a * N + b] = idx(ax + bx, ay + by, az + bz);
        }
    DCHK(df, df->d_cdiff.upload(df->cdiff));
    DCHK(df, df->d_cadd.upload(df->cadd));
    // shells of the combined basis: kind by position inside the cell
    std::vector<int> sfirst, sncomp, scell;
    if (ctx_shells(df->comb, sfirst, sncomp, scell)) return df->fail(TCG_E_STATE, "combined shells");
    std::vector<int> ao_sh, aux_sh, dum_of_cell(N, -1);
    for (size_t s = 0; s < sfirst.size(); ++s) {
        const int loc = sfirst[s] % nfc;
        if (loc < nf) ao_sh.push_back((int)s);
        else if (loc < nf + nau) aux_sh.push_back((int)s);
        else dum_of_cell[scell[s]] = (int)s;
    }
    auto aux_index = [&](int s, int x) {          // (p, RP) of component x of aux shell s
        const int f = sfirst[s] + x, c = f / nfc;
        return (f % nfc - nf) + nau * c;
    };
    auto ao_index = [&](int s, int x) {
        const int f = sfirst[s] + x;
        return (f % nfc) + nf * (f / nfc);
    };
    std::vector<long long> off;
    std::vector<double> vals;
    auto blocks = [&](const std::vector<int> &q4) -> int {
        const int n = (int)q4.size() / 4;
        off.assign(n + 1, 0);
        size_t tot = 0;
        for (int x = 0; x < n; ++x)
            tot += (size_t)sncomp[q4[4 * x]] * sncomp[q4[4 * x + 1]] * sncomp[q4[4 * x + 2]] * sncomp[q4[4 * x + 3]];
        vals.assign(tot, 0.0);
        return n ? tcg_eri_blocks(df->comb, n, q4.data(), vals.data(), off.data()) : 0;
    };
    // ---- metric (P0|Q R) and Qa ----
    const int naux = nau * N;
    std::vector<double> Vm((size_t)nau * naux, 0.0);
    {
        std::vector<int> q4;
        for (int sp : aux_sh)
            if (scell[sp] == 0)
                for (int sq : aux_sh) {
                    q4.push_back(sp);
                    q4.push_back(dum_of_cell[0]);
                    q4.push_back(sq);
                    q4.push_back(dum_of_cell[scell[sq]]);
                }
        if ((rc = blocks(q4))) return df->fail(rc, std::string("metric: ") + tcg_last_error(df->comb));
        for (size_t x = 0; x < q4.size() / 4; ++x) {
            const int sp = q4[4 * x], sq = q4[4 * x + 2];
            for (int b = 0; b < sncomp[sq]; ++b)
                for (int a = 0; a < sncomp[sp]; ++a)
                    Vm[(size_t)aux_index(sp, a) + (size_t)nau * aux_index(sq, b)] = vals[off[x] + a + (size_t)sncomp[sp] * b];
        }
    }
    std::vector<double> Qa(naux);
    for (int c = 0; c < N; ++c)
        for (int p = 0; p < nau; ++p) Qa[(size_t)c * nau + p] = std::sqrt(std::fabs(Vm[(size_t)p + (size_t)nau * p]));
    // ---- V(k) and its factor f(k) on the GPU ----
    std::vector<double> cphi, sphi;
    phase_tables(df->ncl, N, 1.0, cphi, sphi);
    DevBuf<double> dC, dS, dVm, dVr, dVi, dfr, dfi, ddmin;
    DevBuf<int> drank, dmate;
    std::vector<int> mate(N);
    for (int k = 0; k < N; ++k) mate[k] = df->cdiff[k];          // -k (cdiff row 0: 0 - k)
    DCHK(df, dmate.upload(mate));
    DCHK(df, dC.upload(cphi));
    DCHK(df, dS.upload(sphi));
    DCHK(df, dVm.upload(Vm));
    DCHK(df, dVr.alloc((size_t)nau * nau * N));
    DCHK(df, dVi.alloc((size_t)nau * nau * N));
    // Vm as (nau*nau) x N (rows (p, q), columns RQ) times e^{+ik RQ}
    DCHK(df, launch_dgemm_nn(nau * nau, N, N, 1.0, dVm.p, nau * nau, 0, dC.p, N, 0, 0.0, dVr.p, nau * nau, 0, 1, df->st));
    DCHK(df, launch_dgemm_nn(nau * nau, N, N, 1.0, dVm.p, nau * nau, 0, dS.p, N, 0, 0.0, dVi.p, nau * nau, 0, 1, df->st));
    DCHK(df, dfr.alloc((size_t)nau * nau * N));
    DCHK(df, dfi.alloc((size_t)nau * nau * N));
    DCHK(df, drank.alloc(N));
    DCHK(df, ddmin.alloc(N));
    DevBuf<double2> dwork;
    if (nau > 64) DCHK(df, dwork.alloc(2 * (size_t)nau * nau * N));
    DCHK(df, launch_metric_factor(dVr.p, dVi.p, nau, N, dmate.p, kMetricTol, dwork.p, dfr.p, dfi.p, drank.p, ddmin.p,
                                  df->st));
    df->rank.assign(N, 0);
    std::vector<double> dm(N);
    DCHK(df, hipMemcpyAsync(df->rank.data(), drank.p, sizeof(int) * N, hipMemcpyDeviceToHost, df->st));
    DCHK(df, hipMemcpyAsync(dm.data(), ddmin.p, sizeof(double) * N, hipMemcpyDeviceToHost, df->st));
    DCHK(df, hipStreamSynchronize(df->st));
    df->dmin = *std::min_element(dm.begin(), dm.end());
    const double t1 = now_ms_df();
    // ---- Schwarz of the AO pairs (mu0 nu|mu0 nu) ----
    std::vector<int> ao0;
    for (int s : ao_sh)
        if (scell[s] == 0) ao0.push_back(s);
    std::vector<double> Qsh(ao0.size() * ao_sh.size(), 0.0);
    {
        std::vector<int> q4;
        for (int sa : ao0)
            for (int sb : ao_sh) {
                q4.push_back(sa);
                q4.push_back(sb);
                q4.push_back(sa);
                q4.push_back(sb);
            }
        if ((rc = blocks(q4))) return df->fail(rc, std::string("Schwarz: ") + tcg_last_error(df->comb));
        for (size_t x = 0; x < q4.size() / 4; ++x) {
            const int na = sncomp[q4[4 * x]], nb = sncomp[q4[4 * x + 1]];
            double m = 0.0;
            for (int a = 0; a < na; ++a)
                for (int b = 0; b < nb; ++b)
                    m = std::max(m, std::sqrt(std::fabs(vals[off[x] + a + (size_t)na * (b + (size_t)nb * (a + (size_t)na * b))])));
            Qsh[x] = m;
        }
    }
    // ---- (mu0 nu|P) ----
    const size_t nx = (size_t)nf * df->nb, m1 = nx * nau;
    std::vector<double> T3(nx * naux, 0.0);
    {
        std::vector<double> qsp(aux_sh.size(), 0.0);
        for (size_t s = 0; s < aux_sh.size(); ++s)
            for (int x = 0; x < sncomp[aux_sh[s]]; ++x) qsp[s] = std::max(qsp[s], Qa[aux_index(aux_sh[s], x)]);
        std::vector<int> q4;
        for (size_t ia = 0; ia < ao0.size(); ++ia)
            for (size_t ib = 0; ib < ao_sh.size(); ++ib)
                for (size_t ip = 0; ip < aux_sh.size(); ++ip) {
                    if (Qsh[ia * ao_sh.size() + ib] * qsp[ip] < kDfSchwarz) continue;
                    const int sp = aux_sh[ip];
                    q4.push_back(ao0[ia]);
                    q4.push_back(ao_sh[ib]);
                    q4.push_back(sp);
                    q4.push_back(dum_of_cell[scell[sp]]);
                }
        if ((rc = blocks(q4))) return df->fail(rc, std::string("3-index: ") + tcg_last_error(df->comb));
        for (size_t x = 0; x < q4.size() / 4; ++x) {
            const int sa = q4[4 * x], sb = q4[4 * x + 1], sp = q4[4 * x + 2];
            const int na = sncomp[sa], nb = sncomp[sb];
            for (int p = 0; p < sncomp[sp]; ++p)
                for (int b = 0; b < nb; ++b)
                    for (int a = 0; a < na; ++a)
                        T3[(size_t)ao_index(sa, a) + (size_t)nf * ao_index(sb, b) + nx * aux_index(sp, p)] =
                            vals[off[x] + a + (size_t)na * (b + (size_t)nb * p)];
        }
    }
    const double t2 = now_ms_df();
    // ---- B = T3 V^-1/2 as a convolution over cells ----
    {
        DevBuf<double> dT3, dAr, dAi, dWr, dWi;
        DCHK(df, dT3.upload(T3));
        DCHK(df, dAr.alloc(m1 * N));
        DCHK(df, dAi.alloc(m1 * N));
        // T3 is a (nfuc nBas nau) x N matrix (aux index (p, RP)): Z(k) = sum_RP T3(RP) e^{ik RP}
        DCHK(df, launch_dgemm_nn((int)m1, N, N, 1.0, dT3.p, (long)m1, 0, dC.p, N, 0, 0.0, dAr.p, (long)m1, 0, 1, df->st));
        DCHK(df, launch_dgemm_nn((int)m1, N, N, 1.0, dT3.p, (long)m1, 0, dS.p, N, 0, 0.0, dAi.p, (long)m1, 0, 1, df->st));
        dT3.release();
        // W(k) = Z(k) f(k): batched over k, Z(k) is (nfuc nBas) x nau
        DCHK(df, dWr.alloc(m1 * N));
        DCHK(df, dWi.alloc(m1 * N));
        DCHK(df, zgemm_split((int)nx, nau, nau, dAr.p, dAi.p, (long)nx, (long long)m1, dfr.p, dfi.p, nau,
                             (long long)nau * nau, dWr.p, dWi.p, (long)nx, (long long)m1, N, false, df->st));
        // df_B(R') = 1/N sum_k [Re W(k) cos(k R') + Im W(k) sin(k R')]   (cos, sin symmetric in k, R)
        DCHK(df, df->d_B.alloc(m1 * N));
        DCHK(df, launch_dgemm_nn((int)m1, N, N, 1.0 / N, dWr.p, (long)m1, 0, dC.p, N, 0, 0.0, df->d_B.p, (long)m1, 0,
                                 1, df->st));
        DCHK(df, launch_dgemm_nn((int)m1, N, N, 1.0 / N, dWi.p, (long)m1, 0, dS.p, N, 0, 1.0, df->d_B.p, (long)m1, 0,
                                 1, df->st));
        DCHK(df, hipStreamSynchronize(df->st));
    }
    const double t3 = now_ms_df();
    df->times[0] = t1 - t0;
    df->times[1] = t2 - t1;
    df->times[2] = t3 - t2;
    return TCG_OK;
}

int tcg_df_info(tcg_df *df, long long *out, double *dmin) {
    if (!df || !out) return TCG_E_ARG;
    out[0] = df->nau;
    out[1] = (long long)df->nau * df->N;
    int mn = df->nau;
    for (int r : df->rank) mn = std::min(mn, r);
    out[2] = df->nau - mn;                // aux functions dropped (max over k)
    if (dmin) *dmin = df->dmin;
    return TCG_OK;
}

int tcg_df_b(tcg_df *df, double *B) {
    if (!df || !B) return TCG_E_ARG;
    const size_t n = (size_t)df->nf * df->nb * df->nau * df->N;
    DCHK(df, hipMemcpy(B, df->d_B.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return TCG_OK;
}

int tcg_df_jk(tcg_df *df, int nbas, const double *P, double *J, double *K) {
    if (!df) return TCG_E_ARG;
    if (nbas != df->nb || !P || !J || !K) return df->fail(TCG_E_ARG, "tcg_df_jk: bad arguments");
    const int nf = df->nf, nb = df->nb, N = df->N, nau = df->nau, naux = nau * N;
    const size_t nx = (size_t)nf * nb;
    // Ps(x, R) = P(la0 + R, s' + R) for x = (la0, s') (translation of the density rows)
    std::vector<double> Ps(nx * N);
    for (int R = 0; R < N; ++R)
        for (int s = 0; s < nb; ++s) {
            const int sc = s / nf, sl = s % nf;
            const int col = df->cadd[(size_t)sc * N + R] * nf + sl;
            for (int l0 = 0; l0 < nf; ++l0)
                Ps[(size_t)l0 + (size_t)nf * s + nx * R] = P[(size_t)(R * nf + l0) + (size_t)nb * col];
        }
    DevBuf<double> dPs, dG, dgam, dJ;
    DCHK(df, dPs.upload(Ps));
    DCHK(df, dG.alloc((size_t)naux * N));
    DCHK(df, launch_df_gamma(df->d_B.p, dPs.p, nf, nb, nau, N, dG.p, df->st));
    std::vector<double> G((size_t)naux * N), gam(naux, 0.0);
    DCHK(df, hipMemcpyAsync(G.data(), dG.p, sizeof(double) * G.size(), hipMemcpyDeviceToHost, df->st));
    DCHK(df, hipStreamSynchronize(df->st));
    // gamma(q, RQ + R) += G((q, RQ), R)
    for (int R = 0; R < N; ++R)
        for (int rq = 0; rq < N; ++rq)
            for (int q = 0; q < nau; ++q)
                gam[(size_t)df->cadd[(size_t)rq * N + R] * nau + q] += G[(size_t)rq * nau + q + (size_t)naux * R];
    DCHK(df, dgam.upload(gam));
    DCHK(df, dJ.alloc(nx));
    DCHK(df, df->d_jpart.alloc(kDfJChunks * (size_t)nx));  // the J pass's partial sums
    DCHK(df, launch_df_jvec(df->d_B.p, dgam.p, nf, nb, naux, df->d_jpart.p, dJ.p, df->st));
    // K in k space (df_toroidal.f90:893-954): P(k) = sum_R P(0, R) e^{ikR} (its Hermitian part),
    // Kt(kk) = -1/(2N) sum_{k1,Q} M P(k1) M^H with M = Bh_Q(k1, kk - k1) = F(k1, Q, kap = k1 - kk),
    // K(mu0, s0 + R) = 1/N Re sum_kk Kt(kk) e^{-i kk R}.  Only the first nfuc rows of P enter.
    int rc = build_F(df);
    if (rc) return rc;
    std::vector<double> cphi, sphi;
    phase_tables(df->ncl, N, 1.0, cphi, sphi);
    const size_t nff = (size_t)nf * nf;
    std::vector<double2> Pk(nff * N, double2{0.0, 0.0}), Kt(nff * N);
    for (int k = 0; k < N; ++k) {
        double2 *pk = Pk.data() + nff * k;
        for (int R = 0; R < N; ++R) {
            const double c = cphi[R + (size_t)N * k], sn = sphi[R + (size_t)N * k];
            for (int b = 0; b < nf; ++b)
                for (int a = 0; a < nf; ++a) {
                    const double p = P[(size_t)a + (size_t)nb * (R * nf + b)];
                    pk[a + nf * b].x += p * c;
                    pk[a + nf * b].y += p * sn;
                }
        }
        for (int b = 0; b < nf; ++b)
            for (int a = 0; a <= b; ++a) {
                const double2 x = pk[a + nf * b], y = pk[b + nf * a];
                pk[a + nf * b] = double2{0.5 * (x.x + y.x), 0.5 * (x.y - y.y)};
                pk[b + nf * a] = double2{pk[a + nf * b].x, -pk[a + nf * b].y};
            }
    }
    DCHK(df, df->d_Pk.upload(Pk));
    DCHK(df, df->d_Y.alloc(nff * nau * N * N));
    DCHK(df, df->d_Kt.alloc(nff * N));
    DCHK(df, launch_df_kt(df->d_Fr.p, df->d_Fi.p, df->d_Pk.p, nf, nau, N, df->d_cdiff.p, df->d_Y.p, df->d_Kt.p, df->st));
    DCHK(df, hipMemcpyAsync(Kt.data(), df->d_Kt.p, sizeof(double2) * Kt.size(), hipMemcpyDeviceToHost, df->st));
    std::vector<double> Jr(nx);
    DCHK(df, hipMemcpyAsync(Jr.data(), dJ.p, sizeof(double) * nx, hipMemcpyDeviceToHost, df->st));
    DCHK(df, hipStreamSynchronize(df->st));
    std::fill(J, J + (size_t)nb * nb, 0.0);
    std::fill(K, K + (size_t)nb * nb, 0.0);
    for (int c = 0; c < nb; ++c)
        for (int r = 0; r < nf; ++r) J[r + (size_t)nb * c] = Jr[r + (size_t)nf * c];
    for (int R = 0; R < N; ++R)
        for (int b = 0; b < nf; ++b)
            for (int a = 0; a < nf; ++a) {
                double v = 0.0;
                for (int k = 0; k < N; ++k) {
                    const double2 kt = Kt[nff * k + a + nf * b];
                    v += kt.x * cphi[R + (size_t)N * k] + kt.y * sphi[R + (size_t)N * k];
                }
                K[a + (size_t)nb * (R * nf + b)] = v / N;
            }
    return TCG_OK;
}

static int df_kmp2(tcg_df *df, int nbas, const double *coeff_sao, const double *e, const int *occ, bool fp32,
                   double *out) {
    if (!df) return TCG_E_ARG;
    if (nbas != df->nb || !coeff_sao || !e || !occ || !out) return df->fail(TCG_E_ARG, "tcg_df_kmp2: bad arguments");
    const int nf = df->nf, nb = df->nb, N = df->N, nau = df->nau, naux = nau * N;
    const double t0 = now_ms_df();
    // k blocks and the (ov|ov) ranges (src/module/cc_dims.f90:68-88)
    std::vector<double2> Ck((size_t)N * nf * nf);
    int o_hi = 0, omin = nf;
    for (int k = 0; k < N; ++k) {
        int nok = 0;
        for (int p = 0; p < nf; ++p) {
            const int o = occ[(size_t)k * nf + p];
            if (o != 0 && o != 2) return df->fail(TCG_E_ARG, "tcg_df_kmp2: occupation must be 0 or 2");
            if (o == 2) {
                if (p != nok) return df->fail(TCG_E_ARG, "tcg_df_kmp2: occupied bands are not the lowest ones");
                ++nok;
            }
        }
        o_hi = std::max(o_hi, nok);
        omin = std::min(omin, nok);
        for (int col = 0; col < nf; ++col)
            for (int row = 0; row < nf; ++row) {
                const size_t g = (size_t)(k * nf + row) + (size_t)nb * (k * nf + col);
                Ck[(size_t)k * nf * nf + row + (size_t)nf * col] = double2{coeff_sao[2 * g], coeff_sao[2 * g + 1]};
            }
    }
    const int v_lo = omin, nv = nf - v_lo;
    if (o_hi == 0 || nv <= 0) return df->fail(TCG_E_ARG, "tcg_df_kmp2: no occupied or no virtual bands");
    int rc = build_F(df);
    if (rc) return rc;
    DevBuf<double2> dCk, dL, dE;
    DCHK(df, dCk.upload(Ck));
    const size_t nL = (size_t)o_hi * nv * nau * N * N;
    const size_t nE = (size_t)o_hi * nv * o_hi * nv * N * N * N;
    // the k-MP2 sums (kMP2_calc, src/Energies/SAO/MP2/MP2_SAO.f90:6-151)
    std::vector<double> ek((size_t)N * nf);
    for (size_t i = 0; i < ek.size(); ++i) ek[i] = e[i];
    std::vector<int> oc(occ, occ + (size_t)N * nf);
    DevBuf<double> de, dpart;
    DevBuf<int> docc;
    DevBuf<unsigned long long> dnz;
    DCHK(df, de.upload(ek));
    DCHK(df, docc.upload(oc));
    const int nblk = 1024;
    if (fp32) {
        // FP32: L, the (ov|ov) GEMMs on the FP32 matrix cores, compensated FP32 sums
        DevBuf<float2> dL32, dE32;
        DevBuf<float> dp32;
        DCHK(df, dL32.alloc(nL));
        DCHK(df, dE32.alloc(nE));
        DCHK(df, dp32.alloc(4 * nblk));
        DCHK(df, dnz.alloc(1));
        DCHK(df, hipMemsetAsync(dnz.p, 0, sizeof(unsigned long long), df->st));
        DCHK(df, launch_df_mp2_f32(df->d_Fr.p, df->d_Fi.p, dCk.p, nf, nau, N, o_hi, v_lo, df->d_cdiff.p, df->d_cadd.p,
                                   de.p, docc.p, dL32.p, dE32.p, dp32.p, nblk, dnz.p, df->st));
        std::vector<float> part(4 * nblk);
        unsigned long long nzero = 0;
        DCHK(df, hipMemcpyAsync(part.data(), dp32.p, sizeof(float) * 4 * nblk, hipMemcpyDeviceToHost, df->st));
        DCHK(df, hipMemcpyAsync(&nzero, dnz.p, sizeof nzero, hipMemcpyDeviceToHost, df->st));
        DCHK(df, hipStreamSynchronize(df->st));
        // the blocks' (hi, lo) pairs meet in FP64 on the host
        double E2d = 0.0, E2x = 0.0;
        for (int b = 0; b < nblk; ++b) {
            E2d += (double)part[4 * b] + (double)part[4 * b + 1];
            E2x += (double)part[4 * b + 2] + (double)part[4 * b + 3];
        }
        out[0] = 2.0 * E2d - E2x;
        out[1] = 2.0 * E2d;
        out[2] = -E2x;
        out[3] = (double)nzero;
        out[4] = now_ms_df() - t0;
        return TCG_OK;
    }
    DCHK(df, dL.alloc(nL));
    DCHK(df, launch_df_l(df->d_Fr.p, df->d_Fi.p, dCk.p, nf, nau, N, o_hi, v_lo, df->d_cdiff.p, dL.p, df->st));
    DCHK(df, dE.alloc(nE));
    DCHK(df, launch_df_eri(dL.p, nau, N, o_hi, nv, df->d_cdiff.p, df->d_cadd.p, dE.p, df->st));
    DCHK(df, dpart.alloc(2 * nblk));
    DCHK(df, dnz.alloc(1));
    DCHK(df, hipMemsetAsync(dnz.p, 0, sizeof(unsigned long long), df->st));
    DCHK(df, launch_kmp2_sums(dE.p, de.p, docc.p, nf, N, df->ncl[0], df->ncl[1], df->ncl[2], o_hi, v_lo, dpart.p, nblk,
                              dnz.p, df->st));
    std::vector<double> part(2 * nblk);
    unsigned long long nzero = 0;
    DCHK(df, hipMemcpyAsync(part.data(), dpart.p, sizeof(double) * 2 * nblk, hipMemcpyDeviceToHost, df->st));
    DCHK(df, hipMemcpyAsync(&nzero, dnz.p, sizeof nzero, hipMemcpyDeviceToHost, df->st));
    DCHK(df, hipStreamSynchronize(df->st));
    double E2d = 0.0, E2x = 0.0;
    for (int b = 0; b < nblk; ++b) {
        E2d += part[2 * b];
        E2x += part[2 * b + 1];
    }
    out[0] = 2.0 * E2d - E2x;
    out[1] = 2.0 * E2d;
    out[2] = -E2x;
    out[3] = (double)nzero;
    out[4] = now_ms_df() - t0;
    return TCG_OK;
}

int tcg_df_kmp2(tcg_df *df, int nbas, const double *coeff_sao, const double *e, const int *occ, double *out) {
    return df_kmp2(df, nbas, coeff_sao, e, occ, false, out);
}

int tcg_df_kmp2_fp32(tcg_df *df, int nbas, const double *coeff_sao, const double *e, const int *occ, double *out) {
    return df_kmp2(df, nbas, coeff_sao, e, occ, true, out);
}

int tcg_df_times(tcg_df *df, double *out) {
    if (!df || !out) return TCG_E_ARG;
    for (int i = 0; i < 3; ++i) out[i] = df->times[i];
    return TCG_OK;
}

const char *tcg_df_last_error(tcg_df *df) { return df ? df->err.c_str() : "null DF context"; }

int tcg_df_destroy(tcg_df *df) {
    if (!df) return TCG_E_ARG;
    if (df->comb) tcg_destroy(df->comb);
    if (df->st) (void)hipStreamDestroy(df->st);
    delete df;
    return TCG_OK;
}

}  // extern "C"
