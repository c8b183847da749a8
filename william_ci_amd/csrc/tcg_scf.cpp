// Host side of the k-space (Bloch, "SAO") RHF on the GPU: the AO <-> SAO
// transforms, the block eigensolver and the whole RHF_SAO iteration
// (src/Energies/SAO/RHF_SAO.f90:1-553) with every matrix resident on the device.
// One iteration is J/K (stored or integral-direct), the k-space Fock blocks,
// error vector, energy sums and DIIS products, one read-back of a few scalars
// for the convergence test and the DIIS weights, then the per-k eigenproblems,
// occupations and the new density.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/tcg.h"
#include "tcg_devbuf.hpp"
#include "tcg_device.hpp"

namespace tcg {
void ctx_set_error(tcg_ctx *c, const char *msg);
}

using namespace tcg;

namespace {

constexpr int kStallIt = 5;           // stall rule of RHF_SAO.f90:303-320
constexpr double kStallConv = 1e-6;
constexpr double kStallDE = 1e-10;

double wall_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Lat {
    int ncl[3] = {1, 1, 1};
    int N = 1, nf = 0, nb = 0;
};

int fail(tcg_ctx *c, int code, const std::string &msg) {
    ctx_set_error(c, msg.c_str());
    return code;
}

#define SCHK(ctx, expr)                                                                         \
    do {                                                                                        \
        const hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                                   \
            return fail(ctx, TCG_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int lattice_of(tcg_ctx *c, int nbas, Lat &L, const char *who) {
    CtxView v{};
    if (ctx_view(c, &v) != TCG_OK || v.nbas == 0) return fail(c, TCG_E_STATE, std::string(who) + ": no basis set");
    for (int d = 0; d < 3; ++d) L.ncl[d] = v.ncell[d];
    L.N = L.ncl[0] * L.ncl[1] * L.ncl[2];
    L.nf = v.nfuc;
    L.nb = v.nbas;
    if (nbas != L.nb) return fail(c, TCG_E_ARG, std::string(who) + ": nbas differs from the basis set");
    if (L.nf * L.N != L.nb) return fail(c, TCG_E_LATTICE, std::string(who) + ": nbas != nfuc N_cell");
    if (L.nf > kMaxBlock)
        return fail(c, TCG_E_ARG, std::string(who) + ": nfuc above " + std::to_string(kMaxBlock) + " (k block in LDS)");
    return TCG_OK;
}

// phase(k,R) = exp(+i k.R), the angle reduced as in SAO_Transfer.f90:50-57
std::vector<double2> phase_table(const Lat &L) {
    const int N = L.N, nx = L.ncl[0], ny = L.ncl[1], nz = L.ncl[2];
    std::vector<double2> ph((size_t)N * N);
    const double two_pi = 2.0 * M_PI;
    for (int k = 0; k < N; ++k) {
        const int kx = k % nx, ky = (k / nx) % ny, kz = k / (nx * ny);
        for (int R = 0; R < N; ++R) {
            const int rx = R % nx, ry = (R / nx) % ny, rz = R / (nx * ny);
            double th = two_pi * ((double)kx * rx / nx + (double)ky * ry / ny + (double)kz * rz / nz);
            th -= two_pi * std::round(th / two_pi);
            ph[(size_t)k * N + R] = double2{std::cos(th), std::sin(th)};
        }
    }
    return ph;
}

// block k of a full (nbas, nbas) complex matrix (re, im pairs, column-major)
void gather_blocks(const Lat &L, const double *full, std::vector<double2> &blk) {
    const int nf = L.nf, nb = L.nb;
    blk.assign((size_t)L.N * nf * nf, double2{0.0, 0.0});
    for (int k = 0; k < L.N; ++k)
        for (int b = 0; b < nf; ++b)
            for (int a = 0; a < nf; ++a) {
                const size_t g = (size_t)(k * nf + a) + (size_t)nb * (k * nf + b);
                blk[(size_t)k * nf * nf + a + (size_t)nf * b] = double2{full[2 * g], full[2 * g + 1]};
            }
}

void scatter_blocks(const Lat &L, const std::vector<double2> &blk, double *full, bool zero_rest) {
    const int nf = L.nf, nb = L.nb;
    if (zero_rest) std::memset(full, 0, sizeof(double) * 2 * (size_t)nb * nb);
    for (int k = 0; k < L.N; ++k)
        for (int b = 0; b < nf; ++b)
            for (int a = 0; a < nf; ++a) {
                const size_t g = (size_t)(k * nf + a) + (size_t)nb * (k * nf + b);
                const double2 v = blk[(size_t)k * nf * nf + a + (size_t)nf * b];
                full[2 * g] = v.x;
                full[2 * g + 1] = v.y;
            }
}

// Solves the DIIS system of DIIS_extrapolation (src/Energies/DIIS_extrapolation.f90:1-52):
// [B -1; -1 0] [w; l] = [0; -1].  False (history reset) when the matrix is not
// finite, singular or its reciprocal 1-norm condition number is below 1e-14.
bool diis_solve(const std::vector<double> &B, int m, std::vector<double> &w) {
    const int n = m + 1;
    std::vector<double> A((size_t)n * n, 0.0), inv((size_t)n * n, 0.0);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) A[i * n + j] = B[(size_t)i * m + j];
    for (int i = 0; i < m; ++i) A[i * n + m] = A[m * n + i] = -1.0;
    for (double x : A)
        if (!std::isfinite(x)) return false;
    double anorm = 0.0;                          // 1-norm: largest column sum
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += std::fabs(A[i * n + j]);
        anorm = std::max(anorm, s);
    }
    for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
    std::vector<double> G = A;                   // Gauss-Jordan with partial pivoting
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(G[r * n + col]) > std::fabs(G[piv * n + col])) piv = r;
        if (G[piv * n + col] == 0.0) return false;
        if (piv != col)
            for (int j = 0; j < n; ++j) {
                std::swap(G[piv * n + j], G[col * n + j]);
                std::swap(inv[piv * n + j], inv[col * n + j]);
            }
        const double d = 1.0 / G[col * n + col];
        for (int j = 0; j < n; ++j) {
            G[col * n + j] *= d;
            inv[col * n + j] *= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const double f = G[r * n + col];
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j) {
                G[r * n + j] -= f * G[col * n + j];
                inv[r * n + j] -= f * inv[col * n + j];
            }
        }
    }
    double inorm = 0.0;
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += std::fabs(inv[i * n + j]);
        inorm = std::max(inorm, s);
    }
    const double rcond = 1.0 / (anorm * inorm);
    if (!std::isfinite(rcond) || rcond < 1e-14) return false;
    w.assign(m, 0.0);
    for (int i = 0; i < m; ++i) w[i] = -inv[i * n + m];      // b = (0, .., 0, -1)
    return true;
}

// Device buffers of the SCF loop, kept in the context between calls (grown as
// needed), with the pinned read-back area, the J/K timing events and the
// phase table of the lattice they were built for.
struct ScfWork {
    DevBuf<double> dTt, dVt, dH, dP, dJ, dK, dC, dCerr, dmax, dpart, dsc, de, dw, dsmin, dtmp;
    DevBuf<double> dPprev, dDP, dJacc, dKacc;    // incremental direct builds: last full-built density, J, K
    DevBuf<double2> dph, dSk, dXk, dPk, dErrH, dFH, dFx, dCk;
    DevBuf<int> ddead, docc;
    double *hs = nullptr;
    hipEvent_t ej0 = nullptr, ej1 = nullptr;
    int ph_key[4] = {0, 0, 0, 0};                // (nx, ny, nz) of the phase table
    ~ScfWork() {
        if (hs) (void)hipHostFree(hs);
        if (ej0) (void)hipEventDestroy(ej0);
        if (ej1) (void)hipEventDestroy(ej1);
    }
};

int scf_work(tcg_ctx *c, ScfWork *&w) {
    std::shared_ptr<void> &slot = ctx_scf_work(c);
    if (!slot) {
        auto sw = std::make_shared<ScfWork>();
        SCHK(c, hipHostMalloc((void **)&sw->hs, sizeof(double) * 32, hipHostMallocDefault));
        SCHK(c, hipEventCreate(&sw->ej0));
        SCHK(c, hipEventCreate(&sw->ej1));
        slot = sw;
    }
    w = static_cast<ScfWork *>(slot.get());
    return TCG_OK;
}

}  // namespace

extern "C" {

int tcg_ao_to_sao(tcg_ctx *c, int nbas, const double *M_ao, double *M_sao) {
    if (!c) return TCG_E_ARG;
    if (!M_ao || !M_sao) return fail(c, TCG_E_ARG, "tcg_ao_to_sao: null array");
    Lat L;
    int rc = lattice_of(c, nbas, L, "tcg_ao_to_sao");
    if (rc) return rc;
    const int nf = L.nf, nb = L.nb, N = L.N;
    hipStream_t st = ctx_stream(c);
    std::vector<double> rows((size_t)nf * nb);             // unit-cell rows, leading dimension nf
    for (int col = 0; col < nb; ++col)
        for (int a = 0; a < nf; ++a) rows[a + (size_t)nf * col] = M_ao[a + (size_t)nb * col];
    DevBuf<double> dR;
    DevBuf<double2> dph, dK;
    SCHK(c, dR.upload(rows));
    SCHK(c, dph.upload(phase_table(L)));
    SCHK(c, dK.alloc((size_t)N * nf * nf));
    RowSources src{};
    src.n = 1;
    src.p[0] = dR.p;
    src.sa[0] = 1;
    src.sc[0] = nf;
    src.f[0] = 1.0;
    SCHK(c, launch_lattice_fwd(src, dph.p, N, nf, dK.p, st));
    std::vector<double2> blk((size_t)N * nf * nf);
    SCHK(c, hipMemcpyAsync(blk.data(), dK.p, sizeof(double2) * blk.size(), hipMemcpyDeviceToHost, st));
    SCHK(c, hipStreamSynchronize(st));
    scatter_blocks(L, blk, M_sao, true);
    return TCG_OK;
}

int tcg_sao_to_ao(tcg_ctx *c, int nbas, const double *M_sao, double *M_ao) {
    if (!c) return TCG_E_ARG;
    if (!M_ao || !M_sao) return fail(c, TCG_E_ARG, "tcg_sao_to_ao: null array");
    Lat L;
    int rc = lattice_of(c, nbas, L, "tcg_sao_to_ao");
    if (rc) return rc;
    const int nf = L.nf, nb = L.nb, N = L.N;
    hipStream_t st = ctx_stream(c);
    std::vector<double2> blk;
    gather_blocks(L, M_sao, blk);
    DevBuf<double2> dph, dK;
    DevBuf<double> dC, dM;
    SCHK(c, dK.upload(blk));
    SCHK(c, dph.upload(phase_table(L)));
    SCHK(c, dC.alloc((size_t)N * nf * nf));
    SCHK(c, dM.alloc((size_t)nb * nb));
    SCHK(c, launch_lattice_inv(dK.p, dph.p, N, nf, 1, dC.p, nullptr, nullptr, st));
    SCHK(c, launch_circulant(dC.p, nf, L.ncl, dM.p, st));
    SCHK(c, hipMemcpyAsync(M_ao, dM.p, sizeof(double) * (size_t)nb * nb, hipMemcpyDeviceToHost, st));
    SCHK(c, hipStreamSynchronize(st));
    return TCG_OK;
}

int tcg_heev_blocks(tcg_ctx *c, int nbas, double *A_sao, double *e) {
    if (!c) return TCG_E_ARG;
    if (!A_sao || !e) return fail(c, TCG_E_ARG, "tcg_heev_blocks: null array");
    Lat L;
    int rc = lattice_of(c, nbas, L, "tcg_heev_blocks");
    if (rc) return rc;
    const int nf = L.nf, N = L.N;
    hipStream_t st = ctx_stream(c);
    std::vector<double2> blk;
    gather_blocks(L, A_sao, blk);
    DevBuf<double2> dA, dV;
    DevBuf<double> de;
    SCHK(c, dA.upload(blk));
    SCHK(c, dV.alloc(blk.size()));
    SCHK(c, de.alloc((size_t)N * nf));
    SCHK(c, launch_fock_diag(dA.p, nullptr, nullptr, N, nf, de.p, dV.p, st));
    SCHK(c, hipMemcpyAsync(blk.data(), dV.p, sizeof(double2) * blk.size(), hipMemcpyDeviceToHost, st));
    SCHK(c, hipMemcpyAsync(e, de.p, sizeof(double) * (size_t)N * nf, hipMemcpyDeviceToHost, st));
    SCHK(c, hipStreamSynchronize(st));
    scatter_blocks(L, blk, A_sao, false);
    return TCG_OK;
}

int tcg_rhf_sao(tcg_ctx *c, int source, int nbas, int nocc, const double *S, const double *T, const double *V,
                int max_scf, double thresh, int diis, double *e_out, double *coeff_sao, int *occ_out, double *P_out,
                double *out) {
    if (!c) return TCG_E_ARG;
    if (!S || !T || !V || !e_out || !out) return fail(c, TCG_E_ARG, "tcg_rhf_sao: null array");
    Lat L;
    int rc = lattice_of(c, nbas, L, "tcg_rhf_sao");
    if (rc) return rc;
    const int nf = L.nf, nb = L.nb, N = L.N;
    if (nocc < 1 || nocc > nb) return fail(c, TCG_E_ARG, "tcg_rhf_sao: nocc outside 1..nbas");
    if (max_scf < 1 || !(thresh > 0.0)) return fail(c, TCG_E_ARG, "tcg_rhf_sao: max_scf < 1 or thresh <= 0");
    if (source < 0 || source > 2)
        return fail(c, TCG_E_ARG, "tcg_rhf_sao: source must be 0 (stored), 1 (direct) or 2 (direct, incremental)");
    const double t_start = wall_ms();
    hipStream_t st = ctx_stream(c);
    const size_t nn = (size_t)nb * nb, len = (size_t)N * nf * nf;
    // TCG_SCF_LOG=1: one line per iteration on stderr (the reference's SCF table)
    static const bool log_it = [] {
        const char *e = std::getenv("TCG_SCF_LOG");
        return e && e[0] == '1';
    }();

    // --- set-up: S_k, X_k, T^T, V^T, unit-cell rows of Hc = T + V ---
    ScfWork *w = nullptr;
    rc = scf_work(c, w);
    if (rc) return rc;
    auto &dTt = w->dTt, &dVt = w->dVt, &dH = w->dH, &dP = w->dP, &dJ = w->dJ, &dK = w->dK, &dC = w->dC;
    auto &dCerr = w->dCerr, &dmax = w->dmax, &dpart = w->dpart, &dsc = w->dsc, &de = w->de, &dw = w->dw;
    auto &dsmin = w->dsmin, &dtmp = w->dtmp;
    auto &dph = w->dph, &dSk = w->dSk, &dXk = w->dXk, &dPk = w->dPk, &dErrH = w->dErrH, &dFH = w->dFH;
    auto &dFx = w->dFx, &dCk = w->dCk;
    auto &ddead = w->ddead, &docc = w->docc;
    // every buffer at its size first (a grown buffer drains the device, so none is
    // resized while this call's work is queued)
    SCHK(c, dtmp.alloc(nn));
    SCHK(c, dTt.alloc(nn));
    SCHK(c, dVt.alloc(nn));
    SCHK(c, dH.alloc((size_t)nf * nb));
    SCHK(c, dP.alloc(nn));
    SCHK(c, dSk.alloc(len));
    SCHK(c, dXk.alloc(len));
    SCHK(c, ddead.alloc(nb));
    SCHK(c, dsmin.alloc(N));
    SCHK(c, dJ.alloc(nn));
    SCHK(c, dK.alloc(nn));
    SCHK(c, dPk.alloc(len));
    SCHK(c, dErrH.alloc(len * kMaxDiis));
    SCHK(c, dFH.alloc(len * kMaxDiis));
    SCHK(c, dFx.alloc(len));
    SCHK(c, dCk.alloc(len));
    SCHK(c, dC.alloc(len));
    SCHK(c, dCerr.alloc(len));
    SCHK(c, dmax.alloc(N));                                 // one maximum per lattice vector
    SCHK(c, dpart.alloc(4 * (size_t)energy_parts_blocks(nb)));
    SCHK(c, dsc.alloc(32));
    SCHK(c, de.alloc(nb));
    SCHK(c, dw.alloc(nb));
    SCHK(c, docc.alloc(nb));
    SCHK(c, dph.alloc((size_t)N * N));
    if (source == 2) {
        SCHK(c, w->dPprev.alloc(nn));
        SCHK(c, w->dDP.alloc(nn));
        SCHK(c, w->dJacc.alloc((size_t)nf * nb));
        SCHK(c, w->dKacc.alloc((size_t)nf * nb));
        ctx_screen_reset(c);
    }
    const int full_every = ctx_screen_full_every(c);
    if (w->ph_key[0] != L.ncl[0] || w->ph_key[1] != L.ncl[1] || w->ph_key[2] != L.ncl[2]) {
        const std::vector<double2> ph = phase_table(L);
        SCHK(c, hipMemcpy(dph.p, ph.data(), sizeof(double2) * ph.size(), hipMemcpyHostToDevice));
        for (int d = 0; d < 3; ++d) w->ph_key[d] = L.ncl[d];
    }
    double *const hsp = w->hs;
    hipEvent_t ej0 = w->ej0, ej1 = w->ej1;
    SCHK(c, hipMemcpyAsync(dtmp.p, T, sizeof(double) * nn, hipMemcpyHostToDevice, st));
    SCHK(c, launch_transpose(dtmp.p, nb, dTt.p, st));
    // V goes through dP (free until the guess)
    SCHK(c, hipMemcpyAsync(dP.p, V, sizeof(double) * nn, hipMemcpyHostToDevice, st));
    SCHK(c, launch_transpose(dP.p, nb, dVt.p, st));
    SCHK(c, launch_rows_sum(dtmp.p, dP.p, nf, nb, dH.p, st));
    // dtmp is reused for S (stream order: after the transpose and the row sums)
    SCHK(c, hipMemcpyAsync(dtmp.p, S, sizeof(double) * nn, hipMemcpyHostToDevice, st));
    RowSources srcS{};
    srcS.n = 1;
    srcS.p[0] = dtmp.p;
    srcS.sa[0] = 1;
    srcS.sc[0] = nb;
    srcS.f[0] = 1.0;
    SCHK(c, launch_lattice_fwd(srcS, dph.p, N, nf, dSk.p, st));
    SCHK(c, launch_overlap_x(dSk.p, N, nf, dXk.p, ddead.p, dsmin.p, st));

    auto new_density = [&](const double2 *Fk, int frac) -> int {
        SCHK(c, launch_fock_diag(Fk, dXk.p, ddead.p, N, nf, de.p, dCk.p, st));
        SCHK(c, launch_occupy(de.p, nb, nocc, frac, dw.p, docc.p, dsc.p + 24, st));
        SCHK(c, launch_density_k(dCk.p, dw.p, N, nf, dPk.p, st));
        SCHK(c, launch_lattice_inv(dPk.p, dph.p, N, nf, 1, dC.p, nullptr, nullptr, st));
        SCHK(c, launch_circulant(dC.p, nf, L.ncl, dP.p, st));
        return TCG_OK;
    };

    // --- core guess: X^H Hc(k) X per k, the level at the cut filled evenly
    //     (guess_AO_RHF_SAO + guess_occupy_cut) ---
    RowSources srcF{};
    srcF.n = 1;
    srcF.p[0] = dH.p;
    srcF.sa[0] = 1;
    srcF.sc[0] = nf;
    srcF.f[0] = 1.0;
    SCHK(c, launch_lattice_fwd(srcF, dph.p, N, nf, dFx.p, st));
    rc = new_density(dFx.p, 1);
    if (rc) return rc;

    // --- SCF iterations ---
    std::vector<int> order;                  // DIIS history slots, most recent first
    std::vector<double> Bm;                  // their error products, order.size()^2
    double E = 0.0, E_old = 0.0, conv = 1.0, t_jk = 0.0;
    double hist[kStallIt];
    for (double &h : hist) h = HUGE_VAL;
    int n_stable = 0, it = 0, ld = nb;
    bool converged = false, stalled = false;
    double parts[4] = {0, 0, 0, 0};
    while (it < max_scf) {
        ++it;
        SCHK(c, hipEventRecord(ej0, st));
        if (source == 2) {
            // incremental Fock build: a full direct J/K every full_every iterations; in
            // between, J and K of the density change P - P_prev with density screening,
            // added to the last J and K (the direct J/K is linear in P)
            const size_t njk = (size_t)nf * nb;
            if ((it - 1) % full_every == 0) {
                rc = ctx_jk_device(c, 1, dP.p, dJ.p, dK.p, &ld);
                if (rc) return rc;
            } else {
                SCHK(c, launch_axpby(dP.p, w->dPprev.p, w->dDP.p, nn, 1.0, -1.0, st));
                rc = ctx_jk_device(c, 2, w->dDP.p, dJ.p, dK.p, &ld);
                if (rc) return rc;
                SCHK(c, launch_axpby(dJ.p, w->dJacc.p, dJ.p, njk, 1.0, 1.0, st));
                SCHK(c, launch_axpby(dK.p, w->dKacc.p, dK.p, njk, 1.0, 1.0, st));
            }
            SCHK(c, hipMemcpyAsync(w->dPprev.p, dP.p, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
            SCHK(c, hipMemcpyAsync(w->dJacc.p, dJ.p, sizeof(double) * njk, hipMemcpyDeviceToDevice, st));
            SCHK(c, hipMemcpyAsync(w->dKacc.p, dK.p, sizeof(double) * njk, hipMemcpyDeviceToDevice, st));
        } else {
            rc = ctx_jk_device(c, source, dP.p, dJ.p, dK.p, &ld);
            if (rc) return rc;
        }
        SCHK(c, hipEventRecord(ej1, st));
        // slot of this iteration's F and error: a free one, or the oldest
        int slot = 0;
        if ((int)order.size() < kMaxDiis) {
            for (slot = 0; slot < kMaxDiis; ++slot)
                if (std::find(order.begin(), order.end(), slot) == order.end()) break;
        } else {
            slot = order.back();
        }
        double2 *Fk = dFH.p + len * slot, *Ek = dErrH.p + len * slot;
        RowSources src{};
        src.n = 3;
        src.p[0] = dH.p;
        src.sa[0] = 1;
        src.sc[0] = nf;
        src.p[1] = dJ.p;
        src.sa[1] = 1;
        src.sc[1] = ld;
        src.p[2] = dK.p;
        src.sa[2] = 1;
        src.sc[2] = ld;
        src.f[0] = src.f[1] = src.f[2] = 1.0;
        SCHK(c, launch_lattice_fwd(src, dph.p, N, nf, Fk, st));
        SCHK(c, launch_scf_error(Fk, dPk.p, dSk.p, dXk.p, N, nf, Ek, st));
        int nmax = 0;
        SCHK(c, launch_lattice_inv(Ek, dph.p, N, nf, 0, dCerr.p, dmax.p, &nmax, st));
        SCHK(c, launch_energy(dP.p, dTt.p, dVt.p, dJ.p, dK.p, ld, nf, nb, dpart.p, dmax.p, nmax, (double)N, dsc.p, st));
        const int nkeep = std::min<int>((int)order.size(), kMaxDiis - 1);
        DiisSlots sl{};
        sl.m = 1 + nkeep;
        sl.slot[0] = slot;
        for (int j = 0; j < nkeep; ++j) sl.slot[1 + j] = order[j];
        if (diis) SCHK(c, launch_diis_dots(Ek, dErrH.p, (long long)len, sl, dsc.p + 8, st));
        SCHK(c, hipMemcpyAsync(hsp, dsc.p, sizeof(double) * 32, hipMemcpyDeviceToHost, st));
        SCHK(c, hipStreamSynchronize(st));
        float jms = 0.0f;
        SCHK(c, hipEventElapsedTime(&jms, ej0, ej1));
        t_jk += jms;
        for (int q = 0; q < 4; ++q) parts[q] = hsp[q];
        E = parts[0] + parts[1] + parts[2] + parts[3];
        if (it > 1) conv = hsp[4];
        if (log_it) {
            std::fprintf(stderr, "tcg_rhf_sao: SCF %3d  E_el = %.12f  conv = %.3e  (%.1f ms)\n", it, E, conv,
                         wall_ms() - t_start);
            std::fflush(stderr);
        }
        if (conv < thresh) {
            converged = true;
            break;
        }
        if (it > 2 && std::fabs(E - E_old) < std::max(kStallDE, 2e-10 * std::fabs(E))) ++n_stable;
        else n_stable = 0;
        if (n_stable >= kStallIt && conv < kStallConv && conv > 0.5 * hist[it % kStallIt]) {
            converged = true;
            stalled = true;
            break;
        }
        hist[it % kStallIt] = conv;
        E_old = E;
        const double2 *Fuse = Fk;
        if (diis) {
            // prepend this iteration's error to the history (DIIS_extrapolation :20-24)
            const int m = sl.m;
            std::vector<double> Bn((size_t)m * m);
            for (int j = 0; j < m; ++j) Bn[j] = Bn[(size_t)m * j] = hsp[8 + j];
            const int mo = (int)order.size();
            for (int i = 1; i < m; ++i)
                for (int j = 1; j < m; ++j) Bn[(size_t)m * i + j] = Bm[(size_t)mo * (i - 1) + (j - 1)];
            order.insert(order.begin(), slot);
            order.resize(m);
            Bm.swap(Bn);
            std::vector<double> w;
            if (diis_solve(Bm, m, w)) {
                DiisSlots cs{};
                cs.m = m;
                for (int j = 0; j < m; ++j) {
                    cs.slot[j] = order[j];
                    cs.w[j] = w[j];
                }
                SCHK(c, launch_diis_combine(dFH.p, (long long)len, cs, dFx.p, st));
                Fuse = dFx.p;
            } else {
                order.clear();               // reset, F unchanged
                Bm.clear();
            }
        }
        rc = new_density(Fuse, 0);
        if (rc) return rc;
    }

    // --- results ---
    std::vector<double2> Ck(len);
    SCHK(c, hipMemcpyAsync(e_out, de.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    if (coeff_sao) SCHK(c, hipMemcpyAsync(Ck.data(), dCk.p, sizeof(double2) * len, hipMemcpyDeviceToHost, st));
    if (occ_out) SCHK(c, hipMemcpyAsync(occ_out, docc.p, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
    if (P_out) SCHK(c, hipMemcpyAsync(P_out, dP.p, sizeof(double) * nn, hipMemcpyDeviceToHost, st));
    SCHK(c, hipMemcpyAsync(hsp + 24, dsc.p + 24, sizeof(double), hipMemcpyDeviceToHost, st));
    SCHK(c, hipStreamSynchronize(st));
    if (coeff_sao) scatter_blocks(L, Ck, coeff_sao, true);
    out[0] = E;
    for (int q = 0; q < 4; ++q) out[1 + q] = parts[q];
    out[5] = it;
    out[6] = conv;
    out[7] = converged ? 1.0 : 0.0;
    out[8] = stalled ? 1.0 : 0.0;
    out[9] = hsp[24];
    out[10] = t_jk;
    out[11] = wall_ms() - t_start;
    return TCG_OK;
}

}  // extern "C"
