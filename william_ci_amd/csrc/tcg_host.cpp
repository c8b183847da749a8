// Host driver and C ABI of the engine (include/tcg.h).
//
// tcg_eri_build follows the steps of the reference's ERI driver
// (src/Toroidal/ERI_toroidal.f90:36-562), regular-lattice path:
//   1. t-tables (nodes, weights, c_n exp(-C)I_n(C) per periodic direction)
//   2. shells, cell tables, representative shell pairs (first shell in the
//      unit cell) with their primitive pairs and Fourier vectors
//   3. Schwarz: diagonal blocks (SI SJ|SI SJ) on the GPU, replicated over
//      the lattice translations
//   4. primitive Schwarz table (keyword prim_screen, always on)
//   5. translation x 8-fold unique shell quartets, Schwarz screened
//   6. quartets on the GPU, scattered into the SAO/normal layout
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <atomic>
#include <thread>
#include <type_traits>
#include <utility>
#include <memory>
#include <vector>

#include "../../include/tcg.h"
#include "tcg_devbuf.hpp"
#include "tcg_device.hpp"
#include "tcg_engine.hpp"
#include "tcg_math.hpp"
#include "tcg_quad_tables.inc"

using namespace tcg;

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct tcg_ctx {
    int ndp = 1;
    double L[3] = {0, 0, 0}, a[3] = {0, 0, 0};
    int ncl[3] = {1, 1, 1};
    int ncell = 1;
    int nq = 32, nqp = 32;
    int fast = 1;                // 1: nq-point mapped Gauss-Legendre (fast_quad); 0: literal 64/128 table
    std::string err;
    // basis
    int nbas = 0, nfuc = 0;
    std::vector<double> center, expo, coef;
    std::vector<int> lxyz, poff;
    // shells / lattice
    std::vector<Shell> sh;
    int nsh = 0, nshuc = 0;
    std::vector<int> relc;
    // tables
    int ncap = 0;
    std::vector<double> tnode, tweight;
    // pair data (the primitive-pair records themselves are built on the GPU)
    std::vector<ShellPair> sp;
    // device
    hipStream_t st = nullptr;
    DevBuf<double> d_tab, d_tnode, d_tw, d_coef, d_herm, d_blk, d_Q, d_Qsh, d_eri, d_P, d_J, d_K, d_jwork;
    DevBuf<int> d_tablim, d_lxyz, d_poff, d_relc;
    DevBuf<long long> d_tsum;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    bool time_rows = false;    // record an event pair around every row MFMA launch
    std::vector<std::pair<hipEvent_t, hipEvent_t>> row_evs;   // ... read after the build's final sync
    int row_evs_used = 0;
    DevBuf<Shell> d_sh;
    DevBuf<PrimPair> d_pp;
    DevBuf<ShellPair> d_sp;
    DevBuf<double2> d_vec;
    DevBuf<Quartet> d_q4;        // the unique quartet list (ERI build, direct J/K)
    DevBuf<Quartet> d_q4blk;     // quartets of a block request (Schwarz pass, tcg_eri_blocks, self-check)
    DevBuf<unsigned long long> d_cnt, d_qmax;
    // row pipeline
    DevBuf<long long> d_scount, d_rcount, d_soff, d_roff;
    DevBuf<Surv> d_surv;
    DevBuf<RowRec> d_rows;
    DevBuf<unsigned> d_keys, d_keys2;
    DevBuf<int> d_vals, d_order, d_cls;
    DevBuf<unsigned> d_clscnt;
    DevBuf<double> d_rowval, d_realval;
    DevBuf<int> d_sq;
    DevBuf<char> d_tmp;
    long long last_surv = 0, last_rows = 0;
    double rows_ms = 0.0;        // row MFMA kernel time of the last run_quartets (time_rows)
    double row_budget = 24e9;    // bytes of row buffers per chunk
    bool row_budget_set = false; // TCG_ROW_BUDGET given (tests: small chunks on small systems)
    DevBuf<long long> d_soffc, d_roffc;
    // column lists of the resident ERI for the sparse J/K
    bool cols_ready = false;
    DevBuf<long long> d_list_d, d_list_b, d_off_d, d_off_b;
    size_t eri_count = 0;
    int eri_n1 = 0;
    bool eri_physicist = false;
    int cols_phys = 0;           // storage order the column lists were built for
    long long cols_nz = 0;       // entries of each column list
    std::shared_ptr<void> scf_work;   // device buffers of the SCF loop (tcg_scf.cpp)
    // k-point MP2 work buffers (kept between calls, grown as needed)
    DevBuf<double> km_cos, km_sin, km_cosn, km_sinn, km_e, km_part, km_cos2, km_sin2, km_Ar, km_Ai, km_Br, km_Bi;
    DevBuf<double2> km_Ck, km_kmo, km_TA, km_TB;
    DevBuf<int> km_occ;
    DevBuf<long long> km_imap;
    DevBuf<unsigned char> km_fl;
    DevBuf<unsigned long long> km_nz;
    long long stats[8] = {0};
    double times[8] = {0};
    int max_ncomb = 1, max_blk = 1;
    int sp_contract = 0;         // the s/p contraction kernels take the quartets of s and p shells
    int sp_min_quartets = 16384; // ... of quartet lists at least this long
    bool prepared = false;
    int nq4 = 0;                 // unique shell quartets in d_q4 (after setup_quartets; -1: count on the device)
    int q4_cap = 0;              // largest main enumeration so far (the capacity a deferred one writes into)
    std::vector<int> rule_nq;    // node count of each rule in the tables (rule 0 = the run's rule)
    int nqp_all = 32;            // node stride of the tables (largest padded node count)
    DevBuf<int> d_rule_nq;
    // learned per-quartet orders
    std::vector<float> mlp_params, mlp_norm;           // network, feature mean/std and limits
    std::vector<int> learned_orders;                   // index -> mapped Gauss-Legendre order
    float learned_margin = 1.0f;
    double learned_skip = 0.0;         // skip threshold on the learned bound (0: no skipping)
    double learned_skip_cap = 0.0;     // ... only below this Schwarz product (0: no cap)
    float learned_margin_b = 1.5f;     // margin on the bound head, log10 units
    DevBuf<Quartet> d_q4keep;          // the quartets a learned build evaluates (skipped ones removed)
    bool direct_learned = false;       // the direct J/K runs the learned list below
    DevBuf<Quartet> d_q4dl;            // ... its quartets (skipped ones removed) and their rules
    DevBuf<int> d_qrdl;
    DevBuf<float> d_lb, d_lbdl;        // log10 learned bound per quartet (all / the direct list)
    int nq4dl = 0;
    // density screening of the direct J/K (incremental Fock builds, tcg_set_direct_screen)
    double screen_tau = 1e-14;         // skip when bound x largest density element < tau
    int screen_full_every = 50;        // a full (unscreened) build every this many iterations
    int screen_learned = 0;            // 1: the bound is min(Schwarz, learned) when the learned list runs
    DevBuf<double> d_Dsh;
    DevBuf<unsigned char> d_keep;
    DevBuf<long long> d_sidx, d_snsel;
    DevBuf<Quartet> d_q4scr;
    DevBuf<int> d_qrscr;
    long long screen_kept = 0, screen_total = 0;   // quartets of the last direct build, kept / listed
    long long screen_cum[3] = {0, 0, 0};           // since the last reset: kept, listed, builds
    DevBuf<int> d_qrule_keep;
    DevBuf<float> d_mlp, d_feat, d_pred;
    DevBuf<int> d_qrule;
    DevBuf<double> d_expo;
    DevBuf<double> d_center, d_S1, d_T1;   // AO centres; S and T of the last tcg_one_electron
    DevBuf<VecJob> d_jobs;       // pair-vector jobs of the last preparation
    DevBuf<PairBlock> d_pblk;    // shell-pair blocks of the GPU pair set-up
    DevBuf<VecOut> d_vo;
    DevBuf<double2> d_dco;
    DevBuf<double> d_iv;
    long long learned_counts[16] = {0};
    bool learned_on = false;
    DevBuf<unsigned char> d_ood;
    DevBuf<float> d_mlpnorm;
    DevBuf<unsigned long long> d_hist;
    bool direct_ready = false;   // d_q4 / d_Q / primitive Schwarz valid for the direct J/K
    // pinned staging area of the host-to-device uploads: copies leave it
    // asynchronously on st; a region is reused only after st has drained
    char *stage = nullptr;
    size_t stage_cap = 0, stage_used = 0;
    // inputs the device t-tables were built from (rules, ncap, lattice): the
    // tables are rebuilt only when these change
    std::vector<double> tab_key;
    // second stream: zeroing of the ERI array while the set-up runs on st
    hipStream_t st2 = nullptr;
    hipEvent_t ev_fill = nullptr;

    // TCG_PROFILE=1: host wall-clock marks of the build's sub-steps, printed to stderr
    bool prof = false;
    std::vector<std::pair<const char *, double>> marks;
    void mark(const char *what) {
        if (prof) marks.emplace_back(what, now_ms());
    }

    int fail(int code, const char *fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
};

#define HIPCHK(ctx, expr)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return (ctx)->fail(TCG_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// v -> b on c->st through the pinned staging area (the host copy into the area
// is synchronous, so v may change as soon as this returns).  When the area is
// full the stream is drained and the area restarts at its beginning.
template <class T>
static hipError_t upload_async(tcg_ctx *c, DevBuf<T> &b, const std::vector<T> &v) {
    hipError_t e = hipSuccess;
    if (!b.p || v.size() > b.n) {
        e = hipStreamSynchronize(c->st);      // no copy into the old buffer is in flight
        if (e == hipSuccess) e = b.alloc(v.size());
        if (e != hipSuccess) return e;
    }
    if (v.empty()) return hipSuccess;
    const size_t bytes = sizeof(T) * v.size();
    if (c->stage_used + bytes > c->stage_cap) {
        e = hipStreamSynchronize(c->st);
        if (e != hipSuccess) return e;
        c->stage_used = 0;
        if (bytes > c->stage_cap) {
            if (c->stage) (void)hipHostFree(c->stage);
            c->stage = nullptr;
            c->stage_cap = 0;
            const size_t cap = std::max<size_t>(bytes, size_t(32) << 20);
            e = hipHostMalloc((void **)&c->stage, cap, hipHostMallocDefault);
            if (e != hipSuccess) return e;
            c->stage_cap = cap;
        }
    }
    std::memcpy(c->stage + c->stage_used, v.data(), bytes);
    e = hipMemcpyAsync(b.p, c->stage + c->stage_used, bytes, hipMemcpyHostToDevice, c->st);
    c->stage_used += (bytes + 255) & ~size_t(255);
    return e;
}

static int lsum3(const int *l) { return l[0] + l[1] + l[2]; }

// ---------------------------------------------------------------------------
// 1. t-tables (src/module/clifford_tools.f90:620-709 with ncap from ERI_toroidal.f90:94-104)
// ---------------------------------------------------------------------------
// Nodes and weights of one rule (fast: nq mapped Gauss-Legendre points, else
// the literal 64/128 table), padded to nqp.
static void rule_nodes(int fast, int nq, int nqp, std::vector<double> &tnode, std::vector<double> &tweight) {
    tnode.assign(nqp, 0.0);
    tweight.assign(nqp, 0.0);
    if (fast) {
        mapped_gauss_legendre(nq, tnode.data(), tweight.data());
    } else {
        // the literal tables of gauss_legendre_points (src/module/precompute_t.f90:495-891)
        const double *t = nq == 64 ? kQuad64_t : kQuad128_t;
        const double *w = nq == 64 ? kQuad64_w : kQuad128_w;
        std::copy(t, t + nq, tnode.begin());
        std::copy(w, w + nq, tweight.begin());
    }
}

// Tables of every rule the quartets may use: rule 0 = the run's rule
// (fast_quad / nquad_point), rules 1.. = the learned orders; all padded to
// the largest node count c->nqp_all.  Layout [rule][d][n][node] (t-table),
// [rule][d][node tile] (last nonzero order), [rule][d][n] (flop count of a row of last order n),
// [rule][node] (nodes, weights).  The Bessel columns are independent: they
// are spread over host threads.
static int build_tables(tcg_ctx *c, std::vector<double> &tab, std::vector<int> &tablim, std::vector<long long> &tsum,
                        std::vector<double> &tnode, std::vector<double> &tweight) {
    int lmax = 0;
    double emax = 0.0, amin = 1e300;
    for (int i = 0; i < c->nbas; ++i) {
        for (int d = 0; d < c->ndp; ++d) lmax = std::max(lmax, c->lxyz[3 * i + d]);
        for (int p = c->poff[i]; p < c->poff[i + 1]; ++p) emax = std::max(emax, c->expo[p]);
    }
    for (int d = 0; d < c->ndp; ++d) amin = std::min(amin, c->a[d]);
    c->ncap = nmax_est(4.0 * emax / (amin * amin)) + 2 * lmax + 1;
    c->rule_nq.assign(1, c->nq);
    for (int n : c->learned_orders) c->rule_nq.push_back(n);
    int nqp = c->nqp;
    for (int n : c->rule_nq) nqp = std::max(nqp, ((n + kNodeTile - 1) / kNodeTile) * kNodeTile);
    c->nqp_all = nqp;
    const int nr = (int)c->rule_nq.size(), ndp = c->ndp, ncap = c->ncap, rows = ncap + 5, nh = nqp / kNodeTile;
    // the tables depend on the rules, ncap and the lattice only; like the
    // reference's setup_t_tables (clifford_tools.f90:634-636) they are kept
    // from one build to the next while these stay the same
    std::vector<double> key = {(double)c->fast, (double)nqp, (double)ncap, (double)ndp, c->a[0], c->a[1], c->a[2]};
    for (int n : c->rule_nq) key.push_back((double)n);
    if (key == c->tab_key && c->d_tab.p) return 1;
    c->tab_key = key;
    tnode.assign((size_t)nr * nqp, 0.0);
    tweight.assign((size_t)nr * nqp, 0.0);
    for (int r = 0; r < nr; ++r) {
        std::vector<double> tn, tw;
        rule_nodes(r == 0 ? c->fast : 1, c->rule_nq[r], nqp, tn, tw);
        std::copy(tn.begin(), tn.end(), tnode.begin() + (size_t)r * nqp);
        std::copy(tw.begin(), tw.end(), tweight.begin() + (size_t)r * nqp);
        if (r == 0) {
            c->tnode = tn;
            c->tweight = tw;
        }
    }
    tab.assign((size_t)nr * ndp * rows * nqp, 0.0);
    std::vector<int> tnmax((size_t)nr * ndp * nqp, -1);
    // columns (rule, d, node), each a scaled Bessel vector up to ncap
    std::vector<int> cols;
    for (int r = 0; r < nr; ++r)
        for (int d = 0; d < ndp; ++d)
            for (int k = 0; k < c->rule_nq[r]; ++k) cols.push_back((r * ndp + d) * nqp + k);
    std::atomic<size_t> next(0);
    auto work = [&] {
        std::vector<double> col(ncap + 1);
        for (size_t i = next++; i < cols.size(); i = next++) {
            const int code = cols[i], k = code % nqp, rd = code / nqp, d = rd % ndp, r = rd / ndp;
            const double t = tnode[(size_t)r * nqp + k];
            const double C = 2.0 * t * t / (c->a[d] * c->a[d]);
            bessel_scaled_vector(ncap, C, col.data());
            int nmax = ncap;   // last order kept: first below 1e-16 relative, + 5
            for (int n = 0; n <= ncap; ++n)
                if (col[n] < 1e-16 * col[0]) { nmax = std::min(ncap, n + 5); break; }
            for (int n = 0; n <= nmax; ++n) tab[((size_t)rd * rows + n) * nqp + k] = (n == 0 ? 1.0 : 2.0) * col[n];
            tnmax[(size_t)rd * nqp + k] = nmax;
        }
    };
    const unsigned nthr = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < nthr && i < cols.size(); ++i) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    tablim.assign((size_t)nr * ndp * nh, -1);
    for (int rd = 0; rd < nr * ndp; ++rd)
        for (int k = 0; k < nqp; ++k) {
            int &lim = tablim[(size_t)rd * nh + k / kNodeTile];
            lim = std::max(lim, tnmax[(size_t)rd * nqp + k]);
        }
    // flop count of a row of last order n: sum over the rule's nodes of min(n, t_nmax(k)) + 1
    tsum.assign((size_t)nr * ndp * (ncap + 1), 0);
    for (int rd = 0; rd < nr * ndp; ++rd) {
        const int nq = c->rule_nq[rd / ndp];
        for (int n = 0; n <= ncap; ++n) {
            long long s = 0;
            for (int k = 0; k < nq; ++k) s += std::min(n, tnmax[(size_t)rd * nqp + k]) + 1;
            tsum[(size_t)rd * (ncap + 1) + n] = s;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// 2. shells and lattice (ERI_toroidal.f90:106-156)
// ---------------------------------------------------------------------------
static int build_shells(tcg_ctx *c) {
    c->sh.clear();
    for (int i = 0; i < c->nbas;) {
        const int lt = lsum3(&c->lxyz[3 * i]);
        int nc = (lt + 1) * (lt + 2) / 2;
        bool same = i + nc <= c->nbas;
        for (int j = i + 1; same && j < i + nc; ++j) {
            same = lsum3(&c->lxyz[3 * j]) == lt && c->center[3 * j] == c->center[3 * i] &&
                   c->center[3 * j + 1] == c->center[3 * i + 1] && c->center[3 * j + 2] == c->center[3 * i + 2] &&
                   c->poff[j + 1] - c->poff[j] == c->poff[i + 1] - c->poff[i];
            for (int p = 0; same && p < c->poff[i + 1] - c->poff[i]; ++p)
                same = c->expo[c->poff[j] + p] == c->expo[c->poff[i] + p];
        }
        if (!same) return c->fail(TCG_E_ARG, "AO %d: Cartesian components of a shell must be consecutive", i);
        Shell s{};
        s.first_ao = i;
        s.ncomp = nc;
        s.l = lt;
        s.nprim = c->poff[i + 1] - c->poff[i];
        s.poff = c->poff[i];
        for (int d = 0; d < 3; ++d) s.xyz[d] = c->center[3 * i + d];
        if (lt == 1)
            for (int x = 0; x < nc; ++x)
                for (int d = 0; d < 3; ++d)
                    if (c->lxyz[3 * (i + x) + d] == 1) s.dirmap |= (unsigned)x << (2 * d);
        if (lt > 2) return c->fail(TCG_E_ARG, "shell at AO %d: l = %d; this engine handles s, p, d", i, lt);
        if (s.nprim > kMaxPrim) return c->fail(TCG_E_ARG, "shell at AO %d: %d primitives (max %d)", i, s.nprim, kMaxPrim);
        c->sh.push_back(s);
        i += nc;
    }
    c->nsh = (int)c->sh.size();
    c->nshuc = 0;
    for (auto &s : c->sh)
        if (s.first_ao < c->nfuc) c->nshuc++;
    bool ok = c->nsh == c->nshuc * c->ncell && c->nbas == c->nfuc * c->ncell;
    for (int S = 0; ok && S < c->nsh - c->nshuc; ++S)
        ok = c->sh[S + c->nshuc].ncomp == c->sh[S].ncomp && c->sh[S + c->nshuc].first_ao == c->sh[S].first_ao + c->nfuc;
    if (!ok || c->ncell > 4096)
        return c->fail(TCG_E_LATTICE, "the shells do not repeat in every cell (nsh=%d, per cell %d, cells %d)", c->nsh,
                       c->nshuc, c->ncell);
    for (int S = 0; S < c->nsh; ++S) {
        c->sh[S].cell = S / c->nshuc;
        c->sh[S].loc = S % c->nshuc;
    }
    const int nc = c->ncell, n0 = c->ncl[0], n1 = c->ncl[1], n2 = c->ncl[2];
    c->relc.resize((size_t)nc * nc);
    for (int c1 = 0; c1 < nc; ++c1)
        for (int c2 = 0; c2 < nc; ++c2) {
            int dx = ((c2 % n0 - c1 % n0) % n0 + n0) % n0;
            int dy = (((c2 / n0) % n1 - (c1 / n0) % n1) % n1 + n1) % n1;
            int dz = ((c2 / (n0 * n1) - c1 / (n0 * n1)) % n2 + n2) % n2;
            c->relc[(size_t)c1 * nc + c2] = dx + n0 * (dy + n1 * dz);
        }
    return 0;
}

// representative pair of (S1, S2): S1 shifted to the unit cell
static inline int rep_pair(const tcg_ctx *c, int S1, int S2) {
    const int c1 = c->sh[S1].cell, c2 = c->sh[S2].cell;
    const int S2r = c->relc[(size_t)c1 * c->ncell + c2] * c->nshuc + c->sh[S2].loc;
    return c->sh[S1].loc * c->nsh + S2r;
}

// ---------------------------------------------------------------------------
// Representative shell pairs (S1 in the unit cell, S2 anywhere) and where the
// GPU pair set-up (tcg_setup.hip::k_pair_setup) writes their primitive-pair
// records, Fourier-vector jobs, vectors, Bessel values, coefficients and
// Hermite coefficients.  Integer work only: every offset follows from the
// shells' angular momenta and primitive counts, the vectors and Bessel values
// being laid out for the longest length (ncap).
// ---------------------------------------------------------------------------
struct PairLayout {
    size_t npp = 0, njobs = 0, nvo = 0, nvec = 0, niv = 0, nherm = 0, ndco = 0;
};

static int build_pair_blocks(tcg_ctx *c, std::vector<PairBlock> &blk, PairLayout &L) {
    c->sp.assign((size_t)c->nshuc * c->nsh, ShellPair{0, 0});
    blk.clear();
    blk.reserve((size_t)c->nshuc * c->nsh);
    L = PairLayout{};
    const int ndp = c->ndp, nreal = 3 - ndp, kE = 2 * kMaxL + 1;
    for (int S1 = 0; S1 < c->nshuc; ++S1)
        for (int S2 = 0; S2 < c->nsh; ++S2) {
            const Shell &A = c->sh[S1], &B = c->sh[S2];
            const size_t nij = (size_t)A.nprim * B.nprim, ncomb = (size_t)(A.l + 1) * (B.l + 1);
            ShellPair &spr = c->sp[(size_t)S1 * c->nsh + S2];
            spr.first = (int)L.npp;
            spr.count = (int)nij;
            PairBlock pb;
            pb.S1 = S1;
            pb.S2 = S2;
            pb.first = (int)L.npp;
            pb.job0 = (int)L.njobs;
            pb.vo0 = (int)L.nvo;
            pb.vec0 = (int)L.nvec;
            pb.iv0 = (long long)L.niv;
            pb.herm0 = (int)L.nherm;
            pb.dco0 = (int)L.ndco;
            blk.push_back(pb);
            L.npp += nij;
            L.njobs += nij * ndp;
            L.nvo += nij * ndp * ncomb;
            L.nvec += nij * ndp * ncomb * (size_t)(c->ncap + 1);
            L.niv += nij * ndp * (size_t)(c->ncap + A.l + B.l + 1);
            L.nherm += nij * nreal * ncomb * kE;
            L.ndco += nij * ndp * ncomb * (size_t)(2 * (A.l + B.l) + 1);
        }
    if (L.nvec > (size_t)0x7fffffff || L.nherm > (size_t)0x7fffffff || L.ndco > (size_t)0x7fffffff ||
        L.nvo > (size_t)0x7fffffff)
        return c->fail(TCG_E_NOMEM, "pair vectors exceed 2^31 elements (%zu)", L.nvec);
    return 0;
}

// ---------------------------------------------------------------------------
// device launch helper
// ---------------------------------------------------------------------------
static EriArgs base_args(tcg_ctx *c) {
    EriArgs A{};
    A.ndp = c->ndp;
    A.nq = c->nq;                  // rule 0
    A.nqp = c->nqp_all;            // node stride of every rule
    A.rule_nq = c->d_rule_nq.p;
    A.qrule = nullptr;
    A.ncap = c->ncap;
    A.nm_bits = 1;
    while ((1 << A.nm_bits) <= c->ncap) ++A.nm_bits;   // a row's order is at most ncap
    A.fac_const = 2.0 / std::sqrt(kPi);
    for (int d = 0; d < 3; ++d) {
        A.a[d] = c->a[d];
        A.L[d] = c->L[d];
    }
    for (int d = 0; d < c->ndp; ++d) A.fac_const *= c->L[d] * c->L[d];
    A.tab = c->d_tab.p;
    A.tab_lim = c->d_tablim.p;
    A.tsum = c->d_tsum.p;
    A.tnode = c->d_tnode.p;
    A.tweight = c->d_tw.p;
    A.sh = c->d_sh.p;
    A.ao_lxyz = c->d_lxyz.p;
    A.coef = c->d_coef.p;
    A.ao_poff = c->d_poff.p;
    A.pp = c->d_pp.p;
    A.sp = c->d_sp.p;
    A.vec = c->d_vec.p;
    A.herm = c->d_herm.p;
    A.screen_prefactor = 1;
    A.screen_prim = 0;
    A.max_ncomb = c->max_ncomb;
    A.max_blk = c->max_blk;
    {
        int nc = 1;
        for (const Shell &sh : c->sh) nc = std::max(nc, sh.ncomp);
        A.sub_words = nc * nc;
    }
    int max_cf = 1;
    for (const Shell &sh : c->sh) max_cf = std::max(max_cf, sh.ncomp * sh.nprim);
    A.part_words = 160 + 4 * max_cf;
    A.sp_contract = c->sp_contract;
    A.sp_min_quartets = c->sp_min_quartets;
    A.nbas = c->nbas;
    A.nfuc = c->nfuc;
    A.ncell = c->ncell;
    A.relc = c->d_relc.p;
    A.counters = c->d_cnt.p;
    return A;
}

// Quartets of A.q4[0..A.nq4) through the row pipeline (tcg_rows.hip):
// survivors -> rows -> sort by (direction, length) -> MFMA -> contraction.
// One chunk of quartets (A.q4[0..A.nq4)) whose survivor/row offsets so/ro
// start at 0: survivors, rows, sort, MFMA rows, real rows, contraction.
// padded: ns, nr are upper bounds (the exact counts stay on the device); the
// row keys past the last row sort last, their order entries are -1.
static int run_chunk(tcg_ctx *c, const EriArgs &A, const long long *so, const long long *ro, long long ns,
                     long long nr, bool padded = false) {
    HIPCHK(c, c->d_surv.alloc(ns));
    HIPCHK(c, c->d_rows.alloc(nr));
    HIPCHK(c, c->d_keys.alloc(nr));
    HIPCHK(c, c->d_keys2.alloc(nr));
    HIPCHK(c, c->d_vals.alloc(nr));
    HIPCHK(c, c->d_order.alloc(nr));
    HIPCHK(c, c->d_rowval.alloc((size_t)nr * A.nqp));
    if (padded) {
        HIPCHK(c, launch_fill_u32(c->d_keys.p, nr, 0xffffffffu, c->st));
        HIPCHK(c, launch_fill_u32(reinterpret_cast<unsigned *>(c->d_vals.p), nr, 0xffffffffu, c->st));
    }
    HIPCHK(c, launch_surv(A, 1, nullptr, nullptr, so, c->d_surv.p, c->st));
    HIPCHK(c, launch_rows_emit(A, so, ro, c->d_surv.p, c->d_rows.p, c->d_keys.p, c->d_vals.p, nr, c->st));
    c->mark("  surv+emit launched");
    int rbits = 0;
    while ((1 << rbits) < (int)c->rule_nq.size()) ++rbits;
    const int end_bit = A.nm_bits + 2 + (A.qrule ? rbits : 0);
    if (end_bit > 32) return c->fail(TCG_E_ARG, "row sort key needs %d bits (ncap %d, %d rules)", end_bit, c->ncap,
                                     (int)c->rule_nq.size());
    size_t tb = sort_temp_bytes((int)nr, end_bit);
    HIPCHK(c, c->d_tmp.alloc(tb));
    HIPCHK(c, sort_pairs(c->d_tmp.p, tb, c->d_keys.p, c->d_keys2.p, c->d_vals.p, c->d_order.p, (int)nr, end_bit,
                         c->st));
    hipEvent_t rev1 = nullptr;
    if (c->time_rows) {
        if (c->row_evs_used == (int)c->row_evs.size()) {
            std::pair<hipEvent_t, hipEvent_t> e{nullptr, nullptr};
            HIPCHK(c, hipEventCreate(&e.first));
            HIPCHK(c, hipEventCreate(&e.second));
            c->row_evs.push_back(e);
        }
        const auto &e = c->row_evs[c->row_evs_used++];
        HIPCHK(c, hipEventRecord(e.first, c->st));
        rev1 = e.second;
    }
    HIPCHK(c, launch_rows_mfma(A, c->d_rows.p, c->d_order.p, nr, c->d_rowval.p, c->st));
    if (rev1) HIPCHK(c, hipEventRecord(rev1, c->st));
    const int nreal = 3 - A.ndp;
    if (nreal > 0) HIPCHK(c, c->d_realval.alloc((size_t)(nr / A.ndp) * nreal * A.nqp));
    // the block of every quartet goes to the ERI array, or (direct) into J and K
    HIPCHK(c, c->d_cls.alloc(9 * (size_t)A.nq4));          // one list per s/p form and one of the rest
    HIPCHK(c, c->d_clscnt.alloc(9));
    HIPCHK(c, launch_contract(A, so, ro, c->d_surv.p, c->d_rowval.p, c->d_realval.p, c->d_cls.p, c->d_clscnt.p,
                              c->st));
    return 0;
}

// TCG_PROFILE set: the marks of one call, ms since the previous mark, to stderr
static void print_marks(tcg_ctx *c, const char *what, double t0) {
    if (!c->prof) return;
    c->mark("end");
    double prev = t0;
    fprintf(stderr, "%s profile (ms since previous mark):", what);
    for (auto &m : c->marks) {
        fprintf(stderr, " [%s %.3f]", m.first, m.second - prev);
        prev = m.second;
    }
    fprintf(stderr, "\n");
    c->marks.clear();
}

// Sum of the row MFMA launches' event times of this build (after its final sync)
static int read_row_times(tcg_ctx *c) {
    c->rows_ms = 0.0;
    for (int i = 0; i < c->row_evs_used; ++i) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->row_evs[i].first, c->row_evs[i].second));
        c->rows_ms += ms;
    }
    c->row_evs_used = 0;
    return 0;
}

// Quartets of A.q4[0..A.nq4) through the row pipeline (tcg_rows.hip):
// survivors -> rows -> sort by (direction, length) -> MFMA -> contraction,
// in chunks of at most row_cap rows so that the row buffers stay bounded.
// run_quartets' answer when a deferred quartet list outgrew its capacity
constexpr int kRecount = 1;

// Quartets of A.q4[0..A.nq4) through the row pipeline.
// ns_ub, nr_ub >= 0: upper bounds of the survivor and row counts, known on the
// host; the pipeline then runs on buffers of those sizes without reading the
// exact counts back (no host synchronisation; short lists such as the Schwarz
// passes).
// A.nq4_dev set: A.nq4 is the capacity of the quartet list and its length is
// on the device (a deferred enumeration); it is read back with the counts,
// and kRecount is returned when it exceeds the capacity.
static int run_quartets(tcg_ctx *c, EriArgs A, long long ns_ub = -1, long long nr_ub = -1) {
    int nq4 = A.nq4;
    c->last_surv = c->last_rows = 0;
    c->rows_ms = 0.0;
    if (nq4 <= 0) return 0;
    HIPCHK(c, c->d_scount.alloc(nq4 + 1));
    HIPCHK(c, c->d_rcount.alloc(nq4 + 1));
    HIPCHK(c, c->d_soff.alloc(nq4 + 1));
    HIPCHK(c, c->d_roff.alloc(nq4 + 1));
    HIPCHK(c, hipMemsetAsync(c->d_scount.p + nq4, 0, sizeof(long long), c->st));
    HIPCHK(c, hipMemsetAsync(c->d_rcount.p + nq4, 0, sizeof(long long), c->st));
    HIPCHK(c, launch_surv(A, 0, c->d_scount.p, c->d_rcount.p, nullptr, nullptr, c->st));
    size_t tb = scan_temp_bytes(nq4 + 1);
    HIPCHK(c, c->d_tmp.alloc(tb));
    HIPCHK(c, scan_exclusive(c->d_tmp.p, tb, c->d_scount.p, c->d_soff.p, nq4 + 1, c->st));
    HIPCHK(c, scan_exclusive(c->d_tmp.p, tb, c->d_rcount.p, c->d_roff.p, nq4 + 1, c->st));
    if (ns_ub >= 0 && nr_ub >= 0 && nr_ub <= (1ll << 24) && !A.nq4_dev) {
        if (nr_ub == 0) return 0;
        return run_chunk(c, A, c->d_soff.p, c->d_roff.p, ns_ub > 0 ? ns_ub : 1, nr_ub, true);
    }
    long long ns = 0, nr = 0;
    unsigned long long nq4d = 0;
    HIPCHK(c, hipMemcpyAsync(&ns, c->d_soff.p + nq4, sizeof ns, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(&nr, c->d_roff.p + nq4, sizeof nr, hipMemcpyDeviceToHost, c->st));
    if (A.nq4_dev) HIPCHK(c, hipMemcpyAsync(&nq4d, A.nq4_dev, sizeof nq4d, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (A.nq4_dev) {
        if (nq4d > (unsigned long long)nq4) return kRecount;   // the list outgrew its capacity
        nq4 = (int)nq4d;                                       // the counts past it are zero
        A.nq4 = nq4;
        A.nq4_dev = nullptr;
        c->nq4 = nq4;
        c->stats[0] = nq4;
    }
    c->last_surv = ns;
    c->last_rows = nr;
    c->mark("  counts");
    if (nr == 0) return 0;
    // rows per chunk: row record, sort keys/values, node values, real-direction values
    const int nreal = 3 - A.ndp;
    const double row_bytes = 48.0 + 8.0 * A.nqp * (1.0 + (double)nreal / A.ndp);
    const long long row_floor = c->row_budget_set ? 1024 : (1ll << 20);
    const long long row_cap =
        std::min<long long>(1ll << 30, std::max<long long>(row_floor, (long long)(c->row_budget / row_bytes)));
    if (nr <= row_cap) return run_chunk(c, A, c->d_soff.p, c->d_roff.p, ns, nr);
    // chunk boundaries: first quartet whose row offset reaches k * row_cap
    const int nchunk = (int)((nr + row_cap - 1) / row_cap);
    DevBuf<int> d_bounds;
    HIPCHK(c, d_bounds.alloc(nchunk + 1));
    HIPCHK(c, launch_chunk_bounds(c->d_roff.p, nq4, row_cap, nchunk, d_bounds.p, c->st));
    std::vector<int> bounds(nchunk + 1);
    HIPCHK(c, hipMemcpyAsync(bounds.data(), d_bounds.p, sizeof(int) * (nchunk + 1), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, c->d_soffc.alloc(nq4 + 1));
    HIPCHK(c, c->d_roffc.alloc(nq4 + 1));
    for (int k = 0; k < nchunk; ++k) {
        const int q0 = bounds[k], q1 = bounds[k + 1];
        if (q1 <= q0) continue;
        HIPCHK(c, launch_rebase(c->d_soff.p + q0, q1 - q0 + 1, c->d_soffc.p, c->st));
        HIPCHK(c, launch_rebase(c->d_roff.p + q0, q1 - q0 + 1, c->d_roffc.p, c->st));
        long long nsc = 0, nrc = 0;
        HIPCHK(c, hipMemcpyAsync(&nsc, c->d_soffc.p + (q1 - q0), sizeof nsc, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipMemcpyAsync(&nrc, c->d_roffc.p + (q1 - q0), sizeof nrc, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        if (nrc == 0) continue;
        EriArgs B = A;
        B.q4 = A.q4 + q0;
        B.nq4 = q1 - q0;
        if (A.qrule) B.qrule = A.qrule + q0;
        const int rc = run_chunk(c, B, c->d_soffc.p, c->d_roffc.p, nsc, nrc);
        if (rc) return rc;
    }
    return 0;
}

// blocks of q4 (each at its q.out) into d_blk; copied to *out unless out is null
// Blocks of the listed quartets into d_blk (and out).  The list goes to its own
// buffer: the unique quartet list of the direct J/K and of the build stays valid.
static int run_blocks(tcg_ctx *c, const std::vector<Quartet> &q4, size_t nout, std::vector<double> *out, int prefactor) {
    HIPCHK(c, upload_async(c, c->d_q4blk, q4));
    if (nout > c->d_blk.n) HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, c->d_blk.alloc(nout));
    HIPCHK(c, launch_fill(c->d_blk.p, nout, 0.0, c->st));     // blocks without survivors stay zero
    EriArgs A = base_args(c);
    A.q4 = c->d_q4blk.p;
    A.nq4 = (int)q4.size();
    A.blk_out = c->d_blk.p;
    A.screen_prefactor = prefactor;
    // survivor and row bounds: every primitive quartet of the listed quartets
    long long ns_ub = 0, nr_ub = 0;
    for (const Quartet &q : q4) {
        const long long n = q.prim >= 0 ? 1 : (long long)c->sp[q.bra].count * c->sp[q.ket].count;
        long long ncomb = 1;
        for (int x = 0; x < 4; ++x) ncomb *= c->sh[q.s[x]].l + 1;
        ns_ub += n;
        nr_ub += n * c->ndp * ncomb;
    }
    int rc = run_quartets(c, A, ns_ub, nr_ub);
    if (rc) return rc;
    if (!out) return 0;
    out->resize(nout);
    HIPCHK(c, hipMemcpyAsync(out->data(), c->d_blk.p, sizeof(double) * nout, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

// tables and pair data (steps 1-2), uploaded once per basis
static int prepare(tcg_ctx *c) {
    if (c->prepared) return 0;
    // ---- 1-2: tables and pair data ----
    std::vector<double> tab, rtnode, rtw;
    std::vector<int> tablim;
    std::vector<long long> tsum;
    const double tt0 = now_ms();
    const bool tables_kept = build_tables(c, tab, tablim, tsum, rtnode, rtw) == 1;
    c->mark("tables");
    c->times[7] = now_ms() - tt0;
    std::vector<PairBlock> pblk;
    PairLayout PL;
    int rc = build_pair_blocks(c, pblk, PL);
    if (rc) return rc;
    c->mark("pair layout (host)");
    c->max_ncomb = 1;
    for (auto &s : c->sh) c->max_ncomb = std::max(c->max_ncomb, s.l + 1);
    c->max_ncomb = c->max_ncomb * c->max_ncomb * c->max_ncomb * c->max_ncomb;
    c->max_blk = 1;
    for (auto &s : c->sh) c->max_blk = std::max(c->max_blk, s.ncomp);
    c->max_blk = c->max_blk * c->max_blk * c->max_blk * c->max_blk;
    // (TCG_GENERAL_CONTRACT set: every quartet through the general contraction)
    c->sp_contract = getenv("TCG_GENERAL_CONTRACT") ? 0 : 1;
    // (TCG_SP_MIN_QUARTETS: the shortest quartet list for the s/p kernels)
    if (const char *e = getenv("TCG_SP_MIN_QUARTETS")) c->sp_min_quartets = atoi(e);
    for (auto &s : c->sh)
        if (s.l == 1 && s.ncomp != 3) c->sp_contract = 0;
    if (!tables_kept) {
        HIPCHK(c, upload_async(c, c->d_tab, tab));
        HIPCHK(c, upload_async(c, c->d_tablim, tablim));
        HIPCHK(c, upload_async(c, c->d_tsum, tsum));
        HIPCHK(c, upload_async(c, c->d_tnode, rtnode));
        HIPCHK(c, upload_async(c, c->d_tw, rtw));
        HIPCHK(c, upload_async(c, c->d_rule_nq, c->rule_nq));
    }
    HIPCHK(c, upload_async(c, c->d_sh, c->sh));
    HIPCHK(c, upload_async(c, c->d_lxyz, c->lxyz));
    HIPCHK(c, upload_async(c, c->d_coef, c->coef));
    HIPCHK(c, upload_async(c, c->d_expo, c->expo));
    HIPCHK(c, upload_async(c, c->d_center, c->center));
    HIPCHK(c, upload_async(c, c->d_poff, c->poff));
    HIPCHK(c, upload_async(c, c->d_sp, c->sp));
    HIPCHK(c, upload_async(c, c->d_pblk, pblk));
    HIPCHK(c, upload_async(c, c->d_relc, c->relc));
    c->mark("uploads");
    {
        // primitive pairs, jobs, Fourier and Hermite coefficients on the GPU, then the
        // Bessel values and Fourier vectors
        const bool grow = PL.npp > c->d_pp.n || PL.njobs > c->d_jobs.n || PL.nvo > c->d_vo.n ||
                          PL.ndco > c->d_dco.n || PL.nherm > c->d_herm.n || PL.niv > c->d_iv.n ||
                          PL.nvec + (size_t)c->ncap + 64 > c->d_vec.n;
        if (grow) HIPCHK(c, hipStreamSynchronize(c->st));      // no kernel still reads the old buffers
        HIPCHK(c, c->d_pp.alloc(PL.npp));
        HIPCHK(c, c->d_jobs.alloc(PL.njobs));
        HIPCHK(c, c->d_vo.alloc(PL.nvo));
        HIPCHK(c, c->d_dco.alloc(PL.ndco));
        HIPCHK(c, c->d_herm.alloc(PL.nherm));
        HIPCHK(c, c->d_iv.alloc(PL.niv));
        // (the row kernel reads up to ncap + 16 entries past a row's first order, so the
        // vectors end in a padding that keeps those reads inside the allocation)
        HIPCHK(c, c->d_vec.alloc(PL.nvec + (size_t)c->ncap + 64));
        HIPCHK(c, launch_pair_setup(c->d_pblk.p, (int)pblk.size(), c->d_sh.p, c->d_expo.p, c->d_coef.p, c->d_poff.p,
                                    c->ndp, c->ncap, c->L, c->a, c->d_pp.p, c->d_jobs.p, c->d_vo.p, c->d_dco.p,
                                    c->d_herm.p, c->st));
        HIPCHK(c, launch_pair_vectors(c->d_jobs.p, (int)PL.njobs, c->d_vo.p, (int)PL.nvo, c->d_dco.p, c->d_iv.p,
                                      c->d_vec.p, c->st));
    }
    c->mark("pair set-up + vectors (launched)");
    HIPCHK(c, c->d_cnt.alloc(8));
    c->prepared = true;
    return 0;
}

// ---------------------------------------------------------------------------
// One-electron integrals (src/Toroidal/{1D,2D,3D}/*_toroidal.f90).
// Periodic direction, pair of Clifford primitives with barycentre P,
// u = a(x - P):  g_A^i g_B^l = kx a^-(i+l) sin^i(u+phA) sin^l(u+phB) e^{A(cos u - 1)}
// and  Int_0^L e^{iku} e^{A(cos u - 1)} dx = L exp(-A) I_|k|(A), so every
// overlap-type factor is L a^-(i+l) sum_k d_k exp(-A)I_|k|(A) with d_k the
// Fourier coefficients of the trigonometric factor.  The kinetic factor uses
//   d2 g^l/dx2 = l(l-1) g^(l-2) - l^2 a^2 g^l - 2 beta (2l+1) cos(a(x-B)) g^l + 4 beta^2 g^(l+2)
// (src/module/clifford_tools.f90:295-330).  Real directions: Hermite E_0
// overlap and m(m-1)S(m-2) - 2beta(2m+1)S(m) + 4beta^2 S(m+2).
// ---------------------------------------------------------------------------
// The final quartet pass with learned orders: survivors of the whole list,
// their features and the network's predictions, one rule per shell quartet
// (the largest asked for, plain rule outside the training range), then the
// row pipeline once per rule on the quartets of that rule.
// The learned rule of every quartet of A.q4[0..A.nq4) (features of the surviving
// primitive quartets, the network, the largest order per shell quartet): c->d_qrule,
// per-rule counts in c->learned_counts.  With skipping on, or gather set, the kept
// quartets and their rules are gathered into q4_out / rule_out (capacity nq4) and
// A is pointed at them; otherwise A.qrule = c->d_qrule.
static int learned_rules(tcg_ctx *c, EriArgs &A, bool gather, DevBuf<Quartet> &q4_out, DevBuf<int> &rule_out,
                         DevBuf<float> *lb_out = nullptr) {
    const int nq4 = A.nq4;
    const int nord = (int)c->learned_orders.size();
    for (int r = 0; r <= nord + 1 && r < 16; ++r) c->learned_counts[r] = 0;
    if (nq4 <= 0) return 0;
    HIPCHK(c, c->d_scount.alloc(nq4 + 1));
    HIPCHK(c, c->d_rcount.alloc(nq4 + 1));
    HIPCHK(c, c->d_soff.alloc(nq4 + 1));
    HIPCHK(c, c->d_roff.alloc(nq4 + 1));
    HIPCHK(c, hipMemsetAsync(c->d_scount.p + nq4, 0, sizeof(long long), c->st));
    HIPCHK(c, hipMemsetAsync(c->d_rcount.p + nq4, 0, sizeof(long long), c->st));
    HIPCHK(c, launch_surv(A, 0, c->d_scount.p, c->d_rcount.p, nullptr, nullptr, c->st));
    size_t tb = scan_temp_bytes(nq4 + 1);
    HIPCHK(c, c->d_tmp.alloc(tb));
    HIPCHK(c, scan_exclusive(c->d_tmp.p, tb, c->d_scount.p, c->d_soff.p, nq4 + 1, c->st));
    long long ns = 0;
    HIPCHK(c, hipMemcpyAsync(&ns, c->d_soff.p + nq4, sizeof ns, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, c->d_qrule.alloc(nq4));
    if (ns > 0) {
        if (ns > 0x7fffffffLL) return c->fail(TCG_E_NOMEM, "learned orders: too many primitive quartets (%lld)", ns);
        HIPCHK(c, c->d_surv.alloc(ns));
        HIPCHK(c, launch_surv(A, 1, nullptr, nullptr, c->d_soff.p, c->d_surv.p, c->st));
        HIPCHK(c, c->d_feat.alloc((size_t)ns * 6));
        HIPCHK(c, c->d_pred.alloc((size_t)ns * 2));
        HIPCHK(c, c->d_ood.alloc(ns));
        HIPCHK(c, c->d_sq.alloc(ns));
        HIPCHK(c, launch_surv_owner(c->d_soff.p, nq4, c->d_sq.p, c->st));
        HIPCHK(c, launch_features(A, c->d_sq.p, ns, c->d_surv.p, c->d_expo.p, c->d_mlpnorm.p, c->d_feat.p,
                                  c->d_ood.p, c->st));
        HIPCHK(c, launch_mlp_fwd(c->d_mlp.p, c->d_feat.p, (int)ns, c->d_pred.p, c->st));
    }
    const bool skipping = c->learned_skip > 0.0;
    const float log_skip = skipping ? (float)std::log10(c->learned_skip) : -3e30f;
    if (lb_out) HIPCHK(c, c->d_lb.alloc(nq4));
    HIPCHK(c, launch_rules(nq4, c->d_soff.p, c->d_pred.p, c->d_ood.p, c->learned_margin, nord, log_skip,
                           c->learned_margin_b, A.q4, c->d_Qsh.p, c->nsh, c->learned_skip_cap, c->d_qrule.p,
                           lb_out ? c->d_lb.p : nullptr, c->st));
    // quartets per rule (rule 0 = the plain rule; slot 15 = kRuleSkip & 15: the skipped ones)
    HIPCHK(c, c->d_hist.alloc(16));
    HIPCHK(c, launch_fill((double *)c->d_hist.p, 16, 0.0, c->st));
    HIPCHK(c, launch_rule_hist(nq4, c->d_qrule.p, c->d_hist.p, c->st));
    unsigned long long hist[16];
    HIPCHK(c, hipMemcpyAsync(hist, c->d_hist.p, sizeof hist, hipMemcpyDeviceToHost, c->st));
    // one pass over all quartets, each with its rule (rows sorted by rule, direction, length)
    A.qrule = c->d_qrule.p;
    if (skipping || gather) {
        // the skipped quartets leave the list (their elements stay zero in the array)
        DevBuf<unsigned char> keep;
        DevBuf<long long> idx, nsel;
        HIPCHK(c, keep.alloc(nq4));
        HIPCHK(c, idx.alloc(nq4));
        HIPCHK(c, nsel.alloc(1));
        HIPCHK(c, launch_rule_keep(nq4, c->d_qrule.p, keep.p, c->st));
        const size_t tb2 = select_temp_bytes(nq4);
        HIPCHK(c, c->d_tmp.alloc(tb2));
        HIPCHK(c, select_flagged(c->d_tmp.p, tb2, keep.p, idx.p, nsel.p, nq4, c->st));
        HIPCHK(c, q4_out.alloc(nq4));
        HIPCHK(c, rule_out.alloc(nq4));
        if (lb_out) HIPCHK(c, lb_out->alloc(nq4));
        HIPCHK(c, launch_gather_list(A.q4, c->d_qrule.p, lb_out ? c->d_lb.p : nullptr, idx.p,
                                     reinterpret_cast<const unsigned long long *>(nsel.p), nq4, q4_out.p, rule_out.p,
                                     lb_out ? lb_out->p : nullptr, c->st));
        long long kept = 0;
        HIPCHK(c, hipMemcpyAsync(&kept, nsel.p, sizeof kept, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        A.q4 = q4_out.p;
        A.nq4 = (int)kept;
        A.nq4_dev = nullptr;
        A.qrule = rule_out.p;
    } else {
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    for (int r = 0; r < nord; ++r) c->learned_counts[r] = (long long)hist[r + 1];
    c->learned_counts[nord] = (long long)hist[0];
    c->learned_counts[nord + 1] = (long long)hist[15];
    if (gather) {
        // the direct list keeps its rules: the per-survivor features, predictions and
        // owners are not needed again (tens of GB on large lattices)
        c->d_feat.release();
        c->d_pred.release();
        c->d_ood.release();
        c->d_sq.release();
        c->d_lb.release();
        c->d_qrule.release();
    }
    return 0;
}

// A learned ERI build: the rules, then one pass over the kept quartets, each with
// its rule (rows sorted by rule, direction, length)
static int learned_pass(tcg_ctx *c, EriArgs A) {
    int rc = learned_rules(c, A, false, c->d_q4keep, c->d_qrule_keep);
    if (rc) return rc;
    return A.nq4 > 0 ? run_quartets(c, A) : 0;
}

// Step 5 (ERI_toroidal.f90:342-401): unique shell quartets (translation x 8
// permutations), Schwarz screened, on the GPU, appended into the capacity of
// an earlier main enumeration (c->q4_cap).  defer, with such a capacity in
// place: the count stays on the device (c->nq4 = -1, count in d_cnt[3]);
// run_quartets reads it with the survivor counts and answers kRecount when it
// overflowed.  Otherwise counted here, and redone when it overflows.
static int enumerate_quartets(tcg_ctx *c, bool defer) {
    const int nsh = c->nsh, nshuc = c->nshuc, nc = c->ncell;
    if (defer && c->q4_cap > 0 && c->d_q4.p && c->d_q4.n >= (size_t)c->q4_cap) {
        HIPCHK(c, hipMemsetAsync(c->d_cnt.p + 3, 0, sizeof(unsigned long long), c->st));
        HIPCHK(c, launch_enum_quartets(c->d_sh.p, c->d_relc.p, c->d_Qsh.p, c->d_qmax.p, nshuc, nsh, nc, c->d_cnt.p + 3,
                                       c->d_q4.p, c->d_q4.n, c->st));
        c->nq4 = -1;
        return 0;
    }
    unsigned long long nq4 = 0;
    for (int pass = 0; pass < 2; ++pass) {
        const unsigned long long cap = c->d_q4.p ? c->d_q4.n : 0;
        HIPCHK(c, hipMemsetAsync(c->d_cnt.p + 3, 0, sizeof(unsigned long long), c->st));
        HIPCHK(c, launch_enum_quartets(c->d_sh.p, c->d_relc.p, c->d_Qsh.p, c->d_qmax.p, nshuc, nsh, nc, c->d_cnt.p + 3,
                                       cap ? c->d_q4.p : nullptr, cap, c->st));
        HIPCHK(c, hipMemcpyAsync(&nq4, c->d_cnt.p + 3, sizeof nq4, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        if (nq4 > (unsigned long long)0x7fffffff) return c->fail(TCG_E_NOMEM, "too many shell quartets (%llu)", nq4);
        if (nq4 <= cap) break;
        HIPCHK(c, c->d_q4.alloc(nq4));
    }
    c->stats[0] = (long long)nq4;
    c->nq4 = (int)nq4;
    c->q4_cap = (int)std::max<unsigned long long>(nq4, (unsigned long long)c->q4_cap);
    return 0;
}

// Steps 3-5 of the driver (ERI_toroidal.f90:200-401): Schwarz values from
// the diagonal blocks, the primitive Schwarz table, then the unique shell
// quartets on the GPU (c->d_q4[0..c->nq4)).  Shared by the stored build and
// the direct J/K (eri_direct.f90:76-225 does the same set-up).
static int setup_quartets(tcg_ctx *c, double &t2, double &t3, bool defer = false) {
    // ---- 3-4: Schwarz from the diagonal blocks (ERI_toroidal.f90:195-238),
    // replicated over translations, and the primitive Schwarz table (:300-338):
    // both lists of diagonal quartets in one pass through the pipeline; the
    // kernel prefactor screens apply to the shell blocks only (q.prim < 0) ----
    const int nsh = c->nsh, nshuc = c->nshuc, nb = c->nbas, nf = c->nfuc, nc = c->ncell;
    std::vector<Quartet> q4;
    size_t nout = 0;
    for (int SI = 0; SI < nshuc; ++SI)
        for (int SJ = 0; SJ < nsh; ++SJ) {
            Quartet q{};
            q.s[0] = SI; q.s[1] = SJ; q.s[2] = SI; q.s[3] = SJ;
            q.bra = q.ket = rep_pair(c, SI, SJ);
            q.out = (int)nout;
            q.prim = -1;
            q4.push_back(q);
            const int m = c->sh[SI].ncomp * c->sh[SJ].ncomp;
            nout += (size_t)m * m;
        }
    const size_t nshell_q = q4.size();
    for (int SI = 0; SI < nshuc; ++SI)
        for (int SJ = 0; SJ < nsh; ++SJ) {
            const int rp = rep_pair(c, SI, SJ);
            const ShellPair &s = c->sp[rp];
            for (int p = 0; p < s.count; ++p) {
                Quartet q{};
                q.s[0] = SI; q.s[1] = SJ; q.s[2] = SI; q.s[3] = SJ;
                q.bra = q.ket = rp;
                q.out = (int)nout;
                q.prim = p;
                q4.push_back(q);
                const int m = c->sh[SI].ncomp * c->sh[SJ].ncomp;
                nout += (size_t)m * m;
            }
        }
    // blocks stay on the GPU; Q, Qsh, the largest Qsh and the primitive ps are built there
    c->mark("schwarz list (host)");
    int rc = run_blocks(c, q4, nout, nullptr, 1);
    if (rc) return rc;
    c->mark("schwarz blocks");
    const size_t nbb = (size_t)nb * nb, nss = (size_t)nsh * nsh;
    if (nbb > c->d_Q.n || nss > c->d_Qsh.n || !c->d_qmax.p) HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, c->d_Q.alloc(nbb));
    HIPCHK(c, c->d_Qsh.alloc(nss));
    HIPCHK(c, c->d_qmax.alloc(1));
    HIPCHK(c, launch_fill(c->d_Q.p, nbb, 0.0, c->st));
    HIPCHK(c, hipMemsetAsync(c->d_qmax.p, 0, sizeof(unsigned long long), c->st));
    HIPCHK(c, launch_schwarz_tables(c->d_q4blk.p, (int)q4.size(), (int)nshell_q, c->d_sh.p, c->d_sp.p, c->d_blk.p,
                                    c->d_relc.p, nf, nb, nc, nsh, c->d_Q.p, c->d_Qsh.p, c->d_pp.p, c->d_qmax.p, c->st));
    HIPCHK(c, hipMemsetAsync(c->d_cnt.p, 0, 8 * sizeof(unsigned long long), c->st));
    t2 = now_ms();
    c->mark("schwarz tables");

    const int rc2 = enumerate_quartets(c, defer);
    if (rc2) return rc2;
    t3 = now_ms();
    c->mark("quartet list");
    return 0;
}

// V(i, j), i in the unit cell, j anywhere, on the GPU (other rows zero);
// host_work runs on the host while the kernel executes.
// Device copy of the AO list for the one-electron kernels
static OneStArgs one_args(tcg_ctx *c) {
    OneStArgs g{};
    g.ndp = c->ndp;
    g.nf = c->nfuc;
    g.nb = c->nbas;
    for (int d = 0; d < 3; ++d) {
        g.a[d] = c->a[d];
        g.L[d] = c->L[d];
    }
    g.center = c->d_center.p;
    g.lxyz = c->d_lxyz.p;
    g.poff = c->d_poff.p;
    g.expo = c->d_expo.p;
    g.coef = c->d_coef.p;
    return g;
}

template <class F>
static int nuclear_rows(tcg_ctx *c, int natom, const double *xyz, const int *Z, std::vector<double> &V, F host_work,
                        bool fill = false) {
    const int nb = c->nbas, nf = c->nfuc;
    const size_t nn = (size_t)nb * nb;
    std::vector<int> ao_shell(nb), pair_rep((size_t)nf * nb);
    for (int sI = 0; sI < c->nsh; ++sI)
        for (int x = 0; x < c->sh[sI].ncomp; ++x) ao_shell[c->sh[sI].first_ao + x] = sI;
    for (int j = 0; j < nb; ++j)
        for (int i = 0; i < nf; ++i) pair_rep[(size_t)i + (size_t)nf * j] = rep_pair(c, ao_shell[i], ao_shell[j]);
    std::vector<double> nxyz(xyz, xyz + 3 * (size_t)natom);
    std::vector<int> nz(Z, Z + natom);
    DevBuf<double> d_nxyz, d_V;
    DevBuf<int> d_nz, d_aosh, d_prep;
    HIPCHK(c, d_nxyz.upload(nxyz));
    HIPCHK(c, d_nz.upload(nz));
    HIPCHK(c, d_aosh.upload(ao_shell));
    HIPCHK(c, d_prep.upload(pair_rep));
    HIPCHK(c, d_V.alloc(nn));
    HIPCHK(c, launch_fill(d_V.p, nn, 0.0, c->st));
    EriArgs A = base_args(c);
    int maxpp = 1;
    for (int rp : pair_rep) maxpp = std::max(maxpp, c->sp[rp].count);
    HIPCHK(c, launch_nuclear(A, d_nxyz.p, d_nz.p, natom, d_aosh.p, d_prep.p, maxpp, d_V.p, c->st));
    if (fill) HIPCHK(c, launch_translate_fill(d_V.p, nf, nb, c->ncell, c->d_relc.p, c->st));
    if constexpr (!std::is_same<F, std::nullptr_t>::value) host_work();
    V.assign(nn, 0.0);
    HIPCHK(c, hipMemcpyAsync(V.data(), d_V.p, sizeof(double) * nn, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return 0;
}

// Column lists of the resident ERI for the sparse J/K: the non-zero columns
// ERI(:,b,c,d), listed by d and by b (built once per tensor and storage order).
static int stored_cols(tcg_ctx *c, const double *src, int n1, int nbas, int physicist) {
    if (c->cols_ready && c->cols_phys == physicist) return TCG_OK;
    // non-zero columns ERI(:,b,c,d), listed by d and by b (once per tensor)
    const long long ncol = (long long)nbas * nbas * nbas;
    DevBuf<unsigned char> flags;
    DevBuf<int> cnt;
    DevBuf<unsigned> keyb, keyb2;
    DevBuf<long long> nsel, tmp_list;
    HIPCHK(c, flags.alloc(ncol));
    HIPCHK(c, cnt.alloc(2 * (size_t)nbas + 2));
    HIPCHK(c, keyb.alloc(ncol));
    HIPCHK(c, nsel.alloc(1));
    HIPCHK(c, hipMemsetAsync(cnt.p, 0, sizeof(int) * (2 * (size_t)nbas + 2), c->st));
    HIPCHK(c, launch_col_flags(src, n1, nbas, physicist, flags.p, cnt.p, cnt.p + nbas + 1, keyb.p, c->st));
    size_t tb = std::max(select_temp_bytes(ncol), scan_int_temp_bytes(nbas + 1));
    HIPCHK(c, c->d_tmp.alloc(tb));
    HIPCHK(c, tmp_list.alloc(ncol));
    HIPCHK(c, select_flagged(c->d_tmp.p, tb, flags.p, tmp_list.p, nsel.p, ncol, c->st));
    long long nz = 0;
    HIPCHK(c, hipMemcpyAsync(&nz, nsel.p, sizeof nz, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, c->d_off_d.alloc(nbas + 1));
    HIPCHK(c, c->d_off_b.alloc(nbas + 1));
    HIPCHK(c, scan_int_exclusive(c->d_tmp.p, tb, cnt.p, c->d_off_d.p, nbas + 1, c->st));
    HIPCHK(c, scan_int_exclusive(c->d_tmp.p, tb, cnt.p + nbas + 1, c->d_off_b.p, nbas + 1, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, c->d_list_d.alloc(nz > 0 ? nz : 1));
    HIPCHK(c, c->d_list_b.alloc(nz > 0 ? nz : 1));
    if (nz > 0) {
        HIPCHK(c, hipMemcpyAsync(c->d_list_d.p, tmp_list.p, sizeof(long long) * nz, hipMemcpyDeviceToDevice, c->st));
        // keys b of the selected columns, then a stable sort by b
        HIPCHK(c, keyb2.alloc(nz));
        DevBuf<unsigned> kb_sel;
        HIPCHK(c, kb_sel.alloc(nz));
        HIPCHK(c, launch_gather_keys(keyb.p, c->d_list_d.p, nz, kb_sel.p, c->st));
        const int bits = (int)std::ceil(std::log2((double)nbas + 1.0)) + 1;
        size_t tb2 = sort_ll_temp_bytes(nz, bits);
        HIPCHK(c, c->d_tmp.alloc(tb2));
        HIPCHK(c, sort_pairs_ll(c->d_tmp.p, tb2, kb_sel.p, keyb2.p, c->d_list_d.p, c->d_list_b.p, nz, bits, c->st));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->cols_ready = true;
    c->cols_phys = physicist;
    c->cols_nz = nz;
    return TCG_OK;
}

// J and K rows of the unit cell (nfuc x nbas, leading dimension nfuc) from the
// integral-direct build for the density dP on the device.
static int direct_jk_device(tcg_ctx *c, const double *dP, double *dJ, double *dK, double tau = 0.0) {
    if (!c->direct_ready || !c->prepared) return c->fail(TCG_E_STATE, "direct J/K: tcg_direct_init not called");
    const size_t njk = (size_t)c->nfuc * c->nbas;
    HIPCHK(c, launch_fill(dJ, njk, 0.0, c->st));
    HIPCHK(c, launch_fill(dK, njk, 0.0, c->st));
    HIPCHK(c, hipMemsetAsync(c->d_cnt.p, 0, 8 * sizeof(unsigned long long), c->st));
    if (c->nq4 < 0) {                           // a deferred enumeration left its count on the device
        const int rc0 = enumerate_quartets(c, false);
        if (rc0) return rc0;
    }
    EriArgs A = base_args(c);
    A.q4 = c->d_q4.p;
    A.nq4 = c->nq4;
    const float *lb = nullptr;
    if (c->direct_learned) {                    // per-quartet learned rules (tcg_direct_init_learned)
        A.q4 = c->d_q4dl.p;
        A.nq4 = c->nq4dl;
        A.qrule = c->d_qrdl.p;
        if (c->screen_learned) lb = c->d_lbdl.p;
    }
    if (tau > 0.0 && A.nq4 > 0) {
        // density screening: the quartets whose bound times the largest element of the
        // density blocks they contract with is below tau leave this build's list
        const int nshuc = c->nshuc, nsh = c->nsh, n = A.nq4;
        HIPCHK(c, c->d_Dsh.alloc((size_t)nshuc * nsh));
        HIPCHK(c, c->d_keep.alloc(n));
        HIPCHK(c, c->d_sidx.alloc(n));
        HIPCHK(c, c->d_snsel.alloc(1));
        HIPCHK(c, c->d_q4scr.alloc(n));
        if (A.qrule) HIPCHK(c, c->d_qrscr.alloc(n));
        HIPCHK(c, launch_pair_dmax(dP, c->nbas, c->d_sh.p, nshuc, nsh, c->d_Dsh.p, c->st));
        HIPCHK(c, launch_density_keep(A.q4, n, c->d_sh.p, c->d_Qsh.p, nsh, nshuc, c->d_relc.p, c->ncell, c->d_Dsh.p, lb,
                                      tau, c->d_keep.p, c->st));
        const size_t tb = select_temp_bytes(n);
        HIPCHK(c, c->d_tmp.alloc(tb));
        HIPCHK(c, select_flagged(c->d_tmp.p, tb, c->d_keep.p, c->d_sidx.p, c->d_snsel.p, n, c->st));
        HIPCHK(c, launch_gather_list(A.q4, A.qrule, nullptr, c->d_sidx.p,
                                     reinterpret_cast<const unsigned long long *>(c->d_snsel.p), n, c->d_q4scr.p,
                                     A.qrule ? c->d_qrscr.p : nullptr, nullptr, c->st));
        long long kept = 0;
        HIPCHK(c, hipMemcpyAsync(&kept, c->d_snsel.p, sizeof kept, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        c->screen_total = n;
        c->screen_kept = kept;
        A.q4 = c->d_q4scr.p;
        A.nq4 = (int)kept;
        if (A.qrule) A.qrule = c->d_qrscr.p;
    } else {
        c->screen_total = c->screen_kept = A.nq4;
    }
    c->screen_cum[0] += c->screen_kept;
    c->screen_cum[1] += c->screen_total;
    c->screen_cum[2] += 1;
    A.screen_prim = 1;
    A.Qao = c->d_Q.p;
    A.Pden = dP;
    A.Jacc = dJ;
    A.Kacc = dK;
    const int rc = A.nq4 > 0 ? run_quartets(c, A) : 0;
    if (rc) return rc;
    // counters of this J/K (tcg_eri_stats): quartets, survivors, rows, flops
    unsigned long long cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(cnt, c->d_cnt.p, sizeof cnt, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->stats[0] = A.nq4;
    c->stats[1] = c->last_surv;
    c->stats[2] = c->last_rows;
    c->stats[3] = (long long)(cnt[2] + cnt[4]);
    c->stats[4] = (long long)cnt[4];
    return 0;
}

// Read access for the density-fitting driver (tcg_df.cpp), same library.
namespace tcg {
int ctx_view(const tcg_ctx *c, CtxView *v) {
    if (!c || !v) return TCG_E_ARG;
    v->ndp = c->ndp;
    for (int d = 0; d < 3; ++d) {
        v->L[d] = c->L[d];
        v->ncell[d] = c->ncl[d];
    }
    v->nbas = c->nbas;
    v->nfuc = c->nfuc;
    v->fast = c->fast;
    v->nq = c->nq;
    v->center = c->center.data();
    v->lxyz = c->lxyz.data();
    v->poff = c->poff.data();
    v->expo = c->expo.data();
    v->coef = c->coef.data();
    return TCG_OK;
}

// The resident ERI tensor (tcg_eri_build) and the engine's stream, for the
// allocation-free J/K (tcg_noalloc.hip).
int ctx_resident_eri(const tcg_ctx *c, const double **eri, size_t *count, int *n1, int *physicist, int *nbas,
                     hipStream_t *st) {
    if (!c || !c->d_eri.p || c->eri_count == 0) return TCG_E_STATE;
    *eri = c->d_eri.p;
    *count = c->eri_count;
    *n1 = c->eri_n1;
    *physicist = c->eri_physicist ? 1 : 0;
    *nbas = c->nbas;
    *st = c->st;
    return TCG_OK;
}

void ctx_set_error(tcg_ctx *c, const char *msg) {
    if (c) c->err = msg;
}

hipStream_t ctx_stream(const tcg_ctx *c) { return c ? c->st : nullptr; }

std::shared_ptr<void> &ctx_scf_work(tcg_ctx *c) { return c->scf_work; }

void ctx_screen_reset(tcg_ctx *c) {
    for (long long &v : c->screen_cum) v = 0;
}

int ctx_screen_full_every(const tcg_ctx *c) { return c->screen_full_every; }

int ctx_jk_device(tcg_ctx *c, int source, const double *dP, double *dJ, double *dK, int *ld) {
    if (!c || !dP || !dJ || !dK || !ld) return TCG_E_ARG;
    if (source == 0) {
        const int nb = c->nbas, n1 = c->eri_n1;
        if (!c->d_eri.p || n1 < c->nfuc || n1 > nb || c->eri_count != (size_t)n1 * nb * nb * nb)
            return c->fail(TCG_E_STATE, "stored J/K: no resident ERI of this basis (call tcg_eri_build)");
        const int physicist = c->eri_physicist ? 1 : 0;
        const int rc = stored_cols(c, c->d_eri.p, n1, nb, physicist);
        if (rc) return rc;
        HIPCHK(c, launch_jk_sparse(c->d_eri.p, dP, c->d_list_d.p, c->d_off_d.p, c->d_list_b.p, c->d_off_b.p, n1, nb,
                                   physicist, c->cols_nz, dJ, dK, c->st));
        *ld = nb;
        return TCG_OK;
    }
    if (source == 1) {
        *ld = c->nfuc;
        return direct_jk_device(c, dP, dJ, dK);
    }
    if (source == 2) {                         // integral-direct with density screening at tau
        *ld = c->nfuc;
        return direct_jk_device(c, dP, dJ, dK, c->screen_tau);
    }
    return c->fail(TCG_E_ARG, "J/K source must be 0 (stored) or 1 (direct)");
}

int ctx_shells(tcg_ctx *c, std::vector<int> &first_ao, std::vector<int> &ncomp, std::vector<int> &cell) {
    if (!c || c->nbas == 0) return TCG_E_STATE;
    first_ao.clear();
    ncomp.clear();
    cell.clear();
    for (const Shell &s : c->sh) {
        first_ao.push_back(s.first_ao);
        ncomp.push_back(s.ncomp);
        cell.push_back(s.cell);
    }
    return TCG_OK;
}
}  // namespace tcg

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int tcg_create(int ndp, const double L[3], const int ncell[3], int nquad, tcg_ctx **out) {
    if (!out) return TCG_E_ARG;
    *out = nullptr;
    if (ndp < 1 || ndp > 3 || !L || !ncell || nquad < 4 || nquad > kMaxNQ) return TCG_E_ARG;   // fast rule
    for (int d = 0; d < ndp; ++d)
        if (!(L[d] > 0.0)) return TCG_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return TCG_E_DEVICE;
    tcg_ctx *c = new tcg_ctx();
    c->ndp = ndp;
    for (int d = 0; d < 3; ++d) {
        c->L[d] = L[d];
        c->a[d] = (d < ndp) ? 2.0 * kPi / L[d] : 0.0;
        c->ncl[d] = ncell[d] > 0 ? ncell[d] : 1;
    }
    c->ncell = c->ncl[0] * c->ncl[1] * c->ncl[2];
    c->nq = nquad;
    c->nqp = ((nquad + kNodeTile - 1) / kNodeTile) * kNodeTile;
    if (const char *e = getenv("TCG_ROW_BUDGET")) {
        const double v = atof(e);
        if (v > 0.0) {
            c->row_budget = v;
            c->row_budget_set = true;
        }
    }
    if (hipStreamCreate(&c->st) != hipSuccess) {
        delete c;
        return TCG_E_DEVICE;
    }
    // the allocation-free J/K's scratch array: its address on this device
    if (noalloc_setup() != hipSuccess) {
        (void)hipStreamDestroy(c->st);
        delete c;
        return TCG_E_DEVICE;
    }
    *out = c;
    return TCG_OK;
}

int tcg_set_basis(tcg_ctx *c, int nbas, int nfuc, const double *center, const int *lxyz, const int *prim_off,
                  const double *expo, const double *coef) {
    if (!c) return TCG_E_ARG;
    if (nbas <= 0 || nfuc <= 0 || !center || !lxyz || !prim_off || !expo || !coef)
        return c->fail(TCG_E_ARG, "tcg_set_basis: null or empty argument");
    if (nbas != nfuc * c->ncell) return c->fail(TCG_E_ARG, "nbas (%d) != nfuc (%d) * cells (%d)", nbas, nfuc, c->ncell);
    if (prim_off[0] != 0) return c->fail(TCG_E_ARG, "prim_off[0] must be 0");
    for (int i = 0; i < nbas; ++i)
        if (prim_off[i + 1] <= prim_off[i]) return c->fail(TCG_E_ARG, "AO %d has no primitive", i);
    const int np = prim_off[nbas];
    c->nbas = nbas;
    c->nfuc = nfuc;
    c->center.assign(center, center + 3 * (size_t)nbas);
    c->lxyz.assign(lxyz, lxyz + 3 * (size_t)nbas);
    c->poff.assign(prim_off, prim_off + nbas + 1);
    c->expo.assign(expo, expo + np);
    c->coef.assign(coef, coef + np);
    c->prepared = false;
    c->direct_ready = false;
    for (int p = 0; p < np; ++p)
        if (!(expo[p] >= 0.0) || !std::isfinite(expo[p])) return c->fail(TCG_E_ARG, "exponent %d is negative", p);
    for (int i = 0; i < nbas; ++i)
        for (int p = prim_off[i]; p < prim_off[i + 1]; ++p)
            if (expo[p] == 0.0 && (lxyz[3 * i] + lxyz[3 * i + 1] + lxyz[3 * i + 2]) != 0)
                return c->fail(TCG_E_ARG, "exponent 0 (the constant function) only for s functions");
    return build_shells(c);
}

int tcg_eri_build(tcg_ctx *c, int sao, int physicist, double *eri) {
    if (!c) return TCG_E_ARG;
    if (c->nbas == 0) return c->fail(TCG_E_STATE, "tcg_eri_build before tcg_set_basis");
    const double t0 = now_ms();
    c->prof = getenv("TCG_PROFILE") != nullptr;
    c->marks.clear();
    const int nb = c->nbas, nf = c->nfuc;
    const int n1 = sao ? nf : nb;
    const size_t count = (size_t)n1 * nb * nb * nb;
    // the ERI array is zeroed on the second stream (after the work queued so far on
    // st, which may still read it) while the tables, pairs, Schwarz pass and
    // quartet list are built on st
    if (!c->st2) {
        HIPCHK(c, hipStreamCreate(&c->st2));
        HIPCHK(c, hipEventCreateWithFlags(&c->ev_fill, hipEventDisableTiming));
    }
    HIPCHK(c, c->d_eri.alloc(count));
    c->cols_ready = false;
    HIPCHK(c, hipEventRecord(c->ev_fill, c->st));
    HIPCHK(c, hipStreamWaitEvent(c->st2, c->ev_fill, 0));
    HIPCHK(c, launch_fill(c->d_eri.p, count, 0.0, c->st2));
    HIPCHK(c, hipEventRecord(c->ev_fill, c->st2));
    c->mark("fill launched");
    int rc = prepare(c);
    if (rc) return rc;
    HIPCHK(c, hipMemsetAsync(c->d_cnt.p, 0, 3 * sizeof(unsigned long long), c->st));
    const double t1 = now_ms();

    double t2 = 0.0, t3 = 0.0;
    rc = setup_quartets(c, t2, t3, !c->learned_on);
    if (rc) return rc;
    const int nq4 = c->nq4;                     // -1: deferred (the count is on the device)

    // ---- 6: quartets on the GPU, scattered into the ERI array ----
    HIPCHK(c, hipStreamWaitEvent(c->st, c->ev_fill, 0));
    EriArgs A = base_args(c);
    A.q4 = c->d_q4.p;
    A.nq4 = nq4;
    A.screen_prim = 1;
    A.eri = c->d_eri.p;
    A.n1 = n1;
    A.physicist = physicist;
    A.Qao = c->d_Q.p;
    if (!c->ev0) {
        HIPCHK(c, hipEventCreate(&c->ev0));
        HIPCHK(c, hipEventCreate(&c->ev1));
        HIPCHK(c, hipEventCreate(&c->ev2));
        HIPCHK(c, hipEventCreate(&c->ev3));
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->st));
    c->time_rows = true;
    c->row_evs_used = 0;
    if (c->learned_on) {
        rc = learned_pass(c, A);
    } else {
        if (c->nq4 < 0) {                       // deferred enumeration: its count comes with the survivors'
            A.nq4 = (int)c->d_q4.n;
            A.nq4_dev = c->d_cnt.p + 3;
        }
        rc = run_quartets(c, A);
        if (rc == kRecount) {                   // the list outgrew its capacity: count, enlarge, redo
            if ((rc = enumerate_quartets(c, false))) return rc;
            A.q4 = c->d_q4.p;                   // (enlarged: a new buffer)
            A.nq4 = c->nq4;
            A.nq4_dev = nullptr;
            rc = run_quartets(c, A);
        }
    }
    c->time_rows = false;
    if (rc) return rc;
    c->mark("quartets launched");
    HIPCHK(c, hipEventRecord(c->ev1, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->mark("quartets done");
    float kms = 0.f, rms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&kms, c->ev0, c->ev1));
    c->times[5] = kms;
    (void)rms;
    {
        const int rc6 = read_row_times(c);
        if (rc6) return rc6;
    }
    c->times[6] = c->rows_ms;
    // no SAO: rows of the other cells by translation (ERI_toroidal.f90:549-552)
    if (!sao) HIPCHK(c, launch_translate_rows(c->d_eri.p, nb, nf, c->ncell, c->d_relc.p, c->st));
    c->eri_physicist = physicist != 0;
    const double t4 = now_ms();
    unsigned long long cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    HIPCHK(c, hipMemcpy(cnt, c->d_cnt.p, sizeof cnt, hipMemcpyDeviceToHost));
    c->stats[1] = c->last_surv;
    c->stats[2] = c->last_rows;
    c->stats[3] = (long long)(cnt[2] + cnt[4]);
    c->stats[4] = (long long)cnt[4];
    c->eri_count = count;
    c->eri_n1 = n1;
    if (eri) HIPCHK(c, hipMemcpy(eri, c->d_eri.p, sizeof(double) * count, hipMemcpyDeviceToHost));
    c->times[0] = t1 - t0;
    c->times[1] = t2 - t1;
    c->times[2] = t3 - t2;
    c->times[3] = t4 - t3;
    c->times[4] = now_ms() - t0;
    print_marks(c, "tcg_eri_build", t0);
    return TCG_OK;
}

int tcg_jk_stored(tcg_ctx *c, int n1, int nbas, const double *P, const double *eri, double *J, double *K) {
    return tcg_jk_stored_order(c, n1, nbas, P, eri, 0, J, K);
}

int tcg_jk_stored_order(tcg_ctx *c, int n1, int nbas, const double *P, const double *eri, int phys_arg, double *J,
                        double *K) {
    if (!c) return TCG_E_ARG;
    if (!P || !J || !K || nbas <= 0 || n1 <= 0 || n1 > nbas) return c->fail(TCG_E_ARG, "tcg_jk_stored: bad arguments");
    const size_t count = (size_t)n1 * nbas * nbas * nbas;
    const double *src = c->d_eri.p;
    if (eri) {
        HIPCHK(c, c->d_eri.alloc(count));
        HIPCHK(c, hipMemcpy(c->d_eri.p, eri, sizeof(double) * count, hipMemcpyHostToDevice));
        c->cols_ready = false;
        c->eri_count = count;
        c->eri_n1 = n1;
        src = c->d_eri.p;
    } else if (!src || c->eri_count != count || c->eri_n1 != n1) {
        return c->fail(TCG_E_STATE, "tcg_jk_stored: no resident ERI of this shape (call tcg_eri_build)");
    }
    // a caller's array follows the caller's order; the resident one the order it was built in
    const int physicist = eri ? (phys_arg != 0) : (c->eri_physicist ? 1 : 0);
    if (eri) c->eri_physicist = physicist != 0;
    const size_t nn = (size_t)nbas * nbas;
    {
        const int rc = stored_cols(c, src, n1, nbas, physicist);
        if (rc) return rc;
    }
    HIPCHK(c, c->d_P.alloc(nn));
    HIPCHK(c, c->d_J.alloc(nn));
    HIPCHK(c, c->d_K.alloc(nn));
    HIPCHK(c, hipMemcpyAsync(c->d_P.p, P, sizeof(double) * nn, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, launch_fill(c->d_J.p, nn, 0.0, c->st));
    HIPCHK(c, launch_fill(c->d_K.p, nn, 0.0, c->st));
    HIPCHK(c, launch_jk_sparse(src, c->d_P.p, c->d_list_d.p, c->d_off_d.p, c->d_list_b.p, c->d_off_b.p, n1, nbas,
                               physicist, c->cols_nz, c->d_J.p, c->d_K.p, c->st));
    HIPCHK(c, hipMemcpyAsync(J, c->d_J.p, sizeof(double) * nn, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(K, c->d_K.p, sizeof(double) * nn, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return TCG_OK;
}

int tcg_direct_init(tcg_ctx *c) {
    if (!c) return TCG_E_ARG;
    if (c->nbas == 0) return c->fail(TCG_E_STATE, "tcg_direct_init before tcg_set_basis");
    int rc = prepare(c);
    if (rc) return rc;
    double t2 = 0.0, t3 = 0.0;
    rc = setup_quartets(c, t2, t3);
    if (rc) return rc;
    c->direct_ready = true;
    c->direct_learned = false;
    c->d_q4dl.release();                 // the learned list of an earlier tcg_direct_init_learned
    c->d_qrdl.release();
    c->d_lbdl.release();
    c->nq4dl = 0;
    for (long long &v : c->screen_cum) v = 0;
    return TCG_OK;
}

int tcg_set_direct_screen(tcg_ctx *c, double tau, int full_every, int use_learned_bound) {
    if (!c) return TCG_E_ARG;
    if (!(tau >= 0.0) || full_every < 1) return c->fail(TCG_E_ARG, "tcg_set_direct_screen: bad arguments");
    c->screen_tau = tau;
    c->screen_full_every = full_every;
    c->screen_learned = use_learned_bound ? 1 : 0;
    return TCG_OK;
}

int tcg_direct_screen_info(tcg_ctx *c, double *tau, int *full_every, long long *counts) {
    if (!c) return TCG_E_ARG;
    if (tau) *tau = c->screen_tau;
    if (full_every) *full_every = c->screen_full_every;
    if (counts) {
        counts[0] = c->screen_kept;
        counts[1] = c->screen_total;
        for (int i = 0; i < 3; ++i) counts[2 + i] = c->screen_cum[i];
    }
    return TCG_OK;
}

int tcg_direct_init_learned(tcg_ctx *c, long long *rule_counts) {
    if (!c) return TCG_E_ARG;
    if (c->learned_orders.empty()) return c->fail(TCG_E_STATE, "tcg_direct_init_learned before tcg_set_learned");
    int rc = tcg_direct_init(c);
    if (rc) return rc;
    if (c->nq4 < 0) {
        rc = enumerate_quartets(c, false);
        if (rc) return rc;
    }
    HIPCHK(c, hipMemsetAsync(c->d_cnt.p, 0, 8 * sizeof(unsigned long long), c->st));
    EriArgs A = base_args(c);
    A.q4 = c->d_q4.p;
    A.nq4 = c->nq4;
    A.screen_prim = 1;
    A.Qao = c->d_Q.p;
    rc = learned_rules(c, A, true, c->d_q4dl, c->d_qrdl, &c->d_lbdl);
    if (rc) return rc;
    c->nq4dl = A.nq4;
    c->direct_learned = true;
    if (rule_counts)
        for (size_t r = 0; r <= c->learned_orders.size() + 1; ++r) rule_counts[r] = c->learned_counts[r];
    return TCG_OK;
}

int tcg_direct_jk(tcg_ctx *c, int nbas, const double *P, double *J, double *K) {
    return tcg_direct_jk_screened(c, nbas, P, 0.0, J, K);
}

int tcg_direct_jk_screened(tcg_ctx *c, int nbas, const double *P, double tau, double *J, double *K) {
    if (!c) return TCG_E_ARG;
    if (!c->direct_ready || !c->prepared) return c->fail(TCG_E_STATE, "tcg_direct_jk: direct_init not called");
    if (nbas != c->nbas || !P || !J || !K || !(tau >= 0.0)) return c->fail(TCG_E_ARG, "tcg_direct_jk: bad arguments");
    const int nb = c->nbas, nf = c->nfuc;
    const size_t nn = (size_t)nb * nb, njk = (size_t)nf * nb;
    HIPCHK(c, c->d_P.alloc(nn));
    HIPCHK(c, c->d_J.alloc(njk));
    HIPCHK(c, c->d_K.alloc(njk));
    HIPCHK(c, hipMemcpyAsync(c->d_P.p, P, sizeof(double) * nn, hipMemcpyHostToDevice, c->st));
    const double t0 = now_ms();
    int rc = direct_jk_device(c, c->d_P.p, c->d_J.p, c->d_K.p, tau);
    if (rc) return rc;
    std::vector<double> Jr(njk), Kr(njk);
    HIPCHK(c, hipMemcpyAsync(Jr.data(), c->d_J.p, sizeof(double) * njk, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(Kr.data(), c->d_K.p, sizeof(double) * njk, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    std::fill(J, J + nn, 0.0);
    std::fill(K, K + nn, 0.0);
    for (int q = 0; q < nb; ++q)
        for (int p = 0; p < nf; ++p) {
            J[p + (size_t)nb * q] = Jr[p + (size_t)nf * q];
            K[p + (size_t)nb * q] = Kr[p + (size_t)nf * q];
        }
    c->times[3] = now_ms() - t0;
    return TCG_OK;
}

int tcg_mp2_gamma(tcg_ctx *c, int nbas, int nocc, const double *e, const double *C, double *out) {
    if (!c) return TCG_E_ARG;
    if (nbas != c->nbas || nocc <= 0 || nocc >= nbas || !e || !C || !out)
        return c->fail(TCG_E_ARG, "tcg_mp2_gamma: bad arguments");
    if (c->eri_n1 != nbas || !c->d_eri.p)
        return c->fail(TCG_E_STATE, "tcg_mp2_gamma needs the full ERI resident (tcg_eri_build with sao = 0)");
    const int N = nbas, o = nocc, v = nbas - nocc;
    const size_t N2 = (size_t)N * N, N3 = N2 * N;
    // occupied / virtual columns and the transposed virtual block, column-major
    std::vector<double> Co((size_t)N * o), Cv((size_t)N * v), CvT((size_t)v * N);
    for (int p = 0; p < o; ++p)
        for (int m = 0; m < N; ++m) Co[m + (size_t)N * p] = C[m + (size_t)N * p];
    for (int a = 0; a < v; ++a)
        for (int m = 0; m < N; ++m) Cv[m + (size_t)N * a] = C[m + (size_t)N * (o + a)];
    for (int m = 0; m < N; ++m)
        for (int a = 0; a < v; ++a) CvT[a + (size_t)v * m] = Cv[m + (size_t)N * a];
    // positions 4, 3, 2, 1 of the stored array contracted in that order.
    // chemist ERI(mu,nu,la,si) = (mu nu|la si): 4 <- occ (j), 3 <- vir (b), 2 <- occ (i), 1 <- vir (a)
    //   X(a,i,b,j) = (ai|bj) = (ia|jb)
    // physicist ERI(mu,la,nu,si) = (mu nu|la si): 4 <- occ (j), 3 <- occ (i), 2 <- vir (b), 1 <- vir (a)
    //   X(a,b,i,j) = (ia|jb)
    const bool phys = c->eri_physicist;
    const int d4 = o, d3 = phys ? o : v, d2 = phys ? v : o, d1 = v;
    DevBuf<double> dCo, dCv, dCvT, dY1, dY2, dY3, dX, de, dpart;
    HIPCHK(c, dCo.upload(Co));
    HIPCHK(c, dCv.upload(Cv));
    HIPCHK(c, dCvT.upload(CvT));
    HIPCHK(c, dY1.alloc(N3 * d4));
    HIPCHK(c, dY2.alloc(N2 * d3 * d4));
    HIPCHK(c, dY3.alloc((size_t)N * d2 * d3 * d4));
    HIPCHK(c, dX.alloc((size_t)d1 * d2 * d3 * d4));
    const double *C4 = dCo.p, *C3 = phys ? dCo.p : dCv.p, *C2 = phys ? dCv.p : dCo.p;
    hipStream_t st = c->st;
    const double t0 = now_ms();
    // Y1(mu,nu,la,q4) = sum_si ERI(mu,nu,la,si) C4(si,q4)
    HIPCHK(c, launch_dgemm_nn((int)N3, d4, N, 1.0, c->d_eri.p, (long)N3, 0, C4, N, 0, 0.0, dY1.p, (long)N3, 0, 1, st));
    // Y2(mu,nu,q3,q4) = sum_la Y1(mu,nu,la,q4) C3(la,q3), batched over q4
    HIPCHK(c, launch_dgemm_nn((int)N2, d3, N, 1.0, dY1.p, (long)N2, (long long)N3, C3, N, 0, 0.0, dY2.p, (long)N2,
                              (long long)(N2 * d3), d4, st));
    // Y3(mu,q2,q3,q4) = sum_nu Y2(mu,nu,q3,q4) C2(nu,q2), batched over (q3,q4)
    HIPCHK(c, launch_dgemm_nn(N, d2, N, 1.0, dY2.p, N, (long long)N2, C2, N, 0, 0.0, dY3.p, N, (long long)N * d2,
                              (long long)d3 * d4, st));
    // X(q1,q2,q3,q4) = sum_mu Cv(mu,q1) Y3(mu,q2,q3,q4)
    HIPCHK(c, launch_dgemm_nn(d1, d2 * d3 * d4, N, 1.0, dCvT.p, d1, 0, dY3.p, N, 0, 0.0, dX.p, d1, 0, 1, st));
    std::vector<double> ev(e, e + nbas);
    HIPCHK(c, de.upload(ev));
    const int nblk = 1024;
    HIPCHK(c, dpart.alloc(2 * nblk));
    // strides of (i, a, j, b) in X
    long long si, sa, sj, sb;
    if (!phys) { sa = 1; si = v; sb = (long long)v * o; sj = (long long)v * o * v; }
    else { sa = 1; sb = v; si = (long long)v * v; sj = (long long)v * v * o; }
    HIPCHK(c, launch_mp2_sums(dX.p, de.p, o, v, si, sa, sj, sb, dpart.p, nblk, st));
    std::vector<double> part(2 * nblk);
    HIPCHK(c, hipMemcpyAsync(part.data(), dpart.p, sizeof(double) * 2 * nblk, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    double E2d = 0.0, E2x = 0.0;
    for (int b = 0; b < nblk; ++b) {
        E2d += part[2 * b];
        E2x += part[2 * b + 1];
    }
    out[0] = 2.0 * E2d - E2x;
    out[1] = 2.0 * E2d;
    out[2] = -E2x;
    out[3] = now_ms() - t0;
    return TCG_OK;
}

// The k blocks of the coefficients, orbital energies and occupations of a k-point
// MP2 call, the (ov|ov) ranges set_cc_dims sets (src/module/cc_dims.f90:68-88) and
// the phase tables e^{+i k.R} (cells and k points x fastest).
struct KmpInput {
    std::vector<double2> Ck;                    // [k][nf x nf] column-major blocks
    std::vector<double> ek, cosm, sinm, cosn, sinn;
    std::vector<int> oc;
    int o_hi = 0, v_lo = 0;
};

static int kmp2_input(tcg_ctx *c, const char *who, const double *coeff_sao, const double *e, const int *occ,
                      KmpInput &K) {
    const int nf = c->nfuc, N = c->ncell, nb = c->nbas;
    K.Ck.assign((size_t)N * nf * nf, double2{0.0, 0.0});
    K.ek.assign((size_t)N * nf, 0.0);
    K.oc.assign((size_t)N * nf, 0);
    int o_hi = 0, omin = nf;
    double offblk = 0.0, cmax = 0.0;
    for (int col = 0; col < nb; ++col)
        for (int row = 0; row < nb; ++row) {
            const double re = coeff_sao[2 * ((size_t)row + (size_t)nb * col)];
            const double im = coeff_sao[2 * ((size_t)row + (size_t)nb * col) + 1];
            const double m = std::hypot(re, im);
            cmax = std::max(cmax, m);
            if (row / nf != col / nf) offblk = std::max(offblk, m);
            else K.Ck[(size_t)(row / nf) * nf * nf + (row % nf) + (size_t)nf * (col % nf)] = double2{re, im};
        }
    if (offblk > 1e-10 * std::max(1.0, cmax))
        return c->fail(TCG_E_ARG, "%s: coefficients not k-block ordered (off-block %.3e)", who, offblk);
    for (int k = 0; k < N; ++k) {
        int nok = 0;
        for (int p = 0; p < nf; ++p) {
            K.ek[(size_t)k * nf + p] = e[(size_t)k * nf + p];
            K.oc[(size_t)k * nf + p] = occ[(size_t)k * nf + p];
            const int o = occ[(size_t)k * nf + p];
            if (o != 0 && o != 2) return c->fail(TCG_E_ARG, "%s: occupation must be 0 or 2", who);
            if (o == 2) {
                if (p != nok) return c->fail(TCG_E_ARG, "%s: occupied bands are not the lowest ones in a k block", who);
                ++nok;
            }
        }
        o_hi = std::max(o_hi, nok);
        omin = std::min(omin, nok);
    }
    K.o_hi = o_hi;
    K.v_lo = omin;                             // first virtual band (0-based)
    if (o_hi == 0 || K.v_lo >= nf) return c->fail(TCG_E_ARG, "%s: no occupied or no virtual bands", who);
    K.cosm.resize((size_t)N * N);
    K.sinm.resize((size_t)N * N);
    K.cosn.resize((size_t)N * N);
    K.sinn.resize((size_t)N * N);
    const int nx = c->ncl[0], ny = c->ncl[1], nz = c->ncl[2];
    for (int k = 0; k < N; ++k)
        for (int R = 0; R < N; ++R) {
            const double th = 2.0 * kPi * ((double)(k % nx) * (R % nx) / nx + (double)((k / nx) % ny) * ((R / nx) % ny) / ny +
                                           (double)(k / (nx * ny)) * (R / (nx * ny)) / nz);
            K.cosm[R + (size_t)N * k] = std::cos(th);
            K.sinm[R + (size_t)N * k] = std::sin(th);
            K.cosn[R + (size_t)N * k] = std::cos(th) / N;
            K.sinn[R + (size_t)N * k] = std::sin(th) / N;
        }
    return 0;
}

// Orbital factors of the fused transforms: TA = conj C_k(a, m) over the occupied
// bands, TB = C_k(a, v_lo + m) over the virtual ones, [(k nf + a) M + m].
static void kmp2_factors(const KmpInput &K, int nf, int N, std::vector<double2> &TA, std::vector<double2> &TB) {
    const int o_hi = K.o_hi, nv = nf - K.v_lo;
    TA.resize((size_t)N * nf * o_hi);
    TB.resize((size_t)N * nf * nv);
    for (int k = 0; k < N; ++k)
        for (int a = 0; a < nf; ++a) {
            for (int m = 0; m < o_hi; ++m) {
                const double2 v = K.Ck[(size_t)k * nf * nf + a + (size_t)nf * m];
                TA[((size_t)k * nf + a) * o_hi + m] = double2{v.x, -v.y};
            }
            for (int m = 0; m < nv; ++m)
                TB[((size_t)k * nf + a) * nv + m] = K.Ck[(size_t)k * nf * nf + a + (size_t)nf * (K.v_lo + m)];
        }
}

// kMP2_calc on the (ov|ov) blocks in km_kmo: out[0..3] as tcg_kmp2
static int kmp2_sums(tcg_ctx *c, int o_hi, int v_lo, double *out) {
    const int nf = c->nfuc, N = c->ncell;
    const int nblk = 1024;
    HIPCHK(c, c->km_part.alloc(2 * nblk));
    HIPCHK(c, c->km_nz.alloc(1));
    HIPCHK(c, launch_fill((double *)c->km_nz.p, 1, 0.0, c->st));
    HIPCHK(c, launch_kmp2_sums(c->km_kmo.p, c->km_e.p, c->km_occ.p, nf, N, c->ncl[0], c->ncl[1], c->ncl[2], o_hi, v_lo,
                               c->km_part.p, nblk, c->km_nz.p, c->st));
    std::vector<double> part(2 * nblk);
    unsigned long long nzero = 0;
    HIPCHK(c, hipMemcpyAsync(part.data(), c->km_part.p, sizeof(double) * 2 * nblk, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(&nzero, c->km_nz.p, sizeof nzero, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    double E2d = 0.0, E2x = 0.0;
    for (int b = 0; b < nblk; ++b) {
        E2d += part[2 * b];
        E2x += part[2 * b + 1];
    }
    out[0] = 2.0 * E2d - E2x;
    out[1] = 2.0 * E2d;
    out[2] = -E2x;
    out[3] = (double)nzero;
    return 0;
}

int tcg_kmp2(tcg_ctx *c, int nbas, const double *coeff_sao, const double *e, const int *occ, double *out) {
    if (!c) return TCG_E_ARG;
    if (nbas != c->nbas || !coeff_sao || !e || !occ || !out) return c->fail(TCG_E_ARG, "tcg_kmp2: bad arguments");
    if (c->eri_n1 != c->nfuc || !c->d_eri.p || c->eri_physicist)
        return c->fail(TCG_E_STATE, "tcg_kmp2 needs the chemist SAO ERI resident (tcg_eri_build with sao = 1)");
    const int nf = c->nfuc, N = c->ncell, nb = c->nbas;
    KmpInput KI;
    {
        const int rc = kmp2_input(c, "tcg_kmp2", coeff_sao, e, occ, KI);
        if (rc) return rc;
    }
    const std::vector<double2> &Ck = KI.Ck;
    const std::vector<double> &cosm = KI.cosm, &sinm = KI.sinm, &cosn = KI.cosn, &sinn = KI.sinn, &ek = KI.ek;
    const std::vector<int> &oc = KI.oc;
    const int o_hi = KI.o_hi, v_lo = KI.v_lo;
    const int nx = c->ncl[0], ny = c->ncl[1], nz = c->ncl[2];
    const size_t nbl = (size_t)nb, M1 = (size_t)nf * nbl * nbl * nf;     // rows (a, j2, j3, d)
    DevBuf<double> X1r, X1i, X2r, X2i, Sr, Si;
    auto &dcos = c->km_cos, &dsin = c->km_sin, &dcosn = c->km_cosn, &dsinn = c->km_sinn, &de = c->km_e;
    auto &dpart = c->km_part;
    auto &dCk = c->km_Ck, &dekmo = c->km_kmo;
    auto &docc = c->km_occ;
    auto &dnz = c->km_nz;
    HIPCHK(c, upload_async(c, dcos, cosm));
    HIPCHK(c, upload_async(c, dsin, sinm));
    HIPCHK(c, upload_async(c, dcosn, cosn));
    HIPCHK(c, upload_async(c, dsinn, sinn));
    HIPCHK(c, upload_async(c, dCk, Ck));
    HIPCHK(c, upload_async(c, de, ek));
    HIPCHK(c, upload_async(c, docc, oc));
    const int nv = nf - v_lo;
    const size_t nkmo = (size_t)o_hi * nv * o_hi * nv * N * N * N;
    hipStream_t st = c->st;
    const double t0 = now_ms();
    if (N <= 64) {
        // Each lattice transform is fused with the orbital contraction of its index pair
        // (k_dft_mo), occupied index first, so that every later pass runs on the
        // reduced tensor:
        //   A: (c R3) -> (r k3), conj C_k3(c, r)   Y_A(a b R2, r k3, j4)
        //   B: (d R4) -> (s k4), C_k4(d, s)        Y_B(a b R2 r k3, s k4)
        //   C: (b R2) -> (q k2), C_k2(b, q) / N    Y_C(a, q k2, r k3 s k4)
        //   D: a -> p, conj C_k1(a, p), k1 = k4 - k3 + k2
        // with e^{-i k3 R3}, e^{+i k4 R4}, e^{+i k2 R2} (the reference's phase convention).
        // The bra pair (a in cell 0, b in cell R2) is non-zero only for R2 near cell 0:
        // the passes run over those cells alone (from the non-zero column lists).
        {
            const int rc = stored_cols(c, c->d_eri.p, nf, nb, 0);
            if (rc) return rc;
        }
        auto &dfl = c->km_fl;
        HIPCHK(c, dfl.alloc(nb));
        HIPCHK(c, hipMemsetAsync(dfl.p, 0, nb, st));
        HIPCHK(c, launch_list_j2_flags(c->d_list_d.p, c->cols_nz, nb, dfl.p, st));
        std::vector<unsigned char> fl(nb);
        HIPCHK(c, hipMemcpyAsync(fl.data(), dfl.p, nb, hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        std::vector<int> cells2;
        for (int R = 0; R < N; ++R) {
            bool any = false;
            for (int b2 = 0; b2 < nf; ++b2) any |= fl[(size_t)R * nf + b2] != 0;
            if (any) cells2.push_back(R);
        }
        if (cells2.empty()) cells2.push_back(0);               // an all-zero tensor: E_c = 0
        const int n2 = (int)cells2.size();
        std::vector<double> cos2((size_t)n2 * N), sin2((size_t)n2 * N);
        for (int k = 0; k < N; ++k)
            for (int i = 0; i < n2; ++i) {
                cos2[i + (size_t)n2 * k] = cosm[cells2[i] + (size_t)N * k];
                sin2[i + (size_t)n2 * k] = sinm[cells2[i] + (size_t)N * k];
            }
        std::vector<double2> TA, TB;
        kmp2_factors(KI, nf, N, TA, TB);
        const long long IA = (long long)nf * nf * n2, IB = IA * o_hi * N;
        std::vector<long long> imap(IA);
        for (long long i = 0; i < IA; ++i) {
            const long long a = i % nf, b2 = (i / nf) % nf, R2 = cells2[i / ((long long)nf * nf)];
            imap[i] = a + (long long)nf * (b2 + (long long)nf * R2);
        }
        auto &dTA = c->km_TA, &dTB = c->km_TB;
        auto &dcos2 = c->km_cos2, &dsin2 = c->km_sin2;
        auto &dimap = c->km_imap;
        HIPCHK(c, upload_async(c, dTA, TA));
        HIPCHK(c, upload_async(c, dTB, TB));
        HIPCHK(c, upload_async(c, dcos2, cos2));
        HIPCHK(c, upload_async(c, dsin2, sin2));
        HIPCHK(c, upload_async(c, dimap, imap));
        const size_t nYA = (size_t)IA * o_hi * N * nb, nYB = (size_t)IB * nv * N;
        const size_t nYC = (size_t)nf * nv * N * o_hi * N * nv * N;
        auto &Ar = c->km_Ar, &Ai = c->km_Ai, &Br = c->km_Br, &Bi = c->km_Bi;
        HIPCHK(c, Ar.alloc(std::max(nYA, nYC)));
        HIPCHK(c, Ai.alloc(std::max(nYA, nYC)));
        HIPCHK(c, Br.alloc(nYB));
        HIPCHK(c, Bi.alloc(nYB));
        DftMoArgs d{};
        // A: rows (a, b, R2 in cells2) gathered from the tensor, outer j4
        d.Xr = c->d_eri.p;
        d.Xi = nullptr;
        d.I = IA;
        d.O = nb;
        d.imap = dimap.p;
        d.sJ = (long long)nf * nb;
        d.sO = (long long)nf * nb * nb;
        d.nf = nf;
        d.N = N;
        d.NR = N;
        d.M = o_hi;
        d.cosm = dcos.p;
        d.sinm = dsin.p;
        d.sgn = -1.0;
        d.scale = 1.0;
        d.T = dTA.p;
        d.Yr = Ar.p;
        d.Yi = Ai.p;
        HIPCHK(c, launch_dft_mo(d, st));
        // B: rows (a b R2, r k3), j4 = d + nf R4 at stride IB
        d.Xr = Ar.p;
        d.Xi = Ai.p;
        d.I = IB;
        d.O = 1;
        d.imap = nullptr;
        d.sJ = IB;
        d.sO = IB * nf * N;
        d.M = nv;
        d.sgn = 1.0;
        d.T = dTB.p;
        d.Yr = Br.p;
        d.Yi = Bi.p;
        HIPCHK(c, launch_dft_mo(d, st));
        // C: rows a, j2 = b + nf R2 over cells2 at stride nf, outer (r k3, s k4) at stride IA
        d.Xr = Br.p;
        d.Xi = Bi.p;
        d.I = nf;
        d.O = (long long)o_hi * N * nv * N;
        d.sJ = nf;
        d.sO = IA;
        d.NR = n2;
        d.cosm = dcos2.p;
        d.sinm = dsin2.p;
        d.scale = 1.0 / N;
        d.Yr = Ar.p;
        d.Yi = Ai.p;
        HIPCHK(c, launch_dft_mo(d, st));
        HIPCHK(c, dekmo.alloc(nkmo));
        HIPCHK(c, launch_kmo_last(Ar.p, Ai.p, dCk.p, nf, N, nx, ny, nz, o_hi, v_lo, dekmo.p, st));
    } else {
    // lattice transforms as real GEMMs on the whole SAO tensor, then the quarter transforms
    const size_t nX = M1 * N;                  // = nf^2 nb^2 (all elements, k in place of R)
    HIPCHK(c, X1r.alloc(nX));
    HIPCHK(c, X1i.alloc(nX));
    // 1) R4 -> k4: X1(a,j2,j3,d,k4) = sum_R4 ERI(a,j2,j3,d R4) e^{+i k4 R4}
    HIPCHK(c, launch_dgemm_nn((int)M1, N, N, 1.0, c->d_eri.p, (long)M1, 0, dcos.p, N, 0, 0.0, X1r.p, (long)M1, 0, 1, st));
    HIPCHK(c, launch_dgemm_nn((int)M1, N, N, 1.0, c->d_eri.p, (long)M1, 0, dsin.p, N, 0, 0.0, X1i.p, (long)M1, 0, 1, st));
    // 2) R3 -> k3 (e^{-i k3 R3}), batched over (d, k4): rows (a, j2, c), X2(a,j2,c,k3,d,k4)
    HIPCHK(c, X2r.alloc(nX));
    HIPCHK(c, X2i.alloc(nX));
    {
        const long long rows = (long long)nf * nb * nf, bs = (long long)nf * nb * nb, nbatch = (long long)nf * N;
        // re = X1r cos + X1i sin ; im = X1i cos - X1r sin
        HIPCHK(c, launch_dgemm_nn((int)rows, N, N, 1.0, X1r.p, (long)rows, bs, dcos.p, N, 0, 0.0, X2r.p, (long)rows, bs, nbatch, st));
        HIPCHK(c, launch_dgemm_nn((int)rows, N, N, 1.0, X1i.p, (long)rows, bs, dsin.p, N, 0, 1.0, X2r.p, (long)rows, bs, nbatch, st));
        HIPCHK(c, launch_dgemm_nn((int)rows, N, N, 1.0, X1i.p, (long)rows, bs, dcos.p, N, 0, 0.0, X2i.p, (long)rows, bs, nbatch, st));
        HIPCHK(c, launch_dgemm_nn((int)rows, N, N, -1.0, X1r.p, (long)rows, bs, dsin.p, N, 0, 1.0, X2i.p, (long)rows, bs, nbatch, st));
    }
    X1r.release();
    X1i.release();
    // 3) R2 -> k2 (e^{+i k2 R2} / N), batched over (c, k3, d, k4): rows (a, b), S(a,b,k2,c,k3,d,k4)
    HIPCHK(c, Sr.alloc(nX));
    HIPCHK(c, Si.alloc(nX));
    {
        const long long rows = (long long)nf * nf, bs = (long long)nf * nb, nbatch = (long long)nf * N * nf * N;
        HIPCHK(c, launch_dgemm_nn((int)rows, N, N, 1.0, X2r.p, (long)rows, bs, dcosn.p, N, 0, 0.0, Sr.p, (long)rows, bs, nbatch, st));
        HIPCHK(c, launch_dgemm_nn((int)rows, N, N, -1.0, X2i.p, (long)rows, bs, dsinn.p, N, 0, 1.0, Sr.p, (long)rows, bs, nbatch, st));
        HIPCHK(c, launch_dgemm_nn((int)rows, N, N, 1.0, X2r.p, (long)rows, bs, dsinn.p, N, 0, 0.0, Si.p, (long)rows, bs, nbatch, st));
        HIPCHK(c, launch_dgemm_nn((int)rows, N, N, 1.0, X2i.p, (long)rows, bs, dcosn.p, N, 0, 1.0, Si.p, (long)rows, bs, nbatch, st));
    }
    X2r.release();
    X2i.release();
    // 4) quarter transforms with the k-block coefficients, (ov|ov) only
    HIPCHK(c, dekmo.alloc(nkmo));
    HIPCHK(c, launch_kmo(Sr.p, Si.p, dCk.p, nf, N, nx, ny, nz, o_hi, v_lo, dekmo.p, st));
    }
    // 5) kMP2_calc
    {
        const int rc = kmp2_sums(c, o_hi, v_lo, out);
        if (rc) return rc;
    }
    out[4] = now_ms() - t0;
    return TCG_OK;
}

// k-point MP2 without the stored tensor: the SAO integrals are built slab by slab
// of the last index (j4 = d + nf R4 over a set of cells R4), each slab through
// the first two fused transforms (R3 -> k3 with the occupied index, R4 -> k4 with
// the virtual one, the latter summed over the slabs), then the last two transforms
// and the sums once.  Linear in j4, so the result is tcg_kmp2's on the stored tensor.
int tcg_kmp2_direct(tcg_ctx *c, int nbas, const double *coeff_sao, const double *e, const int *occ, int slab_cells,
                    double *out) {
    if (!c) return TCG_E_ARG;
    if (nbas != c->nbas || !coeff_sao || !e || !occ || !out || slab_cells < 0)
        return c->fail(TCG_E_ARG, "tcg_kmp2_direct: bad arguments");
    if (!c->direct_ready || !c->prepared) return c->fail(TCG_E_STATE, "tcg_kmp2_direct: tcg_direct_init not called");
    const int nf = c->nfuc, N = c->ncell, nb = c->nbas;
    if (N > 64) return c->fail(TCG_E_ARG, "tcg_kmp2_direct: at most 64 cells (%d)", N);
    KmpInput KI;
    {
        const int rc = kmp2_input(c, "tcg_kmp2_direct", coeff_sao, e, occ, KI);
        if (rc) return rc;
    }
    const int o_hi = KI.o_hi, v_lo = KI.v_lo, nv = nf - v_lo;
    hipStream_t st = c->st;
    const double t0 = now_ms();
    if (c->nq4 < 0) {                           // a deferred enumeration left its count on the device
        const int rc0 = enumerate_quartets(c, false);
        if (rc0) return rc0;
    }
    // bra cells R2 with any stored element (a in cell 0, b in cell R2): those whose
    // pair Schwarz value passes the element screen against the largest one
    std::vector<double> Q((size_t)nb * nb);
    HIPCHK(c, hipMemcpyAsync(Q.data(), c->d_Q.p, sizeof(double) * Q.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    const double qmax = *std::max_element(Q.begin(), Q.end());
    std::vector<int> cells2;
    for (int R = 0; R < N; ++R) {
        bool any = false;
        for (int b = 0; b < nf && !any; ++b)
            for (int a = 0; a < nf && !any; ++a) any = Q[(size_t)a + (size_t)nb * (b + (size_t)nf * R)] * qmax >= kSchwarz;
        if (any) cells2.push_back(R);
    }
    if (cells2.empty()) cells2.push_back(0);
    const int n2 = (int)cells2.size();
    const long long IA = (long long)nf * nf * n2, IB = IA * o_hi * N;
    // The slab builds run the row pipeline in chunks of a smaller row budget (8 GB:
    // more chunks, each still millions of rows), which leaves the memory to the slabs
    struct BudgetScope {
        tcg_ctx *c;
        double saved;
        ~BudgetScope() { c->row_budget = saved; }
    } budget_scope{c, c->row_budget};
    if (!c->row_budget_set) c->row_budget = std::min(c->row_budget, 8e9);
    // cells per slab: the caller's, or what the free device memory holds next to Y_B
    // (summed over the slabs) and the row pipeline of a slab build (its row budget, a
    // quarter more for survivors and sort space); per cell the slab and Y_A.  Y_C
    // takes the slab's memory after the last slab.
    const size_t nYB = (size_t)IB * nv * N, nYC = (size_t)nf * nv * N * o_hi * N * nv * N;
    const size_t per_cell = (size_t)nf * nb * nb * nf * sizeof(double) + (size_t)IA * o_hi * N * nf * 2 * sizeof(double);
    const size_t fixed = 2 * sizeof(double) * nYB + (size_t)(1.25 * c->row_budget) + ((size_t)1 << 32);
    int nR = slab_cells;
    {
        size_t fr = 0, tot = 0;
        HIPCHK(c, hipMemGetInfo(&fr, &tot));
        // the screened direct J/K's scratch lists and the row pipeline's buffers are
        // rebuilt on demand: their memory goes to the slabs when the slabs' cells do
        // not fit beside them (they stay otherwise: allocating many GB right after
        // freeing them can take seconds, here and in the next J/K)
        const size_t want = (size_t)(slab_cells > 0 ? std::min(slab_cells, N) : N);
        if (fr < fixed + want * per_cell) {
            c->d_q4scr.release();
            c->d_qrscr.release();
            c->d_keep.release();
            c->d_sidx.release();
            c->d_surv.release();
            c->d_rows.release();
            c->d_keys.release();
            c->d_keys2.release();
            c->d_vals.release();
            c->d_order.release();
            c->d_rowval.release();
            c->d_realval.release();
            HIPCHK(c, hipDeviceSynchronize());
            HIPCHK(c, hipMemGetInfo(&fr, &tot));
        }
        c->prof = getenv("TCG_PROFILE") != nullptr;
        c->marks.clear();
        c->mark("release scratch");
        if (fr < 2 * sizeof(double) * (nYB + nYC) + ((size_t)1 << 30))
            return c->fail(TCG_E_NOMEM, "tcg_kmp2_direct: the transforms need %.1f GB, %.1f GB free",
                           2e-9 * sizeof(double) * (nYB + nYC), 1e-9 * fr);
        if (nR == 0) {
            const size_t avail = fr > fixed ? fr - fixed : 0;
            nR = (int)std::max<size_t>(1, std::min<size_t>((size_t)N, avail / per_cell));
        }
    }
    nR = std::min(nR, N);
    const int nslab = (N + nR - 1) / nR;
    // the slabs' cells: ordered by the distance of the cell from cell 0 (nearest
    // image), each cell next to its inverse.  A quartet's permutations store into
    // the relative cells of its shell pairs and their inverses, near cell 0 for the
    // quartets that survive the screens, so the first slabs gather most quartets and
    // each quartet reaches few slabs (the slab builds skip the quartets they do not store)
    std::vector<int> corder;
    {
        const int n0 = c->ncl[0], n1 = c->ncl[1], n2 = c->ncl[2];
        const double h[3] = {c->L[0] / n0, c->L[1] / n1, c->L[2] / n2};
        auto coord = [&](int R, int d) {
            const int n = d == 0 ? n0 : d == 1 ? n1 : n2;
            const int x = d == 0 ? R % n0 : d == 1 ? (R / n0) % n1 : R / (n0 * n1);
            return 2 * x <= n ? x : x - n;
        };
        auto inverse = [&](int R) {
            const int x = R % n0, y = (R / n0) % n1, z = R / (n0 * n1);
            return (n0 - x) % n0 + n0 * ((n1 - y) % n1 + n1 * ((n2 - z) % n2));
        };
        std::vector<std::pair<double, int>> cls;
        for (int R = 0; R < N; ++R) {
            if (inverse(R) < R) continue;
            double d2 = 0.0;
            for (int d = 0; d < 3; ++d) d2 += (coord(R, d) * h[d]) * (coord(R, d) * h[d]);
            cls.push_back({d2, R});
        }
        std::sort(cls.begin(), cls.end());
        for (const auto &cl : cls) {
            corder.push_back(cl.second);
            if (inverse(cl.second) != cl.second) corder.push_back(inverse(cl.second));
        }
    }
    std::vector<int> spos((size_t)nslab * N, -1);
    for (int sl = 0; sl < nslab; ++sl)
        for (int r = 0; r < std::min(nR, N - sl * nR); ++r) spos[(size_t)sl * N + corder[sl * nR + r]] = r;
    DevBuf<int> dspos;
    HIPCHK(c, dspos.upload(spos));
    // tables
    std::vector<double2> TA, TB;
    kmp2_factors(KI, nf, N, TA, TB);
    std::vector<double> cos2((size_t)n2 * N), sin2((size_t)n2 * N);
    for (int k = 0; k < N; ++k)
        for (int i = 0; i < n2; ++i) {
            cos2[i + (size_t)n2 * k] = KI.cosm[cells2[i] + (size_t)N * k];
            sin2[i + (size_t)n2 * k] = KI.sinm[cells2[i] + (size_t)N * k];
        }
    // per slab s: cos/sin of its cells corder[s nR .. s nR + nRs), table [r + nRs k] at offset s nR N
    std::vector<double> cosB((size_t)N * N), sinB((size_t)N * N);
    for (int sl = 0; sl < nslab; ++sl) {
        const int R0 = sl * nR, nRs = std::min(nR, N - R0);
        for (int k = 0; k < N; ++k)
            for (int r = 0; r < nRs; ++r) {
                cosB[(size_t)R0 * N + r + (size_t)nRs * k] = KI.cosm[corder[R0 + r] + (size_t)N * k];
                sinB[(size_t)R0 * N + r + (size_t)nRs * k] = KI.sinm[corder[R0 + r] + (size_t)N * k];
            }
    }
    std::vector<long long> imap(IA);
    for (long long i = 0; i < IA; ++i) {
        const long long a = i % nf, b2 = (i / nf) % nf, R2 = cells2[i / ((long long)nf * nf)];
        imap[i] = a + (long long)nf * (b2 + (long long)nf * R2);
    }
    HIPCHK(c, upload_async(c, c->km_cos, KI.cosm));
    HIPCHK(c, upload_async(c, c->km_sin, KI.sinm));
    HIPCHK(c, upload_async(c, c->km_cosn, cosB));
    HIPCHK(c, upload_async(c, c->km_sinn, sinB));
    HIPCHK(c, upload_async(c, c->km_Ck, KI.Ck));
    HIPCHK(c, upload_async(c, c->km_e, KI.ek));
    HIPCHK(c, upload_async(c, c->km_occ, KI.oc));
    HIPCHK(c, upload_async(c, c->km_TA, TA));
    HIPCHK(c, upload_async(c, c->km_TB, TB));
    HIPCHK(c, upload_async(c, c->km_cos2, cos2));
    HIPCHK(c, upload_async(c, c->km_sin2, sin2));
    HIPCHK(c, upload_async(c, c->km_imap, imap));
    HIPCHK(c, hipStreamSynchronize(st));
    c->mark("tables");
    // work buffers: the slab, Y_A of one slab, Y_B summed over the slabs (then Y_C)
    DevBuf<double> slab, Ar, Ai;
    const size_t nslab_el = (size_t)nf * nb * nb * ((size_t)nf * nR);
    const size_t nYA = (size_t)IA * o_hi * N * ((size_t)nf * nR);
    HIPCHK(c, slab.alloc(nslab_el));
    HIPCHK(c, Ar.alloc(nYA));
    HIPCHK(c, Ai.alloc(nYA));
    auto &Br = c->km_Br, &Bi = c->km_Bi;
    HIPCHK(c, Br.alloc(nYB));
    HIPCHK(c, Bi.alloc(nYB));
    c->mark("slab and Y buffers");
    double t_int = 0.0, t_tr = 0.0;
    for (int sl = 0; sl < nslab; ++sl) {
        const int R0 = sl * nR, nRs = std::min(nR, N - R0);
        const double ta = now_ms();
        // ---- the slab: SAO elements whose last index lies in the slab's cells, at
        // j4 = d + nf r for the cell's position r in the slab ----
        HIPCHK(c, launch_fill(slab.p, (size_t)nf * nb * nb * ((size_t)nf * nRs), 0.0, st));
        HIPCHK(c, hipMemsetAsync(c->d_cnt.p, 0, 8 * sizeof(unsigned long long), st));
        EriArgs A = base_args(c);
        A.q4 = c->d_q4.p;
        A.nq4 = c->nq4;
        if (c->direct_learned) {                // the learned list of tcg_direct_init_learned
            A.q4 = c->d_q4dl.p;
            A.nq4 = c->nq4dl;
            A.qrule = c->d_qrdl.p;
        }
        A.screen_prim = 1;
        A.Qao = c->d_Q.p;
        A.eri = slab.p;
        A.n1 = nf;
        A.physicist = 0;
        A.slab_pos = dspos.p + (size_t)sl * N;
        int rc = run_quartets(c, A);
        if (rc) return rc;
        HIPCHK(c, hipStreamSynchronize(st));
        c->mark("slab integrals");
        const double tb = now_ms();
        t_int += tb - ta;
        // ---- A: (c R3) -> (r k3), rows (a b R2 in cells2), outer j4 of the slab ----
        DftMoArgs d{};
        d.Xr = slab.p;
        d.Xi = nullptr;
        d.I = IA;
        d.O = (long long)nf * nRs;
        d.imap = c->km_imap.p;
        d.sJ = (long long)nf * nb;
        d.sO = (long long)nf * nb * nb;
        d.nf = nf;
        d.N = N;
        d.NR = N;
        d.M = o_hi;
        d.cosm = c->km_cos.p;
        d.sinm = c->km_sin.p;
        d.sgn = -1.0;
        d.scale = 1.0;
        d.T = c->km_TA.p;
        d.Yr = Ar.p;
        d.Yi = Ai.p;
        HIPCHK(c, launch_dft_mo(d, st));
        // ---- B: (d R4) -> (s k4) over the slab's cells, summed into Y_B ----
        d.Xr = Ar.p;
        d.Xi = Ai.p;
        d.I = IB;
        d.O = 1;
        d.imap = nullptr;
        d.sJ = IB;
        d.sO = IB * nf * nRs;
        d.NR = nRs;
        d.cosm = c->km_cosn.p + (size_t)R0 * N;
        d.sinm = c->km_sinn.p + (size_t)R0 * N;
        d.M = nv;
        d.sgn = 1.0;
        d.T = c->km_TB.p;
        d.Yr = Br.p;
        d.Yi = Bi.p;
        d.accum = sl > 0;
        HIPCHK(c, launch_dft_mo(d, st));
        HIPCHK(c, hipStreamSynchronize(st));
        c->mark("slab transforms");
        t_tr += now_ms() - tb;
    }
    // Y_C in the slab's memory when it holds both parts (a fresh allocation of that
    // size right after freeing the slab can take seconds), else in Y_A's buffers
    double *Cr = slab.p, *Ci = slab.p + nYC;
    if (slab.n < 2 * nYC) {
        slab.release();
        HIPCHK(c, Ar.alloc(nYC));
        HIPCHK(c, Ai.alloc(nYC));
        Cr = Ar.p;
        Ci = Ai.p;
    }
    c->mark("Y_C buffers");
    const double tc = now_ms();
    // ---- C: (b R2) -> (q k2) / N over cells2; D: a -> p ----
    DftMoArgs d{};
    d.Xr = Br.p;
    d.Xi = Bi.p;
    d.I = nf;
    d.O = (long long)o_hi * N * nv * N;
    d.imap = nullptr;
    d.sJ = nf;
    d.sO = IA;
    d.nf = nf;
    d.N = N;
    d.NR = n2;
    d.M = nv;
    d.cosm = c->km_cos2.p;
    d.sinm = c->km_sin2.p;
    d.sgn = 1.0;
    d.scale = 1.0 / N;
    d.T = c->km_TB.p;
    d.Yr = Cr;
    d.Yi = Ci;
    HIPCHK(c, launch_dft_mo(d, st));
    const size_t nkmo = (size_t)o_hi * nv * o_hi * nv * N * N * N;
    HIPCHK(c, c->km_kmo.alloc(nkmo));
    HIPCHK(c, launch_kmo_last(Cr, Ci, c->km_Ck.p, nf, N, c->ncl[0], c->ncl[1], c->ncl[2], o_hi, v_lo,
                              c->km_kmo.p, st));
    {
        const int rc = kmp2_sums(c, o_hi, v_lo, out);
        if (rc) return rc;
    }
    t_tr += now_ms() - tc;
    out[4] = now_ms() - t0;
    out[5] = t_int;
    out[6] = t_tr;
    out[7] = nslab;
    print_marks(c, "tcg_kmp2_direct", t0);
    return TCG_OK;
}

int tcg_set_learned(tcg_ctx *c, const float *params, const float *norm, const int *orders, int norders, float margin) {
    if (!c) return TCG_E_ARG;
    if (!params || !norm || !orders || norders <= 0 || norders > 15)
        return c->fail(TCG_E_ARG, "tcg_set_learned: bad arguments");
    for (int r = 0; r < norders; ++r)
        if (orders[r] < 4 || orders[r] > kMaxNQ) return c->fail(TCG_E_ARG, "tcg_set_learned: order %d", orders[r]);
    const int np = tcg_mlp_param_count();
    c->mlp_params.assign(params, params + np);
    c->mlp_norm.assign(norm, norm + 24);
    c->learned_orders.assign(orders, orders + norders);
    c->learned_margin = margin;
    c->prepared = false;          // the tables get the learned rules
    c->direct_ready = false;
    HIPCHK(c, c->d_mlp.upload(c->mlp_params));
    HIPCHK(c, c->d_mlpnorm.upload(c->mlp_norm));
    return TCG_OK;
}

int tcg_set_learned_skip(tcg_ctx *c, double thresh, double schwarz_cap, float margin_b) {
    if (!c) return TCG_E_ARG;
    if (!(thresh >= 0.0) || !(schwarz_cap >= 0.0) || !(margin_b >= 0.0f))
        return c->fail(TCG_E_ARG, "tcg_set_learned_skip: bad arguments");
    c->learned_skip = thresh;
    c->learned_skip_cap = schwarz_cap;
    c->learned_margin_b = margin_b;
    return TCG_OK;
}

int tcg_eri_build_learned(tcg_ctx *c, int sao, int physicist, double *eri, long long *rule_counts) {
    if (!c) return TCG_E_ARG;
    if (c->learned_orders.empty()) return c->fail(TCG_E_STATE, "tcg_eri_build_learned before tcg_set_learned");
    c->learned_on = true;
    const int rc = tcg_eri_build(c, sao, physicist, eri);
    c->learned_on = false;
    if (rc) return rc;
    if (rule_counts)
        for (size_t r = 0; r <= c->learned_orders.size() + 1; ++r) rule_counts[r] = c->learned_counts[r];
    return TCG_OK;
}

int tcg_eri_blocks(tcg_ctx *c, int n, const int *shells, double *out, long long *offsets) {
    if (!c) return TCG_E_ARG;
    if (c->nbas == 0) return c->fail(TCG_E_STATE, "tcg_eri_blocks before tcg_set_basis");
    if (n <= 0 || !shells || !out || !offsets) return c->fail(TCG_E_ARG, "tcg_eri_blocks: bad arguments");
    int rc = prepare(c);
    if (rc) return rc;
    std::vector<Quartet> q4(n);
    size_t nout = 0;
    for (int x = 0; x < n; ++x) {
        int s4[4];
        for (int y = 0; y < 4; ++y) {
            s4[y] = shells[4 * x + y];
            if (s4[y] < 0 || s4[y] >= c->nsh) return c->fail(TCG_E_ARG, "tcg_eri_blocks: shell %d out of range", s4[y]);
        }
        // translate the quartet so that its first shell is in the unit cell
        const int c0 = c->sh[s4[0]].cell;
        int t4[4];
        for (int y = 0; y < 4; ++y) t4[y] = c->relc[(size_t)c0 * c->ncell + c->sh[s4[y]].cell] * c->nshuc + c->sh[s4[y]].loc;
        Quartet &Q = q4[x];
        for (int y = 0; y < 4; ++y) Q.s[y] = t4[y];
        Q.bra = rep_pair(c, t4[0], t4[1]);
        Q.ket = rep_pair(c, t4[2], t4[3]);
        const int lK = c->sh[t4[2]].loc;
        for (int d = 0; d < 3; ++d) Q.shift[d] = c->sh[t4[2]].xyz[d] - c->sh[lK].xyz[d];
        Q.out = (int)nout;
        Q.prim = -1;
        offsets[x] = (long long)nout;
        nout += (size_t)c->sh[t4[0]].ncomp * c->sh[t4[1]].ncomp * c->sh[t4[2]].ncomp * c->sh[t4[3]].ncomp;
    }
    offsets[n] = (long long)nout;
    std::vector<double> blocks;
    rc = run_blocks(c, q4, nout, &blocks, 1);
    if (rc) return rc;
    std::copy(blocks.begin(), blocks.end(), out);
    return TCG_OK;
}

int tcg_one_electron(tcg_ctx *c, int natom, const double *xyz, const int *Z, double *S, double *T, double *V) {
    if (!c) return TCG_E_ARG;
    if (c->nbas == 0) return c->fail(TCG_E_STATE, "tcg_one_electron before tcg_set_basis");
    if (natom <= 0 || !xyz || !Z || !S || !T || !V) return c->fail(TCG_E_ARG, "tcg_one_electron: bad arguments");
    int rc = prepare(c);
    if (rc) return rc;
    const int nb = c->nbas, nf = c->nfuc;
    const size_t nn = (size_t)nb * nb;
    // S and T rows of the unit cell on the GPU, the other rows by translation there too;
    // V (its own kernel) after them on the same stream
    HIPCHK(c, c->d_S1.alloc(nn));
    HIPCHK(c, c->d_T1.alloc(nn));
    OneStArgs g = one_args(c);
    g.S = c->d_S1.p;
    g.T = c->d_T1.p;
    int mp = 1;
    for (int i = 0; i < nb; ++i) mp = std::max(mp, c->poff[i + 1] - c->poff[i]);
    DevBuf<double> work;
    HIPCHK(c, work.alloc(2 * (size_t)nf * nb * mp * mp));
    HIPCHK(c, launch_one_st(g, mp, work.p, c->st));
    HIPCHK(c, launch_translate_fill(c->d_S1.p, nf, nb, c->ncell, c->d_relc.p, c->st));
    HIPCHK(c, launch_translate_fill(c->d_T1.p, nf, nb, c->ncell, c->d_relc.p, c->st));
    std::vector<double> Vv;
    rc = nuclear_rows(c, natom, xyz, Z, Vv, nullptr, true);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(S, c->d_S1.p, sizeof(double) * nn, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(T, c->d_T1.p, sizeof(double) * nn, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    std::copy(Vv.begin(), Vv.end(), V);
    return TCG_OK;
}

int tcg_normalize(tcg_ctx *c, int nbas, const double *center, const int *lxyz, const int *prim_off,
                  const double *expo, const double *coef_raw, double *coef_out) {
    if (!c) return TCG_E_ARG;
    if (nbas <= 0 || !center || !lxyz || !prim_off || !expo || !coef_raw || !coef_out)
        return c->fail(TCG_E_ARG, "tcg_normalize: bad arguments");
    const int np = prim_off[nbas];
    if (np <= 0) return c->fail(TCG_E_ARG, "tcg_normalize: no primitives");
    for (int i = 0; i < nbas; ++i) {
        if (prim_off[i + 1] <= prim_off[i]) return c->fail(TCG_E_ARG, "tcg_normalize: AO %d has no primitive", i);
        if (lxyz[3 * i] + lxyz[3 * i + 1] + lxyz[3 * i + 2] > kMaxL)
            return c->fail(TCG_E_ARG, "tcg_normalize: AO %d above l = %d", i, kMaxL);
    }
    for (int p = 0; p < np; ++p)
        if (!(expo[p] > 0.0)) return c->fail(TCG_E_ARG, "tcg_normalize: exponent %d not positive", p);
    DevBuf<double> dc, de, draw, dout;
    DevBuf<int> dl, dp;
    HIPCHK(c, dc.upload(std::vector<double>(center, center + 3 * (size_t)nbas)));
    HIPCHK(c, dl.upload(std::vector<int>(lxyz, lxyz + 3 * (size_t)nbas)));
    HIPCHK(c, dp.upload(std::vector<int>(prim_off, prim_off + nbas + 1)));
    HIPCHK(c, de.upload(std::vector<double>(expo, expo + np)));
    HIPCHK(c, draw.upload(std::vector<double>(coef_raw, coef_raw + np)));
    HIPCHK(c, dout.alloc(np));
    OneStArgs g{};
    g.ndp = c->ndp;
    g.nf = nbas;
    g.nb = nbas;
    for (int d = 0; d < 3; ++d) {
        g.a[d] = c->a[d];
        g.L[d] = c->L[d];
    }
    g.center = dc.p;
    g.lxyz = dl.p;
    g.poff = dp.p;
    g.expo = de.p;
    int mp = 1;
    for (int i = 0; i < nbas; ++i) mp = std::max(mp, prim_off[i + 1] - prim_off[i]);
    DevBuf<double> work;
    HIPCHK(c, work.alloc((size_t)nbas * mp * mp));
    HIPCHK(c, launch_normalize(g, draw.p, dout.p, mp, work.p, c->st));
    HIPCHK(c, hipMemcpyAsync(coef_out, dout.p, sizeof(double) * np, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return TCG_OK;
}

int tcg_set_quadrature(tcg_ctx *c, int fast, int nquad) {
    if (!c) return TCG_E_ARG;
    if (fast ? (nquad < 4 || nquad > kMaxNQ) : (nquad != 64 && nquad != 128))
        return c->fail(TCG_E_ARG, "tcg_set_quadrature: fast=%d with %d points (fast: 4..%d; tables: 64 or 128)", fast,
                       nquad, kMaxNQ);
    if (c->fast != (fast != 0) || c->nq != nquad) {
        c->fast = fast != 0;
        c->nq = nquad;
        c->nqp = ((nquad + kNodeTile - 1) / kNodeTile) * kNodeTile;
        c->prepared = false;
        c->cols_ready = false;
        c->direct_ready = false;
    }
    return TCG_OK;
}

int tcg_get_quadrature(tcg_ctx *c, int *fast, int *nquad, double *t, double *w) {
    if (!c || !fast || !nquad) return TCG_E_ARG;
    *fast = c->fast;
    *nquad = c->nq;
    if (t || w) {
        if (c->nbas == 0) return c->fail(TCG_E_STATE, "tcg_get_quadrature: nodes need tcg_set_basis first");
        const int rc = prepare(c);
        if (rc) return rc;
        if (t) std::copy(c->tnode.begin(), c->tnode.begin() + c->nq, t);
        if (w) std::copy(c->tweight.begin(), c->tweight.begin() + c->nq, w);
    }
    return TCG_OK;
}

int tcg_fast_quad_self_check(tcg_ctx *c, int natom, const double *xyz, const int *Z, int nq_full, double *deri,
                             double *dv, int *fast_on) {
    if (!c) return TCG_E_ARG;
    if (c->nbas == 0) return c->fail(TCG_E_STATE, "tcg_fast_quad_self_check before tcg_set_basis");
    if (natom <= 0 || !xyz || !Z || !deri || !dv || !fast_on || (nq_full != 64 && nq_full != 128))
        return c->fail(TCG_E_ARG, "tcg_fast_quad_self_check: bad arguments");
    *deri = *dv = 0.0;
    *fast_on = c->fast;
    if (!c->fast) return TCG_OK;                  // nothing to check (fast_quad_self_check.f90:60)
    const int nq_fast = c->nq, nsh = c->nsh, nshuc = std::max(c->nshuc, 1), nf = c->nfuc, nb = c->nbas;
    // sample (fast_quad_self_check.f90:113-161): on-site quartets, 50 LCG quartets, V diagonals + 2 x 20 pairs
    unsigned long long seed = 12345ull;
    auto next = [&]() {
        seed = (seed * 1103515245ull + 12345ull) % 2147483648ull;
        return (long long)(seed / 16);
    };
    const int nrand = 50, nv_rand = 20;
    std::vector<Quartet> q4;
    std::vector<size_t> boff;
    size_t nout = 0;
    for (int it = 0; it < nshuc + nrand; ++it) {
        int q[4];
        if (it < nshuc) {
            q[0] = q[1] = q[2] = q[3] = it;
        } else {
            q[0] = (int)(next() % nshuc);
            for (int x = 1; x < 4; ++x) q[x] = (int)(next() % nsh);
        }
        Quartet Q{};
        for (int x = 0; x < 4; ++x) Q.s[x] = q[x];
        Q.bra = rep_pair(c, q[0], q[1]);
        Q.ket = rep_pair(c, q[2], q[3]);
        const int lK = c->sh[q[2]].loc;
        for (int d = 0; d < 3; ++d) Q.shift[d] = c->sh[q[2]].xyz[d] - c->sh[lK].xyz[d];
        Q.out = (int)nout;
        Q.prim = -1;
        q4.push_back(Q);
        boff.push_back(nout);
        nout += (size_t)c->sh[q[0]].ncomp * c->sh[q[1]].ncomp * c->sh[q[2]].ncomp * c->sh[q[3]].ncomp;
    }
    std::vector<std::pair<int, int>> pv;
    for (int ip = 0; ip < nshuc; ++ip) pv.push_back({c->sh[ip].first_ao, c->sh[ip].first_ao});
    for (int ip = 0; ip < 2 * nv_rand; ++ip) {
        const int a = (int)(next() % nf);
        const int b = (int)(next() % (ip < nv_rand ? nf : nb));
        pv.push_back({a, b});
    }
    // both rules
    std::vector<double> b1, b2, V1, V2;
    int rc = prepare(c);
    if (rc) return rc;
    if ((rc = run_blocks(c, q4, nout, &b1, 1))) return rc;
    if ((rc = nuclear_rows(c, natom, xyz, Z, V1, nullptr))) return rc;
    if ((rc = tcg_set_quadrature(c, 0, nq_full))) return rc;
    if ((rc = prepare(c))) return rc;
    if ((rc = run_blocks(c, q4, nout, &b2, 1))) return rc;
    if ((rc = nuclear_rows(c, natom, xyz, Z, V2, nullptr))) return rc;
    double de = 0.0, dvv = 0.0;
    for (size_t x = 0; x < nout; ++x) de = std::max(de, std::fabs(b1[x] - b2[x]));
    for (auto &p : pv) {
        const size_t ix = (size_t)p.first + (size_t)nb * p.second;
        dvv = std::max(dvv, std::fabs(V1[ix] - V2[ix]));
    }
    const bool ok = de <= 1e-10 && dvv <= 1e-10;
    if (ok && (rc = tcg_set_quadrature(c, 1, nq_fast))) return rc;
    *deri = de;
    *dv = dvv;
    *fast_on = ok ? 1 : 0;
    return TCG_OK;
}

int tcg_eri_stats(tcg_ctx *c, long long *out, int n) {
    if (!c || !out) return TCG_E_ARG;
    for (int i = 0; i < n && i < 8; ++i) out[i] = c->stats[i];
    return TCG_OK;
}

int tcg_eri_times(tcg_ctx *c, double *out, int n) {
    if (!c || !out) return TCG_E_ARG;
    for (int i = 0; i < n && i < 8; ++i) out[i] = c->times[i];
    return TCG_OK;
}

int tcg_eri_device(tcg_ctx *c, const double **dptr, long long *count) {
    if (!c || !dptr || !count) return TCG_E_ARG;
    *dptr = c->d_eri.p;
    *count = (long long)c->eri_count;
    return TCG_OK;
}

const char *tcg_last_error(tcg_ctx *c) { return c ? c->err.c_str() : "null context"; }

int tcg_destroy(tcg_ctx *c) {
    if (!c) return TCG_E_ARG;
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->ev3) (void)hipEventDestroy(c->ev3);
    for (auto &e : c->row_evs) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (c->st) (void)hipStreamSynchronize(c->st);
    if (c->st2) (void)hipStreamSynchronize(c->st2);
    if (c->ev_fill) (void)hipEventDestroy(c->ev_fill);
    if (c->st) (void)hipStreamDestroy(c->st);
    if (c->st2) (void)hipStreamDestroy(c->st2);
    if (c->stage) (void)hipHostFree(c->stage);
    delete c;
    return TCG_OK;
}

}  // extern "C"
