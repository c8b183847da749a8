// Launch interface of the gfx950 kernels (tcg_eri.hip) used by the host
// driver (tcg_host.cpp).  Pointers are device addresses.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "tcg_engine.hpp"

struct tcg_ctx;

namespace tcg {

// What the density-fitting driver reads of a context (host pointers).
struct CtxView {
    int ndp, ncell[3], nbas, nfuc, fast, nq;
    double L[3];
    const double *center, *expo, *coef;
    const int *lxyz, *poff;
};
int ctx_view(const tcg_ctx *c, CtxView *v);
int ctx_shells(tcg_ctx *c, std::vector<int> &first_ao, std::vector<int> &ncomp, std::vector<int> &cell);

// Everything the quartet kernel reads.
struct EriArgs {
    int ndp, nq, nqp, ncap;           // periodic directions, nodes, nodes padded to 16, Bessel cap
    int nm_bits;                      // bits of a row's order field in the sort key (ncap + 1 fits)
    double a[3], L[3];
    double fac_const;                 // 2/sqrt(pi) prod_d L_d^2
    // quadrature rules (rule 0 = the run's rule; more with learned per-quartet orders),
    // every rule padded to nqp nodes:
    const double *tab;                // [rule][d][ncap+1+4][nqp] c_n exp(-C)I_n(C), 0 beyond t_nmax(k)
    const int *tab_lim;               // [rule][d][nqp/16]: last n with a nonzero entry in that node tile
    const long long *tsum;            // [rule][d][ncap+1]: sum over the rule's nodes of min(n, t_nmax(k)) + 1
    const double *tnode, *tweight;    // [rule][nqp] (0 in the padding)
    const int *rule_nq;               // [rule] node count (nullptr: one rule of nq nodes)
    const int *qrule;                 // [nq4] rule of each quartet (nullptr: rule 0)
    const Shell *sh;
    const int *ao_lxyz;               // [nbas][3]
    const double *coef;               // normalised coefficients, per AO
    const int *ao_poff;               // [nbas+1]
    const PrimPair *pp;
    const ShellPair *sp;
    const double2 *vec;               // Fourier vectors (sign +1 form)
    const double *herm;
    const Quartet *q4;
    int nq4;
    const unsigned long long *nq4_dev;   // set: nq4 is a capacity, the list's length is on the device
    int screen_prefactor;             // kernel prefactor screens (shell quartets only, q.prim < 0)
    int screen_prim;                  // primitive Schwarz screen (keyword prim_screen)
    int max_ncomb;                    // largest (lA+1)(lB+1)(lC+1)(lD+1) over the quartets
    int max_blk;                      // largest nA nB nC nD over the quartets
    int sub_words;                    // largest nX nY over two shells: a permutation's J or K sub-block
    int part_words;                   // LDS words per wavefront after the block in the general contraction:
                                      // 64 partial sums, a chunk of 64 survivors (96), 4 x max(ncomp nprim)
    int sp_contract;                  // s/p quartets (at most three p shells) by the s/p contraction kernels
    int sp_min_quartets;              // ... in quartet lists of at least this many quartets
    // block mode
    double *blk_out;
    // scatter mode
    double *eri;                      // ERI(n1,nbas,nbas,nbas)
    int n1, nbas, nfuc, ncell, physicist;
    const int *slab_pos;              // set: a slab build; the cell of a last index -> its position among the
                                      // slab's cells (-1: not in the slab); stored at nfuc position + local index
    const double *Qao;                // [nbas*nbas] function-pair Schwarz values
    // direct J/K mode
    const double *Pden;               // density P(nbas,nbas), column-major
    double *Jacc, *Kacc;              // J, K rows of the unit cell, (nfuc,nbas) column-major
    const int *relc;                  // [ncell*ncell]
    // counters
    unsigned long long *counters;     // [2] contraction flops, [3] quartet count, [4] row (MFMA) flops
};

hipError_t launch_fill(double *p, size_t n, double v, hipStream_t st);
hipError_t launch_fill_u32(unsigned *p, size_t n, unsigned v, hipStream_t st);
// address of the allocation-free J/K's scratch array on the current device (once per device)
hipError_t noalloc_setup();
size_t noalloc_scratch_doubles();
hipError_t launch_nuclear(const EriArgs &a, const double *nuc_xyz, const int *nuc_z, int natom, const int *ao_shell,
                          const int *pair_rep, int max_prim_pairs, double *V, hipStream_t st);
hipError_t launch_pair_setup(const PairBlock *blk, int nblk, const Shell *sh, const double *expo, const double *coef,
                             const int *ao_poff, int ndp, int ncap, const double L[3], const double a[3], PrimPair *pp,
                             VecJob *jobs, VecOut *vo, double2 *dco, double *herm, hipStream_t st);
hipError_t launch_pair_vectors(const VecJob *jobs, int njobs, const VecOut *vo, int nvo, const double2 *dco,
                               double *ivbuf, double2 *vec, hipStream_t st);
hipError_t launch_schwarz_tables(const Quartet *q4, int nquart, int nshell, const Shell *sh, const ShellPair *sp,
                                 const double *blk, const int *relc, int nf, int nb, int nc, int nsh, double *Q,
                                 double *Qsh, PrimPair *pp, unsigned long long *qmax_bits, hipStream_t st);
hipError_t launch_enum_quartets(const Shell *sh, const int *relc, const double *Qsh, const unsigned long long *qmax_bits,
                                int nshuc, int nsh, int ncell, unsigned long long *count, Quartet *out,
                                unsigned long long cap, hipStream_t st);
hipError_t launch_surv(const EriArgs &a, int pass, long long *scount, long long *rcount, const long long *soff,
                       Surv *surv, hipStream_t st);
hipError_t launch_rows_emit(const EriArgs &a, const long long *soff, const long long *roff, const Surv *surv,
                            RowRec *rows, unsigned *keys, int *vals, long long nrows, hipStream_t st);
hipError_t launch_rows_mfma(const EriArgs &a, const RowRec *rows, const int *order, long long nrows, double *rowval,
                            hipStream_t st);
size_t contract_lds_bytes(const EriArgs &a);
hipError_t launch_real_rows(const EriArgs &a, const long long *soff, const long long *roff, const Surv *surv,
                            long long ns, double *realval, hipStream_t st);
hipError_t launch_contract_jk(const EriArgs &a, const long long *soff, const long long *roff, const Surv *surv,
                              const double *rowval, const double *realval, const int *list, const unsigned *cnt,
                              hipStream_t st);
// lists: 9 * a.nq4 ints, cnt: 9 counters (work space of the s/p contraction)
// (also forms the real-direction rows, in realval, of the quartets that need them)
hipError_t launch_contract(const EriArgs &a, const long long *soff, const long long *roff, const Surv *surv,
                           const double *rowval, double *realval, int *lists, unsigned *cnt, hipStream_t st);
hipError_t launch_chunk_bounds(const long long *roff, int nq4, long long cap, int nchunk, int *bounds, hipStream_t st);
hipError_t launch_rebase(const long long *src, int n, long long *dst, hipStream_t st);
hipError_t launch_translate_rows(double *eri, int nb, int nf, int ncell, const int *relc, hipStream_t st);
hipError_t launch_dgemm_nn(int M, int N, int K, double alpha, const double *A, long lda, long long sA, const double *B,
                           long ldb, long long sB, double beta, double *C, long ldc, long long sC, long long batch,
                           hipStream_t st);
hipError_t launch_dft_lattice(const double *Xr, const double *Xi, long long I, long long outer, int N, const double *cosm,
                              const double *sinm, double sgn, double scale, double *Yr, double *Yi, hipStream_t st);
hipError_t launch_kmo(const double *Sre, const double *Sim, const double2 *Ck, int nf, int N, int nx, int ny, int nz,
                      int o_hi, int v_lo, double2 *out, hipStream_t st);
hipError_t launch_kmp2_sums(const double2 *ekmo, const double *ek, const int *occ, int nf, int N, int nx, int ny,
                            int nz, int o_hi, int v_lo, double *part, int nblocks, unsigned long long *nzero,
                            hipStream_t st);
hipError_t launch_mp2_sums(const double *X, const double *e, int o, int v, long long si, long long sa, long long sj,
                           long long sb, double *part, int nblocks, hipStream_t st);
// the lattice transform fused with an orbital contraction (tcg_mp2.hip, k_dft_mo)
struct DftMoArgs {
    const double *Xr, *Xi;
    long long I, O;              // inner rows, outer blocks
    const long long *imap;       // inner row -> offset in X (nullptr: i)
    long long sJ, sO;            // strides of j = c + nf R and of o in X
    int nf, N, NR, M;
    const double *cosm, *sinm;   // [R + NR k]
    double sgn, scale;
    const double2 *T;            // [(k nf + c) M + m]
    double *Yr, *Yi;
    int accum;                   // 1: Y += (sums over slabs of the transformed index pair)
};
hipError_t launch_dft_mo(const DftMoArgs &A, hipStream_t st);
hipError_t launch_list_j2_flags(const long long *list, long long n, int nbas, unsigned char *flags, hipStream_t st);
hipError_t launch_kmo_last(const double *Yr, const double *Yi, const double2 *Ck, int nf, int N, int nx, int ny,
                           int nz, int o_hi, int v_lo, double2 *out, hipStream_t st);
// overlap/kinetic rows and the AO normalisation (tcg_oneel.hip)
struct OneStArgs {
    int ndp, nf, nb;
    double a[3], L[3];
    const double *center;     // [nb][3]
    const int *lxyz, *poff;   // [nb][3], [nb + 1]
    const double *expo, *coef;
    double *S, *T;            // (nb, nb) column-major, rows < nf written
};
// chunks of the auxiliary index in the DF J pass (its partial sums: kDfJChunks nf nb doubles)
constexpr int kDfJChunks = 64;
// work: 2 nf nb mp^2 (S, T) and nb mp^2 doubles; mp = most primitives of an AO
hipError_t launch_one_st(const OneStArgs &g, int mp, double *work, hipStream_t st);
hipError_t launch_normalize(const OneStArgs &g, const double *raw, double *out, int mp, double *work,
                            hipStream_t st);
hipError_t launch_translate_fill(double *M, int nf, int nb, int nc, const int *relc, hipStream_t st);
int mlp_param_count();
hipError_t launch_features(const EriArgs &a, const int *owner, long long ns, const Surv *surv, const double *expo,
                           const float *norm, float *feat, unsigned char *ood, hipStream_t st);
hipError_t launch_surv_owner(const long long *soff, int nq4, int *owner, hipStream_t st);
constexpr int kRuleSkip = -1;     // a shell quartet the learned bound leaves out
hipError_t launch_rules(int nq4, const long long *soff, const float *pred, const unsigned char *ood, float margin,
                        int plain, float log_skip, float margin_b, const Quartet *q4, const double *Qsh, int nsh,
                        double cap, int *qrule, float *lbound, hipStream_t st);
// density screening of the direct J/K (tcg_screen.hip)
hipError_t launch_pair_dmax(const double *P, int nb, const Shell *sh, int nshuc, int nsh, double *Dsh,
                            hipStream_t st);
hipError_t launch_density_keep(const Quartet *q4, int nq4, const Shell *sh, const double *Qsh, int nsh, int nshuc,
                               const int *relc, int ncell, const double *Dsh, const float *lbound, double tau,
                               unsigned char *keep, hipStream_t st);
hipError_t launch_gather_list(const Quartet *q4, const int *qrule, const float *lb, const long long *idx,
                              const unsigned long long *n, long long nmax, Quartet *q4_out, int *rule_out,
                              float *lb_out, hipStream_t st);
hipError_t launch_rule_keep(int nq4, const int *qrule, unsigned char *keep, hipStream_t st);
hipError_t launch_gather_quartets(const Quartet *q4, const int *qrule, const long long *idx,
                                  const unsigned long long *n, long long nmax, Quartet *q4_out, int *rule_out,
                                  hipStream_t st);
hipError_t launch_rule_hist(int nq4, const int *qrule, unsigned long long *hist, hipStream_t st);
hipError_t launch_mlp_fwd(const float *P, const float *X, int n, float *Y, hipStream_t st);
hipError_t launch_mlp_step(const float *P, const float *X, const float *T, int n, float inv_total, float w0, float w1,
                           float *G, float *loss, hipStream_t st);
hipError_t launch_mlp_adam(float *P, float *G, float *M, float *V, float lr, float b1, float b2, float eps, float c1,
                           float c2, hipStream_t st);
size_t scan_temp_bytes(int n);
hipError_t scan_exclusive(void *tmp, size_t bytes, const long long *in, long long *out, int n, hipStream_t st);
size_t sort_temp_bytes(int n, int end_bit);
hipError_t sort_pairs(void *tmp, size_t bytes, const unsigned *kin, unsigned *kout, const int *vin, int *vout, int n,
                      int end_bit, hipStream_t st);
size_t select_temp_bytes(long long n);
hipError_t select_flagged(void *tmp, size_t bytes, const unsigned char *flags, long long *out, long long *nsel,
                          long long n, hipStream_t st);
size_t sort_ll_temp_bytes(long long n, int end_bit);
hipError_t sort_pairs_ll(void *tmp, size_t bytes, const unsigned *kin, unsigned *kout, const long long *vin,
                         long long *vout, long long n, int end_bit, hipStream_t st);
size_t scan_int_temp_bytes(int n);
hipError_t scan_int_exclusive(void *tmp, size_t bytes, const int *in, long long *out, int n, hipStream_t st);
hipError_t launch_col_flags(const double *eri, int n1, int nbas, int physicist, unsigned char *flags, int *cnt_d,
                            int *cnt_t, unsigned *key_t, hipStream_t st);
hipError_t launch_jk_sparse(const double *eri, const double *P, const long long *list_d, const long long *off_d,
                            const long long *list_t, const long long *off_t, int n1, int nbas, int physicist,
                            long long nz, double *J, double *K, hipStream_t st);
hipError_t launch_gather_keys(const unsigned *keys, const long long *idx, long long n, unsigned *out, hipStream_t st);
size_t jk_work_doubles(int n1, int nbas);
hipError_t launch_jk_stored(const double *eri, const double *P, double *J, double *K, int n1, int nbas,
                            double *jwork, hipStream_t st);

// ---- k-space RHF algebra (tcg_scf.hip) ----
constexpr int kMaxBlock = 56;     // largest k block (nfuc): three n x n complex LDS buffers
constexpr int kMaxDiis = 8;       // DIIS history (src/module/HF_constant.f90 max_diis)
// unit-cell rows of up to four real matrices, element (a, c) of source s at
// p[s][a sa[s] + c sc[s]], summed with factors f[s]
struct RowSources {
    const double *p[4];
    long long sa[4], sc[4];
    double f[4];
    int n;
};
struct DiisSlots {
    int slot[kMaxDiis];
    double w[kMaxDiis];
    int m;
};
hipError_t launch_lattice_fwd(const RowSources &src, const double2 *ph, int N, int nf, double2 *out, hipStream_t st);
hipError_t launch_lattice_inv(const double2 *Mk, const double2 *ph, int N, int nf, int skip, double *C,
                              double *maxabs, int *nmax, hipStream_t st);
hipError_t launch_circulant(const double *C, int nf, const int ncl[3], double *M, hipStream_t st);
hipError_t launch_transpose(const double *A, int n, double *B, hipStream_t st);
hipError_t launch_rows_sum(const double *A, const double *B, int nf, int nb, double *H, hipStream_t st);
hipError_t launch_overlap_x(const double2 *Sk, int N, int n, double2 *Xk, int *dead, double *smin, hipStream_t st);
hipError_t launch_fock_diag(const double2 *Fk, const double2 *Xk, const int *dead, int N, int n, double *e,
                            double2 *Ck, hipStream_t st);
hipError_t launch_scf_error(const double2 *Fk, const double2 *Pk, const double2 *Sk, const double2 *Xk, int N, int n,
                            double2 *err, hipStream_t st);
hipError_t launch_density_k(const double2 *Ck, const double *w, int N, int n, double2 *Pk, hipStream_t st);
hipError_t launch_occupy(const double *e, int ne, int nocc, int frac, double *w, int *occ, double *sc,
                         hipStream_t st);
int energy_parts_blocks(int nb);
hipError_t launch_energy(const double *P, const double *Tt, const double *Vt, const double *J, const double *K,
                         int ldjk, int nf, int nb, double *part, const double *maxes, int nmax, double ncell,
                         double *sc, hipStream_t st);
hipError_t launch_diis_dots(const double2 *enew, const double2 *hist, long long len, const DiisSlots &sl, double *out,
                            hipStream_t st);
hipError_t launch_diis_combine(const double2 *hist, long long len, const DiisSlots &sl, double2 *out, hipStream_t st);
// J and K rows of the unit cell for a density already on the device (the GPU
// SCF loop): source 0 = the resident ERI tensor (leading dimension nbas, rows
// < n1 written), 1 = integral-direct (leading dimension nfuc).  dJ, dK hold
// nbas^2 doubles.
int ctx_jk_device(tcg_ctx *c, int source, const double *dP, double *dJ, double *dK, int *ld);
void ctx_screen_reset(tcg_ctx *c);
int ctx_screen_full_every(const tcg_ctx *c);
hipError_t launch_axpby(const double *a, const double *b, double *out, size_t n, double alpha, double beta,
                        hipStream_t st);
hipStream_t ctx_stream(const tcg_ctx *c);
// state the SCF loop keeps in the context between calls (released with it)
std::shared_ptr<void> &ctx_scf_work(tcg_ctx *c);

}  // namespace tcg
