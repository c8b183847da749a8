/*
 * tcg.h -- C ABI of the MI355X Clifford-Gaussian (toroidal) integral engine.
 *
 * Drop-in for the integral and energy calls of the reference's Tori1D/2D/3D
 * path.  Plain pointers and sizes only; every array is owned by the caller,
 * FP64, column-major (Fortran) exactly as the reference declares it.  Every
 * entry point returns 0 on success and a negative status otherwise, with a
 * message from tcg_last_error().  One context per torus system; a context is
 * not re-entrant and must be used from one host thread at a time.  All calls
 * are synchronous: results are in the caller's arrays on return.
 *
 * A Fortran caller binds these with ISO_C_BINDING interfaces (bind(C, name=...)).
 */
#ifndef TCG_H
#define TCG_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tcg_ctx tcg_ctx;

enum {
    TCG_OK = 0,
    TCG_E_ARG = -1,        /* invalid argument */
    TCG_E_STATE = -2,      /* call out of order (e.g. no basis yet) */
    TCG_E_DEVICE = -3,     /* HIP runtime error */
    TCG_E_LATTICE = -4,    /* shells do not repeat in every cell */
    TCG_E_NOMEM = -5
};

/* Torus and quadrature.  Replaces the module state the reference reads:
 *   torus_init (Lx,Ly,Lz,ax,ay,az)      src/module/Torus_init.f90:6-7
 *   unitcell_module (nx,ny,nz)          src/module/unitcell.f90:2-4
 *   nquad_point / fast_quad (32 or 40)  src/calculation_type/calc_call.f90:25-48
 * ndp = number of periodic directions (1: Tori1D, 2: Tori2D, 3: Tori3D).
 * L[d] is read for d < ndp; ncell = (nx, ny, nz). */
int tcg_create(int ndp, const double L[3], const int ncell[3], int nquad, tcg_ctx **out);

/* AO list as a structure of arrays.  Replaces type(ERI_function) AO(nBas)
 * (src/module/ERI_mod.f90:5-12, built by src/basis/class.f90:1-70):
 *   center[3*i+d]      centre of AO i
 *   lxyz[3*i+d]        Cartesian powers of AO i (from AO%orbital)
 *   prim_off[i]..prim_off[i+1]-1   its primitives in expo[] / coef[]
 *   coef[]             NORMALISED contraction coefficients (the reference
 *                      passes the normalised copy AOn, src/calculation_type/Tori1D.f90:80-82)
 * nfuc = functions per unit cell. */
int tcg_set_basis(tcg_ctx *ctx, int nbas, int nfuc, const double *center, const int *lxyz,
                  const int *prim_off, const double *expo, const double *coef);

/* Two-electron integrals.  Replaces
 *   ERI_integral_toroidal(ndp,nBas,AO,ERI)   src/Toroidal/ERI_toroidal.f90:36
 * eri: ERI(n1,nBas,nBas,nBas) with n1 = nfuc (sao != 0, keyword SAO) or nBas
 * (rows of the other cells filled by translation, src/Tools/symmetric_matrix.f90:62);
 * chemist order ERI(i,j,k,l) = (ij|kl), or physicist ERI(i,k,j,l) when
 * physicist != 0.  The tensor also stays resident on the GPU for
 * tcg_jk_stored.  eri may be NULL: then it is kept on the GPU only. */
int tcg_eri_build(tcg_ctx *ctx, int sao, int physicist, double *eri);

/* Coulomb and exchange from the stored integrals.  Replaces
 *   hartree_exchange_potential(n1,nBas,P,ERI,J,K)
 *                                   src/Energies/hartree_exchange_potential.f90:30
 * J(mu,nu) = sum P(la,si)(mu nu|la si), K(mu,nu) = -1/2 sum P(la,si)(mu la|si nu),
 * rows mu <= n1 written, the rest zero; P, J, K are (nBas,nBas).
 * eri == NULL uses the tensor resident from tcg_eri_build (in its storage
 * order); otherwise eri is the caller's chemist ERI(n1,nBas,nBas,nBas)
 * (uploaded for this call). */
int tcg_jk_stored(tcg_ctx *ctx, int n1, int nbas, const double *P, const double *eri, double *J, double *K);

/* The same for a caller's array in either order (the reference reads the
 * keyword physicist, src/Energies/hartree_exchange_potential.f90:124-146):
 * physicist != 0 means eri holds ERI(i,k,j,l) = (ij|kl); then
 * J(mu,la) = sum P(nu,si) ERI(mu,nu,la,si).  With eri == NULL the resident
 * tensor is used in the order it was built and physicist is ignored
 * (tcg_jk_stored behaves the same way). */
int tcg_jk_stored_order(tcg_ctx *ctx, int n1, int nbas, const double *P, const double *eri, int physicist,
                        double *J, double *K);

/* Coulomb and exchange from the resident tensor (tcg_eri_build, or the last
 * caller array tcg_jk_stored uploaded), with NO device allocation.  No
 * counterpart in the reference: an extra form of tcg_jk_stored (eri == NULL)
 * for callers that cannot allocate (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, layout, order and
 * result as tcg_jk_stored with eri == NULL (to rounding: the sums meet in a
 * different order).  Its device memory is one static scratch array of 32 MiB
 * per GPU, resolved by tcg_create: nbas^2 + 2 n1 nbas must not exceed 4194304
 * (else TCG_E_ARG).  Calls that overlap on one GPU share that array and race. */
int tcg_jk_stored_noalloc(tcg_ctx *ctx, int n1, int nbas, const double *P, double *J, double *K);

/* Direct J/K: the integrals are recomputed at every call and contracted on
 * the fly, nothing of size nBas^4 is stored.  Replaces
 *   direct_init(ndp,nBas,AO)        src/module/eri_direct.f90:76-225
 *   direct_JK(nBas,P,J,K)           src/module/eri_direct.f90:262-374
 * tcg_direct_init: Schwarz, primitive Schwarz and the unique quartet list
 * (again after every tcg_set_basis / tcg_set_quadrature); the direct calls that
 * follow use the plain rule (a learned list of tcg_direct_init_learned is freed).
 * tcg_direct_jk: J(mu,nu) += E P(la,si), K(mu,si) -= E P(nu,la) / 2 over every
 * stored element E = ERI(mu,nu,la,si) of the SAO layout, each counted once;
 * rows mu <= nfuc written, the rest zero; P, J, K are (nBas,nBas). */
int tcg_direct_init(tcg_ctx *ctx);
int tcg_direct_jk(tcg_ctx *ctx, int nbas, const double *P, double *J, double *K);

/* Gamma-supercell MP2 from the resident full ERI (tcg_eri_build with sao = 0,
 * chemist or physicist).  Replaces
 *   AO_to_MO_ERI_inplace(nBas,c,ERI,physicist)   src/Tools/AO_to_MO/AO_to_MO.f90:20-111
 *   MP2(nBas,nO,e,ERI,ENuc,EHF,physicist_MO)     src/Energies/MP2/MP2.f90:1-87
 * C(nBas,nBas) MO coefficients (columns ascending in e), e(nBas) orbital
 * energies, nocc occupied orbitals.  Only (ia|jb) is formed (four FP64
 * matrix-core GEMMs, no second nBas^4 array); the resident ERI is unchanged.
 * out[0] = E_c(MP2) = 2 E2d - E2x, out[1] = direct part 2 E2d, out[2] =
 * exchange part -E2x (the three lines the reference prints), out[3] = GPU
 * wall time in ms. */
int tcg_mp2_gamma(tcg_ctx *ctx, int nbas, int nocc, const double *e, const double *C, double *out);

/* k-point MP2 from the resident SAO ERI (tcg_eri_build with sao = 1,
 * chemist).  Replaces
 *   transform_ERI_AO_to_kMO(nBas,ERI,coeff_SAO,p_lo,p_hi,q_lo,q_hi,ERI_kMO)
 *                                src/Energies/SAO/MP2/AO_to_MO_SAO.f90:34-283
 *   kMP2_calc(N_cell,blk,p_lo,p_hi,q_lo,q_hi,ERI_kMO,ENuc,EHF,EcMP2)
 *                                src/Energies/SAO/MP2/MP2_SAO.f90:6-151
 * coeff_sao: coeff_SAO(nBas,nBas) complex, column-major, as (re, im) pairs,
 * k-block diagonal (blocks of nfuc in the cell order, x fastest); e(nBas) and
 * occ(nBas) = 2 or 0 per band, block by block (the K_block fields).  The
 * (ov|ov) ranges follow set_cc_dims (src/module/cc_dims.f90:68-88).
 * out[0] = E_c, out[1] = 2 E2d, out[2] = -E2x, out[3] = number of terms left
 * out for |D| < 1e-8, out[4] = GPU wall time in ms. */
int tcg_kmp2(tcg_ctx *ctx, int nbas, const double *coeff_sao, const double *e, const int *occ, double *out);

/* The same k-point MP2 without the stored SAO tensor (integral-direct).  Replaces
 *   transform_ERI_direct_kMO(nBas,coeff_SAO,p_lo,p_hi,q_lo,q_hi,ERI_out,iout)
 *                                src/Energies/SAO/MP2/AO_to_MO_SAO_direct.f90:26-199
 *   direct_slab(SD,nfuc,nBas,ncell,slab)   src/module/eri_direct.f90:816-869
 * followed by kMP2_calc.  Needs tcg_direct_init.  The integrals are built in
 * slabs of the last index, j4 = d + nf R4 for the cells R4 of the slab (the
 * reference slices by unit-cell shell; the transform is linear in j4 either
 * way), each slab transformed over its two ket indices and summed.  A slab
 * takes cells in order of distance from cell 0 (each next to its inverse); per
 * slab the unique quartets that store into its cells are evaluated (with the
 * learned rules after tcg_direct_init_learned).  slab_cells = cells per slab (0:
 * as many as the free device memory holds next to the transforms' buffers and
 * an 8 GB row budget).  At most 64 cells.
 * out[0..4] as tcg_kmp2; out[5] integral time, out[6] transform time (ms),
 * out[7] number of slabs. */
int tcg_kmp2_direct(tcg_ctx *ctx, int nbas, const double *coeff_sao, const double *e, const int *occ,
                    int slab_cells, double *out);

/* k-space (Bloch, "SAO") matrices.  M_sao is the (nBas,nBas) complex matrix
 * of the reference (re, im pairs, column-major), block diagonal in k: block k
 * (cells and k points in one order, x fastest) holds rows and columns
 * k nfuc .. k nfuc + nfuc - 1.
 *
 * tcg_ao_to_sao replaces transform_AO_to_SAO(nBas,AO,SAO)  src/Tools/SAO_Transfer.f90:1-105:
 *   M_sao(k nf + a, k nf + b) = sum_R exp(+i k.R) M_ao(a, b + R nf)
 * (the unit-cell rows of M_ao; all other elements of M_sao are set to zero).
 * tcg_sao_to_ao replaces transform_SAO_to_AO(nBas,SAO,AO)  :108-233:
 *   M_ao(a + R nf, b + R' nf) = Re (1/N) sum_k exp(+i k.(R-R')) M_sao(k nf + a, k nf + b),
 * block elements of modulus below 1e-15 left out (only the diagonal blocks are read).
 * tcg_heev_blocks replaces diagonalize_block_diagonal_complex_para(nBas,nfuc,Nc,A,e)
 *   src/Tools/general_tools.f90:309-339 (clean_and_symmetrize + zheev 'V','U'):
 * every diagonal block is replaced by its eigenvectors (columns, ascending
 * eigenvalues), e(nBas) block by block; elements outside the blocks are not
 * touched.  Eigenvectors are fixed up to a phase per column.  nfuc <= 56. */
int tcg_ao_to_sao(tcg_ctx *ctx, int nbas, const double *M_ao, double *M_sao);
int tcg_sao_to_ao(tcg_ctx *ctx, int nbas, const double *M_sao, double *M_ao);
int tcg_heev_blocks(tcg_ctx *ctx, int nbas, double *A_sao, double *e);

/* The whole k-space RHF iteration on the GPU.  Replaces
 *   RHF_SAO(nBas,nO,S,T,V,Hc,ERI,X,ENuc,EHF,e,coeff_SAO,blk,natom,Zat)
 *                                    src/Energies/SAO/RHF_SAO.f90:1-553
 * with guess_AO_RHF_SAO (core guess, level at the cut filled evenly,
 * src/Energies/SAO/guess_Huckel_SAO.f90:141-192), X_k per block
 * (get_X_from_overlap_complex_blocks, canonical below 1e-6), DIIS on the
 * k-space error X^H (FPS - SPF) X (history 8, reset below rcond 1e-14) and the
 * reference's convergence and stall rules.  Every matrix stays on the GPU;
 * one iteration reads back a few scalars.
 *   source    0: J/K from the resident SAO ERI (tcg_eri_build, either order);
 *             1: integral-direct (tcg_direct_init first);
 *             2: integral-direct, incremental with density screening
 *                (tcg_set_direct_screen)
 *   nocc      nO, doubly occupied orbitals over the supercell
 *   S, T, V   (nBas,nBas) one-electron matrices
 *   max_scf, thresh, diis   maxSCF, the orbital-error threshold, keyword DIIS
 * Out: e(nBas) orbital energies block by block, ascending in each block;
 * coeff_sao (nBas,nBas) complex (re, im pairs) or NULL; occ(nBas) 2 / 0 or
 * NULL; P(nBas,nBas) the last density or NULL;
 *   out[0] E_HF without E_nuc   out[1..4] ET, EV, EJ, EK   out[5] iterations
 *   out[6] orbital error        out[7] converged           out[8] stalled
 *   out[9] HOMO-LUMO gap        out[10] J/K time (ms, GPU) out[11] wall time (ms)
 * Occupations take the nO lowest levels over all k (ties by block, then band).
 * nfuc <= 56. */
int tcg_rhf_sao(tcg_ctx *ctx, int source, int nbas, int nocc, const double *S, const double *T, const double *V,
                int max_scf, double thresh, int diis, double *e, double *coeff_sao, int *occ, double *P,
                double *out);

/* Density fitting (keyword density_fitting, SAO path).  A DF context holds
 * the fitted three-index tensor df_B(nfuc, nBas, naux) of one torus system
 * on the GPU.
 *
 * tcg_df_create replaces
 *   df_init(ndp,n_atoms,atoms,geometry,nBas,AO,iout)   src/module/df_toroidal.f90:352-709
 * (metric, pivoted Cholesky chol_f :1174-1205 with tolerance 1e-10, the
 * 3-index integrals with the df_schwarz 1e-13 screen, B as a cell
 * convolution).  The auxiliary functions come as an AO list like
 * tcg_set_basis's (normalised coefficients), nau_cell (<= 256) per cell, the
 * cells in the order of the orbital basis; the exponent-0 dummy of
 * dummy_at (:837-844) is added by the library.  ctx supplies the torus, the
 * orbital basis and the quadrature rule. */
typedef struct tcg_df tcg_df;
int tcg_df_create(tcg_ctx *ctx, int nau_cell, const double *aux_center, const int *aux_lxyz,
                  const int *aux_prim_off, const double *aux_expo, const double *aux_coef, tcg_df **out);

/* info[0] = nau, info[1] = naux, info[2] = aux functions skipped by the
 * pivoted Cholesky (largest over k); *dmin = smallest pivot kept (the three
 * lines df_init prints, :601-603). */
int tcg_df_info(tcg_df *df, long long *info, double *dmin);

/* Copy of df_B(nfuc, nBas, naux), column-major (module variable df_B). */
int tcg_df_b(tcg_df *df, double *B);

/* Coulomb and exchange from the fitted integrals.  Replaces
 *   df_JK(nBas,P,J,K)                 src/module/df_toroidal.f90:850-957
 * J(mu0,nu) = sum (mu0 nu|la si) P(la,si), K(mu0,si) = -1/2 sum (mu0 nu|la si) P(nu,la),
 * rows of the unit cell filled, the other rows zero (spread_rows :958-971);
 * exchange reads P through its first nfuc rows, as the k-space exchange does. */
int tcg_df_jk(tcg_df *df, int nbas, const double *P, double *J, double *K);

/* k-point MP2 from the fitted integrals.  Replaces
 *   df_kMO(nBas,coeff_SAO,p_lo,p_hi,q_lo,q_hi,ERI_out,iout)   src/module/df_toroidal.f90:981-1050
 *   df_L_kap                                                 :1053-1112
 * followed by kMP2_calc (src/Energies/SAO/MP2/MP2_SAO.f90:6-151); arguments
 * and out[0..4] as tcg_kmp2. */
int tcg_df_kmp2(tcg_df *df, int nbas, const double *coeff_sao, const double *e, const int *occ, double *out);

/* The same in FP32: L (df_L_kap, :1053-1112) with FP32 arithmetic, the
 * (ov|ov) assembly (the zgemm at df_toroidal.f90:1033) as FP32 matrix-core
 * GEMMs with FP32 accumulation (one complex GEMM over Q per momentum transfer
 * k1 - k2), and the kMP2_calc sums in FP32 kept as compensated (hi, lo) FP32
 * pairs.  Arguments and out[0..4] as tcg_df_kmp2; the difference to
 * tcg_df_kmp2 is the FP32 rounding of the integrals (~1e-7 relative). */
int tcg_df_kmp2_fp32(tcg_df *df, int nbas, const double *coeff_sao, const double *e, const int *occ, double *out);

/* out[0] metric + Cholesky, out[1] Schwarz + 3-index integrals, out[2] B (ms). */
int tcg_df_times(tcg_df *df, double *out);
const char *tcg_df_last_error(tcg_df *df);
int tcg_df_destroy(tcg_df *df);

/* One-electron matrices S, T, V (nBas,nBas): rows of the unit cell computed on
 * the GPU (S, T in closed form per direction, V by the t-quadrature), values below
 * 1e-14 zeroed, the other rows filled by translation there.  Replaces
 *   overlap_matrix_toroidal_nD            e.g. src/Toroidal/1D/1D_Overlap_toroidal.f90:23
 *   kinetic_matrix_toroidal_nD            e.g. src/Toroidal/1D/1D_Kinetic_toroidal.f90:102
 *   nuclear_attraction_matrix_toroidal_nD e.g. src/Toroidal/1D/1D_Nuclear_attraction_toroidal.f90:169
 * xyz[3*c+d] nuclear positions, Z[c] charges (the supercell's atoms). */
int tcg_one_electron(tcg_ctx *ctx, int natom, const double *xyz, const int *Z, double *S, double *T, double *V);

/* Normalised contraction coefficients of an AO list (same structure of arrays
 * as tcg_set_basis, raw coefficients in) on the context's torus.  Replaces
 *   normalize_AO_toroidal_nD   e.g. src/basis/normalization/normalize_1D.f90:17-79
 * every primitive normalised to one (closed-form toroidal self-overlap), then
 * the contraction.  Exponents must be positive. */
int tcg_normalize(tcg_ctx *ctx, int nbas, const double *center, const int *lxyz, const int *prim_off,
                  const double *expo, const double *coef_raw, double *coef_out);

/* Quadrature rule of every later call.  Replaces the module state
 *   fast_quad, nquad_point            src/module/precompute_t.f90:9,13
 * and the node source gauss_legendre_points (src/module/precompute_t.f90:484-895):
 *   fast != 0: nquad mapped Gauss-Legendre points t = 0.6 u/(1-u) (4..128;
 *              the reference uses 32, or 40 with nquad_128)  fast_quad_points :906-942
 *   fast == 0: the literal 64- or 128-point table (nquad 64 or 128). */
int tcg_set_quadrature(tcg_ctx *ctx, int fast, int nquad);

/* Current rule; t, w (nquad each, may be NULL) receive its nodes and weights. */
int tcg_get_quadrature(tcg_ctx *ctx, int *fast, int *nquad, double *t, double *w);

/* Start-up self-check of the fast rule.  Replaces
 *   fast_quad_self_check(ndp,nBas,AO,n_atoms,geometry,atoms)
 *                                   src/Toroidal/fast_quad_self_check.f90:30-233
 * Sample: every on-site shell quartet of the unit cell plus 50 LCG quartets
 * (seed 12345), V(i,i) for the first function of each unit-cell shell plus
 * 2 x 20 LCG pairs; each with the current fast rule and with the nq_full
 * (64 or 128) table.  *deri, *dv = largest differences; if both are <= 1e-10
 * the fast rule stays (*fast_on = 1), otherwise the table rule is left in
 * place for the rest of the run (*fast_on = 0).  With the table rule already
 * in place it returns at once with *fast_on = 0. */
int tcg_fast_quad_self_check(tcg_ctx *ctx, int natom, const double *xyz, const int *Z, int nq_full,
                             double *deri, double *dv, int *fast_on);

/* Shell blocks of a list of shell quartets, with the current rule.  Replaces
 *   ERI_shell_block_toroidal(ndp,nA,SA,nB,SB,nC,SC,nD,SD,blk)
 *                                   src/Toroidal/ERI_shell_block_toroidal.f90:28-390
 * shells[4*x .. 4*x+3]: the four shells (0-based, in the order of the AO list)
 * of quartet x; out[offsets[x] ..] = blk(nA,nB,nC,nD) column-major with the
 * kernel's prefactor screens (no Schwarz screen); offsets has n + 1 entries. */
int tcg_eri_blocks(tcg_ctx *ctx, int n, const int *shells, double *out, long long *offsets);

/* Learned per-quartet quadrature orders.  No counterpart in the reference
 * (its rule is one for the whole run).  A network 6 -> 64 -> 64 -> 2 (BF16
 * matrix-core kernels, FP32 accumulation) predicts for every surviving
 * primitive quartet log10 of its largest integral and the index of the
 * mapped Gauss-Legendre order (t = 0.6 u/(1-u)) it needs; a shell quartet
 * takes the largest index over its primitive quartets plus margin (rounded
 * up), and the current (plain) rule when any of its features leaves the
 * training range or the index passes the last order.
 *   params: tcg_mlp_param_count() floats, W1[6][64] b1 W2[64][64] b2 W3[64][2] b3
 *   norm:   24 floats, feature mean[6], std[6], lowest[6], highest[6]
 *   orders: norders (<= 15) ascending node counts (4..128)
 * The features are defined for 1, 2 and 3 periodic directions (chord
 * distances along the periodic ones, the shortest periodic length as L).
 * rule_counts (norders + 2 entries, may be NULL): shell quartets per order,
 * then those on the plain rule, then those skipped (tcg_set_learned_skip).
 *
 * tcg_set_learned_skip: learned Schwarz-style skipping in the next learned
 * builds.  A shell quartet all of whose primitive quartets are in the training
 * range is left out (its elements stay zero) when the bound head's largest
 * log10 |integral| over them, plus log10 of their number and margin_b, is below
 * log10(thresh) and (schwarz_cap > 0) its Schwarz product Qsh(bra) Qsh(ket) is
 * below schwarz_cap: no skipped integral then exceeds schwarz_cap.  The kept
 * quartets pass the reference's screens as before.  thresh = 0 turns skipping
 * off (the default). */
int tcg_set_learned(tcg_ctx *ctx, const float *params, const float *norm, const int *orders, int norders,
                    float margin);
int tcg_set_learned_skip(tcg_ctx *ctx, double thresh, double schwarz_cap, float margin_b);
int tcg_eri_build_learned(tcg_ctx *ctx, int sao, int physicist, double *eri, long long *rule_counts);

/* tcg_direct_init with the learned rules: the unique quartet list of the
 * direct J/K, each quartet with its learned rule (and learned skipping, when
 * set), evaluated once here; every later tcg_direct_jk (and the direct source
 * of tcg_rhf_sao) runs that list.  tcg_direct_init goes back to the plain rule.
 * rule_counts as tcg_eri_build_learned's.  No counterpart in the reference: a
 * form of direct_init (src/module/eri_direct.f90:76-225). */
int tcg_direct_init_learned(tcg_ctx *ctx, long long *rule_counts);

/* Density screening of the integral-direct J/K and the incremental Fock build
 * of tcg_rhf_sao's source 2.  No counterpart in the reference (its direct_JK
 * evaluates every Schwarz-surviving quartet at every iteration): a form of
 * direct_JK (src/module/eri_direct.f90:262-374) for late SCF iterations.
 *   tcg_rhf_sao with source 2: a full direct J/K every full_every iterations
 *   (the first included); in between, J and K of the density change P - P_prev,
 *   added to the previous J and K, where a shell quartet is left out when its
 *   bound times the largest element of the six density-change blocks it
 *   contracts with (pairs AB, CD, AC, AD, BC, BD) is below tau.  The bound is
 *   the Schwarz product Qsh(AB) Qsh(CD), or (use_learned_bound, with the
 *   learned list of tcg_direct_init_learned) the smaller of it and the
 *   network's bound.  Defaults: tau 1e-14, full_every 50, Schwarz bound.
 * tcg_direct_screen_info: tau, full_every and counts[5]: shell quartets
 * evaluated / listed in the last direct J/K, then since the last
 * tcg_direct_init or tcg_rhf_sao: evaluated, listed, number of J/K builds. */
int tcg_set_direct_screen(tcg_ctx *ctx, double tau, int full_every, int use_learned_bound);
/* One direct J/K with that density screening at tau (tau = 0: tcg_direct_jk):
 * the building block of an incremental Fock build in a caller's own SCF loop,
 * called with the density change. */
int tcg_direct_jk_screened(tcg_ctx *ctx, int nbas, const double *P, double tau, double *J, double *K);
int tcg_direct_screen_info(tcg_ctx *ctx, double *tau, int *full_every, long long *counts);

/* The network itself (no context): parameter count, inference, one
 * forward + backward (loss = mean over samples of w0 (y0 - t0)^2 +
 * w1 (y1 - t1)^2; grad receives dloss/dparams) and Adam training on the GPU
 * (epochs over the n samples in shuffled mini-batches of batch; params is
 * updated in place; loss_hist, may be NULL, gets the mean batch loss of
 * every epoch).  X: n x 6 features, T and Y: n x 2, all row-major FP32. */
int tcg_mlp_param_count(void);
int tcg_mlp_forward(const float *params, const float *X, int n, float *Y);
int tcg_mlp_grad(const float *params, const float *X, const float *T, int n, const float w[2], float *grad,
                 float *loss);
int tcg_mlp_train(float *params, const float *X, const float *T, int n, int epochs, int batch, float lr,
                  unsigned seed, const float w[2], float *loss_hist);

/* Counters of the last tcg_eri_build or tcg_direct_jk:
 *   out[0] shell quartets evaluated   out[1] primitive quartets evaluated
 *   out[2] rows (primitive quartet x direction x power combination)
 *   out[3] algorithmic FP64 flops of the quartet kernel: per primitive quartet
 *          ncomb*[sum_{d periodic}(2 sum_k min(nm_d, t_nmax(k,d)) + 10(nm_d+1))
 *                 + sum_{d real} 2 nq (NN+1)] + 3 nq ncomp4
 *   out[4] the periodic-direction part of out[3] (the row MFMA kernel alone)
 * n = number of slots the caller provides (up to 8). */
int tcg_eri_stats(tcg_ctx *ctx, long long *out, int n);

/* Wall times in milliseconds of the last tcg_eri_build sub-steps, the
 * analogue of the reference's tstep table (src/module/clifford_tools.f90:816-852):
 *   out[0] tables + pairs   out[1] Schwarz + primitive screen
 *   out[2] quartet list     out[3] quartet kernel + scatter   out[4] total
 *   out[5] duration of the quartet phase on the GPU (HIP events on the engine's stream)
 *   out[6] duration of the row MFMA kernel alone (HIP events around its launch)
 *   out[7] the t-table part of out[0] (host) */
int tcg_eri_times(tcg_ctx *ctx, double *out, int n);

/* Device address of the resident ERI tensor (for callers that keep data on
 * the GPU); *count = number of doubles. */
int tcg_eri_device(tcg_ctx *ctx, const double **dptr, long long *count);

const char *tcg_last_error(tcg_ctx *ctx);
int tcg_destroy(tcg_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
