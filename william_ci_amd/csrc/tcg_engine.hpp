// Internal data layout of the engine, shared by the host setup (tcg_host.cpp)
// and the gfx950 kernels (tcg_eri.hip).  Everything the kernels read is a
// flat structure of arrays in HBM.
#pragma once
#include <cstdint>

namespace tcg {

constexpr int kMaxL = 4;            // up to g shells
constexpr int kMaxNQ = 128;         // quadrature nodes (the 128-point table), padded to node tiles of 16
constexpr int kNodeTile = 16;       // nodes per MFMA column tile
constexpr int kRowTile = 16;        // rows per MFMA tile
constexpr int kMaxPrim = 16;        // primitives per shell
constexpr double kScreenTol = 2.22e-16;  // prefactor screen (ERI_shell_block_toroidal.f90:42)
constexpr double kPrimTol = 1e-14;       // primitive Schwarz (clifford_tools.f90:74)
constexpr double kSchwarz = 1e-13;       // shell Schwarz (ERI_toroidal.f90:64)

// One primitive pair (i of shell A, j of shell B) of a representative shell
// pair (A in the unit cell).  Coefficient-free prefactor, barycentres,
// Bessel lengths per periodic direction, offsets of its Fourier vectors and
// Hermite coefficients.
struct PrimPair {
    double fac;          // prod_d exp(-2(e1+e2-p_d)/a_d^2) * prod_real exp(-mu R^2)
    double cen[3];       // barycentre per direction (periodic: toroidal barycentre)
    double q;            // e1 + e2
    double cmax;         // max_comp |c_A(i)| * max_comp |c_B(j)|  (shell-wide max per primitive)
    double ps;           // primitive Schwarz value sqrt|(ij|ij)| (max over components)
    int N[3];            // Bessel length per periodic direction
    float A[3];          // its Bessel argument 2 p_d / a_d^2 (the row length of a product of two pairs)
    int voff[3];         // offset (complex elements) of the Fourier vectors of direction d:
                         //   vec[voff + combo*(N+1) + n], combo = ia*(lB+1)+ib
    int eoff;            // offset of Hermite E: herm[eoff + ((dr*(lA+1)+ia)*(lB+1)+ib)*(2*kMaxL+1) + t]
    int pi, pj;          // primitive indices
};

// Representative shell pair (A in the unit cell, B anywhere).
struct ShellPair {
    int first;           // first PrimPair
    int count;           // n_prim(A) * n_prim(B), ordered i major
};

// One shell quartet to evaluate.
struct Quartet {
    double shift[3];     // translation of the ket pair: centre(SK) - centre(representative of SK)
    int s[4];            // global shell indices SI, SJ, SK, SL
    int bra;             // representative pair of (SI, SJ)   (SI is in the unit cell)
    int ket;             // representative pair of (SK, SL) translated to the unit cell
    int out;             // offset of the output block (block mode), or -1
    int prim;            // >= 0: only primitive pair 'prim' of bra and ket (primitive Schwarz pass)
};

// One pair Fourier-vector job: exp(-A) I_n(A) up to N + lA + lB, then the
// vectors of all power pairs (ia <= lA, ib <= lB) at vec[voff + (ia*(lB+1)+ib)*(N+1)].
struct VecJob {
    double A, phA, phB;
    int N, lA, lB, voff;
    long ivoff;          // scratch for the Bessel values
};

// A surviving primitive quartet: global primitive-pair indices and prefactor.
struct Surv {
    int pb, pk;
    double fac;
};

// One row of the quadrature contraction: (primitive quartet, periodic
// direction, power combination).
struct RowRec {
    double c1, s1;       // e^{i phi}, phi = a_d (P_d - Q_d)
    double c4, s4;       // e^{4 i phi}
    double c64, s64;     // e^{64 i phi}
    double sgn;          // a_d^-(ia+ib+ic+id)
    int gofs, hofs;      // offsets of the bra / ket Fourier vectors
    int nm;              // last order: min(N_bra, N_ket)
    int d;               // periodic direction
};

// One representative shell pair (S1 in the unit cell, S2 anywhere) of the
// GPU pair set-up and where its records go: nij = nprim(S1) nprim(S2)
// primitive pairs (i major), each with ndp Fourier-vector jobs of
// (l1+1)(l2+1) vectors.  Vector and Bessel storage is laid out for the
// largest length (ncap + 1 orders per vector, ncap + l1 + l2 + 1 values per
// job), so every offset is known before the lengths are.
struct PairBlock {
    int S1, S2;
    int first;           // first PrimPair
    int job0;            // first VecJob: + ij * ndp + d
    int vo0;             // first VecOut: + (ij * ndp + d) * (l1+1)(l2+1) + combo
    int vec0;            // first vector element: + (ij * ndp + d) * (l1+1)(l2+1)(ncap+1)
    long long iv0;       // first Bessel value: + (ij * ndp + d) * (ncap + l1 + l2 + 1)
    int herm0;           // first Hermite coefficient: + ij * nreal (l1+1)(l2+1)(2 kMaxL + 1)
    int dco0;            // first Fourier coefficient: + VecOut index in the block * (2(l1+l2)+1)
};

// One Fourier vector to assemble: its job (Bessel values), coefficient
// offset (2K+1 complex d_k) and output offset.
struct VecOut {
    int job, dco, voff, K;
};

struct Shell {
    int first_ao;        // first AO
    int ncomp;           // Cartesian components
    int l;               // angular momentum
    int nprim;
    int poff;            // offset into expo/coef of the first AO's primitives
    int cell;            // cell index
    int loc;             // local index inside the cell
    unsigned dirmap;     // p shell: component along direction d in bits 2d..2d+1 (0 otherwise)
    double xyz[3];
};

}  // namespace tcg
