"""k-point MP2 in numpy.  TEST INFRASTRUCTURE ONLY: the checker for the
engine's k-space transform and k-MP2.

  phases        e^{+i k.R}, cells and k points x fastest
                (src/Energies/SAO/MP2/AO_to_MO_SAO.f90:97-113)
  ao_to_kmo     transform_ERI_AO_to_kMO (src/Energies/SAO/MP2/AO_to_MO_SAO.f90:34-283):
                S(a,b,c,d;k2,k3,k4) = 1/N sum_{R2,R3,R4} ERI(a, b R2, c R3, d R4)
                                      e^{+i k2 R2} e^{-i k3 R3} e^{+i k4 R4},
                ERI_kMO(p,q,r,s,k1,k2,k3) = sum conj C_k1(a,p) C_k2(b,q)
                                            conj C_k3(c,r) C_k4(d,s) S,
                k4 = k1 - k2 + k3
  kmp2          kMP2_calc (src/Energies/SAO/MP2/MP2_SAO.f90:6-151): kb = ki - ka + kj,
                terms with |D| < 1e-8 left out and counted,
                E_c = 2 E2d - E2x, E2d = -sum |g|^2 / D, E2x = -sum Re(g conj g_x) / D

Only for small systems (the whole S(a,b,c,d;k2,k3,k4) is held).
"""
import numpy as np


def _coords(ncell):
    nx, ny, nz = ncell
    idx = np.arange(nx * ny * nz)
    return np.stack([idx % nx, (idx // nx) % ny, idx // (nx * ny)], axis=1)


def phases(ncell):
    nx, ny, nz = ncell
    c = _coords(ncell)
    th = 2.0 * np.pi * (np.outer(c[:, 0], c[:, 0]) / nx + np.outer(c[:, 1], c[:, 1]) / ny
                        + np.outer(c[:, 2], c[:, 2]) / nz)
    return np.exp(1j * th)                     # [k, R]


def k_index(ncell, ka, kb, kc, sign_b=-1, sign_c=+1):
    """Index of k_a + sign_b k_b + sign_c k_c (mod lattice)."""
    nx, ny, nz = ncell
    c = _coords(ncell)
    v = c[ka] + sign_b * c[kb] + sign_c * c[kc]
    return (v[..., 0] % nx) + nx * ((v[..., 1] % ny) + ny * (v[..., 2] % nz))


def ao_to_kmo(E, Ck, ncell, o_hi, v_lo):
    """E: SAO ERI as C-ordered E[l,k,j,i] = ERI(i,j,k,l), i < nfuc.
    Ck[k] (nfuc, nfuc): coefficients of block k (columns = bands).
    Returns ERI_kMO[p, q, r, s, k1, k2, k3] (p, r < o_hi; q, s >= v_lo, 0-based)."""
    N = int(np.prod(ncell))
    nf = E.shape[3]
    ph = phases(ncell)
    # ERI(a, j2, j3, j4) with j = band + nfuc R (band fastest): in C order the
    # axes of E.transpose(3,2,1,0) are (a, j2, j3, j4) -> (a, R2, b, R3, c, R4, d)
    T = E.transpose(3, 2, 1, 0).reshape(nf, N, nf, N, nf, N, nf)
    S = np.einsum("aRbTcUd,kR,lT,mU->abcdklm", T, ph, ph.conj(), ph, optimize=True) / N
    # S[a,b,c,d,k2,k3,k4]
    occ = slice(0, o_hi)
    vir = slice(v_lo, nf)
    out = np.zeros((o_hi, nf - v_lo, o_hi, nf - v_lo, N, N, N), dtype=complex)
    for k1 in range(N):
        for k2 in range(N):
            for k3 in range(N):
                k4 = int(k_index(ncell, k1, k2, k3))
                blk = S[:, :, :, :, k2, k3, k4]
                out[:, :, :, :, k1, k2, k3] = np.einsum(
                    "abcd,ap,bq,cr,ds->pqrs", blk, Ck[k1][:, occ].conj(), Ck[k2][:, vir], Ck[k3][:, occ].conj(),
                    Ck[k4][:, vir], optimize=True)
    return out


def kmp2(ekmo, ek, occ, ncell, v_lo):
    """ek[k, band], occ[k, band] in {0, 2}; ekmo from ao_to_kmo.  Returns
    (E_c, 2 E2d, -E2x, number of zero-denominator terms)."""
    N = int(np.prod(ncell))
    E2d = E2x = 0.0
    nzero = 0
    for ki in range(N):
        for kj in range(N):
            for ka in range(N):
                kb = int(k_index(ncell, ki, ka, kj))
                for i in np.nonzero(occ[ki] == 2)[0]:
                    for j in np.nonzero(occ[kj] == 2)[0]:
                        for a in np.nonzero(occ[ka] == 0)[0]:
                            for b in np.nonzero(occ[kb] == 0)[0]:
                                D = ek[ka, a] + ek[kb, b] - ek[ki, i] - ek[kj, j]
                                g = ekmo[i, a - v_lo, j, b - v_lo, ki, ka, kj]
                                if abs(D) < 1e-8:
                                    nzero += 1
                                    continue
                                gx = ekmo[i, b - v_lo, j, a - v_lo, ki, kb, kj]
                                E2d -= (g * g.conjugate()).real / D
                                E2x -= (g * gx.conjugate()).real / D
    return 2.0 * E2d - E2x, 2.0 * E2d, -E2x, nzero
