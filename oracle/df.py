"""Density fitting in numpy.  TEST INFRASTRUCTURE ONLY: the checker for the
engine's DF path (tcg_df_*); imported by tests/ only.

Restates src/module/df_toroidal.f90:
  combined     the integrals' basis per cell: orbital functions, auxiliary
               functions, one exponent-0 s function (dummy_at :837-844)
  integrals    metric Vm(p0, Q R) = (P0 1|Q 1) (:545-566), Qa = sqrt|Vm(p,p)|,
               (mu0 nu|P) = (mu0 nu|P 1) with the df_schwarz screen
               Qsh(SA,SB) max Qa(SP) < 1e-13 (:605-636); every integral from the
               C restatement of ERI_shell_block_toroidal (orc.py)
  chol_f       :1174-1205 -- LAPACK's zpstrf itself (scipy.linalg.lapack),
               L^-1 by triangular solve, f(piv(i), j) = conj(L^-1(j, i))
  df_B         :568-700 -- V(k) = sum_R Vm(:, R) e^{ikR}, Hermitian part, f(k);
               f(-k) = conj f(k) for the second of each pair (k, -k);
               W(k) = [sum_RP T3(RP) e^{ik RP}] f(k), B(R') = 1/N Re sum_k W(k) e^{-ik R'}
  jk           df_JK :850-957 written in real space:
               J(mu0,nu) = sum_Q B^Q(mu0,nu) gamma_Q, gamma_Q = sum B^Q(la,si) P(la,si),
               K(mu0,si) = -1/2 sum_Q sum_{nu,la} B^Q(mu0,nu) P(nu,la) B^Q(la,si),
               B^Q(la0 + R, si) = B(la0, si - R, Q - R); P translated from its first rows
  kmo          df_kMO + df_L_kap :981-1112 (F, L, then 1/N sum_Q L L)

Checks that hold the restatement (tests/test_df.py): the route through k space
equals the real-space fit T3 V^-1 T3 of the whole supercell metric; chol_f is
zpstrf's factor; kmo equals transform_ERI_AO_to_kMO (oracle/kmp2.py) of the
fitted integrals; jk equals the stored-integral J/K of the fitted integrals.
Small systems only (every array is held whole).
"""
from __future__ import annotations

import dataclasses

import numpy as np
import scipy.linalg
import scipy.linalg.lapack as lapack

from . import kmp2 as _kmp2

METRIC_TOL = 1e-10      # metric_tol_aux, df_toroidal.f90:68
DF_SCHWARZ = 1e-13      # df_schwarz, df_toroidal.f90:70


def cell_ops(ncell):
    """cdiff[a, b] = a - b, cadd[a, b] = a + b (cells x fastest)."""
    nx, ny, nz = ncell
    N = nx * ny * nz
    c = _kmp2._coords(ncell)
    a = np.arange(N)
    d = c[a][:, None, :] - c[a][None, :, :]
    s = c[a][:, None, :] + c[a][None, :, :]
    idx = lambda v: (v[..., 0] % nx) + nx * ((v[..., 1] % ny) + ny * (v[..., 2] % nz))
    return idx(d), idx(s)


def combined(sysm, coef, aux):
    """Combined AO list per cell [orbital | auxiliary | dummy]; returns the
    system, its coefficients and (nf, nau)."""
    N = sysm.n_cell
    nf, nau = sysm.nfuc, aux.nbas // N
    center, lxyz, expo, cf, off = [], [], [], [], [0]

    def add(xyz, l, e, c):
        center.append(xyz)
        lxyz.append(l)
        expo.extend(e)
        cf.extend(c)
        off.append(len(expo))

    for R in range(N):
        for i in range(R * nf, (R + 1) * nf):
            o0, o1 = sysm.prim_off[i], sysm.prim_off[i + 1]
            add(sysm.center[i], sysm.lxyz[i], sysm.expo[o0:o1], coef[o0:o1])
        for p in range(R * nau, (R + 1) * nau):
            o0, o1 = aux.prim_off[p], aux.prim_off[p + 1]
            add(aux.center[p], aux.lxyz[p], aux.expo[o0:o1], aux.coef[o0:o1])
        add(aux.center[R * nau], (0, 0, 0), [0.0], [1.0])
    cs = dataclasses.replace(sysm, center=np.asarray(center, dtype=np.float64).reshape(-1, 3),
                             lxyz=np.asarray(lxyz, dtype=np.int32).reshape(-1, 3),
                             prim_off=np.asarray(off, dtype=np.int32), expo=np.asarray(expo, dtype=np.float64),
                             coef=np.asarray(cf, dtype=np.float64))
    return cs, cs.coef.copy(), (nf, nau)


def shells(cs):
    """(first AO, components) of each shell of an AO list (consecutive
    Cartesian components of one l on one centre)."""
    out = []
    i = 0
    while i < cs.nbas:
        l = int(cs.lxyz[i].sum())
        n = (l + 1) * (l + 2) // 2
        out.append((i, n))
        i += n
    return out


def integrals(osys, cs, nf, nau, schwarz=DF_SCHWARZ):
    """Vm (nau, naux), T3 (nf, nb, naux) from an oracle system loaded with the
    combined basis ``cs``."""
    N = cs.n_cell
    nfc = nf + nau + 1
    nb, naux = nf * N, nau * N
    sh = shells(cs)
    kind = lambda f: 0 if f % nfc < nf else (1 if f % nfc < nf + nau else 2)
    aux_i = lambda f: (f % nfc - nf) + nau * (f // nfc)
    ao_i = lambda f: (f % nfc) + nf * (f // nfc)
    ao_sh = [s for s in sh if kind(s[0]) == 0]
    aux_sh = [s for s in sh if kind(s[0]) == 1]
    dummy = {s[0] // nfc: s for s in sh if kind(s[0]) == 2}
    Vm = np.zeros((nau, naux))
    for (fp, np_) in aux_sh:
        if fp // nfc:
            continue
        for (fq, nq) in aux_sh:
            d0, dq = dummy[0], dummy[fq // nfc]
            blk = osys.eri_block(fp, np_, d0[0], 1, fq, nq, dq[0], 1)
            for a in range(np_):
                for b in range(nq):
                    Vm[aux_i(fp + a), aux_i(fq + b)] = blk[a, 0, b, 0]
    Qa = np.sqrt(np.abs(np.tile(np.diag(Vm[:, :nau]), N)))
    T3 = np.zeros((nf, nb, naux))
    for (fa, na) in ao_sh:
        if fa // nfc:
            continue
        for (fb, nbb) in ao_sh:
            # Schwarz factor of the shell pair: largest sqrt|(ab|ab)| (eri_direct.f90:159-178)
            blk_d = osys.eri_block(fa, na, fb, nbb, fa, na, fb, nbb)
            qab = max(np.sqrt(abs(blk_d[a, b, a, b])) for a in range(na) for b in range(nbb))
            for (fp, np_) in aux_sh:
                qp = max(Qa[aux_i(fp + x)] for x in range(np_))
                if qab * qp < schwarz:
                    continue
                blk = osys.eri_block(fa, na, fb, nbb, fp, np_, dummy[fp // nfc][0], 1)
                for a in range(na):
                    for b in range(nbb):
                        for p in range(np_):
                            T3[ao_i(fa + a), ao_i(fb + b), aux_i(fp + p)] = blk[a, b, p, 0]
    return Vm, T3


def chol_f(V, tol=METRIC_TOL):
    """Pivoted Cholesky of a Hermitian V by LAPACK zpstrf; returns f, rank m and
    the last pivot kept L(m,m)^2."""
    n = V.shape[0]
    c, piv, m, info = lapack.zpstrf(np.array(V, dtype=complex, order="F"), tol=tol, lower=1)
    if info < 0:
        raise RuntimeError("zpstrf failed")
    L = np.tril(c[:m, :m])
    Linv = scipy.linalg.solve_triangular(L, np.eye(m), lower=True)
    f = np.zeros((n, n), dtype=complex)
    for j in range(m):
        for i in range(m):
            f[piv[i] - 1, j] = np.conj(Linv[j, i])
    return f, m, (L[m - 1, m - 1].real ** 2 if m else 0.0)


def df_B(Vm, T3, ncell, tol=METRIC_TOL):
    """df_B (nf, nb, naux), ranks per k and the smallest pivot kept."""
    N = int(np.prod(ncell))
    nau = Vm.shape[0]
    nf, nb, naux = T3.shape
    ph = _kmp2.phases(ncell)                         # [k, R] = e^{ikR}
    cdiff, _ = cell_ops(ncell)
    fk = np.zeros((N, nau, nau), dtype=complex)
    ranks = np.zeros(N, dtype=int)
    dmin = np.inf
    for k in range(N):
        Mk = sum(Vm[:, R * nau:(R + 1) * nau] * ph[k, R] for R in range(N))
        Mk = 0.5 * (Mk + Mk.conj().T)
        mk = cdiff[0, k]                             # -k
        if mk < k:
            fk[k] = fk[mk].conj()
            ranks[k] = ranks[mk]
            continue
        fk[k], ranks[k], d = chol_f(Mk, tol)
        dmin = min(dmin, d)
    # aux index Q = p + nau RP: T3[..., Q] -> T[x, RP, p]
    T = T3.reshape(nf * nb, N, nau)
    Z = np.einsum("xRp,kR->kxp", T, ph)              # Z(k)
    W = np.einsum("kxp,kpq->kxq", Z, fk)
    B = np.einsum("kxq,kR->xRq", W, ph.conj()).real / N
    return B.reshape(nf, nb, naux), ranks, dmin


def b_full(B, ncell):
    """B^Q(la, si) for every la of the supercell: Bf[Q, la, si]."""
    N = int(np.prod(ncell))
    nf, nb, naux = B.shape
    nau = naux // N
    cdiff, _ = cell_ops(ncell)
    Bf = np.zeros((naux, nb, nb))
    for R in range(N):
        for si in range(nb):
            sis = cdiff[si // nf, R] * nf + si % nf
            for Q in range(naux):
                Qs = Q % nau + nau * cdiff[Q // nau, R]
                Bf[Q, R * nf:(R + 1) * nf, si] = B[:, sis, Qs]
    return Bf


def eri_sao(B, ncell):
    """Fitted integrals in the SAO layout E[l, k, j, i] = (i j|k l), i < nf."""
    Bf = b_full(B, ncell)
    nf = B.shape[0]
    E = np.einsum("Qij,Qkl->ijkl", Bf[:, :nf, :], Bf, optimize=True)
    return E.transpose(3, 2, 1, 0).copy()


def realspace_eri(Vm, T3, ncell):
    """(i j|k l) = sum_PQ (i j|P) [V^-1]_PQ (Q|k l) with the supercell metric:
    the fit done without k points or Cholesky (full rank only)."""
    N = int(np.prod(ncell))
    nau = Vm.shape[0]
    nf, nb, naux = T3.shape
    cdiff, _ = cell_ops(ncell)
    V = np.zeros((naux, naux))
    for RP in range(N):
        for RQ in range(N):
            V[RP * nau:(RP + 1) * nau, RQ * nau:(RQ + 1) * nau] = Vm[:, cdiff[RQ, RP] * nau:(cdiff[RQ, RP] + 1) * nau]
    # 3-index for every row cell: (k0 + R, l | Q) = T3(k0, l - R, Q - R)
    T = np.zeros((nb, nb, naux))
    for R in range(N):
        for l in range(nb):
            ls = cdiff[l // nf, R] * nf + l % nf
            for Q in range(naux):
                Qs = Q % nau + nau * cdiff[Q // nau, R]
                T[R * nf:(R + 1) * nf, l, Q] = T3[:, ls, Qs]
    X = np.linalg.solve(V, T.reshape(nb * nb, naux).T)            # V^-1 (Q|kl)
    E = (T[:nf].reshape(nf * nb, naux) @ X).reshape(nf, nb, nb, nb)
    return E.transpose(3, 2, 1, 0).copy()


def jk(B, P, ncell):
    """J, K (nb x nb, rows >= nf zero) from df_B, as df_JK defines them."""
    N = int(np.prod(ncell))
    nf, nb, naux = B.shape
    cdiff, _ = cell_ops(ncell)
    Bf = b_full(B, ncell)
    gam = np.einsum("Qls,ls->Q", Bf, P)
    J = np.zeros((nb, nb))
    K = np.zeros((nb, nb))
    J[:nf] = np.einsum("mnQ,Q->mn", B, gam)
    Pt = np.zeros((nb, nb))
    for nu in range(nb):
        for la in range(nb):
            Pt[nu, la] = P[nu % nf, cdiff[la // nf, nu // nf] * nf + la % nf]
    X = np.einsum("mnQ,nl->Qml", B, Pt)
    K[:nf] = -0.5 * np.einsum("Qml,Qls->ms", X, Bf)
    return J, K


def kmo(B, Ck, ncell, o_hi, v_lo):
    """ERI_kMO(p, q, r, s, k1, k2, k3) of df_kMO (p, r < o_hi; q, s >= v_lo)."""
    N = int(np.prod(ncell))
    nf, nb, naux = B.shape
    nau = naux // N
    nv = nf - v_lo
    ph = _kmp2.phases(ncell)
    cdiff, cadd = cell_ops(ncell)
    # df_B(a, b0 + nf Rb, q + nau RQ) as axes (a, Rb, b0, RQ, q)
    Bv = B.reshape(nf, N, nf, N, nau)
    L = np.zeros((o_hi, nv, nau, N, N), dtype=complex)
    for kap in range(N):
        Bk = np.einsum("aRbSq,S->aRbq", Bv, ph[kap].conj())        # sum_RQ e^{-i kap RQ}
        F = np.einsum("aRbq,kR->abqk", Bk, ph)                     # sum_Rb e^{ik Rb}
        for k2 in range(N):
            k1 = cdiff[k2, kap]
            L[:, :, :, k1, k2] = np.einsum("ap,bv,abq->pvq", Ck[k1][:, :o_hi].conj(), Ck[k2][:, v_lo:],
                                           F[:, :, :, k2])
    out = np.zeros((o_hi, nv, o_hi, nv, N, N, N), dtype=complex)
    for k1 in range(N):
        for k2 in range(N):
            for k3 in range(N):
                k4 = cadd[cdiff[k1, k2], k3]
                out[:, :, :, :, k1, k2, k3] = np.einsum("pqQ,rsQ->pqrs", L[:, :, :, k1, k2], L[:, :, :, k3, k4]) / N
    return out
