"""Gamma-supercell RHF and MP2 in numpy on a full ERI tensor.  TEST
INFRASTRUCTURE ONLY: the checker for the engine's SCF and MP2.

  rhf_gamma   src/Energies/RHF.f90:1-369 with guess_AO_RHF (src/Energies/guess_RHF.f90:1),
              get_X_from_overlap (src/Energies/Energies_tools.f90:84), DIIS
              (src/Energies/DIIS_extrapolation.f90:1), J/K as in
              src/Energies/hartree_exchange_potential.f90:15-21
  mp2_gamma   AO_to_MO_ERI_inplace + MP2 (src/Tools/AO_to_MO/AO_to_MO.f90:20,
              src/Energies/MP2/MP2.f90:39-68):
              E_c = sum_{ijab} (ia|jb) [2 (ia|jb) - (ib|ja)] / (e_i + e_j - e_a - e_b)

The supercell formulation is independent of the engine's Bloch (k-space)
one; for a closed-shell system both give the same energies.
"""
import numpy as np


def _X(S, thresh=1e-6):
    e, U = np.linalg.eigh(S)
    if np.all(e > thresh):
        return (U / np.sqrt(e)) @ U.T
    X = U / np.sqrt(np.where(e > thresh, e, 1.0))
    X[:, e <= thresh] = 0.0
    return X


def rhf_gamma(ch, S, T, V, nO, enuc, max_scf=100, thresh=1e-9, diis=True):
    """ch[i,j,k,l] = (ij|kl), full supercell.  Returns (E_HF, mo_energies, C, P, iterations)."""
    Hc = T + V
    X = _X(S)
    e, U = np.linalg.eigh(X.T @ Hc @ X)
    C = X @ U
    P = 2.0 * C[:, :nO] @ C[:, :nO].T
    errs, Fs = [], []
    conv = 1.0
    for it in range(1, max_scf + 1):
        J = np.einsum("mnls,ls->mn", ch, P, optimize=True)
        K = -0.5 * np.einsum("mlsn,ls->mn", ch, P, optimize=True)
        F = Hc + J + K
        E = np.sum(P * T.T) + np.sum(P * V.T) + 0.5 * np.sum(P * J.T) + 0.5 * np.sum(P * K.T)
        err = X.T @ (F @ P @ S - S @ P @ F) @ X
        if it > 1:
            conv = np.abs(err).max()
        if conv < thresh:
            break
        if diis:
            errs.insert(0, err.ravel())
            Fs.insert(0, F.ravel())
            del errs[8:], Fs[8:]
            m = len(errs)
            A = np.zeros((m + 1, m + 1))
            A[:m, :m] = np.array(errs) @ np.array(errs).T
            A[:m, m] = A[m, :m] = -1.0
            b = np.zeros(m + 1)
            b[m] = -1.0
            if 1.0 / np.linalg.cond(A, 1) < 1e-14:
                errs.clear()
                Fs.clear()
            else:
                F = (np.linalg.solve(A, b)[:m] @ np.array(Fs)).reshape(F.shape)
        e, U = np.linalg.eigh(X.T @ F @ X)
        C = X @ U
        P = 2.0 * C[:, :nO] @ C[:, :nO].T
    return E + enuc, e, C, P, it


def mp2_gamma(ch, e, C, nO):
    o = C[:, :nO]
    v = C[:, nO:]
    iajb = np.einsum("pqrs,pi,qa,rj,sb->iajb", ch, o, v, o, v, optimize=True)
    eo, ev = e[:nO], e[nO:]
    D = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
    return float(np.sum(iajb * (2.0 * iajb - iajb.transpose(0, 3, 2, 1)) / D))
