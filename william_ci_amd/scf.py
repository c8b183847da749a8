"""Restricted Hartree-Fock on top of the engine: in the Bloch (SAO) basis
(rhf_sao) and in the supercell AO basis (rhf, the Gamma-point form).

Host-side mirror of the reference's SCF driver, which stays in Fortran in the
drop-in; the engine supplies the integrals and the Coulomb/exchange build:

  RHF_SAO                src/Energies/SAO/RHF_SAO.f90:1-553
  RHF                    src/Energies/RHF.f90:1-369 (+ guess_AO_RHF, src/Energies/guess_RHF.f90:1)
  guess_AO_RHF_SAO       src/Energies/SAO/guess_RHF_SAO.f90:1 (+ guess_occupy_cut,
                         src/Energies/SAO/guess_Huckel_SAO.f90:141-192)
  transform_AO_to_SAO /  src/Tools/SAO_Transfer.f90:1,108
  transform_SAO_to_AO
  get_X_from_overlap     src/Energies/Energies_tools.f90:84 (canonical, drop <= 1e-6)
  DIIS_extrapolation     src/Energies/DIIS_extrapolation.f90:1 (history 8, reset rcond < 1e-14)
  NRE                    src/Energies/Nuclear_repulsion.f90:1 (chord distance)
  HF_parameter           src/module/HF_constant.f90 (maxSCF 100, orbital error 1e-9)

rhf_sao runs the dense algebra (per-k nfuc x nfuc eigenproblems, lattice
transforms, DIIS) in numpy around the GPU J/K build; rhf_sao_gpu runs the same
iteration entirely on the GPU (tcg_rhf_sao) and is the product path; rhf_sao
stays as its cross-check and for callers that bring their own J/K.
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import Callable, Optional

import numpy as np

MAX_SCF = 100
THRESH = 1e-9
MAX_DIIS = 8
LINDEP = 1e-6
LINDEP_SHIFT = 1e4


def nuclear_repulsion(sysm) -> float:
    """Sum Z_i Z_j / d_ij with the chord distance along the periodic directions."""
    g, Z = np.asarray(sysm.geom, dtype=np.float64), sysm.Z.astype(np.float64)
    d = g[:, None, :] - g[None, :, :]                       # all pairs at once
    for k in range(sysm.ndp):
        a = 2 * math.pi / sysm.L[k]
        d[:, :, k] = np.sqrt(2.0 * (1.0 - np.cos(a * d[:, :, k])) / (a * a))
    i, j = np.triu_indices(sysm.natom, 1)
    return float(np.sum(Z[i] * Z[j] / np.sqrt(np.sum(d[i, j] ** 2, axis=1))))


def canonical_X(S: np.ndarray, thresh: float = LINDEP) -> np.ndarray:
    """S^-1/2 (Lowdin) when all eigenvalues exceed thresh, else canonical with
    the small-eigenvalue columns zeroed."""
    e, U = np.linalg.eigh(S)
    if np.all(e > thresh):
        return (U / np.sqrt(e)) @ U.conj().T
    X = U / np.sqrt(np.where(e > thresh, e, 1.0))
    X[:, e <= thresh] = 0.0
    return X


def _shift_dropped(X: np.ndarray, Fp: np.ndarray) -> None:
    dead = np.all(X == 0, axis=0)
    idx = np.nonzero(dead)[0]
    Fp[idx, idx] = LINDEP_SHIFT


class Bloch:
    """AO <-> SAO (Bloch) transforms of translation-symmetric supercell matrices."""

    def __init__(self, sysm):
        self.nf = sysm.nfuc
        self.nc = sysm.ncell           # (nx, ny, nz), cells x fastest
        self.N = sysm.n_cell
        nx, ny, nz = self.nc
        cells = np.arange(self.N)
        cx, cy, cz = cells % nx, (cells // nx) % ny, cells // (nx * ny)
        self._dx = (cx[:, None] - cx[None, :]) % nx      # cell differences R - R'
        self._dy = (cy[:, None] - cy[None, :]) % ny
        self._dz = (cz[:, None] - cz[None, :]) % nz

    def to_k(self, M: np.ndarray) -> np.ndarray:
        """M_k(a,b) = sum_R e^{+i k.R} M(a, b + R nf), from the unit-cell rows of M."""
        nf, (nx, ny, nz) = self.nf, self.nc
        B = M[:nf, :].reshape(nf, nz, ny, nx, nf).transpose(1, 2, 3, 0, 4)
        Mk = np.fft.ifftn(B, axes=(0, 1, 2)) * self.N
        return Mk.reshape(self.N, nf, nf)

    def max_abs_ao(self, Mk: np.ndarray) -> float:
        """Largest |element| of to_ao(Mk) without forming the matrix."""
        nf, (nx, ny, nz) = self.nf, self.nc
        return float(np.abs(np.fft.ifftn(Mk.reshape(nz, ny, nx, nf, nf), axes=(0, 1, 2)).real).max())

    def to_ao(self, Mk: np.ndarray) -> np.ndarray:
        """M(a + R nf, b + R' nf) = Re (1/N) sum_k e^{+i k.(R-R')} M_k(a,b)."""
        nf, (nx, ny, nz) = self.nf, self.nc
        C = np.fft.ifftn(Mk.reshape(nz, ny, nx, nf, nf), axes=(0, 1, 2)).real   # C[dR]
        blocks = C[self._dz, self._dy, self._dx]       # (N, N, nf, nf): block (R, R')
        return blocks.transpose(0, 2, 1, 3).reshape(self.N * nf, self.N * nf)


def _occupy_cut(e: np.ndarray, c: np.ndarray, nO: int, deg_tol: float = 1e-8) -> np.ndarray:
    """Density of the nO lowest orbitals, a degenerate level at the cut filled evenly."""
    tol = deg_tol * max(1.0, abs(e[nO - 1]))
    lo = nO - 1
    while lo > 0 and abs(e[lo - 1] - e[nO - 1]) < tol:
        lo -= 1
    hi = nO - 1
    while hi < len(e) - 1 and abs(e[hi + 1] - e[nO - 1]) < tol:
        hi += 1
    if hi == nO - 1:
        return 2.0 * c[:, :nO] @ c[:, :nO].T
    occ = 2.0 * (nO - lo) / (hi - lo + 1)
    w = np.concatenate([np.full(lo, 2.0), np.full(hi - lo + 1, occ)])
    cw = c[:, :hi + 1] * np.sqrt(w)
    return cw @ cw.T


def _occupy_cut_k(ek: np.ndarray, ck: np.ndarray, nO: int, bl: "Bloch", deg_tol: float = 1e-8):
    """_occupy_cut over the union of the k spectra: ek (Nk, nf), ck (Nk, nf, nf)
    columns = orbitals; returns the AO density and its k blocks."""
    flat = ek.ravel()
    order = np.argsort(flat, kind="stable")
    e = flat[order]
    tol = deg_tol * max(1.0, abs(e[nO - 1]))
    lo = nO - 1
    while lo > 0 and abs(e[lo - 1] - e[nO - 1]) < tol:
        lo -= 1
    hi = nO - 1
    while hi < len(e) - 1 and abs(e[hi + 1] - e[nO - 1]) < tol:
        hi += 1
    w = np.zeros(len(e))
    w[order[:lo]] = 2.0
    w[order[lo:hi + 1]] = 2.0 * (nO - lo) / (hi - lo + 1)
    cw = ck * np.sqrt(w.reshape(ek.shape))[:, None, :]
    Pk = cw @ cw.conj().transpose(0, 2, 1)
    return bl.to_ao(Pk), Pk


class _DIIS:
    def __init__(self, n: int):
        self.err = []
        self.F = []
        self.n = n

    def extrapolate(self, err: np.ndarray, F: np.ndarray) -> np.ndarray:
        self.err.insert(0, err.ravel().copy())
        self.F.insert(0, F.ravel().copy())
        del self.err[self.n:], self.F[self.n:]
        m = len(self.err)
        E = np.array(self.err)
        A = np.zeros((m + 1, m + 1))
        A[:m, :m] = (E.conj() @ E.T).real
        A[:m, m] = A[m, :m] = -1.0
        b = np.zeros(m + 1)
        b[m] = -1.0
        rcond = 1.0 / np.linalg.cond(A, 1) if np.all(np.isfinite(A)) else 0.0
        if not np.isfinite(rcond) or rcond < 1e-14:
            self.err.clear()
            self.F.clear()
            return F
        w = np.linalg.solve(A, b)
        return (w[:m] @ np.array(self.F)).reshape(F.shape)


@dataclasses.dataclass
class RHFResult:
    energy: float                   # E_HF including nuclear repulsion
    e_nuc: float
    converged: bool
    iterations: int
    orbital_error: float
    mo_energies: np.ndarray         # per k: (Ncell, nf), ascending
    mo_coef: np.ndarray             # per k: (Ncell, nf, nf) AO-Bloch coefficients
    P: np.ndarray                   # AO density (nBas, nBas)
    occupied: np.ndarray            # (Ncell, nf) bool
    times: dict


def rhf_sao(sysm, S: np.ndarray, T: np.ndarray, V: np.ndarray, jk: Callable[[np.ndarray], tuple],
            diis: bool = True, max_scf: int = MAX_SCF, thresh: float = THRESH,
            log: Optional[Callable[[str], None]] = None) -> RHFResult:
    """RHF_SAO: k-space RHF with the unit-cell rows of J and K from ``jk(P)``."""
    t0 = time.perf_counter()
    nb, nO, Nc, nf = sysm.nbas, sysm.n_occ, sysm.n_cell, sysm.nfuc
    Hc = T + V
    enuc = nuclear_repulsion(sysm)
    bl = Bloch(sysm)
    Sk = bl.to_k(S)
    sk, Uk = np.linalg.eigh(Sk)
    kspace = bool(np.all(sk > LINDEP))
    if kspace:
        # No linear dependence: X = S^-1/2 (Lowdin), F, P, S are block-circulant, so the
        # whole loop runs block by block in k space -- the same numbers as the supercell
        # forms: (X^T err X)(k) = X(k)^H err(k) X(k), its largest element is the largest
        # element of its inverse lattice transform, and Frobenius products (DIIS) are sums
        # over k (Parseval).  The core guess fills a degenerate level at the cut evenly
        # over its whole space, as in the supercell.
        Xk = (Uk / np.sqrt(sk)[:, None, :]) @ Uk.conj().transpose(0, 2, 1)
        Hk = bl.to_k(Hc)
        e0k, U0k = np.linalg.eigh(Xk.conj().transpose(0, 2, 1) @ Hk @ Xk)
        P, Pk = _occupy_cut_k(e0k, Xk @ U0k, nO, bl)
    else:
        X = canonical_X(S)
        # core guess in the supercell, degenerate level at the cut shared evenly
        cp = X.T @ Hc @ X
        _shift_dropped(X, cp)
        e0, U0 = np.linalg.eigh(cp)
        P = _occupy_cut(e0, X @ U0, nO)
        Xk = np.zeros_like(Sk)
        for k in range(Nc):
            Xk[k] = canonical_X(Sk[k])
    dead = [np.nonzero(np.all(Xk[k] == 0, axis=0))[0] for k in range(Nc)]
    dead_k = np.concatenate([np.full(len(d), k) for k, d in enumerate(dead)]).astype(int)
    dead_i = np.concatenate(dead).astype(int) if dead_k.size else np.zeros(0, dtype=int)
    XkH = Xk.conj().transpose(0, 2, 1)
    dii = _DIIS(MAX_DIIS)
    E_old, conv, converged = 0.0, 1.0, False
    n_stable, hist = 0, [np.inf] * 5
    t_jk = 0.0
    it = 0
    ek = np.zeros((Nc, nf))
    Ck = np.zeros((Nc, nf, nf), dtype=complex)
    occ = np.zeros((Nc, nf), dtype=bool)
    while it < max_scf:
        it += 1
        tj = time.perf_counter()
        J, K = jk(P)
        t_jk += time.perf_counter() - tj
        ET = float(np.sum(P * T.T))
        EV = float(np.sum(P * V.T))
        EJ = 0.5 * float(np.sum(P * J.T)) * Nc
        EK = 0.5 * float(np.sum(P * K.T)) * Nc
        E = ET + EV + EJ + EK
        if kspace:
            Fk = bl.to_k(Hc + J + K)
            Mk = Fk @ Pk @ Sk
            err_o = XkH @ (Mk - Mk.conj().transpose(0, 2, 1)) @ Xk
            if it > 1:
                conv = bl.max_abs_ao(err_o)
        else:
            F = Hc + J + K
            F = bl.to_ao(bl.to_k(F))
            M = F @ P @ S
            err = M - M.T                     # S P F = (F P S)^T: F, P, S are symmetric
            err_o = X.T @ err @ X
            if it > 1:
                conv = float(np.abs(err_o).max())
        if log:
            log(f"  SCF {it:3d}  E = {E + enuc:.12f}  conv = {conv:.3e}")
        if conv < thresh:
            converged = True
            break
        if it > 2 and abs(E - E_old) < max(1e-10, 2e-10 * abs(E)):
            n_stable += 1
        else:
            n_stable = 0
        if n_stable >= 5 and conv < 1e-6 and conv > 0.5 * hist[it % 5]:
            converged = True
            break
        hist[it % 5] = conv
        E_old = E
        if kspace:
            if diis:
                Fk = dii.extrapolate(err_o, Fk)
        else:
            if diis:
                F = dii.extrapolate(err_o, F)
            Fk = bl.to_k(F)
        Fp = XkH @ Fk @ Xk                           # all k blocks at once
        Fp[dead_k, dead_i, dead_i] = LINDEP_SHIFT
        ek, Up = np.linalg.eigh(Fp)
        Ck = Xk @ Up
        order = np.argsort(ek.ravel(), kind="stable")[:nO]
        occ[:] = False
        occ.ravel()[order] = True
        Cocc = Ck * occ[:, None, :]
        Pk = 2.0 * Cocc @ Cocc.conj().transpose(0, 2, 1)
        P = bl.to_ao(Pk)
    return RHFResult(energy=E + enuc, e_nuc=enuc, converged=converged, iterations=it, orbital_error=conv,
                     mo_energies=ek.copy(), mo_coef=Ck.copy(), P=P, occupied=occ.copy(),
                     times={"scf_s": time.perf_counter() - t0, "jk_s": t_jk})


def rhf_sao_gpu(sysm, eng, S: np.ndarray, T: np.ndarray, V: np.ndarray, source: str = "stored",
                diis: bool = True, max_scf: int = MAX_SCF, thresh: float = THRESH) -> RHFResult:
    """RHF_SAO with the whole iteration on the GPU (engine.rhf_sao, tcg_rhf_sao):
    the same guess, DIIS and convergence rules as rhf_sao, with J/K from the
    resident SAO ERI (source='stored') or integral-direct ('direct')."""
    t0 = time.perf_counter()
    nf, Nc = sysm.nfuc, sysm.n_cell
    r = eng.rhf_sao(S, T, V, sysm.n_occ, source=source, max_scf=max_scf, thresh=thresh, diis=diis)
    enuc = nuclear_repulsion(sysm)
    C = r["coeff_sao"]
    Ck = np.stack([C[k * nf:(k + 1) * nf, k * nf:(k + 1) * nf] for k in range(Nc)])   # copies of the k blocks
    return RHFResult(energy=r["e_hf"] + enuc, e_nuc=enuc, converged=r["converged"], iterations=r["iterations"],
                     orbital_error=r["conv"], mo_energies=r["e"].reshape(Nc, nf), mo_coef=Ck, P=r["P"],
                     occupied=(r["occ"] == 2).reshape(Nc, nf),
                     times={"scf_s": time.perf_counter() - t0, "jk_s": r["jk_ms"] * 1e-3,
                            "gpu_loop_s": r["wall_ms"] * 1e-3, "gap": r["gap"], "stalled": r["stalled"]})


@dataclasses.dataclass
class RHFGammaResult:
    energy: float                   # E_HF including nuclear repulsion
    e_nuc: float
    converged: bool
    iterations: int
    orbital_error: float
    mo_energies: np.ndarray         # (nBas,), ascending
    mo_coef: np.ndarray             # (nBas, nBas), columns = orbitals
    P: np.ndarray
    times: dict


def rhf(sysm, S: np.ndarray, T: np.ndarray, V: np.ndarray, jk: Callable[[np.ndarray], tuple],
        diis: bool = True, max_scf: int = MAX_SCF, thresh: float = THRESH,
        log: Optional[Callable[[str], None]] = None) -> RHFGammaResult:
    """RHF in the supercell AO basis (src/Energies/RHF.f90:1-369), J and K of
    every row from ``jk(P)`` (the stored full ERI, hartree_exchange_potential
    with n1 = nBas).  Core guess (guess_AO_RHF, src/Energies/guess_RHF.f90),
    energy before the convergence test, orbital error X^T (FPS - SPF) X,
    DIIS on that error, removed directions shifted to the top."""
    t0 = time.perf_counter()
    nO = sysm.n_occ
    Hc = T + V
    enuc = nuclear_repulsion(sysm)
    X = canonical_X(S)
    cp = X.T @ Hc @ X
    _shift_dropped(X, cp)
    e, U = np.linalg.eigh(cp)
    C = X @ U
    P = 2.0 * C[:, :nO] @ C[:, :nO].T
    dii = _DIIS(MAX_DIIS)
    conv, converged, E = 1.0, False, 0.0
    t_jk = 0.0
    it = 0
    while it < max_scf:
        it += 1
        tj = time.perf_counter()
        J, K = jk(P)
        t_jk += time.perf_counter() - tj
        F = Hc + J + K
        E = float(np.sum(P * T.T) + np.sum(P * V.T) + 0.5 * np.sum(P * J.T) + 0.5 * np.sum(P * K.T))
        err = X.T @ (F @ P @ S - S @ P @ F) @ X
        if it > 1:
            conv = float(np.abs(err).max())
        if log:
            log(f"{it:3d} E = {E + enuc:.12f} conv = {conv:.3e}")
        if conv < thresh:
            converged = True
            break
        if diis:
            F = dii.extrapolate(err, F)
        Fp = X.T @ F @ X
        _shift_dropped(X, Fp)
        e, U = np.linalg.eigh(Fp)
        C = X @ U
        P = 2.0 * C[:, :nO] @ C[:, :nO].T
    return RHFGammaResult(energy=E + enuc, e_nuc=enuc, converged=converged, iterations=it, orbital_error=conv,
                          mo_energies=e, mo_coef=C, P=P,
                          times={"total_s": time.perf_counter() - t0, "jk_s": t_jk})
