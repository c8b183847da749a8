"""Density fitting: the auxiliary set and the GPU DF context (tcg_df).

Host-side mirror of the reference's module df_toroidal
(src/module/df_toroidal.f90):

* ``etb_aux``      follows ``etb_atom`` (:205-285), keyword ``df_etb``: per
  orbital l the exponent range over primitives with max |c| > 1e-3, l capped
  by the element's row (1 for H-Be, 2 for B-Ca, ...), for every
  L = li + lj <= 2 l_max an even-tempered uncontracted set
  e_min_L beta^i, n_L = ceil(ln((e_max_L + e_min_L)/e_min_L) / ln beta),
  largest exponent first.
* ``read_aux``     the file ``AuxBasis`` (keyword ``df_auxfile``,
  ``read_shells_file`` :119-191): the Dalton format of ``Basis``.
* ``aux_list``     the auxiliary AO list of the supercell (``classification``
  on the same atoms, normalised like the orbital basis, :478-488).
* ``DF``           df_init / df_JK / df_kMO through the C ABI (tcg_df_*).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import os
from typing import Dict

import numpy as np

from . import basis as _basis
from . import system as _system
from .engine import EngineError, lib

_vp = ctypes.c_void_p
_ci = ctypes.c_int
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
# even-tempered H set of the DF-MP2 workload: s 8 exponents 0.1-40, p 5 over 0.2-10,
# d 3 over 0.3-3, geometric, uncontracted (41 Cartesian functions per H)
AUX_H_ETB41 = os.path.join(DATA, "AuxBasis_H_etb41")


def _row_lmax(z: int) -> int:
    if z <= 0:
        return 0
    if z <= 4:
        return 1
    if z <= 20:
        return 2
    if z <= 57:
        return 3
    return 4


def etb_aux(orbital: Dict[int, _system.AtomBasis], beta: float = 2.0) -> Dict[int, _system.AtomBasis]:
    """Even-tempered auxiliary set per element (etb_atom, df_toroidal.f90:205-285)."""
    out = {}
    for z, ab in orbital.items():
        emin, emax, lmax = {}, {}, -1
        for l, (ex, co) in ab.shells.items():
            for k in range(len(ex)):
                if co.shape[1] and np.max(np.abs(co[k])) > 1e-3:
                    emin[l] = min(emin.get(l, 1e99), ex[k])
                    emax[l] = max(emax.get(l, 0.0), ex[k])
                    lmax = max(lmax, l)
        if lmax < 0:
            raise ValueError(f"element {z}: no orbital basis functions")
        lmax = min(lmax, _row_lmax(z))
        if 2 * lmax > 2:
            raise ValueError("the generated auxiliary set needs l > 2 (this engine handles s, p, d)")
        shells = {}
        for LL in range(2 * lmax + 1):
            eL_max, eL_min = -1.0, math.inf
            for li in range(lmax + 1):
                lj = LL - li
                if lj < 0 or lj > lmax or li not in emax or lj not in emax:
                    continue
                eL_max = max(eL_max, 2.0 * math.sqrt(emax[li] * emax[lj]))
                eL_min = min(eL_min, 2.0 * math.sqrt(emin[li] * emin[lj]))
            if eL_max < 0:
                continue
            n = math.ceil(math.log((eL_max + eL_min) / eL_min) / math.log(beta))
            if n <= 0:
                continue
            ex = np.array([eL_min * beta ** (n - 1 - i) for i in range(n)])
            shells[LL] = (ex, np.eye(n))
        out[z] = _system.AtomBasis(shells=shells)
    return out


def read_aux(path_or_text: str) -> Dict[int, _system.AtomBasis]:
    """An ``AuxBasis`` file (Dalton format, read_shells_file :119-191)."""
    text = path_or_text
    if "\n" not in path_or_text and os.path.exists(path_or_text):
        with open(path_or_text) as f:
            text = f.read()
    return _system.read_basis(text)


def aux_list(sysm, aux: Dict[int, _system.AtomBasis]):
    """Auxiliary AO list of the supercell, normalised: a TorusSystem whose AO
    arrays are the auxiliary functions (cells in the orbital basis' order)."""
    center, lxyz, off, expo, coef = _system.classification(sysm.geom, sysm.Z, aux)
    axs = dataclasses.replace(sysm, center=center, lxyz=lxyz, prim_off=off, expo=expo, coef=coef)
    if axs.nbas % sysm.n_cell:
        raise ValueError("auxiliary functions do not repeat in every cell")
    return axs.with_coef(_basis.normalize(axs))


def _bind():
    L = lib()
    if getattr(L, "_df_bound", False):
        return L
    L.tcg_df_create.argtypes = [_vp, _ci, _dp, _ip, _ip, _dp, _dp, ctypes.POINTER(_vp)]
    L.tcg_df_info.argtypes = [_vp, _lp, ctypes.POINTER(ctypes.c_double)]
    L.tcg_df_b.argtypes = [_vp, _dp]
    L.tcg_df_jk.argtypes = [_vp, _ci, _dp, _dp, _dp]
    L.tcg_df_kmp2.argtypes = [_vp, _ci, _dp, _dp, _ip, _dp]
    L.tcg_df_kmp2_fp32.argtypes = [_vp, _ci, _dp, _dp, _ip, _dp]
    L.tcg_df_times.argtypes = [_vp, _dp]
    L.tcg_df_last_error.argtypes = [_vp]
    L.tcg_df_last_error.restype = ctypes.c_char_p
    L.tcg_df_destroy.argtypes = [_vp]
    L._df_bound = True
    return L


class DF:
    """df_init on the GPU for the engine's torus system and basis.  ``aux``
    is an auxiliary AO list from ``aux_list``."""

    def __init__(self, eng, aux):
        L = _bind()
        s = eng.s
        self.s = s
        self.nau = aux.nbas // s.n_cell
        h = _vp()
        rc = L.tcg_df_create(eng.h, self.nau, np.ascontiguousarray(aux.center).ravel(),
                             np.ascontiguousarray(aux.lxyz, dtype=np.int32).ravel(),
                             np.ascontiguousarray(aux.prim_off, dtype=np.int32), np.ascontiguousarray(aux.expo),
                             np.ascontiguousarray(aux.coef, dtype=np.float64), ctypes.byref(h))
        self.h = h
        if rc != 0:
            msg = L.tcg_df_last_error(h).decode() if h.value else "bad arguments"
            raise EngineError(f"tcg_df_create failed ({rc}): {msg}")

    def _check(self, rc, what):
        if rc != 0:
            raise EngineError(f"{what} failed ({rc}): {_bind().tcg_df_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            _bind().tcg_df_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        out = np.zeros(3, dtype=np.int64)
        dmin = ctypes.c_double(0.0)
        self._check(_bind().tcg_df_info(self.h, out, ctypes.byref(dmin)), "tcg_df_info")
        return {"nau": int(out[0]), "naux": int(out[1]), "dropped": int(out[2]), "smallest_pivot": dmin.value}

    def times(self):
        out = np.zeros(3)
        self._check(_bind().tcg_df_times(self.h, out), "tcg_df_times")
        return {"metric_ms": out[0], "three_index_ms": out[1], "b_ms": out[2]}

    def B(self) -> np.ndarray:
        """df_B as a C-ordered array B[Q, nu, mu0] (Fortran df_B(mu0, nu, Q))."""
        s = self.s
        out = np.zeros(s.nfuc * s.nbas * self.nau * s.n_cell)
        self._check(_bind().tcg_df_b(self.h, out), "tcg_df_b")
        return out.reshape(self.nau * s.n_cell, s.nbas, s.nfuc)

    def jk(self, P: np.ndarray):
        """J, K (ordinary numpy, rows >= nfuc zero) from the fitted integrals."""
        nb = self.s.nbas
        Pf = np.ascontiguousarray(np.asarray(P, dtype=np.float64).T)
        J = np.zeros(nb * nb)
        K = np.zeros(nb * nb)
        self._check(_bind().tcg_df_jk(self.h, nb, Pf.ravel(), J, K), "tcg_df_jk")
        return J.reshape(nb, nb).T.copy(), K.reshape(nb, nb).T.copy()

    def kmp2(self, mo_coef: np.ndarray, mo_energies: np.ndarray, occupied: np.ndarray, fp32: bool = False):
        """k-point DF-MP2; arguments and result as Engine.kmp2.  fp32: the
        three-index transform and the (ov|ov) assembly in FP32 (FP32 matrix-core
        GEMMs, FP32 accumulation) with compensated FP32 energy sums
        (tcg_df_kmp2_fp32)."""
        s = self.s
        nb, nf = s.nbas, s.nfuc
        N = nb // nf
        C = np.zeros((nb, nb), dtype=complex)
        for k in range(N):
            C[k * nf:(k + 1) * nf, k * nf:(k + 1) * nf] = mo_coef[k]
        cf = np.ascontiguousarray(C.T).view(np.float64).ravel()
        e = np.ascontiguousarray(np.asarray(mo_energies, dtype=np.float64).reshape(N * nf))
        occ = np.ascontiguousarray(np.where(np.asarray(occupied).reshape(N * nf), 2, 0).astype(np.int32))
        out = np.zeros(5)
        fn = _bind().tcg_df_kmp2_fp32 if fp32 else _bind().tcg_df_kmp2
        self._check(fn(self.h, nb, cf, e, occ, out), "tcg_df_kmp2_fp32" if fp32 else "tcg_df_kmp2")
        return {"ec": out[0], "direct": out[1], "exchange": out[2], "zero_denominators": int(out[3]),
                "gpu_ms": out[4]}
