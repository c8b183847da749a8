"""ctypes access to the C restatement of the reference path
(oracle/tcg_oracle.c -> oracle/_build/libtcgoracle.so).

Test infrastructure only: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the engine.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libtcgoracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_ci = ctypes.c_int
_cd = ctypes.c_double


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_iv_scaled.argtypes = [_ci, _cd]
        L.orc_iv_scaled.restype = _cd
        L.orc_bessel_vector.argtypes = [_ci, _cd, _dp]
        L.orc_bessel_vector_fast.argtypes = [_ci, _cd, _dp]
        L.orc_nmax_est.argtypes = [_cd]
        L.orc_nmax_est.restype = _ci
        L.orc_fast_quad_points.argtypes = [_ci, _dp, _dp]
        L.orc_set_nodes.argtypes = [_ci, _dp, _dp]
        L.orc_init.argtypes = [_ci, _dp, _ip, _ci, _ci]
        L.orc_set_basis.argtypes = [_ci, _dp, _ip, _ip, _dp, _dp]
        L.orc_setup_tables.restype = _ci
        L.orc_get_tables.argtypes = [_dp, _dp, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_bary.argtypes = [_ci, _cd, _cd, _cd, _cd, ctypes.POINTER(_cd), ctypes.POINTER(_cd)]
        L.orc_eri_block.argtypes = [_ci] * 8 + [_dp]
        L.orc_eri_full.argtypes = [_ci, _ci, _dp, _lp]
        L.orc_jk.argtypes = [_ci, _ci, _ci, _dp, _dp, _dp, _dp]
        L.orc_counters.argtypes = [_lp]
        _lib = L
    return _lib


def iv_scaled(v: int, x: float) -> float:
    return lib().orc_iv_scaled(v, x)


def bessel_vector(N: int, A: float, fast: bool = False) -> np.ndarray:
    out = np.zeros(N + 1)
    (lib().orc_bessel_vector_fast if fast else lib().orc_bessel_vector)(N, A, out)
    return out


def fast_quad_points(n: int):
    t = np.zeros(n)
    w = np.zeros(n)
    lib().orc_fast_quad_points(n, t, w)
    return t, w


def nmax_est(A: float) -> int:
    return lib().orc_nmax_est(A)


class OracleSystem:
    """A torus system loaded into the oracle.  ``coef`` must be the normalised
    contraction coefficients (the engine's and the reference's inputs)."""

    def __init__(self, sysm, coef: np.ndarray, nq: int = 32, nodes=None):
        """nodes = (t, w): an explicit rule (the 64/128-point tables) instead
        of the nq-point fast rule."""
        self.s = sysm
        L = lib()
        L.orc_init(sysm.ndp, np.array(sysm.L, dtype=np.float64), np.array(sysm.ncell, dtype=np.int32),
                   sysm.nfuc, nq)
        if nodes is not None:
            t, w = (np.ascontiguousarray(x, dtype=np.float64) for x in nodes)
            nq = len(t)
            if L.orc_set_nodes(nq, t, w) != 0:
                raise ValueError(f"oracle: {nq} nodes not supported")
        L.orc_set_basis(sysm.nbas, np.ascontiguousarray(sysm.center), np.ascontiguousarray(sysm.lxyz),
                        np.ascontiguousarray(sysm.prim_off), np.ascontiguousarray(sysm.expo),
                        np.ascontiguousarray(coef, dtype=np.float64))
        self.ncap = L.orc_setup_tables()
        self.nq = nq

    def tables(self):
        nq, nd, nc = self.nq, self.s.ndp, self.ncap
        t = np.zeros(nq)
        w = np.zeros(nq)
        tb = np.zeros(nd * nq * (nc + 1))
        tn = np.zeros(nd * nq, dtype=np.int32)
        lib().orc_get_tables(t, w, tb.ctypes.data, tn.ctypes.data)
        return t, w, tb.reshape(nd, nq, nc + 1), tn.reshape(nd, nq)

    def bary(self, d: int, e1: float, e2: float, x1: float, x2: float):
        pe, cen = _cd(0), _cd(0)
        lib().orc_bary(d, e1, e2, x1, x2, ctypes.byref(pe), ctypes.byref(cen))
        return pe.value, cen.value

    def eri_block(self, a, na, b, nb, c, nc, d, nd):
        out = np.zeros(na * nb * nc * nd)
        rc = lib().orc_eri_block(a, na, b, nb, c, nc, d, nd, out)
        if rc:
            raise RuntimeError(f"orc_eri_block rc={rc}")
        return out.reshape(nd, nc, nb, na).transpose(3, 2, 1, 0).copy()

    def eri_full(self, sao: bool = True, physicist: bool = False):
        """Returns (E, stats); E[l,k,j,i] = ERI(i,j,k,l) in the reference's
        Fortran layout ERI(n1,nBas,nBas,nBas); stats = (shell quartets, primitive quartets)."""
        nb = self.s.nbas
        n1 = self.s.nfuc if sao else nb
        out = np.zeros(n1 * nb * nb * nb)
        st = np.zeros(4, dtype=np.int64)
        rc = lib().orc_eri_full(int(sao), int(physicist), out, st)
        if rc:
            raise RuntimeError(f"orc_eri_full rc={rc}")
        return out.reshape(nb, nb, nb, n1), st

    def jk(self, P: np.ndarray, E: np.ndarray, physicist: bool = False):
        """J, K (rows <= n1) from the stored ERI; P, J, K as ordinary (row=mu) numpy matrices."""
        nb = self.s.nbas
        n1 = E.shape[3]
        Pf = np.ascontiguousarray(P.T)          # column-major
        J = np.zeros(nb * nb)
        K = np.zeros(nb * nb)
        lib().orc_jk(n1, nb, int(physicist), Pf.ravel(), np.ascontiguousarray(E).ravel(), J, K)
        return J.reshape(nb, nb).T.copy(), K.reshape(nb, nb).T.copy()
