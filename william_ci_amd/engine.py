"""Python binding of the engine's C ABI (include/tcg.h), with the names and
argument meaning of the reference's Fortran entry points it replaces:

  ERI_integral_toroidal       src/Toroidal/ERI_toroidal.f90:36
  hartree_exchange_potential  src/Energies/hartree_exchange_potential.f90:30

The engine is the HIP library william_ci_amd/libtcg.so.  There is no CPU
fallback: if the library or a GPU is missing, every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtcg.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_ci = ctypes.c_int
_vp = ctypes.c_void_p


class EngineError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EngineError(f"engine library missing: {LIB_PATH} (build with python -m william_ci_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        L.tcg_create.argtypes = [_ci, _dp, _ip, _ci, ctypes.POINTER(_vp)]
        L.tcg_set_basis.argtypes = [_vp, _ci, _ci, _dp, _ip, _ip, _dp, _dp]
        L.tcg_eri_build.argtypes = [_vp, _ci, _ci, _vp]
        L.tcg_jk_stored.argtypes = [_vp, _ci, _ci, _dp, _vp, _dp, _dp]
        L.tcg_jk_stored_order.argtypes = [_vp, _ci, _ci, _dp, _vp, _ci, _dp, _dp]
        L.tcg_mp2_gamma.argtypes = [_vp, _ci, _ci, _dp, _dp, _dp]
        L.tcg_kmp2.argtypes = [_vp, _ci, _dp, _dp, _ip, _dp]
        L.tcg_kmp2_direct.argtypes = [_vp, _ci, _dp, _dp, _ip, _ci, _dp]
        L.tcg_set_direct_screen.argtypes = [_vp, ctypes.c_double, _ci, _ci]
        L.tcg_direct_jk_screened.argtypes = [_vp, _ci, _dp, ctypes.c_double, _dp, _dp]
        L.tcg_direct_screen_info.argtypes = [_vp, _dp, _ip, np.ctypeslib.ndpointer(dtype=np.int64)]
        L.tcg_one_electron.argtypes = [_vp, _ci, _dp, _ip, _dp, _dp, _dp]
        L.tcg_eri_stats.argtypes = [_vp, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"), _ci]
        L.tcg_eri_times.argtypes = [_vp, _dp, _ci]
        L.tcg_eri_device.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_longlong)]
        L.tcg_jk_stored_noalloc.argtypes = [_vp, _ci, _ci, _dp, _dp, _dp]
        L.tcg_direct_init.argtypes = [_vp]
        L.tcg_direct_jk.argtypes = [_vp, _ci, _dp, _dp, _dp]
        L.tcg_normalize.argtypes = [_vp, _ci, _dp, _ip, _ip, _dp, _dp, _dp]
        L.tcg_ao_to_sao.argtypes = [_vp, _ci, _dp, _dp]
        L.tcg_sao_to_ao.argtypes = [_vp, _ci, _dp, _dp]
        L.tcg_heev_blocks.argtypes = [_vp, _ci, _dp, _dp]
        L.tcg_rhf_sao.argtypes = [_vp, _ci, _ci, _ci, _dp, _dp, _dp, _ci, ctypes.c_double, _ci, _dp, _dp, _ip, _dp,
                                  _dp]
        L.tcg_set_quadrature.argtypes = [_vp, _ci, _ci]
        L.tcg_get_quadrature.argtypes = [_vp, ctypes.POINTER(_ci), ctypes.POINTER(_ci), _vp, _vp]
        L.tcg_fast_quad_self_check.argtypes = [_vp, _ci, _dp, _ip, _ci, ctypes.POINTER(ctypes.c_double),
                                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_ci)]
        L.tcg_last_error.argtypes = [_vp]
        L.tcg_last_error.restype = ctypes.c_char_p
        L.tcg_destroy.argtypes = [_vp]
        _lib = L
    return _lib


class Engine:
    """One torus system on the GPU (a tcg_ctx)."""

    def __init__(self, sysm, coef: np.ndarray, nquad: int = 32):
        self.s = sysm
        L = lib()
        h = _vp()
        rc = L.tcg_create(sysm.ndp, np.array(sysm.L, dtype=np.float64), np.array(sysm.ncell, dtype=np.int32),
                          nquad, ctypes.byref(h))
        if rc != 0 or not h.value:
            raise EngineError(f"tcg_create failed ({rc}): no HIP device or bad torus")
        self.h = h
        self._check(L.tcg_set_basis(h, sysm.nbas, sysm.nfuc, np.ascontiguousarray(sysm.center).ravel(),
                                    np.ascontiguousarray(sysm.lxyz, dtype=np.int32).ravel(),
                                    np.ascontiguousarray(sysm.prim_off, dtype=np.int32),
                                    np.ascontiguousarray(sysm.expo), np.ascontiguousarray(coef, dtype=np.float64)),
                    "tcg_set_basis")

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().tcg_last_error(self.h).decode()
            raise EngineError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().tcg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- ERI_integral_toroidal -------------------------------------------------
    def eri_build(self, sao: bool = True, physicist: bool = False, to_host: bool = True):
        """Builds the ERI; returns it as the reference stores it (Fortran
        ERI(n1,nBas,nBas,nBas)) viewed C-ordered as E[l,k,j,i] = ERI(i,j,k,l),
        or None when to_host is False (the tensor then stays on the GPU only)."""
        nb = self.s.nbas
        n1 = self.s.nfuc if sao else nb
        out = np.empty(n1 * nb * nb * nb) if to_host else None
        self._check(lib().tcg_eri_build(self.h, int(sao), int(physicist),
                                        out.ctypes.data if to_host else None), "tcg_eri_build")
        return out.reshape(nb, nb, nb, n1) if to_host else None

    def stats(self):
        out = np.zeros(8, dtype=np.int64)
        lib().tcg_eri_stats(self.h, out, 8)
        return {"shell_quartets": int(out[0]), "prim_quartets": int(out[1]), "rows": int(out[2]),
                "flops": int(out[3]), "row_flops": int(out[4])}

    def times(self):
        out = np.zeros(8)
        lib().tcg_eri_times(self.h, out, 8)
        return {"tables_pairs_ms": out[0], "schwarz_ms": out[1], "quartet_list_ms": out[2],
                "quartets_ms": out[3], "total_ms": out[4], "kernel_ms": out[5], "rows_kernel_ms": out[6],
                "t_tables_ms": out[7]}

    def eri_device(self):
        p = _vp()
        n = ctypes.c_longlong()
        lib().tcg_eri_device(self.h, ctypes.byref(p), ctypes.byref(n))
        return p.value, n.value

    # -- quadrature rule (fast_quad / nquad_point) ------------------------------
    def set_quadrature(self, fast: bool, nquad: int):
        """fast: nquad mapped Gauss-Legendre points; else the 64/128 table."""
        self._check(lib().tcg_set_quadrature(self.h, int(fast), int(nquad)), "tcg_set_quadrature")

    def quadrature(self):
        """(fast, nquad, t, w) of the rule in place."""
        f, n = _ci(), _ci()
        self._check(lib().tcg_get_quadrature(self.h, ctypes.byref(f), ctypes.byref(n), None, None),
                    "tcg_get_quadrature")
        t = np.zeros(n.value)
        w = np.zeros(n.value)
        self._check(lib().tcg_get_quadrature(self.h, ctypes.byref(f), ctypes.byref(n), t.ctypes.data,
                                             w.ctypes.data), "tcg_get_quadrature")
        return bool(f.value), n.value, t, w

    def fast_quad_self_check(self, nq_full: int = 64):
        """fast_quad_self_check: sample ERIs and V with the fast rule and the
        nq_full table; keeps the fast rule only if both agree to 1e-10."""
        s = self.s
        de, dv, on = ctypes.c_double(), ctypes.c_double(), _ci()
        self._check(lib().tcg_fast_quad_self_check(self.h, s.natom, np.ascontiguousarray(s.geom).ravel(),
                                                   np.ascontiguousarray(s.Z, dtype=np.int32), int(nq_full),
                                                   ctypes.byref(de), ctypes.byref(dv), ctypes.byref(on)),
                    "tcg_fast_quad_self_check")
        return {"eri_max_diff": de.value, "v_max_diff": dv.value, "fast_quad": bool(on.value)}

    # -- overlap / kinetic / nuclear_attraction_matrix_toroidal_nD -------------
    def one_electron(self):
        """S, T, V of the supercell as ordinary (row-major) numpy matrices."""
        s = self.s
        nb = s.nbas
        S = np.zeros(nb * nb)
        T = np.zeros(nb * nb)
        V = np.zeros(nb * nb)
        self._check(lib().tcg_one_electron(self.h, s.natom, np.ascontiguousarray(s.geom).ravel(),
                                           np.ascontiguousarray(s.Z, dtype=np.int32), S, T, V), "tcg_one_electron")
        f = lambda M: M.reshape(nb, nb).T                  # the library's column-major arrays (views)
        return f(S), f(T), f(V)

    def refresh_basis(self, coef):
        """Re-sends the AO list (drops every table derived from it)."""
        s = self.s
        self._check(lib().tcg_set_basis(self.h, s.nbas, s.nfuc, np.ascontiguousarray(s.center).ravel(),
                                        np.ascontiguousarray(s.lxyz, dtype=np.int32).ravel(),
                                        np.ascontiguousarray(s.prim_off, dtype=np.int32),
                                        np.ascontiguousarray(s.expo), np.ascontiguousarray(coef, dtype=np.float64)),
                    "tcg_set_basis")

    # -- hartree_exchange_potential ---------------------------------------------
    def jk_stored(self, P: np.ndarray, n1: int | None = None, eri: np.ndarray | None = None,
                  physicist: bool = False):
        """J, K as ordinary row-major numpy matrices (rows > n1 zero), from the
        resident ERI (eri None, in the order it was built) or from a host ERI
        in the reference layout (chemist, or physicist when asked)."""
        nb = self.s.nbas
        n1 = self.s.nfuc if n1 is None else n1
        Pf = np.ascontiguousarray(np.asarray(P, dtype=np.float64).T)
        J = np.zeros(nb * nb)
        K = np.zeros(nb * nb)
        Ea = None
        if eri is not None:
            # the caller's array in the reference's memory order (Fortran
            # ERI(n1,nBas,nBas,nBas) = C-ordered E[l,k,j,i]); the bound copy
            # lives until the call has returned
            Ea = np.ascontiguousarray(eri, dtype=np.float64)
            if Ea.size != n1 * nb ** 3:
                raise EngineError(f"jk_stored: eri has {Ea.size} elements, expected n1*nBas^3 = {n1 * nb ** 3}")
        ep = None if Ea is None else Ea.ctypes.data
        self._check(lib().tcg_jk_stored_order(self.h, n1, nb, Pf.ravel(), ep, int(physicist), J, K),
                    "tcg_jk_stored")
        del Ea
        return J.reshape(nb, nb).T.copy(), K.reshape(nb, nb).T.copy()

    def jk_stored_noalloc(self, P: np.ndarray, n1: int | None = None):
        """jk_stored from the resident tensor through tcg_jk_stored_noalloc (no
        device allocation; J/K rows and P in the library's static scratch)."""
        nb = self.s.nbas
        n1 = self.s.nfuc if n1 is None else n1
        Pf = np.ascontiguousarray(np.asarray(P, dtype=np.float64).T)
        J = np.zeros(nb * nb)
        K = np.zeros(nb * nb)
        self._check(lib().tcg_jk_stored_noalloc(self.h, n1, nb, Pf.ravel(), J, K), "tcg_jk_stored_noalloc")
        return J.reshape(nb, nb).T.copy(), K.reshape(nb, nb).T.copy()

    # -- AO_to_MO_ERI_inplace + MP2 ---------------------------------------------
    def mp2_gamma(self, e: np.ndarray, C: np.ndarray, nocc: int):
        """Gamma-supercell MP2 from the resident full ERI (eri_build(sao=False)).
        C: (nBas, nBas) MO coefficients (ordinary numpy, columns = orbitals,
        ascending e).  Returns dict(ec, direct, exchange, gpu_ms)."""
        nb = self.s.nbas
        Cf = np.ascontiguousarray(np.asarray(C, dtype=np.float64).T)   # column-major C(mu,p)
        out = np.zeros(4)
        self._check(lib().tcg_mp2_gamma(self.h, nb, int(nocc), np.ascontiguousarray(e, dtype=np.float64),
                                         Cf.ravel(), out), "tcg_mp2_gamma")
        return {"ec": out[0], "direct": out[1], "exchange": out[2], "gpu_ms": out[3]}

    # -- transform_ERI_AO_to_kMO + kMP2_calc --------------------------------------
    def _kblocks(self, mo_coef, mo_energies, occupied):
        """coeff_SAO (column-major (re, im) pairs, k-block diagonal), e and occ
        (2 / 0) in the k-block order of the C ABI."""
        s = self.s
        nb, nf = s.nbas, s.nfuc
        N = nb // nf
        C = np.zeros((nb, nb), dtype=complex)
        for k in range(N):
            C[k * nf:(k + 1) * nf, k * nf:(k + 1) * nf] = mo_coef[k]
        cf = np.ascontiguousarray(C.T).view(np.float64).ravel()           # column-major, (re, im) pairs
        e = np.ascontiguousarray(np.asarray(mo_energies, dtype=np.float64).reshape(N * nf))
        occ = np.ascontiguousarray(np.where(np.asarray(occupied).reshape(N * nf), 2, 0).astype(np.int32))
        return cf, e, occ

    def kmp2(self, mo_coef: np.ndarray, mo_energies: np.ndarray, occupied: np.ndarray):
        """k-point MP2 from the resident SAO ERI (eri_build(sao=True)).
        mo_coef[k] (nfuc x nfuc, columns = bands), mo_energies[k, band],
        occupied[k, band] (bool), k ordered as the cells (x fastest) -- the
        fields of scf.rhf_sao's result.  Returns dict(ec, direct, exchange,
        zero_denominators, gpu_ms)."""
        nb = self.s.nbas
        cf, e, occ = self._kblocks(mo_coef, mo_energies, occupied)
        out = np.zeros(5)
        self._check(lib().tcg_kmp2(self.h, nb, cf, e, occ, out), "tcg_kmp2")
        return {"ec": out[0], "direct": out[1], "exchange": out[2], "zero_denominators": int(out[3]),
                "gpu_ms": out[4]}

    def kmp2_direct(self, mo_coef: np.ndarray, mo_energies: np.ndarray, occupied: np.ndarray, slab_cells: int = 0):
        """k-point MP2 without the stored tensor (tcg_kmp2_direct, after
        direct_init): the SAO integrals in slabs of slab_cells cells of the last
        index (0: sized from the free device memory).  Same arguments and keys as
        kmp2, plus integral_ms, transform_ms and slabs."""
        nb, nf = self.s.nbas, self.s.nfuc
        cf, e, occ = self._kblocks(mo_coef, mo_energies, occupied)
        out = np.zeros(8)
        self._check(lib().tcg_kmp2_direct(self.h, nb, cf, e, occ, int(slab_cells), out), "tcg_kmp2_direct")
        return {"ec": out[0], "direct": out[1], "exchange": out[2], "zero_denominators": int(out[3]),
                "gpu_ms": out[4], "integral_ms": out[5], "transform_ms": out[6], "slabs": int(out[7])}

    def set_direct_screen(self, tau: float = 1e-14, full_every: int = 50, learned_bound: bool = False):
        """Density screening of the incremental direct J/K (tcg_set_direct_screen)."""
        self._check(lib().tcg_set_direct_screen(self.h, float(tau), int(full_every), 1 if learned_bound else 0),
                    "tcg_set_direct_screen")

    def direct_screen_info(self):
        """Quartets evaluated / listed in the last direct J/K and since the last
        reset (tcg_direct_screen_info)."""
        tau = np.zeros(1)
        fe = np.zeros(1, dtype=np.int32)
        c = np.zeros(5, dtype=np.int64)
        self._check(lib().tcg_direct_screen_info(self.h, tau, fe, c), "tcg_direct_screen_info")
        return {"tau": float(tau[0]), "full_every": int(fe[0]), "kept": int(c[0]), "listed": int(c[1]),
                "kept_total": int(c[2]), "listed_total": int(c[3]), "builds": int(c[4])}

    # -- direct_init / direct_JK -------------------------------------------------
    def direct_init(self):
        """Schwarz tables and the unique quartet list for the direct J/K."""
        self._check(lib().tcg_direct_init(self.h), "tcg_direct_init")

    def direct_jk(self, P: np.ndarray, tau: float = 0.0):
        """J, K (row-major numpy, rows > nfuc zero) with the integrals
        recomputed on the GPU and contracted on the fly; tau > 0: with density
        screening at tau (tcg_direct_jk_screened)."""
        nb = self.s.nbas
        Pf = np.ascontiguousarray(np.asarray(P, dtype=np.float64).T)
        J = np.zeros(nb * nb)
        K = np.zeros(nb * nb)
        if tau > 0.0:
            self._check(lib().tcg_direct_jk_screened(self.h, nb, Pf.ravel(), float(tau), J, K),
                        "tcg_direct_jk_screened")
        else:
            self._check(lib().tcg_direct_jk(self.h, nb, Pf.ravel(), J, K), "tcg_direct_jk")
        return J.reshape(nb, nb).T.copy(), K.reshape(nb, nb).T.copy()

    # -- k-space (SAO) algebra ----------------------------------------------------
    @staticmethod
    def _cplx_out(buf: np.ndarray, nb: int) -> np.ndarray:
        return buf.view(np.complex128).reshape(nb, nb).T.copy()      # column-major (re, im) pairs

    @staticmethod
    def _cplx_in(M: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray(np.asarray(M, dtype=np.complex128).T).view(np.float64).ravel()

    def ao_to_sao(self, M: np.ndarray) -> np.ndarray:
        """transform_AO_to_SAO: the k-block-diagonal (nBas, nBas) complex matrix."""
        nb = self.s.nbas
        out = np.zeros(2 * nb * nb)
        Mf = np.ascontiguousarray(np.asarray(M, dtype=np.float64).T).ravel()
        self._check(lib().tcg_ao_to_sao(self.h, nb, Mf, out), "tcg_ao_to_sao")
        return self._cplx_out(out, nb)

    def sao_to_ao(self, M: np.ndarray) -> np.ndarray:
        """transform_SAO_to_AO: the real (nBas, nBas) supercell matrix of the k blocks of M."""
        nb = self.s.nbas
        out = np.zeros(nb * nb)
        self._check(lib().tcg_sao_to_ao(self.h, nb, self._cplx_in(M), out), "tcg_sao_to_ao")
        return out.reshape(nb, nb).T.copy()

    def heev_blocks(self, A: np.ndarray):
        """diagonalize_block_diagonal_complex_para: (A with its k blocks replaced by
        their eigenvectors, e block by block ascending)."""
        nb = self.s.nbas
        buf = self._cplx_in(A).copy()
        e = np.zeros(nb)
        self._check(lib().tcg_heev_blocks(self.h, nb, buf, e), "tcg_heev_blocks")
        return self._cplx_out(buf, nb), e

    def rhf_sao(self, S, T, V, nocc: int, source: str = "stored", max_scf: int = 100, thresh: float = 1e-9,
                diis: bool = True) -> dict:
        """RHF_SAO with every iteration on the GPU (tcg_rhf_sao).  source: 'stored'
        (the resident SAO ERI), 'direct' (after direct_init) or 'direct_incremental'
        (direct, incremental Fock builds with density screening, set_direct_screen)."""
        nb = self.s.nbas
        src = {"stored": 0, "direct": 1, "direct_incremental": 2}[source]
        f = lambda M: np.ascontiguousarray(np.asarray(M, dtype=np.float64).T).ravel()   # no copy for F-ordered M
        e = np.zeros(nb)
        C = np.zeros(2 * nb * nb)
        occ = np.zeros(nb, dtype=np.int32)
        P = np.zeros(nb * nb)
        out = np.zeros(12)
        self._check(lib().tcg_rhf_sao(self.h, src, nb, int(nocc), f(S), f(T), f(V), int(max_scf), float(thresh),
                                      1 if diis else 0, e, C, occ, P, out), "tcg_rhf_sao")
        return {"e_hf": out[0], "ET": out[1], "EV": out[2], "EJ": out[3], "EK": out[4],
                "iterations": int(out[5]), "conv": out[6], "converged": bool(out[7]), "stalled": bool(out[8]),
                "gap": out[9], "jk_ms": out[10], "wall_ms": out[11],
                "e": e, "coeff_sao": C.view(np.complex128).reshape(nb, nb).T, "occ": occ,
                "P": P.reshape(nb, nb).T}


def normalized_coef(sysm) -> np.ndarray:
    """normalize_AO_toroidal_nD on the GPU (tcg_normalize): the normalised
    contraction coefficients of the system's AO list."""
    L = lib()
    h = _vp()
    rc = L.tcg_create(sysm.ndp, np.array(sysm.L, dtype=np.float64), np.array(sysm.ncell, dtype=np.int32), 32,
                      ctypes.byref(h))
    if rc != 0:
        raise EngineError(f"tcg_create failed ({rc})")
    try:
        out = np.zeros(len(sysm.coef))
        rc = L.tcg_normalize(h, sysm.nbas, np.ascontiguousarray(sysm.center, dtype=np.float64).ravel(),
                             np.ascontiguousarray(sysm.lxyz, dtype=np.int32).ravel(),
                             np.ascontiguousarray(sysm.prim_off, dtype=np.int32),
                             np.ascontiguousarray(sysm.expo, dtype=np.float64),
                             np.ascontiguousarray(sysm.coef, dtype=np.float64), out)
        if rc != 0:
            raise EngineError(f"tcg_normalize: {L.tcg_last_error(h).decode()}")
        return out
    finally:
        L.tcg_destroy(h)


def ERI_integral_toroidal(sysm, coef, physicist: bool = False, sao: bool = True, nquad: int = 32):
    """ERI(n1,nBas,nBas,nBas) of the normalised AO list (reference layout, as E[l,k,j,i])."""
    eng = Engine(sysm, coef, nquad)
    try:
        return eng.eri_build(sao=sao, physicist=physicist)
    finally:
        eng.close()
