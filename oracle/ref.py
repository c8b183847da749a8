"""ctypes access to oracle/_ref/libtcgref.so: the reference's own Fortran
integral routines (built by oracle/Makefile.ref from the reference sources).

Test infrastructure only.  Used by the tests to pin the C restatement
(oracle/tcg_oracle.c) and the GPU engine to the reference itself.  The
library exists only where it was built; ``available()`` says whether it did.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_ref", "libtcgref.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_ci = ctypes.c_int


_LIB_FAST_PATH = os.path.join(_HERE, "_ref_fast", "libtcgref.so")


def available() -> bool:
    return os.path.exists(_LIB_PATH)


def fast_available() -> bool:
    return os.path.exists(_LIB_FAST_PATH)


def use_fast_build() -> None:
    """Loads the reference built with its own optimisation flags
    (oracle/_ref_fast) instead of the -O2 checker build; for timing only.
    Must come before the first call of lib()."""
    global _LIB_PATH, _lib
    if _lib is not None and _LIB_PATH != _LIB_FAST_PATH:
        raise RuntimeError("the reference library is already loaded")
    _LIB_PATH = _LIB_FAST_PATH


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(_LIB_PATH)
        L.ref_init.argtypes = [_ci, _dp, _ip, _ci, _ci, _ci, _ci, _ci]
        L.ref_set_basis.argtypes = [_ci, _dp, _ip, _ip, _dp, _dp]
        L.ref_get_coef.argtypes = [_dp]
        L.ref_normalize.argtypes = []
        L.ref_setup_tables.argtypes = [ctypes.POINTER(_ci)]
        L.ref_get_nodes.argtypes = [ctypes.POINTER(_ci), _dp, _dp, _ip]
        L.ref_eri_block.argtypes = [_ci] * 8 + [_dp]
        L.ref_eri_full.argtypes = [_dp]
        L.ref_one_electron.argtypes = [_ci, _dp, _ip, _dp, _dp, _dp]
        L.ref_nre.argtypes = [_ci, _dp, _ip, ctypes.POINTER(ctypes.c_double)]
        L.ref_jk.argtypes = [_ci, _ci, _dp, _dp, _dp, _dp]
        L.ref_fast_quad_self_check.argtypes = [_ci, _dp, _ip, _ci, ctypes.c_char_p, _dp]
        L.ref_direct_jk.argtypes = [_dp, _dp, _dp]
        L.ref_mp2.argtypes = [_ci, _ci, _dp, _dp, _dp, ctypes.c_char_p]
        L.ref_kmp2.argtypes = [_ci, np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS"), _dp, _ip,
                               _dp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]
        _cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
        L.ref_ao_to_sao.argtypes = [_ci, _dp, _cp]
        L.ref_sao_to_ao.argtypes = [_ci, _cp, _dp]
        L.ref_clean_and_symmetrize.argtypes = [_ci, _cp]
        L.ref_details.argtypes = [_ci, _ci, _dp, _dp, _dp, _dp, ctypes.c_char_p]
        _lib = L
    return _lib


def clean_and_symmetrize(A: np.ndarray) -> np.ndarray:
    """The reference's clean_and_symmetrize (src/Tools/SAO_Transfer.f90:235) on a
    copy of the complex square matrix A (row-major numpy)."""
    n = A.shape[0]
    buf = np.ascontiguousarray(np.asarray(A, dtype=np.complex128).T).ravel()
    lib().ref_clean_and_symmetrize(n, buf)
    return buf.reshape(n, n).T.copy()


class RefSystem:
    """One torus system loaded into the reference's module state."""

    def __init__(self, sysm, nq: int = 32, fast_quad: bool = True, sao: bool = True,
                 physicist: bool = False, normalize: bool = True):
        self.s = sysm
        self.sao = sao
        L = lib()
        Lv = np.array(sysm.L, dtype=np.float64)
        ncl = np.array(sysm.ncell, dtype=np.int32)
        L.ref_init(sysm.ndp, Lv, ncl, sysm.nfuc, nq, int(fast_quad), int(sao), int(physicist))
        # the reference's AO x/y/z are stored per AO; coordinates as Fortran (3, nbas)
        xyz = np.ascontiguousarray(sysm.center, dtype=np.float64)
        lx = np.ascontiguousarray(sysm.lxyz, dtype=np.int32)
        L.ref_set_basis(sysm.nbas, xyz, lx, sysm.nprim(), sysm.expo, sysm.coef)
        if normalize:
            L.ref_normalize()
        self.coef = np.zeros_like(sysm.coef)
        L.ref_get_coef(self.coef)
        n = _ci(0)
        L.ref_setup_tables(ctypes.byref(n))
        self.ncap = n.value

    def nodes(self):
        t = np.zeros(256)
        w = np.zeros(256)
        tn = np.zeros(256 * 3, dtype=np.int32)
        nq = _ci(0)
        lib().ref_get_nodes(ctypes.byref(nq), t, w, tn)
        q = nq.value
        return t[:q].copy(), w[:q].copy(), tn[:q * self.s.ndp].reshape(self.s.ndp, q).copy()

    def eri_block(self, a, na, b, nb, c, nc, d, nd):
        """Shell block; a..d are 0-based first AOs. Returns blk[na,nb,nc,nd]."""
        out = np.zeros(na * nb * nc * nd)
        lib().ref_eri_block(a + 1, na, b + 1, nb, c + 1, nc, d + 1, nd, out)
        return out.reshape(nd, nc, nb, na).transpose(3, 2, 1, 0).copy()

    def eri_full(self):
        """ERI as the reference stores it: Fortran ERI(n1,nBas,nBas,nBas).

        Returned as a C-ordered array E[l,k,j,i] (the same memory), i.e.
        E[l,k,j,i] = ERI(i,j,k,l)."""
        nb = self.s.nbas
        n1 = self.s.nfuc if self.sao else nb
        out = np.zeros(n1 * nb * nb * nb)
        lib().ref_eri_full(out)
        return out.reshape(nb, nb, nb, n1)

    def details(self, dirpath, S, T, V, E):
        """The reference's details_integrals(_SAO) (keyword Details) writing into
        dirpath; S, T, V row-major numpy, E as eri_full returns it (E[l,k,j,i])."""
        nb = self.s.nbas
        f = lambda M: np.ascontiguousarray(np.asarray(M, dtype=np.float64).T).ravel()
        lib().ref_details(nb, int(self.sao), f(S), f(T), f(V), np.ascontiguousarray(E, dtype=np.float64).ravel(),
                          str(dirpath).encode())

    def one_electron(self):
        nb = self.s.nbas
        S = np.zeros(nb * nb)
        T = np.zeros(nb * nb)
        V = np.zeros(nb * nb)
        geom = np.asfortranarray(self.s.geom).ravel(order="F").copy()
        lib().ref_one_electron(self.s.natom, geom, self.s.Z.astype(np.int32), S, T, V)
        f = lambda M: M.reshape(nb, nb).T.copy()
        return f(S), f(T), f(V)

    def direct_jk(self, P):
        """direct_init + direct_JK on the density P (row-major numpy, symmetric);
        J, K returned as row-major numpy matrices (rows > nfuc zero)."""
        nb = self.s.nbas
        Pf = np.ascontiguousarray(np.asarray(P, dtype=np.float64).T)
        J = np.zeros(nb * nb)
        K = np.zeros(nb * nb)
        lib().ref_direct_jk(Pf.ravel(), J, K)
        return J.reshape(nb, nb).T.copy(), K.reshape(nb, nb).T.copy()

    def mp2(self, E_full, e, C, nocc):
        """The reference's AO_to_MO_ERI_inplace + MP2 on the full chemist ERI
        (C-ordered E[l,k,j,i] = ERI(i,j,k,l), as eri_full returns it with
        sao=False), orbital energies e and coefficients C (row-major numpy,
        columns = orbitals).  Returns the correlation energy it prints."""
        import re
        import tempfile
        nb = self.s.nbas
        eri = np.ascontiguousarray(E_full, dtype=np.float64).ravel().copy()
        Cf = np.ascontiguousarray(np.asarray(C, dtype=np.float64).T).ravel()
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "mp2.out")
            lib().ref_mp2(nb, int(nocc), np.ascontiguousarray(e, dtype=np.float64), Cf, eri, path.encode())
            text = open(path).read()
        m = re.search(r"MP2 correlation energy =\s*([-0-9.E+]+)", text)
        return float(m.group(1))

    def kmp2(self, E_sao, Ck, ek, occ):
        """The reference's transform_ERI_AO_to_kMO + kMP2_calc on the SAO ERI
        (C-ordered E[l,k,j,i] = ERI(i,j,k,l), i < nfuc), per-k coefficients
        Ck[k] (nfuc x nfuc, columns = bands), energies ek[k, band] and
        occupations occ[k, band] (2 or 0).  Returns E_c(MP2)."""
        import tempfile
        nb, nf = self.s.nbas, self.s.nfuc
        N = nb // nf
        csao = np.zeros((nb, nb), dtype=complex)
        for k in range(N):
            csao[k * nf:(k + 1) * nf, k * nf:(k + 1) * nf] = Ck[k]
        cf = np.ascontiguousarray(csao.T)              # Fortran column-major
        eri = np.ascontiguousarray(E_sao, dtype=np.float64).ravel()
        ec = ctypes.c_double(0.0)
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "kmp2.out")
            lib().ref_kmp2(nb, cf.ravel(), np.ascontiguousarray(np.asarray(ek, dtype=np.float64).ravel()),
                           np.ascontiguousarray(np.asarray(occ, dtype=np.int32).ravel()), eri, path.encode(),
                           ctypes.byref(ec))
        return ec.value

    def fast_quad_self_check(self, nq128: bool = False):
        """The reference's fast_quad_self_check on this system (the fast
        rule must be in place).  Returns its report (sample sizes, largest
        ERI and V differences) and the rule it leaves in place."""
        import re
        import tempfile
        geom = np.asfortranarray(self.s.geom).ravel(order="F").copy()
        out = np.zeros(2)
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "selfcheck.out")
            lib().ref_fast_quad_self_check(self.s.natom, geom, self.s.Z.astype(np.int32), int(nq128),
                                           path.encode(), out)
            text = open(path).read()
        m = re.search(r"self-check:\s*(\d+) ERI shell quartets, largest difference\s*([-0-9.E+]+)\s*;"
                      r"\s*nuclear attraction \(sample of\s*(\d+)\)\s*([-0-9.E+]+)", text)
        return {"n_quartets": int(m.group(1)), "eri_max_diff": float(m.group(2)), "n_v": int(m.group(3)),
                "v_max_diff": float(m.group(4)), "fast_quad": bool(out[0]), "nquad": int(out[1]), "text": text}

    def ao_to_sao(self, M):
        """transform_AO_to_SAO on the real (nBas, nBas) matrix M (row-major numpy)."""
        nb = self.s.nbas
        out = np.zeros(nb * nb, dtype=np.complex128)
        lib().ref_ao_to_sao(nb, np.ascontiguousarray(np.asarray(M, dtype=np.float64).T).ravel(), out)
        return out.reshape(nb, nb).T.copy()

    def sao_to_ao(self, M):
        """transform_SAO_to_AO on the complex (nBas, nBas) matrix M (row-major numpy)."""
        nb = self.s.nbas
        out = np.zeros(nb * nb)
        lib().ref_sao_to_ao(nb, np.ascontiguousarray(np.asarray(M, dtype=np.complex128).T).ravel(), out)
        return out.reshape(nb, nb).T.copy()

    def nre(self) -> float:
        geom = np.asfortranarray(self.s.geom).ravel(order="F").copy()
        e = ctypes.c_double(0.0)
        lib().ref_nre(self.s.natom, geom, self.s.Z.astype(np.int32), ctypes.byref(e))
        return e.value
