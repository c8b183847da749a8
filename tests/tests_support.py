"""Helpers shared by the tests (not part of the engine).

``normalised_coef`` normalises the AO list the way
src/basis/normalization/normalize_1D.f90:17-79 does (each primitive to one,
then the contraction to one), with the toroidal overlap computed
independently of the reference's Bessel-ratio formulas: the periodic factor
by the trapezoid rule on the circle (spectrally exact for these smooth
periodic integrands), the real factors by the Gaussian moments
(2m-1)!!/(2p)^m sqrt(pi/p).
"""
import math

import numpy as np


def _dfact(n):
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def self_overlap_matrix(sysm, ao):
    """S_ij = <g_i|g_j> for the primitives i, j of AO ``ao`` (same centre)."""
    lx = sysm.lxyz[ao]
    ex = sysm.expo[sysm.prim_off[ao]:sysm.prim_off[ao + 1]]
    n = len(ex)
    S = np.ones((n, n))
    for d in range(3):
        if d < sysm.ndp:
            L = sysm.L[d]
            a = 2 * math.pi / L
            emax = 2 * ex.max()
            npts = int(max(512, 40 * L * math.sqrt(emax)))
            u = a * (np.arange(npts) * L / npts)           # centre at 0 (translation invariant)
            base = (np.sin(u) / a) ** (2 * lx[d])
            for i in range(n):
                for j in range(n):
                    f = base * np.exp(-2 * (ex[i] + ex[j]) / a ** 2 * (1 - np.cos(u)))
                    S[i, j] *= f.sum() * L / npts
        else:
            m = lx[d]
            for i in range(n):
                for j in range(n):
                    p = ex[i] + ex[j]
                    S[i, j] *= _dfact(2 * m - 1) / (2 * p) ** m * math.sqrt(math.pi / p)
    return S


def normalised_coef(sysm):
    coef = sysm.coef.copy()
    for ao in range(sysm.nbas):
        o0, o1 = sysm.prim_off[ao], sysm.prim_off[ao + 1]
        S = self_overlap_matrix(sysm, ao)
        c = coef[o0:o1] / np.sqrt(np.diag(S))
        c = c / math.sqrt(c @ S @ c)
        coef[o0:o1] = c
    return coef
