"""Normalisation of the AO list (host-side input preparation).

Mirrors normalize_AO_toroidal_{1,2,3}D (src/basis/normalization/normalize_1D.f90:17-79):
each primitive is normalised to one, then the contraction.  The toroidal
self-overlap is evaluated exactly: the periodic factor
    Int_0^L (sin(a u)/a)^(2l) exp(-2(e_i+e_j)/a^2 (1 - cos(a u))) du
by the trapezoid rule on the circle (exact to rounding for these smooth
periodic integrands once the grid resolves them), the real factors by the
Gaussian moments (2m-1)!! / (2p)^m sqrt(pi/p).
"""
from __future__ import annotations

import math

import numpy as np


def _dfact(n: int) -> float:
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def _self_overlap(sysm, ao: int) -> np.ndarray:
    lx = sysm.lxyz[ao]
    ex = sysm.expo[sysm.prim_off[ao]:sysm.prim_off[ao + 1]]
    n = len(ex)
    S = np.ones((n, n))
    esum = ex[:, None] + ex[None, :]
    for d in range(3):
        if d < sysm.ndp:
            L = sysm.L[d]
            a = 2 * math.pi / L
            npts = int(max(512, 40 * L * math.sqrt(2 * ex.max())))
            u = a * (np.arange(npts) * (L / npts))
            base = (np.sin(u) / a) ** (2 * lx[d])
            one_m_cos = 1.0 - np.cos(u)
            for i in range(n):
                for j in range(n):
                    S[i, j] *= np.sum(base * np.exp(-2 * esum[i, j] / a ** 2 * one_m_cos)) * (L / npts)
        else:
            m = lx[d]
            S *= np.vectorize(lambda p: _dfact(2 * m - 1) / (2 * p) ** m * math.sqrt(math.pi / p))(esum)
    return S


def normalize(sysm) -> np.ndarray:
    """Normalised contraction coefficients of the AO list (same layout as sysm.coef)."""
    coef = np.array(sysm.coef, dtype=np.float64, copy=True)
    cache = {}
    for ao in range(sysm.nbas):
        o0, o1 = sysm.prim_off[ao], sysm.prim_off[ao + 1]
        key = (tuple(sysm.lxyz[ao]), tuple(sysm.expo[o0:o1]), tuple(sysm.coef[o0:o1]))
        if key not in cache:
            S = _self_overlap(sysm, ao)
            c = coef[o0:o1] / np.sqrt(np.diag(S))
            cache[key] = c / math.sqrt(c @ S @ c)
        coef[o0:o1] = cache[key]
    return coef
