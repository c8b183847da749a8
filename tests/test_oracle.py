"""Pins the oracle (oracle/tcg_oracle.c) before anything is compared with it.

Checks, per restated function:
  * iv_scaled, Bessel vectors  -> scipy.special.ive (textbook), and the reference
  * fast_quad nodes/weights    -> numpy Gauss-Legendre mapped by t = 0.6u/(1-u)
  * barycentre                 -> the phasor identity e1 e^{iaA} + e2 e^{iaB} = p e^{iaP}
  * shell-quartet kernel       -> a->0 limit (ordinary Gaussian integrals, Richardson in 1/L^2),
                                  brute-force periodic quadrature of tiny quartets,
                                  the reference's own kernel (oracle/_ref) where built
  * ERI driver                 -> 8-fold symmetry, translation invariance,
                                  the reference's own driver (oracle/_ref) where built
  * J/K                        -> einsum restatement of J = (mn|ls)P_ls, K = -1/2 (ml|sn)P_ls
"""
import dataclasses
import math

import numpy as np
import pytest
from scipy import integrate, special

from conftest import needs_ref
from tests_support import normalised_coef
from oracle import orc
from william_ci_amd import system as tsys


def test_iv_scaled_vs_scipy():
    worst = 0.0
    for v in range(0, 13):
        for x in np.geomspace(1e-8, 1e5, 400):
            b = special.ive(v, x)
            if b > 1e-250:
                worst = max(worst, abs(orc.iv_scaled(v, x) - b) / b)
    assert worst < 1e-12


@pytest.mark.parametrize("A", [1e-3, 0.7, 5.0, 31.0, 250.0, 3000.0])
@pytest.mark.parametrize("fast", [False, True])
def test_bessel_vectors_vs_scipy(A, fast):
    N = orc.nmax_est(A)
    v = orc.bessel_vector(N, A, fast=fast)
    ref = special.ive(np.arange(N + 1), A)
    # the fast variant starts its recurrence at N itself, so only orders well
    # above the truncation tail carry full relative accuracy
    big = ref > (1e-8 if fast else 1e-250) * ref[0]
    rel = np.abs(v[big] - ref[big]) / ref[big]
    assert rel.max() < 1e-11, rel.max()
    assert np.abs(v - ref).max() < 1e-14 * ref[0]


def test_nmax_estimate_is_below_1e17():
    for A in [0.01, 1.0, 10.0, 100.0, 1e4]:
        N = orc.nmax_est(A)
        assert special.ive(N, A) < 1e-16 * special.ive(0, A)


@pytest.mark.parametrize("n", [32, 40])
def test_fast_quad_points_vs_legendre(n):
    t, w = orc.fast_quad_points(n)
    u, wu = np.polynomial.legendre.leggauss(n)
    u = (u + 1) / 2
    wu = wu / 2
    assert np.allclose(t, 0.6 * u / (1 - u), rtol=1e-13, atol=0)
    assert np.allclose(w, wu * 0.6 / (1 - u) ** 2, rtol=1e-12, atol=0)
    # the rule integrates 1/(1+t^2) on [0, inf) to pi/2
    assert abs(np.sum(w / (1 + t * t)) - math.pi / 2) < 1e-9


def _mini_system(L, centres, lxyz, expos, coefs=None, ndp=1):
    n = len(centres)
    off = np.arange(n + 1, dtype=np.int32)
    return tsys.TorusSystem(kind="Tori1D", ndp=ndp, L=tuple(L), ncell=(1, 1, 1), spacing=0.0, keywords=[],
                            geom=np.asarray(centres, float), Z=np.ones(n, np.int32), nauc=n,
                            center=np.asarray(centres, float), lxyz=np.asarray(lxyz, np.int32), prim_off=off,
                            expo=np.asarray(expos, float),
                            coef=np.ones(n) if coefs is None else np.asarray(coefs, float))


def test_barycentre_phasor_identity():
    s = _mini_system((18.0, 0, 0), [[0, 0, 0]], [[0, 0, 0]], [1.0])
    o = orc.OracleSystem(s, s.coef)
    a = 2 * math.pi / 18.0
    rng = np.random.default_rng(3)
    for _ in range(200):
        e1, e2 = rng.uniform(0.05, 20, 2)
        x1, x2 = rng.uniform(0, 18.0, 2)
        if abs(abs(x1 - x2) - 9.0) < 1e-3:
            continue
        p, P = o.bary(1, e1, e2, x1, x2)
        z = e1 * np.exp(1j * a * x1) + e2 * np.exp(1j * a * x2)
        assert abs(p - abs(z)) < 1e-12 * (e1 + e2)
        assert abs(np.exp(1j * a * P) - z / abs(z)) < 1e-10
    # antipodal branch: p = |e1-e2|, centre = the tighter primitive's centre
    p, P = o.bary(1, 3.0, 1.0, 1.0, 10.0)
    assert p == 2.0 and P == 1.0


def _gauss_ssss(al, A, be, B, ga, C, de, D):
    """Ordinary (unnormalised) Gaussian (ss|ss)."""
    A, B, C, D = map(np.asarray, (A, B, C, D))
    p, q = al + be, ga + de
    P = (al * A + be * B) / p
    Q = (ga * C + de * D) / q
    T = p * q / (p + q) * np.sum((P - Q) ** 2)
    F0 = 1.0 if T < 1e-14 else 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
    return (2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * math.exp(-al * be / p * np.sum((A - B) ** 2))
            * math.exp(-ga * de / q * np.sum((C - D) ** 2)) * F0)


def _torus_block(L, cen, lx, ex):
    s = _mini_system((L, 0, 0), cen, lx, ex)
    o = orc.OracleSystem(s, s.coef)
    return o.eri_block(0, 1, 1, 1, 2, 1, 3, 1)[0, 0, 0, 0]


@pytest.mark.slow
def test_kernel_large_torus_limit():
    """a -> 0: the Clifford Gaussians become ordinary Gaussians; the error is O(a^2),
    removed by Richardson extrapolation from L and 2L."""
    cen = [[0.3, 0.1, -0.2], [1.1, -0.4, 0.5], [-0.6, 0.8, 0.0], [0.2, 0.0, 1.2]]
    ex = [1.3, 0.7, 0.9, 1.8]
    s_lx = [[0, 0, 0]] * 4
    v1 = _torus_block(400.0, cen, s_lx, ex)
    v2 = _torus_block(800.0, cen, s_lx, ex)
    ext = (4 * v2 - v1) / 3
    ref = _gauss_ssss(ex[0], cen[0], ex[1], cen[1], ex[2], cen[2], ex[3], cen[3])
    assert abs(ext - ref) < 2e-8 * abs(ref), (v1, v2, ext, ref)
    # p_x (periodic) and p_y (real) on the first centre: (1/2alpha) d/dA of (ss|ss)
    for comp, ax in (([1, 0, 0], 0), ([0, 1, 0], 1)):
        lx = [comp] + s_lx[1:]
        w1 = _torus_block(400.0, cen, lx, ex)
        w2 = _torus_block(800.0, cen, lx, ex)
        extp = (4 * w2 - w1) / 3
        h = 1e-4
        cp = [list(c) for c in cen]
        cm = [list(c) for c in cen]
        cp[0][ax] += h
        cm[0][ax] -= h
        d = (_gauss_ssss(ex[0], cp[0], ex[1], cen[1], ex[2], cen[2], ex[3], cen[3]) -
             _gauss_ssss(ex[0], cm[0], ex[1], cen[1], ex[2], cen[2], ex[3], cen[3])) / (2 * h)
        refp = d / (2 * ex[0])
        assert abs(extp - refp) < 1e-7 * abs(ref), (comp, extp, refp)


def _brute_periodic(L, cen, lx, ex, nx=256):
    """2/sqrt(pi) Int dt Int Int dx1 dx2 rho_ab(x1) rho_cd(x2) exp(-t^2 chord^2) * Iy * Iz,
    x by the trapezoid rule on the circle (spectral), t and the lag s by adaptive quadrature."""
    a = 2 * math.pi / L
    x = np.arange(nx) * L / nx

    def g(i, xx):
        u = a * (xx - cen[i][0])
        return (np.sin(u) / a) ** lx[i][0] * np.exp(-2 * ex[i] / a ** 2 * (1 - np.cos(u)))

    def corr(sv):
        rab = g(0, x) * g(1, x)
        rcd = g(2, x + sv) * g(3, x + sv)
        return np.sum(rab * rcd) * L / nx

    p, q = ex[0] + ex[1], ex[2] + ex[3]

    def real_dir(t, d):
        A, B, C, D = (cen[i][d] for i in range(4))
        P = (ex[0] * A + ex[1] * B) / p
        Q = (ex[2] * C + ex[3] * D) / q
        den = p * q + (p + q) * t * t
        return (math.pi / math.sqrt(den) * math.exp(-p * q * t * t / den * (P - Q) ** 2) *
                math.exp(-ex[0] * ex[1] / p * (A - B) ** 2) * math.exp(-ex[2] * ex[3] / q * (C - D) ** 2))

    def per_dir(t):
        f = lambda sv: corr(sv) * math.exp(-t * t * 2 * (1 - math.cos(a * sv)) / a ** 2)
        # the kernel is periodic and peaked at s = 0 (and L)
        r1, _ = integrate.quad(f, 0, L / 2, limit=400, epsabs=0, epsrel=1e-12, points=[0.0])
        r2, _ = integrate.quad(f, L / 2, L, limit=400, epsabs=0, epsrel=1e-12, points=[L])
        return r1 + r2

    val, _ = integrate.quad(lambda t: per_dir(t) * real_dir(t, 1) * real_dir(t, 2), 0, np.inf,
                            limit=400, epsabs=0, epsrel=1e-11)
    return 2 / math.sqrt(math.pi) * val


@pytest.mark.slow
@pytest.mark.parametrize("lx0", [[0, 0, 0], [1, 0, 0]])
def test_kernel_vs_bruteforce_periodic(lx0):
    L = 6.0
    cen = [[0.3, 0.1, -0.2], [1.9, -0.4, 0.5], [4.6, 0.8, 0.0], [5.5, 0.0, 0.3]]
    ex = [1.3, 0.7, 0.9, 1.8]
    lx = [lx0, [0, 0, 0], [0, 0, 0], [0, 0, 0]]
    s = _mini_system((L, 0, 0), cen, lx, ex)
    o = orc.OracleSystem(s, s.coef)
    v = o.eri_block(0, 1, 1, 1, 2, 1, 3, 1)[0, 0, 0, 0]
    ref = _brute_periodic(L, cen, lx, ex)
    assert abs(v - ref) < 1e-9 * max(abs(ref), 1e-3), (v, ref)


@pytest.fixture(scope="module")
def h10():
    s = tsys.workload("h10_ring")
    return s, normalised_coef(s)


def test_eri_full_symmetries(h10):
    s, coef = h10
    o = orc.OracleSystem(s, coef)
    E, st = o.eri_full(sao=False)
    ch = E.transpose(3, 2, 1, 0)           # ch[i,j,k,l] = (ij|kl)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.abs(ch - ch.transpose(perm)).max() < 1e-14
    nf = s.nfuc
    sh = np.roll(np.arange(s.nbas), -3 * nf)  # translate every index by 3 cells
    assert np.abs(ch - ch[np.ix_(sh, sh, sh, sh)]).max() < 1e-14
    diag = np.einsum("ijij->ij", ch)
    assert diag.min() > -1e-15
    # (the Schwarz inequality is NOT an invariant here: the chord-distance kernel
    #  1/|r|_E on the torus is not positive definite, and it is violated by ~1e-7)
    # SAO layout holds the first nfuc rows of the same tensor
    Es, _ = o.eri_full(sao=True)
    assert np.array_equal(Es, E[:, :, :, :nf])


def test_jk_vs_einsum(h10):
    s, coef = h10
    o = orc.OracleSystem(s, coef)
    E, _ = o.eri_full(sao=False)
    ch = E.transpose(3, 2, 1, 0)
    rng = np.random.default_rng(0)
    P = rng.standard_normal((s.nbas, s.nbas))
    P = P + P.T
    J, K = o.jk(P, E)
    assert np.allclose(J, np.einsum("mnls,ls->mn", ch, P), atol=1e-12)
    assert np.allclose(K, -0.5 * np.einsum("mlsn,ls->mn", ch, P), atol=1e-12)
    Jp, Kp = o.jk(P, np.ascontiguousarray(ch.transpose(3, 1, 2, 0)), physicist=True)
    assert np.allclose(Jp, J, atol=1e-12) and np.allclose(Kp, K, atol=1e-12)


@needs_ref
def test_normaliser_matches_reference(h10):
    from oracle.ref import RefSystem
    s, coef = h10
    assert np.abs(RefSystem(s).coef - coef).max() < 1e-13


@needs_ref
def test_oracle_matches_reference_h10(h10):
    from oracle.ref import RefSystem
    s, _ = h10
    r = RefSystem(s)
    o = orc.OracleSystem(s, r.coef)
    for blk in [(0, 1, 5, 1, 7, 3, 2, 3), (2, 3, 2, 3, 2, 3, 2, 3), (0, 1, 45, 1, 30, 1, 17, 3)]:
        assert np.abs(r.eri_block(*blk) - o.eri_block(*blk)).max() < 1e-15
    E1 = r.eri_full()
    E2, _ = o.eri_full()
    assert np.abs(E1 - E2).max() < 1e-15


@needs_ref
@pytest.mark.slow
@pytest.mark.parametrize("case", ["1d-d", "2d", "3d-d"])
def test_oracle_matches_reference_2d3d(case):
    from oracle.ref import RefSystem
    s = {"1d-d": lambda: tsys.chain_1d(6, 1.8, 10.8, basis=tsys.H_BASIS_D),
         "2d": lambda: tsys.lattice_2d_h2(3, 3.6, 10.8),
         "3d-d": lambda: tsys.lattice_3d_h2(2, 4.0, 8.0)}[case]()
    r = RefSystem(s)
    o = orc.OracleSystem(s, r.coef)
    nf = s.nfuc
    shells = [(0, 1), (1, 1), (2, 3)] + ([(5, 6)] if nf // s.nauc > 5 else [])
    for (a, na) in shells:
        for (b, nb) in shells:
            for (c, nc) in shells[:2]:
                for (d, nd) in shells:
                    blk = (a, na, b + nf, nb, c, nc, d + nf, nd)
                    assert np.abs(r.eri_block(*blk) - o.eri_block(*blk)).max() < 1e-15
    if case != "3d-d":
        E1 = r.eri_full()
        E2, _ = o.eri_full()
        assert np.abs(E1 - E2).max() < 1e-14
