"""Density fitting on the CPU: the auxiliary sets and the numpy restatement
(oracle/df.py) of src/module/df_toroidal.f90.

The reference's DF needs LAPACK (zpstrf, zheev, zgemm), which this image has
for Fortran only inside scipy, so the reference DF is not built; the
restatement is held instead by
  * the even-tempered rule worked out by hand for the H basis (etb_atom);
  * chol_f = LAPACK zpstrf (scipy's binding) checked against the defining
    properties: P^T V P = L L^H, f^H V f = 1 on the kept set, pivots in
    decreasing order, rank and stop rule on a rank-deficient metric;
  * the k-space route of df_init (V(k), f(k), convolution over cells) equal
    to the real-space fit T3 V^-1 T3 of the whole supercell metric;
  * df_kMO's L route equal to transform_ERI_AO_to_kMO of the fitted integrals;
  * df_JK's J, K equal to the stored-integral J/K of the fitted integrals;
  * the Coulomb-metric fit never exceeding the exact (mu nu|mu nu);
  * the exponent-0 dummy integrals of the restatement equal to the
    reference's ERI_shell_block_toroidal (oracle/_ref), one block per forked
    process.
"""
import math
import os

import numpy as np
import pytest

from conftest import needs_ref
from tests_support import normalised_coef
from william_ci_amd import df as wdf
from william_ci_amd import system as tsys


def _ring4():
    return tsys.chain_1d(4, 1.8, 7.2)


@pytest.fixture(scope="module")
def ring4_fit():
    from oracle import df as odf
    from oracle import orc
    s = _ring4()
    coef = normalised_coef(s)
    aux = wdf.aux_list(s, wdf.etb_aux(tsys.read_basis(tsys.H_BASIS)))
    cs, ccoef, (nf, nau) = odf.combined(s, coef, aux)
    o = orc.OracleSystem(cs, ccoef, nq=32)
    Vm, T3 = odf.integrals(o, cs, nf, nau)
    B, ranks, dmin = odf.df_B(Vm, T3, s.ncell)
    return dict(s=s, coef=coef, aux=aux, cs=cs, ccoef=ccoef, nf=nf, nau=nau, Vm=Vm, T3=T3, B=B, ranks=ranks)


def test_etb_h_basis():
    """etb_atom for H (s 13.01..0.122, p 0.727; l_max 1): L=0 from 2 sqrt(e e)
    = 0.244..26.02 -> n = ceil(log2(26.264/0.244)) = 7; L=1 from
    2 sqrt(0.122 0.727)..2 sqrt(13.01 0.727) -> 4; L=2 from 2 x 0.727 -> 1."""
    aux = wdf.etb_aux(tsys.read_basis(tsys.H_BASIS))[1].shells
    assert sorted(aux) == [0, 1, 2]
    np.testing.assert_allclose(aux[0][0], 0.244 * 2.0 ** np.arange(6, -1, -1), rtol=1e-15)
    emin1 = 2 * math.sqrt(0.122 * 0.727)
    np.testing.assert_allclose(aux[1][0], emin1 * 2.0 ** np.arange(3, -1, -1), rtol=1e-15)
    np.testing.assert_allclose(aux[2][0], [1.454], rtol=1e-15)
    for l in aux:
        assert np.array_equal(aux[l][1], np.eye(len(aux[l][0])))
    # 7 s + 4 x 3 p + 1 x 6 d
    s = _ring4()
    assert wdf.aux_list(s, wdf.etb_aux(tsys.read_basis(tsys.H_BASIS))).nbas == 25 * 4


def test_aux_file_h_etb41():
    """The committed AuxBasis of the DF-MP2 workload: s 8 over 0.1-40, p 5 over
    0.2-10, d 3 over 0.3-3, geometric, uncontracted; 41 functions per H."""
    aux = wdf.read_aux(wdf.AUX_H_ETB41)[1].shells
    for l, (n, lo, hi) in {0: (8, 0.1, 40.0), 1: (5, 0.2, 10.0), 2: (3, 0.3, 3.0)}.items():
        ex, co = aux[l]
        assert len(ex) == n and np.array_equal(co, np.eye(n))
        np.testing.assert_allclose([ex[0], ex[-1]], [hi, lo], rtol=1e-10)
        r = ex[:-1] / ex[1:]
        np.testing.assert_allclose(r, r[0], rtol=1e-9)
    s = tsys.workload("h64_chain")
    assert wdf.aux_list(s, wdf.read_aux(wdf.AUX_H_ETB41)).nbas == 41 * 64


def test_chol_f_properties():
    from oracle import df as odf
    rng = np.random.default_rng(5)
    n = 9
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    V = A @ A.conj().T + 0.1 * np.eye(n)
    f, m, dmin = odf.chol_f(V)
    assert m == n
    np.testing.assert_allclose(f.conj().T @ V @ f, np.eye(n), atol=1e-12)
    np.testing.assert_allclose(f @ f.conj().T, np.linalg.inv(V), atol=1e-12)
    # rank 5 metric: 5 kept, f f^H is the inverse on the range
    A5 = A[:, :5]
    V5 = A5 @ A5.conj().T
    f5, m5, _ = odf.chol_f(V5)
    assert m5 == 5
    np.testing.assert_allclose(f5.conj().T @ V5 @ f5, np.diag([1.0] * 5 + [0.0] * 4), atol=1e-10)
    np.testing.assert_allclose(V5 @ f5 @ f5.conj().T @ V5, V5, atol=1e-9)
    # first pivot = the largest diagonal; the stop rule compares the Schur diagonal with tol
    from scipy.linalg import lapack
    c, piv, r, info = lapack.zpstrf(np.array(V, order="F"), tol=1e-10, lower=1)
    assert piv[0] - 1 == int(np.argmax(V.diagonal().real))
    d = np.abs(np.diag(c)[:r]) ** 2
    assert np.all(np.diff(d) <= 1e-12 * d[0])
    assert abs(dmin - d[-1]) < 1e-12 * d[0]
    _, m_hi, _ = odf.chol_f(V, tol=d[3] * 1.0000001)
    assert m_hi == 3


def test_kspace_fit_equals_realspace_fit(ring4_fit):
    from oracle import df as odf
    F = ring4_fit
    assert np.all(F["ranks"] == F["nau"])
    E1 = odf.eri_sao(F["B"], F["s"].ncell)
    E2 = odf.realspace_eri(F["Vm"], F["T3"], F["s"].ncell)
    assert np.abs(E1 - E2).max() < 1e-12 * np.abs(E2).max()


def test_fit_below_exact_and_close(ring4_fit):
    """Coulomb-metric fitting: (mu nu|mu nu)_fit <= (mu nu|mu nu) for every
    product, and the fit is close to the exact integrals."""
    from oracle import df as odf
    from oracle import orc
    F = ring4_fit
    s = F["s"]
    E = orc.OracleSystem(s, F["coef"], nq=32).eri_full(sao=True)[0]
    Ef = odf.eri_sao(F["B"], s.ncell)
    nf, nb = s.nfuc, s.nbas
    diag_exact = np.array([[E[j, i, j, i] for j in range(nb)] for i in range(nf)])
    diag_fit = np.array([[Ef[j, i, j, i] for j in range(nb)] for i in range(nf)])
    assert np.all(diag_fit <= diag_exact + 1e-12)
    assert np.abs(Ef - E).max() < 5e-3


def test_jk_equals_stored_jk_of_fit(ring4_fit):
    from oracle import df as odf
    from oracle import orc
    F = ring4_fit
    s = F["s"]
    nf, nb, N = s.nfuc, s.nbas, s.n_cell
    rng = np.random.default_rng(3)
    # a translation-invariant symmetric density: P(mu0 + R, nu + R) = P(mu0, nu)
    P0 = rng.normal(size=(nf, nb))
    P = np.zeros((nb, nb))
    from oracle.df import cell_ops
    cdiff, cadd = cell_ops(s.ncell)
    for R in range(N):
        for nu in range(nb):
            P[R * nf:(R + 1) * nf, cadd[nu // nf, R] * nf + nu % nf] = P0[:, nu]
    P = 0.5 * (P + P.T)
    J, K = odf.jk(F["B"], P, s.ncell)
    Ef = odf.eri_sao(F["B"], s.ncell)
    o = orc.OracleSystem(s, F["coef"], nq=32)
    Jr, Kr = o.jk(P, Ef)
    assert np.abs(J[:nf] - Jr[:nf]).max() < 1e-12
    assert np.abs(K[:nf] - Kr[:nf]).max() < 1e-12


def test_kmo_equals_transform_of_fit(ring4_fit):
    from oracle import df as odf
    from oracle import kmp2 as okm
    F = ring4_fit
    s = F["s"]
    nf, N = s.nfuc, s.n_cell
    rng = np.random.default_rng(9)
    Ck = []
    for _ in range(N):
        q, _ = np.linalg.qr(rng.normal(size=(nf, nf)) + 1j * rng.normal(size=(nf, nf)))
        Ck.append(q)
    o_hi, v_lo = 2, 1
    got = odf.kmo(F["B"], Ck, s.ncell, o_hi, v_lo)
    want = okm.ao_to_kmo(odf.eri_sao(F["B"], s.ncell), Ck, s.ncell, o_hi, v_lo)
    assert np.abs(got - want).max() < 1e-12


@needs_ref
def test_dummy_blocks_match_reference(ring4_fit):
    """(P 1|Q 1) and (a b|P 1) of the restatement against the reference's
    shell-block routine, in this process (the same routines built with
    AddressSanitizer show no out-of-bounds access on these blocks)."""
    from oracle import df as odf
    from oracle import orc, ref
    F = ring4_fit
    cs, nf, nau = F["cs"], F["nf"], F["nau"]
    nfc = nf + nau + 1
    sh = odf.shells(cs)
    aux_sh = [x for x in sh if nf <= x[0] % nfc < nf + nau]
    ao_sh = [x for x in sh if x[0] % nfc < nf]
    dum = {x[0] // nfc: x[0] for x in sh if x[0] % nfc == nfc - 1}
    cases = []
    for (fp, n1) in aux_sh[:12:3]:
        for (fq, n2) in aux_sh[1::11]:
            cases.append((fp, n1, dum[fp // nfc], 1, fq, n2, dum[fq // nfc], 1))
    for (fa, na) in ao_sh[:3]:
        for (fb, nb_) in ao_sh[::4]:
            for (fp, n1) in aux_sh[2::13]:
                cases.append((fa, na, fb, nb_, fp, n1, dum[fp // nfc], 1))
    o = orc.OracleSystem(cs, F["ccoef"], nq=32)
    want = [o.eri_block(*c) for c in cases]
    r = ref.RefSystem(cs, nq=32, fast_quad=True, sao=True, normalize=False)
    for c, w in zip(cases, want):
        b = r.eri_block(*c)
        assert np.abs(b - w).max() <= 1e-13 * max(1.0, np.abs(w).max()), c


@needs_ref
def test_dummy_blocks_are_the_small_exponent_limit(ring4_fit):
    """(a b|P 1) is the limit of (a b|P g_eps) as the dummy's exponent eps -> 0
    (g_eps -> 1 uniformly on the torus): the reference's blocks with eps =
    1e-5, 1e-6, 1e-7 approach the exponent-0 block linearly in eps."""
    import dataclasses
    from oracle import df as odf
    from oracle import ref
    F = ring4_fit
    cs, nf, nau = F["cs"], F["nf"], F["nau"]
    nfc = nf + nau + 1
    sh = odf.shells(cs)
    aux_sh = [x for x in sh if nf <= x[0] % nfc < nf + nau]
    ao_sh = [x for x in sh if x[0] % nfc < nf]
    dum = {x[0] // nfc: x[0] for x in sh if x[0] % nfc == nfc - 1}
    cases = [(fa, na, fb, nb_, fp, n1, dum[fp // nfc], 1)
             for (fa, na) in ao_sh[:2] for (fb, nb_) in ao_sh[::5] for (fp, n1) in aux_sh[2::13]]

    def blocks(eps):
        expo = cs.expo.copy()
        for g in dum.values():
            expo[cs.prim_off[g]] = eps
        r = ref.RefSystem(dataclasses.replace(cs, expo=expo), nq=32, fast_quad=True, sao=True, normalize=False)
        return [r.eri_block(*c) for c in cases]

    b0 = blocks(0.0)
    diffs = []
    for eps in (1e-5, 1e-6, 1e-7):
        be = blocks(eps)
        diffs.append(max(float(np.abs(x - y).max()) for x, y in zip(be, b0)))
    scale = max(float(np.abs(x).max()) for x in b0)
    print("max |(ab|P g_eps) - (ab|P 1)|:", diffs, "scale", scale)
    assert diffs[2] < 1e-5 * scale
    assert 5.0 < diffs[0] / diffs[1] < 20.0 and 5.0 < diffs[1] / diffs[2] < 20.0
