"""GPU parity of the ERI build (tcg_eri_build) against the oracle.

Bar: |ERI_gpu - ERI_oracle| <= 1e-10 per integral (FP64 quadrature; the
oracle matches the reference bit for bit on these inputs), the same set of
stored (nonzero) elements up to Schwarz-threshold flips of negligible
integrals, and the same quartet counts as the reference's driver.
"""
import numpy as np
import pytest

from oracle import orc
from tests_support import normalised_coef
from william_ci_amd import system as tsys

pytestmark = pytest.mark.gpu

CASES = {
    "h10_ring": lambda: tsys.workload("h10_ring"),
    "h2_one_cell": lambda: tsys.chain_1d(1, 10.0, 10.0, basis=tsys.H_BASIS),
    "h6_d_1d": lambda: tsys.chain_1d(6, 1.8, 10.8, basis=tsys.H_BASIS_D),
    "h2_3x3_2d": lambda: tsys.lattice_2d_h2(3, 3.6, 10.8),
    "h2_2x2x2_d_3d": lambda: tsys.lattice_3d_h2(2, 4.0, 8.0),
}


def _compare(s, tol=1e-10):
    from william_ci_amd.engine import Engine
    coef = normalised_coef(s)
    print(f"case nbas={s.nbas} ndp={s.ndp} cells={s.ncell}", flush=True)
    eng = Engine(s, coef)
    Eg = eng.eri_build()
    st_g = eng.stats()
    eng.close()
    Eo, st_o = orc.OracleSystem(s, coef).eri_full()
    diff = np.abs(Eg - Eo)
    assert diff.max() <= tol, diff.max()
    flips = (Eg != 0) != (Eo != 0)
    # elements stored by one side only: Schwarz-threshold flips and symmetry zeros that
    # one side computes as round-off (1e-35 in 3D) -- all negligible
    assert np.abs(Eo[flips]).max(initial=0.0) < 1e-12 and np.abs(Eg[flips]).max(initial=0.0) < 1e-12
    print(f"max |diff| {diff.max():.3e}, flips {int(flips.sum())}, stored {int((Eo != 0).sum())}", flush=True)
    assert st_g["shell_quartets"] == st_o[0]
    assert abs(st_g["prim_quartets"] - st_o[1]) <= 1e-3 * st_o[1]
    return Eg, Eo


@pytest.mark.parametrize("case", list(CASES))
def test_eri_matches_oracle(case):
    _compare(CASES[case]())


def test_eri_h64_chain_matches_oracle():
    _compare(tsys.workload("h64_chain"))


def test_jk_stored_matches_oracle():
    from william_ci_amd.engine import Engine
    s = tsys.workload("h10_ring")
    coef = normalised_coef(s)
    eng = Engine(s, coef)
    E = eng.eri_build()
    rng = np.random.default_rng(1)
    P = rng.standard_normal((s.nbas, s.nbas))
    P = P + P.T
    J, K = eng.jk_stored(P)
    Jo, Ko = orc.OracleSystem(s, coef).jk(P, E)
    assert np.abs(J - Jo).max() < 1e-11 * max(1.0, np.abs(Jo).max())
    assert np.abs(K - Ko).max() < 1e-11 * max(1.0, np.abs(Ko).max())
    # host-supplied ERI path gives the same
    J2, K2 = eng.jk_stored(P, eri=E)
    assert np.abs(J2 - J).max() < 1e-13 and np.abs(K2 - K).max() < 1e-13
    eng.close()


@pytest.mark.parametrize("name", ["h10_ring", "h2_3x3_2d"])
def test_chunked_pipeline_matches(name, monkeypatch):
    """The quartet list split into many row chunks (a tiny row budget) gives
    the same ERIs as one pass, and the direct J/K too."""
    from william_ci_amd.engine import Engine
    s = CASES[name]()
    coef = normalised_coef(s)
    print(f"case {name} one pass vs chunked", flush=True)
    eng = Engine(s, coef)
    E1 = eng.eri_build()
    rows = eng.stats()["rows"]
    eng.close()
    monkeypatch.setenv("TCG_ROW_BUDGET", str(max(2e5, rows * 1200.0 / 7)))   # about 7 chunks
    eng = Engine(s, coef)
    E2 = eng.eri_build()
    P = np.eye(s.nbas)
    eng.direct_init()
    J2, K2 = eng.direct_jk(P)
    J1, K1 = eng.jk_stored(P)
    eng.close()
    print(f"  rows {rows}  max |diff| {np.abs(E1 - E2).max():.2e}")
    assert np.abs(E1 - E2).max() <= 1e-14
    assert np.abs(J1 - J2).max() < 1e-12 and np.abs(K1 - K2).max() < 1e-12


@pytest.mark.parametrize("sao", [True, False])
def test_jk_stored_physicist_order(sao):
    """J/K from a physicist-ordered tensor (hartree_exchange_potential's
    physicist branch) equal the chemist ones."""
    from william_ci_amd.engine import Engine
    s = tsys.workload("h10_ring")
    coef = normalised_coef(s)
    print(f"case h10_ring physicist J/K sao={sao}", flush=True)
    n1 = s.nfuc if sao else s.nbas
    rng = np.random.default_rng(4)
    P = rng.standard_normal((s.nbas, s.nbas))
    P = P + P.T
    eng = Engine(s, coef)
    eng.eri_build(sao=sao, physicist=False, to_host=False)
    Jc, Kc = eng.jk_stored(P, n1=n1)
    Ep = eng.eri_build(sao=sao, physicist=True)
    Jp, Kp = eng.jk_stored(P, n1=n1)
    Jh, Kh = eng.jk_stored(P, n1=n1, eri=Ep, physicist=True)
    eng.close()
    Jo, Ko = orc.OracleSystem(s, coef).jk(P, Ep, physicist=True)
    for J, K in ((Jp, Kp), (Jh, Kh), (Jo, Ko)):
        assert np.abs(J - Jc).max() < 1e-12 * np.abs(Jc).max()
        assert np.abs(K - Kc).max() < 1e-12 * np.abs(Kc).max()


@pytest.mark.parametrize("name", ["h10_ring", "h6_d_1d", "h2_3x3_2d", "h2_2x2x2_d_3d"])
def test_sp_contraction_matches_general(name, monkeypatch):
    """The lanes-over-nodes contraction of the s/p quartets (one to three p
    shells, uniform p coefficients) gives the integrals of the general
    contraction: the same sums in another order, so equal to round-off."""
    from william_ci_amd.engine import Engine
    s = CASES[name]()
    coef = normalised_coef(s)
    print(f"case {name} s/p form vs general", flush=True)
    monkeypatch.setenv("TCG_SP_MIN_QUARTETS", "0")       # the s/p kernels also on these short lists
    eng = Engine(s, coef)
    E1 = eng.eri_build()
    eng.close()
    monkeypatch.setenv("TCG_GENERAL_CONTRACT", "1")
    eng = Engine(s, coef)
    E2 = eng.eri_build()
    eng.close()
    d = np.abs(E1 - E2).max()
    print(f"  max |diff| {d:.2e}", flush=True)
    assert d <= 1e-13 * max(1.0, np.abs(E2).max())
