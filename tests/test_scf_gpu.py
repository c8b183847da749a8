"""GPU parity of the one-electron integrals and of the RHF that consumes the
engine's integrals and Coulomb/exchange.

  * S, T, V (tcg_one_electron) against the reference's own routines
    (oracle/_ref) to 1e-10, and S against an independent quadrature.
  * E_RHF from the engine (Bloch-basis SCF, GPU J/K) against the oracle's
    Gamma-supercell RHF on the oracle's ERIs: |dE| <= 1e-8 Eh.
"""
import numpy as np
import pytest

from conftest import needs_ref
from oracle import orc, rhf
from william_ci_amd import scf
from william_ci_amd import system as tsys
from william_ci_amd.basis import normalize

pytestmark = pytest.mark.gpu

SYSTEMS = {
    "h10_ring": lambda: tsys.workload("h10_ring"),
    "h6_d_1d": lambda: tsys.chain_1d(6, 1.8, 10.8, basis=tsys.H_BASIS_D),
    "h2_3x3_2d": lambda: tsys.lattice_2d_h2(3, 3.6, 10.8),
    "h2_2x2x2_d_3d": lambda: tsys.lattice_3d_h2(2, 4.0, 8.0),
}


def _engine(s):
    from william_ci_amd.engine import Engine
    return Engine(s, normalize(s))


@needs_ref
@pytest.mark.parametrize("case", list(SYSTEMS))
def test_one_electron_matches_reference(case):
    from oracle.ref import RefSystem
    s = SYSTEMS[case]()
    print(f"case {case} nbas={s.nbas}", flush=True)
    eng = _engine(s)
    S, T, V = eng.one_electron()
    eng.close()
    Sr, Tr, Vr = RefSystem(s).one_electron()
    for name, A, B in (("S", S, Sr), ("T", T, Tr), ("V", V, Vr)):
        d = np.abs(A - B).max()
        print(f"  {name}: max |diff| {d:.2e} (max {np.abs(B).max():.3f})", flush=True)
        assert d < 1e-10, (name, d)


@needs_ref
@pytest.mark.parametrize("case", list(SYSTEMS))
def test_normalization_matches_reference(case):
    """tcg_normalize (closed-form self-overlaps on the GPU) against the
    reference's normalize_AO_toroidal_nD and the independent quadrature of
    tests_support."""
    from oracle.ref import RefSystem
    from tests_support import normalised_coef
    from william_ci_amd.engine import normalized_coef
    s = SYSTEMS[case]()
    print(f"case {case} nbas={s.nbas}", flush=True)
    c = normalized_coef(s)
    cr = RefSystem(s).coef
    cq = normalised_coef(s)
    d = np.abs(c - cr).max() / np.abs(cr).max()
    print(f"  vs reference {d:.2e}  vs quadrature {np.abs(c - cq).max() / np.abs(cq).max():.2e}", flush=True)
    assert d < 1e-12
    assert np.abs(c - cq).max() < 1e-10 * np.abs(cq).max()


def test_overlap_diagonal_and_symmetry():
    s = tsys.workload("h10_ring")
    eng = _engine(s)
    S, T, V = eng.one_electron()
    eng.close()
    assert np.abs(np.diag(S) - 1.0).max() < 1e-12
    for M in (S, T, V):
        assert np.abs(M - M.T).max() < 1e-12


@needs_ref
def test_rhf_h10_matches_gamma_oracle():
    """The engine's whole path (its S, T, V, ERI, GPU J/K, k-space SCF and
    E_nuc) against the pinned Gamma-supercell oracle fed only with the
    reference's own S, T, V, full ERI and NRE."""
    from oracle.ref import RefSystem
    s = tsys.workload("h10_ring")
    coef = normalize(s)
    eng = _engine(s)
    S, T, V = eng.one_electron()
    eng.eri_build(to_host=False)
    res = scf.rhf_sao(s, S, T, V, lambda P: eng.jk_stored(P))
    eng.close()
    assert res.converged
    r = RefSystem(s, sao=False)
    Sr, Tr, Vr = r.one_electron()
    E = r.eri_full()
    enuc = r.nre()
    assert abs(res.e_nuc - enuc) < 1e-12
    ch = E.transpose(3, 2, 1, 0).copy()
    Eg, e, C, P, it = rhf.rhf_gamma(ch, Sr, Tr, Vr, s.n_occ, enuc)
    print(f"E_RHF engine {res.energy:.12f}  oracle {Eg:.12f}  diff {res.energy - Eg:.2e}", flush=True)
    assert abs(res.energy - Eg) < 1e-8


@needs_ref
@pytest.mark.parametrize("physicist", [False, True])
@pytest.mark.parametrize("sao", [True, False])
def test_jk_stored_matches_reference(sao, physicist):
    """tcg_jk_stored_order on a host ERI and on the resident tensor against
    the reference's hartree_exchange_potential (oracle/_ref)."""
    from oracle.ref import RefSystem, lib as rlib
    s = tsys.workload("h10_ring")
    nb, nf = s.nbas, s.nfuc
    n1 = nf if sao else nb
    print(f"case h10_ring J/K sao={sao} physicist={physicist}", flush=True)
    r = RefSystem(s, sao=sao, physicist=physicist)
    E = r.eri_full()
    rng = np.random.default_rng(11)
    P = rng.standard_normal((nb, nb))
    P = P + P.T
    Pf = np.ascontiguousarray(P.T).ravel()
    Jr = np.zeros(nb * nb)
    Kr = np.zeros(nb * nb)
    rlib().ref_jk(n1, nb, Pf, np.ascontiguousarray(E).ravel(), Jr, Kr)
    Jr, Kr = Jr.reshape(nb, nb).T, Kr.reshape(nb, nb).T
    from william_ci_amd.engine import Engine
    eng = Engine(s, r.coef)
    J1, K1 = eng.jk_stored(P, n1=n1, eri=E, physicist=physicist)
    # the same logical array held in Fortran order (a bound copy is passed)
    J3, K3 = eng.jk_stored(P, n1=n1, eri=np.asfortranarray(E), physicist=physicist)
    # (same values; the host-array J/K adds with atomics, so the last bits may differ)
    assert np.abs(J3 - J1).max() < 1e-13 * np.abs(J1).max() and np.abs(K3 - K1).max() < 1e-13 * np.abs(J1).max()
    with pytest.raises(Exception, match="expected n1"):
        eng.jk_stored(P, n1=n1, eri=E.ravel()[:-1], physicist=physicist)
    eng.eri_build(sao=sao, physicist=physicist, to_host=False)
    J2, K2 = eng.jk_stored(P, n1=n1)
    eng.close()
    scale = np.abs(Jr).max()
    for name, A, B in (("J host", J1, Jr), ("K host", K1, Kr), ("J resident", J2, Jr), ("K resident", K2, Kr)):
        d = np.abs(A - B).max()
        print(f"  {name}: max |diff| {d:.2e}", flush=True)
        assert d < 1e-11 * scale, name
