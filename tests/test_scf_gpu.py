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


def test_overlap_diagonal_and_symmetry():
    s = tsys.workload("h10_ring")
    eng = _engine(s)
    S, T, V = eng.one_electron()
    eng.close()
    assert np.abs(np.diag(S) - 1.0).max() < 1e-12
    for M in (S, T, V):
        assert np.abs(M - M.T).max() < 1e-12


def test_rhf_h10_matches_gamma_oracle():
    s = tsys.workload("h10_ring")
    coef = normalize(s)
    eng = _engine(s)
    S, T, V = eng.one_electron()
    eng.eri_build(to_host=False)
    res = scf.rhf_sao(s, S, T, V, lambda P: eng.jk_stored(P))
    eng.close()
    assert res.converged
    o = orc.OracleSystem(s, coef)
    E, _ = o.eri_full(sao=False)
    ch = E.transpose(3, 2, 1, 0).copy()
    Eg, e, C, P, it = rhf.rhf_gamma(ch, S, T, V, s.n_occ, scf.nuclear_repulsion(s))
    print(f"E_RHF engine {res.energy:.12f}  oracle {Eg:.12f}  diff {res.energy - Eg:.2e}", flush=True)
    assert abs(res.energy - Eg) < 1e-8
