"""k-point MP2 restatement (oracle/kmp2.py) against the reference's own
transform_ERI_AO_to_kMO + kMP2_calc, and the closed-shell identity
E_kMP2 = E_MP2 of the Gamma supercell on the same RHF solution."""
import numpy as np

from conftest import needs_ref


@needs_ref
def test_kmp2_oracle_matches_reference_and_gamma():
    from oracle import kmp2 as okm
    from oracle import orc, ref
    from oracle import rhf as orhf
    from william_ci_amd import scf
    from william_ci_amd import system as tsys
    s = tsys.workload("h10_ring")
    r = ref.RefSystem(s)
    S, T, V = r.one_electron()
    E = r.eri_full()
    o = orc.OracleSystem(s, r.coef)
    res = scf.rhf_sao(s, S, T, V, lambda P: o.jk(P, E))
    occ = np.where(res.occupied, 2, 0)
    nok = (occ == 2).sum(1)
    ekmo = okm.ao_to_kmo(E, res.mo_coef, s.ncell, int(nok.max()), int(nok.min()))
    ec, _, _, nzero = okm.kmp2(ekmo, res.mo_energies, occ, s.ncell, int(nok.min()))
    er = r.kmp2(E, res.mo_coef, res.mo_energies, occ)
    assert nzero == 0
    assert abs(ec - er) < 1e-13
    # Gamma supercell MP2 with the supercell RHF on the full ERI: the same energy
    Ef, _ = o.eri_full(sao=False)
    ch = Ef.transpose(3, 2, 1, 0)
    _, e, C, _, _ = orhf.rhf_gamma(ch, S, T, V, s.n_occ, r.nre())
    eg = orhf.mp2_gamma(ch, e, C, s.n_occ)
    assert abs(ec - eg) < 1e-9
