"""Pins the energy side of the oracle (oracle/rhf.py and the nuclear repulsion
fed to every E_RHF comparison) to the reference and to what the mathematics
fixes, on the CPU.

  * nuclear repulsion        scf.nuclear_repulsion against the reference's NRE
                             (src/Energies/Nuclear_repulsion.f90:1-59) for 1D, 2D
                             and 3D tori, and its large-L limit (chord -> the
                             Euclidean distance, so E_nuc -> sum Z Z / r)
  * Coulomb / exchange       oracle/orc.py's J/K and the einsum forms of
                             rhf_gamma against the reference's own
                             hartree_exchange_potential
                             (src/Energies/hartree_exchange_potential.f90:30),
                             chemist and physicist order, full and SAO rows
  * rhf_gamma's energy       on the reference's S, T, V and full ERI:
                             - E_elec = 1/2 Tr[P (H + F)] = sum_i (h_ii + e_i)
                               over the occupied orbitals (closed shell), the
                               orbital energies coming from its own Fock matrix;
                             - stationarity: for random occupied-virtual
                               rotations C(s) = C exp(s kappa), E(s) - E(0) is
                               O(s^2) (symmetric difference of the energy
                               functional built from the reference's J/K is
                               ~0, the curvature is not).  A wrong factor in
                               the Fock matrix moves the converged point off
                               the stationary point of the energy.
"""
import math

import numpy as np
import pytest
from scipy.linalg import expm

from conftest import needs_ref
from oracle import orc
from oracle import rhf as orhf
from william_ci_amd import scf
from william_ci_amd import system as tsys

SYSTEMS = {
    "h10_ring": lambda: tsys.workload("h10_ring"),
    "h64_chain": lambda: tsys.workload("h64_chain"),
    "h2_3x3_2d": lambda: tsys.lattice_2d_h2(3, 3.6, 10.8),
    "h2_8x8": lambda: tsys.workload("h2_8x8"),
    "h2_2x2x2_d_3d": lambda: tsys.lattice_3d_h2(2, 4.0, 8.0),
    "h2_4x4x4_d": lambda: tsys.workload("h2_4x4x4_d"),
}


def _ref_nre(s):
    from oracle import ref
    # the reference's NRE reads only the torus lengths of the module state
    ref.lib().ref_init(s.ndp, np.array(s.L, dtype=np.float64), np.array(s.ncell, dtype=np.int32), s.nfuc,
                       32, 1, 1, 0)
    geom = np.asfortranarray(s.geom).ravel(order="F").copy()
    import ctypes
    e = ctypes.c_double(0.0)
    ref.lib().ref_nre(s.natom, geom, s.Z.astype(np.int32), ctypes.byref(e))
    return e.value


@needs_ref
@pytest.mark.parametrize("case", list(SYSTEMS))
def test_nuclear_repulsion_matches_reference(case):
    s = SYSTEMS[case]()
    mine, want = scf.nuclear_repulsion(s), _ref_nre(s)
    print(f"{case}: E_nuc {mine:.14f} reference {want:.14f}")
    assert abs(mine - want) <= 1e-12 * max(1.0, abs(want))


def test_nuclear_repulsion_large_torus_limit():
    # H2 (bond 1.4) in a torus of length L: the chord distance is
    # (2/a) sin(a x / 2) = x (1 - (a x)^2 / 24 + ...) -> x
    for L in (1e3, 1e4):
        for kind, line in (("Tori1D", f"1 {L} {L}"), ("Tori3D", f"1 1 1 {L} {L} {L} {L}")):
            s = tsys.build_system(f"{kind}\n{line}\nSAO\n$$\nH 0 0 0\nH 1.4 0 0\n$$\n", tsys.H_BASIS)
            a = 2 * math.pi / L
            chord = 2.0 / a * math.sin(a * 1.4 / 2)
            # 1 - cos(a x) (the reference's form) loses ~eps / (a x)^2 relative
            tol = 10 * 2.2e-16 / (a * 1.4) ** 2
            assert abs(scf.nuclear_repulsion(s) * chord - 1.0) < tol
            assert abs(scf.nuclear_repulsion(s) - 1 / 1.4) < (a * 1.4) ** 2 / 24


@pytest.fixture(scope="module")
def h10_ref():
    """The reference's S, T, V, full ERI and NRE for the H10 ring."""
    from oracle import ref
    s = tsys.workload("h10_ring")
    r = ref.RefSystem(s, sao=False)
    E = r.eri_full()                      # E[l,k,j,i] = ERI(i,j,k,l)
    S, T, V = r.one_electron()
    enuc = r.nre()
    return s, r, E, (S, T, V), enuc


def _ref_jk(s, P, E, n1, physicist):
    """The reference's hartree_exchange_potential on a Fortran-ordered ERI."""
    from oracle import ref
    nb = s.nbas
    ref.lib().ref_init(s.ndp, np.array(s.L, dtype=np.float64), np.array(s.ncell, dtype=np.int32), s.nfuc,
                       32, 1, int(n1 != nb), int(physicist))
    Pf = np.ascontiguousarray(np.asarray(P, dtype=np.float64).T).ravel()
    J = np.zeros(nb * nb)
    K = np.zeros(nb * nb)
    ref.lib().ref_jk(n1, nb, Pf, np.ascontiguousarray(E, dtype=np.float64).ravel(), J, K)
    return J.reshape(nb, nb).T.copy(), K.reshape(nb, nb).T.copy()


@needs_ref
@pytest.mark.parametrize("physicist", [False, True])
def test_jk_matches_reference(h10_ref, physicist):
    s, r, E, _, _ = h10_ref
    nb, nf = s.nbas, s.nfuc
    rng = np.random.default_rng(7)
    P = rng.standard_normal((nb, nb))
    P = P + P.T
    ch = E.transpose(3, 2, 1, 0)                         # ch[i,j,k,l] = (ij|kl)
    if physicist:
        Eref = np.ascontiguousarray(ch.transpose(3, 1, 2, 0))   # physicist ERI(i,k,j,l) in Fortran memory
    else:
        Eref = E
    Jr, Kr = _ref_jk(s, P, Eref, nb, physicist)
    # the oracle's J/K and the einsum forms rhf_gamma uses
    Jo, Ko = orc.OracleSystem(s, r.coef).jk(P, Eref, physicist=physicist)
    Je = np.einsum("mnls,ls->mn", ch, P, optimize=True)
    Ke = -0.5 * np.einsum("mlsn,ls->mn", ch, P, optimize=True)
    scale = np.abs(Jr).max()
    for name, A in (("oracle J", Jo), ("einsum J", Je)):
        assert np.abs(A - Jr).max() < 1e-12 * scale, name
    for name, A in (("oracle K", Ko), ("einsum K", Ke)):
        assert np.abs(A - Kr).max() < 1e-12 * scale, name
    # the SAO rows: the reference fills rows <= nfuc only
    Esao = np.ascontiguousarray(Eref[..., :nf])
    Js, Ks = _ref_jk(s, P, Esao, nf, physicist)
    assert np.abs(Js[:nf] - Jr[:nf]).max() < 1e-12 * scale and np.abs(Js[nf:]).max() == 0.0
    assert np.abs(Ks[:nf] - Kr[:nf]).max() < 1e-12 * scale and np.abs(Ks[nf:]).max() == 0.0


def _energy(s, C, Hc, E, nO):
    """E_elec of the closed-shell determinant C[:, :nO] with the reference's
    J/K (RHF.f90:198-218: Tr(P H) + 1/2 Tr(P J) + 1/2 Tr(P K))."""
    P = 2.0 * C[:, :nO] @ C[:, :nO].T
    J, K = _ref_jk(s, P, E, s.nbas, False)
    return float(np.sum(P * Hc) + 0.5 * np.sum(P * J) + 0.5 * np.sum(P * K)), P, J, K


@needs_ref
def test_rhf_gamma_energy_is_pinned(h10_ref):
    s, r, E, (S, T, V), enuc = h10_ref
    nb, nO = s.nbas, s.n_occ
    Hc = T + V
    ch = E.transpose(3, 2, 1, 0)
    Eg, e, C, Pc, it = orhf.rhf_gamma(ch, S, T, V, nO, enuc, thresh=1e-11)
    # the orbitals of the last diagonalisation; one more Fock build at them
    E0, P, J, K = _energy(s, C, Hc, E, nO)
    F = Hc + J + K
    # the reported energy at these orbitals (the SCF returns the energy of the
    # density it converged on, which differs from C's by the last 1e-11 step)
    assert abs(Eg - (E0 + enuc)) < 1e-9
    # E_elec = 1/2 Tr[P (H + F)] and, in the MO basis, sum_i (h_ii + e_i)
    assert abs(0.5 * np.sum(P * (Hc + F)) - E0) < 1e-10
    eps = np.einsum("mi,mn,ni->i", C, F, C)
    hmo = np.einsum("mi,mn,ni->i", C, Hc, C)
    assert np.abs(eps - e).max() < 1e-8                # C diagonalises the converged Fock matrix
    assert abs(np.sum(hmo[:nO] + eps[:nO]) - E0) < 1e-10
    # orthonormality in the S metric
    assert np.abs(C.T @ S @ C - np.eye(nb)).max() < 1e-10
    # stationarity under occupied-virtual rotations
    rng = np.random.default_rng(3)
    kap = np.zeros((nb, nb))
    kap[nO:, :nO] = rng.standard_normal((nb - nO, nO))
    kap = kap - kap.T
    kap /= np.abs(kap).max()
    h = 1e-4
    Ep = _energy(s, C @ expm(h * kap), Hc, E, nO)[0]
    Em = _energy(s, C @ expm(-h * kap), Hc, E, nO)[0]
    grad = (Ep - Em) / (2 * h)
    curv = (Ep + Em - 2 * E0) / h ** 2
    # the analytic gradient along kappa: 4 sum_{ai} kappa_ai F^MO_ai
    Fmo = C.T @ F @ C
    g_an = 4.0 * np.sum(kap[nO:, :nO] * Fmo[nO:, :nO])
    print(f"E_HF {Eg:.12f}  dE/ds {grad:.2e} (analytic {g_an:.2e})  d2E/ds2 {curv:.3f}")
    assert abs(grad) < 1e-7 and abs(g_an) < 1e-7
    assert curv > 1e-2                                 # a minimum along kappa, not a flat direction
    # a planted mistake: the Fock matrix with the exchange at full weight does
    # not give a stationary point of this energy
    Fbad = Hc + J + 2 * K
    eb, Ub = np.linalg.eigh(orhf._X(S).T @ Fbad @ orhf._X(S))
    Cb = orhf._X(S) @ Ub
    Ebp = _energy(s, Cb @ expm(h * kap), Hc, E, nO)[0]
    Ebm = _energy(s, Cb @ expm(-h * kap), Hc, E, nO)[0]
    assert abs((Ebp - Ebm) / (2 * h)) > 1e-4


@needs_ref
def test_bloch_rhf_matches_pinned_gamma(h10_ref):
    """rhf_sao (the k-space loop the engine's tests use) against the pinned
    Gamma RHF, both on the reference's integrals; E_nuc from the reference."""
    s, r, E, (S, T, V), enuc = h10_ref
    nf = s.nfuc
    ch = E.transpose(3, 2, 1, 0)
    Eg = orhf.rhf_gamma(ch, S, T, V, s.n_occ, enuc)[0]
    Esao = np.ascontiguousarray(E[..., :nf])
    res = scf.rhf_sao(s, S, T, V, lambda P: _ref_jk(s, P, Esao, nf, False))
    assert res.converged
    assert abs(res.e_nuc - enuc) < 1e-12
    print(f"E_RHF k-space {res.energy:.12f}  Gamma {Eg:.12f}")
    assert abs(res.energy - Eg) < 1e-9
