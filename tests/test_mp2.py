"""Gamma-supercell MP2 (AO_to_MO_ERI_inplace + MP2,
src/Tools/AO_to_MO/AO_to_MO.f90:20-111, src/Energies/MP2/MP2.f90:1-87).

CPU: the numpy restatement (oracle/rhf.py mp2_gamma) equals the reference's
own transform + MP2 on the same full ERI and orbitals; the MP2 sum vanishes
for a density of one orbital pair and is invariant under a rotation of the
occupied (or virtual) orbitals among themselves.
"""
import numpy as np
import pytest

from conftest import needs_ref
from oracle import rhf as orhf


def _random_case(n=8, nocc=3, seed=0):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n, n, n)) * 0.05
    ch = A + A.transpose(1, 0, 2, 3)
    ch = ch + ch.transpose(0, 1, 3, 2)
    ch = ch + ch.transpose(2, 3, 0, 1)            # 8-fold symmetric (ij|kl)
    C, _ = np.linalg.qr(rng.standard_normal((n, n)))
    e = np.sort(np.concatenate([rng.uniform(-2.0, -0.5, nocc), rng.uniform(0.2, 2.0, n - nocc)]))
    return ch, e, C


def test_mp2_invariant_under_occupied_rotation():
    ch, e, C = _random_case()
    nocc = 3
    # equal occupied energies make E_MP2 invariant under occupied rotations
    e = e.copy()
    e[:nocc] = -1.0
    R, _ = np.linalg.qr(np.random.default_rng(5).standard_normal((nocc, nocc)))
    C2 = C.copy()
    C2[:, :nocc] = C[:, :nocc] @ R
    assert abs(orhf.mp2_gamma(ch, e, C, nocc) - orhf.mp2_gamma(ch, e, C2, nocc)) < 1e-13


def test_mp2_single_pair_closed_form():
    # one occupied, one virtual orbital: E = -(ia|ia)^2 / (e_a - e_i) * (2 - 1)
    ch, e, C = _random_case(n=2, nocc=1, seed=2)
    g = np.einsum("pqrs,p,q,r,s->", ch, C[:, 0], C[:, 1], C[:, 0], C[:, 1])
    want = -g * g / (2.0 * (e[1] - e[0]))
    assert abs(orhf.mp2_gamma(ch, e, C, 1) - want) < 1e-15


@needs_ref
def test_oracle_mp2_matches_reference():
    from oracle import ref
    from william_ci_amd import system as tsys
    ch, e, C = _random_case(n=10, nocc=4, seed=7)
    s = tsys.workload("h10_ring")              # only the reference's module state is needed
    r = ref.RefSystem(s, sao=False)
    E = np.ascontiguousarray(ch.transpose(3, 2, 1, 0))      # E[l,k,j,i] = ERI(i,j,k,l)
    import types
    r.s = types.SimpleNamespace(nbas=10)
    got = r.mp2(E, e, C, 4)
    want = orhf.mp2_gamma(ch, e, C, 4)
    assert abs(got - want) < 1e-14 * max(1.0, abs(want)) + 1e-15
