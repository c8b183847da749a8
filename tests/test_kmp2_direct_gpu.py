"""Integral-direct k-point MP2 (tcg_kmp2_direct): the SAO integrals built in
slabs of the last index and transformed slab by slab, without the stored
tensor.

The reference's transform_ERI_direct_kMO (src/Energies/SAO/MP2/AO_to_MO_SAO_direct.f90:26)
calls zgemm and is not built here; it states that its result is that of
transform_ERI_AO_to_kMO on the stored tensor (:7-8).  So the direct path is
held to tcg_kmp2 on the stored SAO tensor (itself pinned to the reference's
transform_ERI_AO_to_kMO + kMP2_calc in test_mp2_gpu.py) on the same orbitals,
for several slab sizes (one cell per slab up to all cells in one), and on the
H10 ring to the reference's compiled k-MP2 directly.  Bar: 1e-11 Eh (the
two paths differ only in the order of the sums over the slabs).
"""
import numpy as np
import pytest

from conftest import needs_ref
from tests_support import normalised_coef
from william_ci_amd import system as tsys

pytestmark = pytest.mark.gpu

SYSTEMS = {"h10_ring": lambda: tsys.workload("h10_ring"),
           "h2_3x3_2d": lambda: tsys.lattice_2d_h2(3, 3.6, 10.8),
           "h2_2x2x2_d_3d": lambda: tsys.lattice_3d_h2(2, 4.0, 8.0)}
CASES = [("h10_ring", 1), ("h10_ring", 3), ("h10_ring", 0), ("h2_3x3_2d", 2), ("h2_3x3_2d", 0),
         ("h2_2x2x2_d_3d", 3)]


def _rhf(s, eng):
    from william_ci_amd import scf
    S, T, V = eng.one_electron()
    eng.eri_build(sao=True, to_host=False)
    return scf.rhf_sao_gpu(s, eng, S, T, V)


@pytest.mark.parametrize("name,slab", CASES, ids=[f"{n}-slab{c}" for n, c in CASES])
def test_direct_kmp2_equals_stored(name, slab):
    from william_ci_amd.engine import Engine
    s = SYSTEMS[name]()
    print(f"case {name} slab_cells={slab} nbas={s.nbas}", flush=True)
    eng = Engine(s, normalised_coef(s))
    r = _rhf(s, eng)
    stored = eng.kmp2(r.mo_coef, r.mo_energies, r.occupied)
    eng.direct_init()
    got = eng.kmp2_direct(r.mo_coef, r.mo_energies, r.occupied, slab_cells=slab)
    eng.close()
    print(f"  E_c stored {stored['ec']:.14f}  direct {got['ec']:.14f}  diff {got['ec'] - stored['ec']:.2e}  "
          f"slabs {got['slabs']}  integrals {got['integral_ms']:.1f} ms  transforms {got['transform_ms']:.1f} ms")
    assert stored["ec"] < 0.0
    assert abs(got["ec"] - stored["ec"]) < 1e-11
    assert abs(got["direct"] - stored["direct"]) < 1e-11
    assert got["zero_denominators"] == stored["zero_denominators"]
    if slab:
        assert got["slabs"] == -(-s.n_cell // slab)


@needs_ref
def test_direct_kmp2_matches_reference():
    from oracle import ref
    from william_ci_amd import scf
    from william_ci_amd.engine import Engine
    s = tsys.workload("h10_ring")
    print("case h10_ring direct k-MP2 vs the reference", flush=True)
    rs = ref.RefSystem(s)
    eng = Engine(s, rs.coef)
    S, T, V = eng.one_electron()
    eng.direct_init()
    res = scf.rhf_sao_gpu(s, eng, S, T, V, source="direct")
    got = eng.kmp2_direct(res.mo_coef, res.mo_energies, res.occupied, slab_cells=4)
    eng.close()
    occ = np.where(res.occupied, 2, 0)
    want = rs.kmp2(rs.eri_full(), res.mo_coef, res.mo_energies, occ)
    print(f"  E_kMP2 direct {got['ec']:.14f}  reference {want:.14f}  diff {got['ec'] - want:.2e}")
    assert abs(got["ec"] - want) < 1e-10
