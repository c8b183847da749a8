"""GPU parity of the direct J/K (tcg_direct_init / tcg_direct_jk, the
reference's direct_init / direct_JK, src/module/eri_direct.f90:76-374).

The direct J/K must equal hartree_exchange_potential applied to the stored
SAO ERI: every stored element contracted exactly once.  Oracle: the C
restatement's stored ERI and J/K, and where the checker library is built the
reference's own direct_JK.  Bar: 1e-11 relative to max |J|, |K| (FP64; the
order of the sums differs).  Both accumulation paths are covered: J/K rows in
LDS (the default at these sizes) and directly in global memory.
"""
import os

import numpy as np
import pytest

from conftest import needs_ref
from oracle import orc
from tests_support import normalised_coef
from william_ci_amd import system as tsys

pytestmark = pytest.mark.gpu

CASES = {
    "h10_ring": lambda: tsys.workload("h10_ring"),
    "h6_d_1d": lambda: tsys.chain_1d(6, 1.8, 10.8, basis=tsys.H_BASIS_D),
    "h2_3x3_2d": lambda: tsys.lattice_2d_h2(3, 3.6, 10.8),
    "h2_2x2x2_d_3d": lambda: tsys.lattice_3d_h2(2, 4.0, 8.0),
}


def _density(nb, seed=3):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((nb, nb))
    return (A + A.T) / nb


def _rel(a, b):
    return np.abs(a - b).max() / max(1e-300, np.abs(b).max())


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("path", ["lds", "global"])
def test_direct_jk_matches_stored(name, path, monkeypatch):
    from william_ci_amd.engine import Engine
    if path == "global":
        monkeypatch.setenv("TCG_JK_GLOBAL", "1")
    s = CASES[name]()
    coef = normalised_coef(s)
    print(f"case {name} ({path}) nbas={s.nbas} ndp={s.ndp}", flush=True)
    P = _density(s.nbas)
    Eo, _ = orc.OracleSystem(s, coef).eri_full()
    Jo, Ko = orc.OracleSystem(s, coef).jk(P, Eo)
    eng = Engine(s, coef)
    eng.direct_init()
    J, K = eng.direct_jk(P)
    J2, K2 = eng.direct_jk(2.0 * P)          # second call on the same set-up
    eng.close()
    print(f"  J rel {_rel(J, Jo):.2e}  K rel {_rel(K, Ko):.2e}")
    assert _rel(J, Jo) < 1e-11 and _rel(K, Ko) < 1e-11
    assert _rel(J2, 2.0 * Jo) < 1e-11 and _rel(K2, 2.0 * Ko) < 1e-11
    assert np.all(J[s.nfuc:] == 0.0) and np.all(K[s.nfuc:] == 0.0)


@needs_ref
@pytest.mark.parametrize("name", ["h10_ring", "h2_3x3_2d"])
def test_direct_jk_matches_reference(name):
    from oracle import ref
    from william_ci_amd.engine import Engine
    s = CASES[name]()
    print(f"case {name} vs reference direct_JK", flush=True)
    r = ref.RefSystem(s)
    P = _density(s.nbas, seed=5)
    Jr, Kr = r.direct_jk(P)
    eng = Engine(s, r.coef)
    eng.direct_init()
    J, K = eng.direct_jk(P)
    eng.close()
    print(f"  J rel {_rel(J, Jr):.2e}  K rel {_rel(K, Kr):.2e}")
    assert _rel(J, Jr) < 1e-11 and _rel(K, Kr) < 1e-11


def test_rhf_direct_equals_stored():
    from william_ci_amd import scf
    from william_ci_amd.engine import Engine
    s = tsys.workload("h10_ring")
    print("case RHF h10_ring direct vs stored", flush=True)
    eng = Engine(s, normalised_coef(s))
    S, T, V = eng.one_electron()
    eng.eri_build(to_host=False)
    e_st = scf.rhf_sao(s, S, T, V, lambda P: eng.jk_stored(P)).energy
    eng.direct_init()
    e_di = scf.rhf_sao(s, S, T, V, lambda P: eng.direct_jk(P)).energy
    eng.close()
    print(f"  E stored {e_st:.12f}  direct {e_di:.12f}  diff {e_di - e_st:.2e}")
    assert abs(e_di - e_st) < 1e-10


def test_direct_jk_unchanged_by_block_requests():
    """A shell-block request (tcg_eri_blocks) between two direct J/K builds leaves
    the direct quartet list alone: the second build equals the first."""
    from william_ci_amd.engine import Engine
    from william_ci_amd.learned import eri_blocks
    s = tsys.lattice_2d_h2(3, 3.6, 10.8)
    print("case h2_3x3_2d direct J/K around a block request", flush=True)
    eng = Engine(s, normalised_coef(s))
    eng.direct_init()
    P = _density(s.nbas, seed=4)
    J1, K1 = eng.direct_jk(P)
    q = np.array([[0, 1, 2, 3], [0, 0, 5, 7], [1, 2, 1, 2]], dtype=np.int32)
    eri_blocks(eng, q)
    J2, K2 = eng.direct_jk(P)
    eng.close()
    assert _rel(J2, J1) < 1e-13 and _rel(K2, K1) < 1e-13
