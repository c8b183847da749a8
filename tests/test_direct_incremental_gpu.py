"""Incremental direct SCF (tcg_rhf_sao source 2, tcg_set_direct_screen):
J and K of the density change with density screening, added to the previous
J and K; and the learned per-quartet orders in the direct J/K
(tcg_direct_init_learned).

Bars:
  * incremental vs the plain direct RHF (every quartet at every iteration, the
    reference's direct_JK): the same converged energy to 1e-8 Eh at the
    default threshold (1e-14 on bound x density change);
  * the screened J/K of a density equals the unscreened one within the
    threshold's reach, and with tau = 0 it is the unscreened one exactly in
    the number of quartets;
  * learned orders in the direct RHF: within 1 microhartree of the plain
    direct RHF on the converged quadrature (the 128-point table), i.e. of the
    integrals the network is trained to reproduce.
"""
import os

import numpy as np
import pytest

from tests_support import normalised_coef
from william_ci_amd import system as tsys

pytestmark = pytest.mark.gpu

SYSTEMS = {"h10_ring": lambda: tsys.workload("h10_ring"),
           "h2_3x3_2d": lambda: tsys.lattice_2d_h2(3, 3.6, 10.8),
           "h2_2x2x2_d_3d": lambda: tsys.lattice_3d_h2(2, 4.0, 8.0)}
MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "william_ci_amd", "data",
                     "quartet_orders.npz")


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_incremental_direct_rhf_matches_plain(name):
    from william_ci_amd import scf
    from william_ci_amd.engine import Engine
    s = SYSTEMS[name]()
    print(f"case {name} incremental direct RHF", flush=True)
    eng = Engine(s, normalised_coef(s))
    S, T, V = eng.one_electron()
    eng.direct_init()
    r0 = scf.rhf_sao_gpu(s, eng, S, T, V, source="direct")
    eng.set_direct_screen()                                   # defaults: tau 1e-14, full_every 50
    r1 = scf.rhf_sao_gpu(s, eng, S, T, V, source="direct_incremental")
    info = eng.direct_screen_info()
    eng.close()
    print(f"  plain {r0.energy:.12f} ({r0.iterations} it)  incremental {r1.energy:.12f} ({r1.iterations} it)  "
          f"quartets {info['kept_total']}/{info['listed_total']} over {info['builds']} builds")
    assert r0.converged and r1.converged
    assert abs(r1.energy - r0.energy) < 1e-8
    assert info["builds"] == r1.iterations and info["kept_total"] <= info["listed_total"]


def test_screened_jk_of_a_density():
    """tcg_direct_jk_screened: with a tiny tau every quartet stays and J/K equal
    the plain direct J/K (to the order of the atomic sums); at tau = 1e-13 on a
    density change of size 1e-8 (late SCF iterations) fewer quartets are
    evaluated, and since each skipped quartet's elements times the density
    blocks they meet stay below tau, J and K move by at most 8 tau per skipped
    quartet (eight permutations)."""
    from william_ci_amd.engine import Engine
    s = SYSTEMS["h2_3x3_2d"]()
    print("case h2_3x3_2d screened J/K", flush=True)
    eng = Engine(s, normalised_coef(s))
    eng.direct_init()
    rng = np.random.default_rng(3)
    P = rng.standard_normal((s.nbas, s.nbas)) * 1e-8
    P = P + P.T
    J0, K0 = eng.direct_jk(P)
    n_all = eng.direct_screen_info()["kept"]
    J1, K1 = eng.direct_jk(P, tau=1e-300)
    i1 = eng.direct_screen_info()
    J2, K2 = eng.direct_jk(P, tau=1e-13)
    i2 = eng.direct_screen_info()
    eng.close()
    scale = max(np.abs(J0).max(), np.abs(K0).max())
    print(f"  quartets: all {n_all}, tau 1e-300 {i1['kept']}, tau 1e-13 {i2['kept']};  "
          f"max |dJ| {np.abs(J2 - J0).max():.2e} |dK| {np.abs(K2 - K0).max():.2e} (scale {scale:.2e})")
    assert i1["kept"] == i1["listed"] == n_all
    assert np.abs(J1 - J0).max() <= 1e-14 * scale and np.abs(K1 - K0).max() <= 1e-14 * scale
    skipped = i2["listed"] - i2["kept"]
    assert 0 < i2["kept"] < i2["listed"]
    assert np.abs(J2 - J0).max() <= 8e-13 * skipped and np.abs(K2 - K0).max() <= 8e-13 * skipped


@pytest.mark.skipif(not os.path.exists(MODEL), reason="no trained model committed")
@pytest.mark.parametrize("name", ["h2_3x3_2d", "h2_2x2x2_d_3d"])
def test_learned_direct_rhf_converged_quadrature(name):
    from william_ci_amd import learned, scf
    from william_ci_amd.engine import Engine
    s = SYSTEMS[name]()
    print(f"case {name} learned direct RHF", flush=True)
    eng = Engine(s, normalised_coef(s))
    S, T, V = eng.one_electron()
    sc = eng.fast_quad_self_check(64)                       # the reference's start-up rule choice
    m = learned.Model.load()
    m.install(eng)
    eng.direct_init()
    r_plain = scf.rhf_sao_gpu(s, eng, S, T, V, source="direct")
    counts = learned.direct_init_learned(eng, m)
    r_l = scf.rhf_sao_gpu(s, eng, S, T, V, source="direct")
    eng.set_direct_screen()
    r_li = scf.rhf_sao_gpu(s, eng, S, T, V, source="direct_incremental")
    eng.set_quadrature(False, 128)
    eng.direct_init()
    r128 = scf.rhf_sao_gpu(s, eng, S, T, V, source="direct")
    eng.close()
    nord = len(m.orders)
    print(f"  rule {'fast' if sc['fast_quad'] else '64-point table'}; orders {counts.tolist()}")
    print(f"  E plain {r_plain.energy:.12f}  learned {r_l.energy:.12f}  learned+incremental {r_li.energy:.12f}  "
          f"128-point table {r128.energy:.12f}")
    assert counts[:nord].sum() + counts[nord] + counts[nord + 1] > 0
    assert abs(r_l.energy - r128.energy) < 1e-6
    assert abs(r_li.energy - r128.energy) < 1e-6
    assert abs(r_plain.energy - r128.energy) < 1e-4          # the plain rule's own quadrature error
