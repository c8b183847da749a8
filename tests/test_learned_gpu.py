"""GPU checks of the learned-order network kernels (tcg_mlp_forward,
tcg_mlp_grad, tcg_mlp_train) against the numpy network (oracle/mlp.py with
BF16 rounding at the same places), and of the learned ERI build.

Bars: outputs within 1e-3 of the output scale; gradients within 2e-2 in
relative norm per parameter block (BF16 operands, FP32 accumulation, a few
roundings that fall on the other side); training lowers the loss of a
smooth target at least tenfold.  The learned build with an infinite margin
is the plain build; with the committed model its ERIs stay within 1e-8 of
the plain rule and the RHF energy within 1 microhartree.
"""
import os

import numpy as np
import pytest

from oracle import mlp as omlp
from tests_support import normalised_coef
from william_ci_amd import system as tsys

pytestmark = pytest.mark.gpu


def _params(seed=0):
    from william_ci_amd import learned
    p = learned.init_params(np.random.default_rng(seed))
    rng = np.random.default_rng(seed + 1)
    o = 6 * 64
    p[o:o + 64] = rng.normal(0, 0.1, 64)               # non-zero biases
    return p


def test_forward_matches_oracle():
    from william_ci_amd import learned
    rng = np.random.default_rng(2)
    p = _params()
    X = rng.normal(size=(1000, 6)).astype(np.float32)
    print("case mlp forward n=1000", flush=True)
    Y = learned.forward(p, X)
    Yo, _ = omlp.forward(p.astype(np.float64), X.astype(np.float64), emulate=True)
    scale = np.abs(Yo).max()
    print(f"  max |dY| {np.abs(Y - Yo).max():.3e} (scale {scale:.3f})")
    assert np.abs(Y - Yo).max() <= 1e-3 * scale


def test_gradient_matches_oracle():
    from william_ci_amd import learned
    rng = np.random.default_rng(3)
    p = _params(4)
    X = rng.normal(size=(517, 6)).astype(np.float32)        # not a multiple of 64
    T = rng.normal(size=(517, 2)).astype(np.float32)
    print("case mlp gradient n=517", flush=True)
    g, loss = learned.grad(p, X, T)
    lo, go = omlp.loss_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64), emulate=True)
    assert abs(loss - lo) <= 1e-4 * lo
    off = 0
    for name, shp in omlp.SHAPES:
        n = int(np.prod(shp))
        a, b = g[off:off + n], go[off:off + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"  {name}: rel {rel:.2e}")
        assert rel < 2e-2, name
        off += n


def test_training_lowers_loss():
    from william_ci_amd import learned
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, size=(20000, 6)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2], np.abs(X[:, 3]) * 3 + X[:, 4]], axis=1).astype(np.float32)
    print("case mlp training n=20000", flush=True)
    p, hist = learned.train(learned.init_params(np.random.default_rng(0)), X, T, epochs=30, batch=512, lr=3e-3,
                            w=(1.0, 1.0))
    print(f"  loss {hist[0]:.4f} -> {hist[-1]:.5f}")
    assert hist[-1] < 0.1 * hist[0]
    # the trained network through the inference kernel agrees with the training loss
    Y = learned.forward(p, X)
    assert np.mean(np.sum((Y - T) ** 2, axis=1)) < 0.2 * hist[0]


def test_learned_build_with_infinite_margin_is_plain():
    from william_ci_amd import learned
    from william_ci_amd.engine import Engine
    s = tsys.workload("h10_ring")
    print("case learned build, margin 100 (all plain)", flush=True)
    eng = Engine(s, normalised_coef(s))
    eng.set_quadrature(False, 64)
    E0 = eng.eri_build()
    m = learned.Model(params=_params(), norm=np.concatenate([np.zeros(6), np.ones(6), -1e30 * np.ones(6),
                                                             1e30 * np.ones(6)]).astype(np.float32), margin=100.0)
    m.install(eng)
    E1, counts = learned.eri_build_learned(eng, m)
    eng.close()
    print(f"  counts {counts.tolist()}  max |diff| {np.abs(E1 - E0).max():.2e}")
    assert counts[len(m.orders)] == counts.sum() and np.abs(E1 - E0).max() <= 1e-15


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(__file__)), "william_ci_amd", "data",
                                                    "quartet_orders.npz")), reason="no trained model committed")
def test_learned_build_energy_within_microhartree():
    from william_ci_amd import learned, scf
    from william_ci_amd.engine import Engine
    s = tsys.workload("h10_ring")
    print("case learned build h10_ring with the trained model", flush=True)
    eng = Engine(s, normalised_coef(s))
    S, T, V = eng.one_electron()
    eng.set_quadrature(False, 64)
    E0 = eng.eri_build()
    e0 = scf.rhf_sao(s, S, T, V, lambda P: eng.jk_stored(P)).energy
    m = learned.Model.load()
    m.install(eng)
    E1, counts = learned.eri_build_learned(eng, m)
    e1 = scf.rhf_sao(s, S, T, V, lambda P: eng.jk_stored(P)).energy
    # spot check: sampled shell quartets against a converged rule (40 mapped points,
    # the 128-point table to ~1e-14); the network's target is 1e-10 per integral
    spot = learned.spot_check(eng, E1, n=256)
    spot0 = learned.spot_check(eng, E0, n=256)
    eng.close()
    print(f"  counts {counts.tolist()}  max |dERI| {np.abs(E1 - E0).max():.2e}  dE {e1 - e0:.2e}  "
          f"spot check {spot:.2e} (plain 64-point table {spot0:.2e})")
    assert np.abs(E1 - E0).max() <= 1e-8
    assert abs(e1 - e0) <= 1e-6
    assert spot <= 1e-10


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(__file__)), "william_ci_amd", "data",
                                                    "quartet_orders.npz")), reason="no trained model committed")
def test_learned_skipping():
    """Skipping by the bound head at 1e-13 inside the Schwarz band below 1e-10:
    the skipped quartets are counted, they leave zeros and change nothing else,
    every element they held is below the cap (the Schwarz inequality), and the
    energy moves by far less than 1 uEh."""
    from william_ci_amd import learned, scf
    from william_ci_amd.engine import Engine
    s = tsys.workload("h64_chain")
    print("case learned skipping h64_chain", flush=True)
    eng = Engine(s, normalised_coef(s))
    S, T, V = eng.one_electron()
    m = learned.Model.load()
    m.install(eng)
    E1, c1 = learned.eri_build_learned(eng, m)
    e1 = scf.rhf_sao(s, S, T, V, lambda P: eng.jk_stored(P)).energy
    learned.set_skip(eng, 1e-13)
    E2, c2 = learned.eri_build_learned(eng, m)
    e2 = scf.rhf_sao(s, S, T, V, lambda P: eng.jk_stored(P)).energy
    learned.set_skip(eng, 0.0)
    eng.close()
    nord = len(m.orders)
    d = np.abs(E2 - E1)
    print(f"  skipped {c2[nord + 1]} of {c2.sum()} shell quartets; max |dERI| {d.max():.2e}; dE {e2 - e1:.2e}")
    assert c1[nord + 1] == 0 and c2.sum() == c1.sum()
    assert np.all(E2[E2 != E1] == 0.0)          # a skipped quartet leaves zeros, the rest is unchanged
    assert d.max() < 1e-10
    assert abs(e2 - e1) < 1e-8


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(__file__)), "william_ci_amd", "data",
                                                    "quartet_orders.npz")), reason="no trained model committed")
def test_learned_build_chunked_equals_one_pass(monkeypatch):
    """Rules stay attached to their quartets when the list is split into chunks."""
    from william_ci_amd import learned
    from william_ci_amd.engine import Engine
    s = tsys.workload("h10_ring")
    coef = normalised_coef(s)
    print("case learned build chunked", flush=True)
    m = learned.Model.load()
    eng = Engine(s, coef)
    m.install(eng)
    E1, c1 = learned.eri_build_learned(eng, m)
    rows = eng.stats()["rows"]
    eng.close()
    monkeypatch.setenv("TCG_ROW_BUDGET", str(max(2e5, rows * 1200.0 / 5)))
    eng = Engine(s, coef)
    m.install(eng)
    E2, c2 = learned.eri_build_learned(eng, m)
    eng.close()
    assert np.array_equal(c1, c2) and np.abs(E1 - E2).max() <= 1e-14


def test_learned_gain_against_converged_quadrature():
    """The learned orders against the plain path at the learned path's accuracy
    (the 128-point table, the converged quadrature the network reproduces) on
    the Tori1D H64 chain: fewer row flops (counted by the row kernel) and a
    shorter ERI build, with E_total = E_RHF + E_c(k-MP2) within 1 microhartree.
    Bars: flops at least 2x fewer, build time at least 1.5x shorter (the bench
    measures about 2.2x), |dE_total| <= 1e-6 Eh."""
    import time
    from william_ci_amd import learned, scf
    from william_ci_amd.engine import Engine
    s = tsys.workload("h64_chain")
    print("case learned vs 128-point table, h64_chain", flush=True)
    coef = normalised_coef(s)
    eng = Engine(s, coef)
    S, T, V = eng.one_electron()

    def timed(build):
        build()                                              # first call: buffers
        t0 = time.perf_counter()
        for _ in range(5):
            eng.refresh_basis(coef)
            build()
        return (time.perf_counter() - t0) / 5

    def energy():
        r = scf.rhf_sao_gpu(s, eng, S, T, V)
        return r.energy + eng.kmp2(r.mo_coef, r.mo_energies, r.occupied)["ec"]

    eng.set_quadrature(False, 128)
    t128 = timed(lambda: eng.eri_build(to_host=False))
    f128 = eng.stats()["row_flops"]
    e128 = energy()
    eng.set_quadrature(False, 64)
    m = learned.Model.load()
    m.install(eng)
    tl = timed(lambda: learned.eri_build_learned(eng, m, to_host=False))
    fl = eng.stats()["row_flops"]
    el = energy()
    eng.close()
    print(f"  128-point: {t128 * 1e3:.2f} ms, {f128:.3e} row flops;  learned: {tl * 1e3:.2f} ms, {fl:.3e} row flops;"
          f"  dE_total {el - e128:.2e}")
    assert f128 >= 2.0 * fl
    assert t128 >= 1.5 * tl
    assert abs(el - e128) <= 1e-6
