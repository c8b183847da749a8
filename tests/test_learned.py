"""The learned-order network and its features (CPU part).

The numpy network (oracle/mlp.py) is pinned by central finite differences
of its own loss; the product's feature formulas by invariances (a common
translation along the torus, the exchange of bra and ket, of the two
functions of a pair) and by their closed form in the large-torus limit.
"""
import math

import numpy as np

from oracle import mlp as omlp

NP = 6 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2


def _params(seed=0):
    rng = np.random.default_rng(seed)
    p = []
    for fi, fo in ((6, 64), (64, 64), (64, 2)):
        p.append(rng.normal(0, math.sqrt(2.0 / fi), size=(fi, fo)).ravel())
        p.append(rng.normal(0, 0.1, size=fo))
    return np.concatenate(p)


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params()
    assert p.size == NP
    X = rng.normal(size=(37, 6))
    T = rng.normal(size=(37, 2))
    loss, g = omlp.loss_grad(p, X, T, emulate=False)
    h = 1e-6
    for k in rng.choice(NP, size=60, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (omlp.loss_grad(pp, X, T, emulate=False)[0] - omlp.loss_grad(pm, X, T, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -3.14159, 1e-30])
    r = omlp.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0                     # tie to even
    assert r[2] == 1.0 + 2.0 ** -6
    assert abs(r[3] / x[3] - 1.0) < 2.0 ** -8


def _feat(e, x, L, ls=(0, 0, 0, 0)):
    from william_ci_amd import learned
    xyz = np.zeros((1, 4, 3))
    xyz[0, :, 0] = x
    return learned.features(np.array([e]), xyz, np.array([sum(ls)]), np.array([L]))[0]


def test_feature_invariances():
    e = (0.3, 1.7, 0.9, 5.0)
    x = (1.0, 4.0, 7.5, 9.0)
    L = 12.0
    f = _feat(e, x, L)
    # common translation along the torus (by less than the wrap of the means)
    assert np.allclose(_feat(e, [v + 2.7 for v in x], L), f, atol=1e-12)
    # bra <-> ket and the two functions of a pair
    assert np.allclose(_feat((e[2], e[3], e[0], e[1]), (x[2], x[3], x[0], x[1]), L), f, atol=1e-12)
    assert np.allclose(_feat((e[1], e[0], e[2], e[3]), (x[1], x[0], x[2], x[3]), L), f, atol=1e-12)


def test_feature_large_torus_limit():
    """L -> infinity: chord -> straight distance, circular means -> Gaussian product centres."""
    e = (0.5, 1.5, 0.8, 0.2)
    x = (0.0, 1.0, 3.0, 4.5)
    f = _feat(e, x, 1e6)
    P, Q = e[0] + e[1], e[2] + e[3]
    rho = P * Q / (P + Q)
    xP = (e[0] * x[0] + e[1] * x[1]) / P
    xQ = (e[2] * x[2] + e[3] * x[3]) / Q
    assert abs(f[0] - math.log10(rho)) < 1e-12
    assert abs(f[2] - math.log10(1 + rho * (xP - xQ) ** 2)) < 1e-6
    lk = -(e[0] * e[1] / P * 1.0 + e[2] * e[3] / Q * 1.5 ** 2) / math.log(10)
    assert abs(f[5] - lk / 10) < 1e-6
