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


def _feat3(e, xyz, L3, ndp, ls=(0, 0, 0, 0)):
    from william_ci_amd import learned
    return learned.features(np.array([e]), np.array([xyz], dtype=np.float64), np.array([sum(ls)]),
                            np.array([L3], dtype=np.float64), ndp)[0]


def test_features_more_periodic_directions():
    """2D/3D tori: with every centre on one line along x, the periodic y and z
    directions add nothing and the features are the Tori1D ones when x has the
    shortest length; the features do not depend on which axis is which, and
    are invariant under a common translation along every periodic direction."""
    e = (0.3, 1.7, 0.9, 5.0)
    x = (1.0, 4.0, 7.5, 9.0)
    f1 = _feat(e, x, 12.0)
    on_x = [(v, 0.5, -1.0) for v in x]
    assert np.allclose(_feat3(e, on_x, (12.0, 20.0, 0.0), 2), f1, atol=1e-12)
    assert np.allclose(_feat3(e, on_x, (12.0, 20.0, 30.0), 3), f1, atol=1e-12)
    pts = [(1.0, 2.0, 3.0), (4.0, 0.5, 7.0), (7.5, 9.0, 1.0), (9.0, 3.0, 5.5)]
    L3 = (12.0, 15.0, 18.0)
    f = _feat3(e, pts, L3, 3)
    perm = (2, 0, 1)                                # axes relabelled
    assert np.allclose(_feat3(e, [tuple(p[i] for i in perm) for p in pts], tuple(L3[i] for i in perm), 3), f,
                       atol=1e-12)
    assert np.allclose(_feat3(e, [(p[0] + 1.1, p[1] - 0.7, p[2] + 2.3) for p in pts], L3, 3), f, atol=1e-12)
    # 2D: the real z direction enters as an ordinary distance; large periodic lengths
    # give the ordinary Gaussian-product features
    big = _feat3(e, pts, (1e6, 1e6, 0.0), 2)
    P, Q = e[0] + e[1], e[2] + e[3]
    cP = [(e[0] * pts[0][d] + e[1] * pts[1][d]) / P for d in range(3)]
    cQ = [(e[2] * pts[2][d] + e[3] * pts[3][d]) / Q for d in range(3)]
    rho = P * Q / (P + Q)
    R2 = sum((cP[d] - cQ[d]) ** 2 for d in range(3))
    assert abs(big[2] - math.log10(1 + rho * R2)) < 1e-6
