"""Learned per-quartet quadrature orders.

No counterpart in the reference, whose t-quadrature is one rule for the
whole run (32/40 mapped Gauss-Legendre points, or the 64/128-point tables
when the start-up self-check rejects them).  Here a small network predicts,
for every primitive quartet, log10 of its largest integral and the index of
the smallest mapped Gauss-Legendre order (t = 0.6 u/(1-u), ORDERS) that
reproduces the 128-point table (Gauss-Legendre to 1e-12) to 1e-10.  (The
64-point table the reference falls back to is itself ~1e-8 away from the
converged integral, so it is not the target.)  The engine then evaluates each shell
quartet with the largest order its primitive quartets ask for (plus a
margin), and keeps the plain rule for quartets outside the training range.

Training data come from the plain quadrature itself: random Tori1D systems
(one cell, L uniform in [5, 120] bohr) of single-primitive s/p/d shells
(exponents log-uniform in [0.05, 50], centres uniform on the torus); every
shell quartet is evaluated with each order and with the 128-point table.  The network (6 -> 64 -> 64 -> 2, BF16 matrix-core kernels with FP32
accumulation, Adam) is trained on the GPU by the engine's own forward and
backward kernels (tcg_mlp_train).

Features (the same formulas as the engine's k_features): see FEATURES.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import os
import time
from typing import Optional

import numpy as np

from . import basis as _basis
from . import engine as _eng
from . import system as _sys

ORDERS = (8, 12, 16, 20, 24, 32, 40, 48)     # mapped Gauss-Legendre node counts, index = class
PLAIN = (False, 64)                          # the reference's rule after its self-check on H chains
LABEL_REF = (False, 128)                     # the labels' reference: the 128-point table (Gauss-Legendre to 1e-12)
LABEL_TOL = 1e-10                            # |I_order - I_reference| per integral
FEATURES = ("log10 rho", "log10 rho L^2", "log10(1 + rho R^2)", "(lA+lB+lC+lD)/4", "log10 emax L^2",
            "max(log10 K_AB K_CD, -30)/10")
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "quartet_orders.npz")

_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


def _lib():
    L = _eng.lib()
    if not getattr(L, "_mlp_bound", False):
        L.tcg_mlp_param_count.restype = ctypes.c_int
        L.tcg_mlp_forward.argtypes = [_fp, _fp, ctypes.c_int, _fp]
        L.tcg_mlp_grad.argtypes = [_fp, _fp, _fp, ctypes.c_int, _fp, _fp, _fp]
        L.tcg_mlp_train.argtypes = [_fp, _fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_uint, _fp, _fp]
        L.tcg_set_learned.argtypes = [ctypes.c_void_p, _fp, _fp, np.ctypeslib.ndpointer(dtype=np.int32),
                                      ctypes.c_int, ctypes.c_float]
        L.tcg_eri_build_learned.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                            np.ctypeslib.ndpointer(dtype=np.int64)]
        L.tcg_set_learned_skip.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_float]
        L.tcg_direct_init_learned.argtypes = [ctypes.c_void_p, np.ctypeslib.ndpointer(dtype=np.int64)]
        L.tcg_eri_blocks.argtypes = [ctypes.c_void_p, ctypes.c_int, np.ctypeslib.ndpointer(dtype=np.int32),
                                     np.ctypeslib.ndpointer(dtype=np.float64), np.ctypeslib.ndpointer(dtype=np.int64)]
        L._mlp_bound = True
    return L


def param_count() -> int:
    return int(_lib().tcg_mlp_param_count())


# -- features ------------------------------------------------------------------
def _chord2(a, dx):
    return 2.0 * (1.0 - np.cos(a * dx)) / (a * a)


def features(e, xyz, lsum, L, ndp=1):
    """e: (n, 4) exponents, xyz: (n, 4, 3) centres, lsum: (n,) total angular
    momentum, L: (n,) torus lengths (one periodic direction, x) or (n, 3)
    lengths of the periodic directions d < ndp (ndp scalar or (n,)).  Chord
    distances along the periodic directions, ordinary ones along the others;
    L in the features is the shortest periodic length.  Returns (n, 6)."""
    e = np.asarray(e, dtype=np.float64)
    n = len(e)
    L = np.asarray(L, dtype=np.float64)
    if L.ndim == 1:
        L3 = np.zeros((n, 3))
        L3[:, 0] = L
    else:
        L3 = L
    nd = np.broadcast_to(np.asarray(ndp), (n,))
    per = np.arange(3)[None, :] < nd[:, None]                    # (n, 3) periodic directions
    a = 2.0 * math.pi / np.where(per, L3, 1.0)
    P, Q = e[:, 0] + e[:, 1], e[:, 2] + e[:, 3]
    rho = P * Q / (P + Q)

    def centre(i, j):
        s_ = e[:, i, None] * np.sin(a * xyz[:, i]) + e[:, j, None] * np.sin(a * xyz[:, j])
        c_ = e[:, i, None] * np.cos(a * xyz[:, i]) + e[:, j, None] * np.cos(a * xyz[:, j])
        real = (e[:, i, None] * xyz[:, i] + e[:, j, None] * xyz[:, j]) / (e[:, i] + e[:, j])[:, None]
        return np.where(per, np.arctan2(s_, c_) / a, real)

    def d2(X, Y):
        dx = X - Y
        return np.sum(np.where(per, _chord2(a, dx), dx * dx), axis=1)

    R2 = d2(centre(0, 1), centre(2, 3))
    dAB = d2(xyz[:, 0], xyz[:, 1])
    dCD = d2(xyz[:, 2], xyz[:, 3])
    lk = -(e[:, 0] * e[:, 1] / P * dAB + e[:, 2] * e[:, 3] / Q * dCD) / math.log(10.0)
    Lv = np.min(np.where(per, L3, np.inf), axis=1)
    return np.stack([np.log10(rho), np.log10(rho * Lv * Lv), np.log10(1.0 + rho * R2), 0.25 * np.asarray(lsum),
                     np.log10(e.max(axis=1) * Lv * Lv), np.maximum(lk, -30.0) / 10.0], axis=1)


# -- training data from the plain quadrature -----------------------------------
def random_torus(rng, nshell: int, ndp: int = 1):
    """A torus of one cell with ndp periodic directions: nshell
    single-primitive shells (l in {0,1,2}), exponents log-uniform in [0.05, 50],
    lengths uniform in [5, 120] per periodic direction, centres uniform on the
    torus along those directions (and uniform in [-2, 2] bohr along the real
    ones, except for Tori1D, whose centres lie on the x axis as in the chains)."""
    Ls = np.zeros(3)
    Ls[:ndp] = rng.uniform(5.0, 120.0, size=ndp)
    ls = rng.choice([0, 1, 2], size=nshell, p=[0.5, 0.3, 0.2])
    ex = np.exp(rng.uniform(math.log(0.05), math.log(50.0), size=nshell))
    xs = np.zeros((nshell, 3))
    for d in range(3):
        if d < ndp:
            xs[:, d] = rng.uniform(0.0, Ls[d], size=nshell)
        elif ndp > 1:
            xs[:, d] = rng.uniform(-2.0, 2.0, size=nshell)
    center, lxyz, expo, shell_of = [], [], [], []
    for s in range(nshell):
        for comp in _sys._CART[int(ls[s])]:
            center.append(tuple(xs[s]))
            lxyz.append(comp)
            expo.append(ex[s])
            shell_of.append(s)
    nb = len(center)
    kind = {1: "Tori1D", 2: "Tori2D", 3: "Tori3D"}[ndp]
    sysm = _sys.TorusSystem(kind=kind, ndp=ndp, L=tuple(float(v) for v in Ls), ncell=(1, 1, 1),
                            spacing=float(Ls[0]), keywords=[], geom=np.zeros((1, 3)), Z=np.array([1]), nauc=1,
                            center=np.array(center, dtype=np.float64), lxyz=np.array(lxyz, dtype=np.int32),
                            prim_off=np.arange(nb + 1, dtype=np.int32), expo=np.array(expo, dtype=np.float64),
                            coef=np.ones(nb))
    shells = {"l": ls, "e": ex, "x": xs, "of": np.array(shell_of)}
    return sysm, shells


def torus_samples(sysm, shells):
    """Features and labels of every canonical shell quartet (i<=j, k<=l,
    (ij) <= (kl)) of one synthetic torus."""
    coef = _basis.normalize(sysm)
    eng = _eng.Engine(sysm, coef, nquad=ORDERS[0])
    try:
        eng.set_quadrature(*LABEL_REF)
        Eref = eng.eri_build()
        errs = []
        for n in ORDERS:
            eng.set_quadrature(True, n)
            errs.append(np.abs(eng.eri_build() - Eref))
    finally:
        eng.close()
    ns = len(shells["l"])
    starts = np.nonzero(np.r_[True, np.diff(shells["of"]) != 0])[0]

    def blockmax(A):
        # per shell 4-tuple maximum; E[l,k,j,i] = (ij|kl) -> result [si, sj, sk, sl]
        for ax in range(4):
            A = np.maximum.reduceat(A, starts, axis=ax)
        return A.transpose(3, 2, 1, 0).ravel()

    big = blockmax(np.abs(Eref))
    order = np.full(ns ** 4, len(ORDERS), dtype=np.int64)
    for r in reversed(range(len(ORDERS))):
        order[blockmax(errs[r]) <= LABEL_TOL] = r
    i, j, k, l = np.unravel_index(np.arange(ns ** 4), (ns, ns, ns, ns))
    keep = (i <= j) & (k <= l) & (i * ns + j <= k * ns + l)
    # the 4-tuple index is ((i*ns + j)*ns + k)*ns + l with i the first function's shell
    q = np.nonzero(keep)[0]
    i, j, k, l = i[q], j[q], k[q], l[q]
    e = np.stack([shells["e"][i], shells["e"][j], shells["e"][k], shells["e"][l]], axis=1)
    xyz = np.zeros((len(q), 4, 3))
    for c, s in enumerate((i, j, k, l)):
        xyz[:, c, :] = shells["x"][s]
    lsum = shells["l"][i] + shells["l"][j] + shells["l"][k] + shells["l"][l]
    X = features(e, xyz, lsum, np.tile(np.asarray(sysm.L, dtype=np.float64), (len(q), 1)), sysm.ndp)
    bound = np.log10(np.maximum(big[q], 1e-30))
    return X, np.stack([bound, order[q].astype(np.float64)], axis=1)


def make_dataset(n_tori: int, nshell: int = 12, seed: int = 12345, log=None, ndps=(1, 2, 3)):
    """Features and labels of n_tori random tori, their periodic dimension
    cycling through ndps."""
    rng = np.random.default_rng(seed)
    Xs, Ts = [], []
    t0 = time.perf_counter()
    for t in range(n_tori):
        sysm, shells = random_torus(rng, nshell, ndps[t % len(ndps)])
        X, T = torus_samples(sysm, shells)
        Xs.append(X)
        Ts.append(T)
        if log and (t + 1) % max(1, n_tori // 10) == 0:
            log(f"dataset: {t + 1}/{n_tori} tori, {sum(len(x) for x in Xs)} quartets, "
                f"{time.perf_counter() - t0:.1f} s")
    return np.concatenate(Xs), np.concatenate(Ts)


# -- the network ----------------------------------------------------------------
def init_params(rng) -> np.ndarray:
    """He-style initialisation of W1[6][64] b1 W2[64][64] b2 W3[64][2] b3."""
    p = []
    for fan_in, fan_out in ((6, 64), (64, 64), (64, 2)):
        p.append(rng.normal(0.0, math.sqrt(2.0 / fan_in), size=(fan_in, fan_out)).ravel())
        p.append(np.zeros(fan_out))
    out = np.concatenate(p).astype(np.float32)
    assert out.size == param_count()
    return out


def forward(params: np.ndarray, X: np.ndarray) -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.zeros((len(X), 2), dtype=np.float32)
    rc = _lib().tcg_mlp_forward(np.ascontiguousarray(params, dtype=np.float32), X, len(X), Y)
    if rc:
        raise _eng.EngineError(f"tcg_mlp_forward failed ({rc})")
    return Y


def grad(params, X, T, w=(0.1, 1.0)):
    g = np.zeros(param_count(), dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    rc = _lib().tcg_mlp_grad(np.ascontiguousarray(params, dtype=np.float32), np.ascontiguousarray(X, dtype=np.float32),
                             np.ascontiguousarray(T, dtype=np.float32), len(X), np.array(w, dtype=np.float32), g, loss)
    if rc:
        raise _eng.EngineError(f"tcg_mlp_grad failed ({rc})")
    return g, float(loss[0])


def train(params, X, T, epochs=20, batch=4096, lr=2e-3, seed=1, w=(0.1, 1.0)):
    p = np.ascontiguousarray(params, dtype=np.float32).copy()
    hist = np.zeros(epochs, dtype=np.float32)
    rc = _lib().tcg_mlp_train(p, np.ascontiguousarray(X, dtype=np.float32), np.ascontiguousarray(T, dtype=np.float32),
                              len(X), epochs, batch, lr, seed, np.array(w, dtype=np.float32), hist)
    if rc:
        raise _eng.EngineError(f"tcg_mlp_train failed ({rc})")
    return p, hist


@dataclasses.dataclass
class Model:
    params: np.ndarray           # float32, param_count()
    norm: np.ndarray             # float32 (24,): mean, std, lowest, highest of the raw features
    orders: tuple = ORDERS
    margin: float = 1.0

    def normalise(self, F):
        return ((F - self.norm[:6]) / self.norm[6:12]).astype(np.float32)

    def save(self, path: str = MODEL_PATH, **extra):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.params, norm=self.norm, orders=np.array(self.orders, dtype=np.int32),
                 margin=np.float32(self.margin), **extra)

    @staticmethod
    def load(path: str = MODEL_PATH) -> "Model":
        d = np.load(path, allow_pickle=False)
        return Model(params=d["params"].astype(np.float32), norm=d["norm"].astype(np.float32),
                     orders=tuple(int(v) for v in d["orders"]), margin=float(d["margin"]))

    def install(self, eng) -> None:
        rc = _lib().tcg_set_learned(eng.h, np.ascontiguousarray(self.params, dtype=np.float32),
                                    np.ascontiguousarray(self.norm, dtype=np.float32),
                                    np.array(self.orders, dtype=np.int32), len(self.orders), self.margin)
        eng._check(rc, "tcg_set_learned")


def fit(X, T, epochs=60, batch=4096, lr=2e-3, seed=1, log=None):
    """Normalise the features, train on the GPU; returns (Model, loss history)."""
    mean = X.mean(axis=0)
    std = X.std(axis=0)
    std[std < 1e-6] = 1.0
    lo, hi = X.min(axis=0), X.max(axis=0)
    span = hi - lo
    norm = np.concatenate([mean, std, lo - 0.02 * span, hi + 0.02 * span]).astype(np.float32)
    m = Model(params=init_params(np.random.default_rng(seed)), norm=norm)
    t0 = time.perf_counter()
    p, hist = train(m.params, m.normalise(X), T, epochs=epochs, batch=batch, lr=lr, seed=seed, w=(0.05, 1.0))
    if log:
        log(f"training: {epochs} epochs on {len(X)} quartets in {time.perf_counter() - t0:.1f} s, "
            f"loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    m.params = p
    return m, hist


def eri_build_learned(eng, model: Model, sao=True, physicist=False, to_host=True):
    """ERI with learned per-quartet orders (the model installed on eng);
    returns (ERI or None, shell quartets per order with the plain rule last)."""
    s = eng.s
    nb = s.nbas
    n1 = s.nfuc if sao else nb
    out = np.empty(n1 * nb * nb * nb) if to_host else None
    counts = np.zeros(len(model.orders) + 2, dtype=np.int64)      # per order, plain rule, skipped
    rc = _lib().tcg_eri_build_learned(eng.h, int(sao), int(physicist), out.ctypes.data if to_host else None, counts)
    eng._check(rc, "tcg_eri_build_learned")
    return (out.reshape(nb, nb, nb, n1) if to_host else None), counts


def direct_init_learned(eng, model: Model):
    """direct_init with the learned rules (model installed on eng): every later
    direct J/K (and the direct RHF) evaluates each quartet with its learned order;
    returns the shell quartets per order, the plain rule, and skipped."""
    counts = np.zeros(len(model.orders) + 2, dtype=np.int64)
    eng._check(_lib().tcg_direct_init_learned(eng.h, counts), "tcg_direct_init_learned")
    return counts


def set_skip(eng, thresh: float, schwarz_cap: float = 1e-10, margin_b: float = 1.5) -> None:
    """Learned Schwarz-style skipping in the next learned builds (tcg_set_learned_skip):
    shell quartets whose predicted largest |integral| (bound head, plus log10 of
    the primitive-quartet count and margin_b) is below thresh, and whose Schwarz
    product is below schwarz_cap, are left out; thresh = 0 turns it off."""
    eng._check(_lib().tcg_set_learned_skip(eng.h, float(thresh), float(schwarz_cap), float(margin_b)),
               "tcg_set_learned_skip")


def eri_blocks(eng, quartets: np.ndarray):
    """Shell blocks (current rule) of quartets (n, 4) of shell indices."""
    q = np.ascontiguousarray(quartets, dtype=np.int32).ravel()
    n = len(q) // 4
    # the largest block size bounds the output
    nmax = 6 ** 4
    out = np.zeros(n * nmax)
    off = np.zeros(n + 1, dtype=np.int64)
    rc = _lib().tcg_eri_blocks(eng.h, n, q, out, off)
    eng._check(rc, "tcg_eri_blocks")
    return out[:off[-1]], off


def shell_starts(sysm):
    """First AO of every shell, with the engine's grouping rule (consecutive
    Cartesian components with the same centre and primitives,
    src/Toroidal/ERI_toroidal.f90:106-125)."""
    starts, i, nb = [], 0, sysm.nbas
    lsum = sysm.lxyz.sum(axis=1)
    while i < nb:
        lt = int(lsum[i])
        nc = (lt + 1) * (lt + 2) // 2
        same = i + nc <= nb
        for j in range(i + 1, min(i + nc, nb)):
            pi, pj = slice(sysm.prim_off[i], sysm.prim_off[i + 1]), slice(sysm.prim_off[j], sysm.prim_off[j + 1])
            same = same and lsum[j] == lt and np.array_equal(sysm.center[j], sysm.center[i]) \
                and np.array_equal(sysm.expo[pi], sysm.expo[pj])
        if not same:
            nc = 1
        starts.append(i)
        i += nc
    return np.array(starts + [nb])


def spot_check(eng, E, n=256, rule=(True, 40), seed=7):
    """Largest difference between the learned SAO ERI E (E[l,k,j,i],
    i < nfuc) and n sampled shell quartets evaluated with a converged rule
    (40 mapped Gauss-Legendre points: the 128-point table to ~1e-14).  The
    rule in place on eng is restored afterwards."""
    s = eng.s
    fast0, n0, _, _ = eng.quadrature()
    st = shell_starts(s)
    nsh = len(st) - 1
    nshuc = int(np.sum(st[:-1] < s.nfuc))
    rng = np.random.default_rng(seed)
    q = np.stack([rng.integers(0, nshuc, n), rng.integers(0, nsh, n), rng.integers(0, nsh, n),
                  rng.integers(0, nsh, n)], axis=1)
    eng.set_quadrature(*rule)
    try:
        vals, off = eri_blocks(eng, q)
    finally:
        eng.set_quadrature(fast0, n0)
    worst = 0.0
    for x in range(n):
        f = [np.arange(st[q[x, y]], st[q[x, y] + 1]) for y in range(4)]
        blk = vals[off[x]:off[x + 1]].reshape(len(f[3]), len(f[2]), len(f[1]), len(f[0]))
        got = E[np.ix_(f[3], f[2], f[1], f[0])]
        # the stored array holds Schwarz-screened zeros where the block is tiny
        worst = max(worst, float(np.max(np.abs(np.where(got != 0.0, got - blk, 0.0)))))
    return worst


def evaluate(model: Model, X, T, margin_b: float = 1.5):
    """Order and bound statistics of the model on samples (X, T): RMSEs, the
    fraction of quartets whose order is under-predicted (after the margin), and
    the fraction whose bound (plus margin_b) falls below the true log10 bound."""
    Y = forward(model.params, model.normalise(X))
    pred = np.ceil(np.maximum(Y[:, 1] + model.margin, 0.0)).astype(int)
    return {"bound_rmse": float(np.sqrt(np.mean((Y[:, 0] - T[:, 0]) ** 2))),
            "order_rmse": float(np.sqrt(np.mean((Y[:, 1] - T[:, 1]) ** 2))),
            "under_orders": float(np.mean(pred < T[:, 1])),
            "under_bound": float(np.mean(Y[:, 0] + margin_b < T[:, 0])), "n": int(len(X))}


def main():
    """Generate the training data with the plain quadrature on tori with 1, 2
    and 3 periodic directions, train on the GPU, check on held-out tori (another
    seed) and save the model (python -m william_ci_amd.learned [n_tori] [path])."""
    import sys
    n_tori = int(sys.argv[1]) if len(sys.argv) > 1 else 1200
    path = sys.argv[2] if len(sys.argv) > 2 else MODEL_PATH
    log = lambda m: print(m, flush=True)
    X, T = make_dataset(n_tori, log=log)
    log(f"orders needed: {np.bincount(T[:, 1].astype(int), minlength=len(ORDERS) + 1).tolist()}")
    model, hist = fit(X, T, log=log)
    train = evaluate(model, X, T)
    log(f"train set: {train}")
    held = {}
    for nd in (1, 2, 3):
        Xh, Th = make_dataset(max(12, n_tori // 12), seed=777 + nd, ndps=(nd,))
        held[nd] = evaluate(model, Xh, Th)
        log(f"held-out Tori{nd}D: {held[nd]}")
    model.save(path, loss=hist, n_samples=np.int64(len(X)), under=np.float64(train["under_orders"]),
               order_hist=np.bincount(T[:, 1].astype(int), minlength=len(ORDERS) + 1),
               heldout=np.array([[held[nd][k] for k in ("bound_rmse", "order_rmse", "under_orders", "under_bound",
                                                        "n")] for nd in (1, 2, 3)]))
    log(f"saved {path}")


if __name__ == "__main__":
    main()
