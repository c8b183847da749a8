"""The per-quartet network 6 -> 64 -> 64 -> 2 (ReLU) in numpy FP64.  TEST
INFRASTRUCTURE ONLY: the checker for the engine's BF16 matrix-core forward
and backward kernels (william_ci_amd/csrc/tcg_mlp.hip).

With emulate=True the operands of every matrix product are rounded to BF16
(round to nearest even) exactly where the kernels round them: the inputs,
the weights, the hidden activations, the output error and the hidden
gradients; bias gradients are sums of the unrounded values.  Parameter
layout: W1[6][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2].
Loss = mean over samples of w0 (y0 - t0)^2 + w1 (y1 - t1)^2.
"""
import numpy as np

SHAPES = (("W1", (6, 64)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 2)), ("b3", (2,)))


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([d[name].ravel() for name, _ in SHAPES])


def bf16(x):
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, X, emulate=True):
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    W = unpack(p)
    h1 = np.maximum(r(X) @ r(W["W1"]) + W["b1"], 0.0)
    h2 = np.maximum(r(h1) @ r(W["W2"]) + W["b2"], 0.0)
    y = r(h2) @ r(W["W3"]) + W["b3"]
    return y, (h1, h2)


def loss_grad(p, X, T, w=(0.1, 1.0), emulate=True):
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    W = unpack(p)
    n = len(X)
    y, (h1, h2) = forward(p, X, emulate)
    wv = np.asarray(w, dtype=np.float64)
    e = y - T
    loss = float(np.sum(wv * e * e) / n)
    dY = 2.0 * wv * e / n
    dH2 = (r(dY) @ r(W["W3"]).T) * (h2 > 0)
    dH1 = (r(dH2) @ r(W["W2"]).T) * (h1 > 0)
    g = {"W3": r(h2).T @ r(dY), "b3": dY.sum(0), "W2": r(h1).T @ r(dH2), "b2": dH2.sum(0),
         "W1": r(X).T @ r(dH1), "b1": dH1.sum(0)}
    return loss, pack(g)
