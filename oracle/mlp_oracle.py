"""FP64 CPU oracle of the learned warm-start network (test infrastructure only).

Network: 16 features -> 64 -> 64 -> 64 (ReLU) -> 2 outputs (kappa_v, kappa_rho),
9,538 parameters, flat parameter order [W1 (16x64, row k = input), b1, W2 (64x64),
b2, W3 (64x64), b3, W4 (64x2), b4].  Loss: mean over the batch of the squared
error summed over the two outputs.

`bf16=True` rounds to bfloat16 exactly where the device kernels store BF16
(inputs, weights of layers 1-3, hidden activations after ReLU), so the device's
FP32-accumulated result can be compared at accumulation-order precision.
The analytic backward pass here is itself pinned by central finite differences
in FP64 (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import numpy as np

NF, NH, NO = 16, 64, 2
SHAPES = [("W1", (NF, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
          ("W3", (NH, NH)), ("b3", (NH,)), ("W4", (NH, NO)), ("b4", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[o:o + n], np.float64).reshape(shp)
        o += n
    return out


def pack(params) -> np.ndarray:
    return np.concatenate([np.asarray(params[n], np.float64).ravel() for n, _ in SHAPES])


def bf16_round(x):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> np.uint64(16)) & np.uint64(1)
    r = ((u + np.uint64(0x7FFF) + lsb) >> np.uint64(16)) << np.uint64(16)
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(a)
    if nan.any():
        out[nan] = np.nan
    return out


def init(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases (the initialisation the library uses is its own;
    this one is for tests)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / shp[0])
            p[name] = rng.uniform(-lim, lim, shp)
        else:
            p[name] = np.zeros(shp)
    return pack(p)


def forward(theta, X, bf16=False):
    P = unpack(theta)
    rd = bf16_round if bf16 else (lambda v: np.asarray(v, np.float64))
    x = rd(X)
    h1p = x @ rd(P["W1"]) + P["b1"]
    h1 = rd(np.maximum(h1p, 0.0))
    h2p = h1 @ rd(P["W2"]) + P["b2"]
    h2 = rd(np.maximum(h2p, 0.0))
    h3p = h2 @ rd(P["W3"]) + P["b3"]
    h3 = rd(np.maximum(h3p, 0.0))
    y = h3 @ P["W4"] + P["b4"]
    return y, (x, h1p, h1, h2p, h2, h3p, h3)


def loss_and_grad(theta, X, T, bf16=False):
    P = unpack(theta)
    rd = bf16_round if bf16 else (lambda v: np.asarray(v, np.float64))
    y, (x, h1p, h1, h2p, h2, h3p, h3) = forward(theta, X, bf16)
    B = X.shape[0]
    d = y - np.asarray(T, np.float64)
    loss = float((d * d).sum() / B)
    dy = 2.0 * d / B
    g = {"W4": h3.T @ dy, "b4": dy.sum(0)}
    dh3 = (dy @ P["W4"].T) * (h3p > 0)
    dh3 = rd(dh3)
    g["W3"] = h2.T @ dh3
    g["b3"] = dh3.sum(0)
    dh2 = rd((dh3 @ rd(P["W3"]).T) * (h2p > 0))
    g["W2"] = h1.T @ dh2
    g["b2"] = dh2.sum(0)
    dh1 = rd((dh2 @ rd(P["W2"]).T) * (h1p > 0))
    g["W1"] = x.T @ dh1
    g["b1"] = dh1.sum(0)
    return loss, pack(g)
