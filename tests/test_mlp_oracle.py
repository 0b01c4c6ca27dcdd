"""Pins the FP64 MLP oracle: analytic gradients against central finite
differences, and the bfloat16 rounding helper against the definition."""
from __future__ import annotations

import numpy as np

import mlp_oracle as M


def test_parameter_count():
    assert M.NPARAM == 9538   # 16*64+64 + 2*(64*64+64) + 64*2+2


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    theta = M.init(1) + rng.normal(0, 0.05, M.NPARAM)
    theta[M.NPARAM - 2:] = 0.3
    X = rng.normal(size=(37, M.NF))
    T = rng.normal(size=(37, M.NO))
    loss, g = M.loss_and_grad(theta, X, T)
    idx = rng.choice(M.NPARAM, 300, replace=False)
    eps = 1e-6
    for k in idx:
        e = np.zeros(M.NPARAM)
        e[k] = eps
        lp, _ = M.loss_and_grad(theta + e, X, T)
        lm, _ = M.loss_and_grad(theta - e, X, T)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_bf16_rounding_definition():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-30], np.float32)
    r = M.bf16_round(x)
    # bfloat16 keeps 8 significand bits: representable values are exact,
    # the tie at 1 + 2^-8 rounds to even (1.0), above it rounds up to 1 + 2^-7
    assert r[0] == 1.0
    assert r[1] == 1.0
    assert r[2] == 1.0078125
    assert M.bf16_round(np.array([1.01171875], np.float32))[0] == 1.015625   # tie 1+3*2^-8 -> even mantissa
    for v in x:
        rv = M.bf16_round(np.array([v], np.float32))[0]
        bits = np.array([rv], np.float32).view(np.uint32)[0]
        assert bits & 0xFFFF == 0
        assert abs(rv - v) <= abs(v) * 2 ** -8


def _block_slices():
    o, out = 0, {}
    for name, shp in M.SHAPES:
        k = int(np.prod(shp))
        out[name] = slice(o, o + k)
        o += k
    return out


def test_bf16_output_layer_gradient_matches_finite_differences():
    """In bf16 mode (the device's rounding points) the loss is piecewise
    constant in the parameters of layers 1-3 -- a step smaller than a bfloat16
    ulp of a weight or activation changes nothing, a larger one jumps -- so
    finite differences cannot check those blocks.  The output layer (W4, b4)
    is applied in FP32 to fixed rounded activations, so there the bf16-mode
    loss is smooth and every one of its 130 gradients is checked by central
    differences."""
    rng = np.random.default_rng(2)
    theta = M.init(3) + rng.normal(0, 0.05, M.NPARAM)
    X = rng.normal(size=(41, M.NF))
    T = rng.normal(size=(41, M.NO))
    _, g = M.loss_and_grad(theta, X, T, bf16=True)
    sl = _block_slices()
    eps = 1e-6
    for k in list(range(sl["W4"].start, sl["W4"].stop)) + list(range(sl["b4"].start, sl["b4"].stop)):
        e = np.zeros(M.NPARAM)
        e[k] = eps
        lp, _ = M.loss_and_grad(theta + e, X, T, bf16=True)
        lm, _ = M.loss_and_grad(theta - e, X, T, bf16=True)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_bf16_hidden_gradients_are_the_fp64_gradients_at_the_rounded_point():
    """For layers 1-3 the bf16-mode backward pass must be the FP64 backward
    pass (pinned by finite differences above) evaluated at the rounded
    inputs and weights, up to the 8-bit rounding of the hidden activations
    and their gradients: every block within 3e-2 of its largest magnitude.  A
    dropped term, a wrong transpose or a missing ReLU mask gives O(1)."""
    rng = np.random.default_rng(4)
    theta = M.init(5) + rng.normal(0, 0.05, M.NPARAM)
    X = rng.normal(size=(256, M.NF))
    T = rng.normal(size=(256, M.NO))
    _, gb = M.loss_and_grad(theta, X, T, bf16=True)
    sl = _block_slices()
    rounded = theta.copy()
    for name in ("W1", "W2", "W3"):
        rounded[sl[name]] = M.bf16_round(theta[sl[name]])
    _, gf = M.loss_and_grad(rounded, M.bf16_round(X), T, bf16=False)
    for name, _ in M.SHAPES:
        a, b = gb[sl[name]], gf[sl[name]]
        e = np.abs(a - b).max() / np.abs(b).max()
        assert e <= 3e-2, (name, e)
