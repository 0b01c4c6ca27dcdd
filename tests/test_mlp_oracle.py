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
