"""dfsph_mlp_grad_noalloc, the allocation-free variant of dfsph_mlp_grad.

It runs the same kernels over a static device scratch array instead of
allocated buffers, so its loss and gradient must equal dfsph_mlp_grad's bit
for bit: checked on two calls in a row, the second with fewer rows (the
array keeps the first call's contents), then on a call at the row limit and
the rejection of one row more.  Runs alone in a fresh process or after the
other GPU tests.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import dam_break_small

pytestmark = pytest.mark.gpu


def _case(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 16)).astype(np.float32)
    t = rng.normal(size=(n, 2)).astype(np.float32)
    return x, t


def test_noalloc_gradient_equals_allocating_entry_point(dfsph_mod):
    D = dfsph_mod
    x0, lo, hi = dam_break_small(8)
    sim = D.Sim(D.make_params(0.01, lo, hi, capacity=x0.shape[0]))
    sim.mlp_init(11)
    th = sim.mlp_get_params()
    rng = np.random.default_rng(4)
    sim.mlp_set_params((th + rng.normal(0, 0.02, th.shape)).astype(np.float32))   # non-zero biases
    for n, seed in ((3000, 1), (130, 2), (D.MLP_GRAD_NOALLOC_MAX_ROWS, 3)):
        print(f"case: mlp_grad_noalloc n={n}", flush=True)
        x, t = _case(n, seed)
        loss_a, g_a = sim.mlp_grad(x, t)
        loss_b, g_b = sim.mlp_grad_noalloc(x, t)
        assert np.all(np.isfinite(g_a)) and np.abs(g_a).max() > 0
        assert loss_b == loss_a, (n, loss_a, loss_b)
        assert np.array_equal(g_a, g_b), (n, int(np.sum(g_a != g_b)))
    x, t = _case(D.MLP_GRAD_NOALLOC_MAX_ROWS + 1, 5)
    with pytest.raises(D.DfsphError) as e:
        sim.mlp_grad_noalloc(x, t)
    assert e.value.code == D.E_INVALID
    # The context still works after the rejection.
    x, t = _case(64, 6)
    assert np.array_equal(sim.mlp_grad(x, t)[1], sim.mlp_grad_noalloc(x, t)[1])
    sim.close()
