"""The LDS-staged walk of the solver-phase sweeps (Jacobi, forces, density
source) against the same sweeps gathering from memory
(DFSPH_FLAG_NO_STAGED_WINDOWS).  Both visit the same neighbours in the same
order with the same arithmetic, so every field is bit-identical, step after
step; the oracle comparisons of test_gpu_parity cover both paths.

Cases: the 32^3 dam break from rest, converged and reference policies, at
workgroups of 256 / 128 / 64 threads (four, two and one waves per LDS
region set); cells of 2h (64 particles a cell: windows beyond the LDS
capacity walk from memory, and runs beyond 32 slots re-test their
candidates); the 1M dam break over 12 steps (developing flow)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import dam_break_small

pytestmark = pytest.mark.gpu

FIELDS = ("FIELD_POSITION", "FIELD_VELOCITY", "FIELD_PRESSURE", "FIELD_PRESSURE_DIV", "FIELD_DENSITY_ADV",
          "FIELD_DENSITY_CHANGE", "FIELD_PRESSURE_ACCEL")


def _run(D, x, lo, hi, r, flags, steps, pol, check_every, **kw):
    sim = D.Sim(D.make_params(r, lo, hi, capacity=x.shape[0], grid=D.GRID_FINE, flags=flags, **kw))
    sim.init_scene(x)
    snaps, stats = [], []
    for s in range(steps):
        st = sim.step(pol)
        stats.append((st["iters_div"], st["iters_density"], st["err_div"], st["err_density"], st["dt"]))
        if (s + 1) % check_every == 0:
            snaps.append([sim.export(getattr(D, f)) for f in FIELDS])
    sim.close()
    return snaps, stats


def _assert_same(D, a, b):
    (sa, ta), (sb, tb) = a, b
    assert ta == tb
    for k, (fa, fb) in enumerate(zip(sa, sb)):
        for name, u, v in zip(FIELDS, fa, fb):
            assert np.array_equal(u, v), f"snapshot {k}: {name} differs"


@pytest.mark.parametrize("block", [256, 128, 64])
@pytest.mark.parametrize("mode", ["converged", "reference"])
def test_staged_walk_bit_identical(dfsph_mod, block, mode):
    D = dfsph_mod
    print(f"case: staged vs memory walk, 32^3 dam break, {mode}, block {block}", flush=True)
    x, lo, hi = dam_break_small(32)
    pol = D.default_policy(D.POLICY_CONVERGED if mode == "converged" else D.POLICY_REFERENCE)
    if mode == "converged":
        pol.warm_start = D.WARM_PREVIOUS
    runs = [_run(D, x, lo, hi, 0.01, fl, 30, pol, 10, block_size=block) for fl in (0, D.FLAG_NO_STAGED_WINDOWS)]
    _assert_same(D, *runs)


def test_staged_walk_overfull_windows(dfsph_mod):
    D = dfsph_mod
    print("case: staged vs memory walk, cells of 2h", flush=True)
    x, lo, hi = dam_break_small(24)
    pol = D.default_policy(D.POLICY_CONVERGED)
    runs = [_run(D, x, lo, hi, 0.01, fl, 12, pol, 4, fine_cell_scale=2.0)
            for fl in (0, D.FLAG_NO_STAGED_WINDOWS)]
    _assert_same(D, *runs)


def test_staged_walk_1m(dfsph_mod):
    D = dfsph_mod
    from dfsph import scenes

    print("case: staged vs memory walk, 1M dam break, 12 steps", flush=True)
    cfg = scenes.load("c2_dam_break_1m")
    x = scenes.positions(cfg)
    pol = scenes.policy(cfg, warm_start=D.WARM_PREVIOUS)
    out = []
    for fl in (0, D.FLAG_NO_STAGED_WINDOWS):
        sim = D.Sim(scenes.params(cfg, flags=fl))
        sim.init_scene(x)
        st = [sim.step(pol) for _ in range(12)]
        out.append(([(s["iters_div"], s["iters_density"], s["err_density"]) for s in st],
                    [[sim.export(getattr(D, f)) for f in FIELDS]]))
        sim.close()
    assert out[0][0] == out[1][0]
    for name, u, v in zip(FIELDS, out[0][1][0], out[1][1][0]):
        assert np.array_equal(u, v), name
