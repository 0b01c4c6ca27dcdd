"""Warm-start safety on the GPU.

The converged policy stops on the average residual.  A warm start (previous
pressure, caller guesses, learned) can leave a few over-pressured particles
that the average hides; they gain huge velocities, the CFL step collapses and
the run blows up.  Warm-started solves therefore converge only once the
largest per-particle residual is within a bound too, and a warm-started solve
that misses that at max_iter is redone from zero pressure.

Bars: a planted absurd guess is iterated out (no redo) and leaves no residual
above the bound; a guess too large to iterate out within max_iter is redone,
and the redone step is bit-identical to the plain (zero-start) step from the
same state; with the bound disabled the guess is accepted; the 1M dam break
with the previous-pressure start stays bounded over 300 steps.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import dam_break_small

pytestmark = pytest.mark.gpu


def _scene(D):
    x, lo, hi = dam_break_small(16)
    sim = D.Sim(D.make_params(0.01, lo, hi, capacity=x.shape[0]))
    sim.init_scene(x)
    return sim, x


def test_outlier_guess_is_iterated_out(dfsph_mod):
    D = dfsph_mod
    print("case: planted outlier guess", flush=True)
    sim, x = _scene(D)
    plain = D.default_policy(D.POLICY_CONVERGED)
    for _ in range(3):
        sim.step(plain)
    guess = np.zeros(x.shape[0], np.float32)
    guess[1234] = 1e7                          # one absurd pressure in the bulk
    sim.set_warm_start(None, guess)
    ext = D.default_policy(D.POLICY_CONVERGED, warm_start=D.WARM_EXTERNAL)
    st = sim.step(ext)
    print(f"  fallback {st['warm_fallback']}, iters {st['iters_density']}, "
          f"max residual {st['max_residual_density']:.3e}, err {st['err_density']:.2e}")
    assert st["warm_fallback"] == 0 and st["converged_density"] == 1
    assert st["max_residual_density"] <= 100 * ext.tol_density
    assert st["err_density"] <= ext.tol_density
    sim.close()


def test_outlier_guess_is_redone_from_zero(dfsph_mod):
    D = dfsph_mod
    print("case: outlier guess beyond max_iter", flush=True)
    plain = D.default_policy(D.POLICY_CONVERGED, max_iter_density=6)
    ref, x = _scene(D)
    for _ in range(3):
        ref.step(plain)
    st_ref = ref.step(plain)
    x_ref = ref.export(D.FIELD_POSITION)
    ref.close()

    sim, _ = _scene(D)
    for _ in range(3):
        sim.step(plain)
    guess = np.zeros(x.shape[0], np.float32)
    guess[1234] = 1e12                         # too large to iterate out in 6 iterations
    sim.set_warm_start(None, guess)
    ext = D.default_policy(D.POLICY_CONVERGED, warm_start=D.WARM_EXTERNAL, max_iter_density=6)
    st = sim.step(ext)
    print(f"  fallback {st['warm_fallback']}, iters {st['iters_density']} vs plain {st_ref['iters_density']}")
    assert st_ref["converged_density"] == 1
    assert st["warm_fallback"] == 1 and st["converged_density"] == 1
    assert np.array_equal(sim.export(D.FIELD_POSITION), x_ref)
    sim.close()


def test_bound_can_be_disabled(dfsph_mod):
    D = dfsph_mod
    print("case: outlier bound disabled", flush=True)
    sim, x = _scene(D)
    plain = D.default_policy(D.POLICY_CONVERGED)
    for _ in range(3):
        sim.step(plain)
    guess = np.zeros(x.shape[0], np.float32)
    guess[1234] = 1e7
    sim.set_warm_start(None, guess)
    ext = D.default_policy(D.POLICY_CONVERGED, warm_start=D.WARM_EXTERNAL, residual_bound=-1.0)
    st = sim.step(ext, check=False)
    print(f"  accepted: fallback {st['warm_fallback']}, max residual {st['max_residual_density']:.3e}, "
          f"err {st['err_density']:.2e}")
    assert st["warm_fallback"] == 0
    assert st["max_residual_density"] > 100 * ext.tol_density
    sim.close()


def test_previous_warm_start_stays_bounded_1m(dfsph_mod):
    D = dfsph_mod
    from dfsph import scenes
    print("case: previous warm start, 1M, 300 steps", flush=True)
    cfg = scenes.load("c2_dam_break_1m")
    sim = D.Sim(scenes.params(cfg))
    sim.init_scene(scenes.positions(cfg))
    pol = scenes.policy(cfg, warm_start=D.WARM_PREVIOUS)
    vmax, fb = 0.0, 0
    for _ in range(300):
        st = sim.step(pol)
        assert st["converged_density"] == 1 and st["err_density"] <= cfg["tol_density"]
        vmax = max(vmax, st["max_speed"])
        fb += st["warm_fallback"]
    print(f"  max speed {vmax:.2f} m/s, {fb} solves redone from zero")
    assert vmax < 15.0
    sim.close()


def test_div_start_selects_the_divergence_start(dfsph_mod):
    """policy.div_start names the divergence solve's start apart from the
    density solve's: naming the start warm_start already implies changes
    nothing (bit for bit over several steps); a different one changes the
    divergence solve only where the starts differ, and every step converges.
    Batches refuse it."""
    D = dfsph_mod
    print("case: div_start on a 32k dam break", flush=True)
    x, lo, hi = dam_break_small(32)
    theta = None

    def run(warm, div_start, steps=6):
        nonlocal theta
        sim = D.Sim(D.make_params(0.01, lo, hi, capacity=x.shape[0]))
        sim.init_scene(x)
        if theta is None:
            sim.mlp_init(3)
            theta = sim.mlp_get_params()
        sim.mlp_set_params(theta)
        pol = D.default_policy(D.POLICY_CONVERGED, warm_start=warm, div_start=div_start)
        stats = [sim.step(pol) for _ in range(steps)]
        out = [sim.export(f) for f in (D.FIELD_POSITION, D.FIELD_VELOCITY, D.FIELD_PRESSURE_DIV)]
        sim.close()
        assert all(s["converged_div"] and s["converged_density"] for s in stats)
        return [(s["iters_div"], s["iters_density"]) for s in stats], out

    same = [(D.WARM_PREVIOUS, D.DIV_START_PREVIOUS), (D.WARM_ZERO, D.DIV_START_ZERO),
            (D.WARM_MLP, D.DIV_START_MLP), (D.WARM_MLP_DENSITY, D.DIV_START_PREVIOUS)]
    for warm, ds in same:
        a_it, a = run(warm, D.DIV_START_DEFAULT)
        b_it, b = run(warm, ds)
        assert a_it == b_it and all(np.array_equal(p, q) for p, q in zip(a, b)), (warm, ds)
    # the density solve warm, the divergence solve from zero: the first step
    # (no previous pressure yet) matches the default; later ones differ
    d_it, d = run(D.WARM_PREVIOUS, D.DIV_START_ZERO)
    p_it, p = run(D.WARM_PREVIOUS, D.DIV_START_DEFAULT)
    assert d_it[0] == p_it[0]
    assert not all(np.array_equal(u, v) for u, v in zip(d, p))
    print(f"  iterations, previous start with the divergence solve from zero: {d_it}; default: {p_it}")
    bad = D.default_policy(D.POLICY_CONVERGED, div_start=4)
    sim = D.Sim(D.make_params(0.01, lo, hi, capacity=x.shape[0]))
    sim.init_scene(x)
    with pytest.raises(D.DfsphError) as e:
        sim.step(bad)
    assert e.value.code == D.E_INVALID
    sim.close()
