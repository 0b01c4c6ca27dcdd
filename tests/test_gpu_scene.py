"""Emitters and preset scenes on the GPU.

Bars: emitted particles sit bit for bit where the oracle's restatement of
shaders/apply_emitter.comp.glsl:29-65 puts them, with the emitter velocity,
flags 1 and the next original indices; emission stops at capacity; the grown
scene steps (and converges) with the new particles; a preset scene (brick +
droplet, initial_fill) steps converged and keeps its droplet flags.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle as O
from conftest import dam_break_small

pytestmark = pytest.mark.gpu


def test_emit_matches_oracle_and_steps(dfsph_mod, tmp_path):
    D = dfsph_mod
    print("case: emitters", flush=True)
    x, lo, hi = dam_break_small(12)
    n = x.shape[0]
    sim = D.Sim(D.make_params(0.01, lo, hi, capacity=n + 100))
    sim.init_scene(x)
    pol = D.default_policy(D.POLICY_CONVERGED)
    for _ in range(2):
        sim.step(pol)
    em = np.array([[1.0, 0.6, 0.5, 0.05, 0.0, -10.0, 0.0, 5000.0],
                   [1.5, 0.7, 0.4, 0.03, -0.5, 0.0, 0.2, 100.0]], np.float32)
    assert sim.emit(em, 64) == 64
    xo, vo = O.emit(em, n, 64)
    assert np.array_equal(sim.export(D.FIELD_POSITION)[n:], xo)
    assert np.array_equal(sim.export(D.FIELD_VELOCITY)[n:], vo)
    ids = sim.export(D.FIELD_ID)
    assert np.array_equal(np.sort(ids), np.arange(n + 64, dtype=np.uint32))
    with pytest.raises(D.DfsphError):                 # per-step fields wait for the next step
        sim.export(D.FIELD_DENSITY)
    st = sim.step(pol)
    assert st["converged_density"] == 1
    assert sim.export(D.FIELD_DENSITY).shape[0] == n + 64
    # the CFL step saw the emitted speeds (10 m/s: 0.4 * 2r / 10 < dt_max = 1e-3)
    assert st["dt"] <= 0.4 * 0.02 / 10.0 * (1 + 1e-6)
    assert sim.emit(em, 1000) == 36                   # capacity n + 100
    assert sim.emit(em, 5) == 0
    sim.step(pol)
    d = sim.capture_save(str(tmp_path), 1)
    rec = np.fromfile(os.path.join(d, "particles.bin"), np.uint32).reshape(-1, 32)
    assert np.all(rec[:n, 22] == 0) and np.all(rec[n:, 22] == 1)
    for bad in (np.zeros((0, 8), np.float32), np.zeros((9, 8), np.float32)):
        with pytest.raises(D.DfsphError):
            sim.emit(bad, 1)
    nan = em.copy()
    nan[0, 0] = np.nan
    with pytest.raises(D.DfsphError):
        sim.emit(nan, 1)
    sim.close()


def test_step_after_emit_rebuilds_cells(dfsph_mod):
    """A step leaves the next step's cell build on the stream; emitting
    particles in between must discard it: the next step's neighbour counts
    equal the oracle's for the grown scene, bit for bit."""
    D = dfsph_mod
    print("case: step after emit rebuilds cells", flush=True)
    x, lo, hi = dam_break_small(12)
    n = x.shape[0]
    sim = D.Sim(D.make_params(0.01, lo, hi, capacity=n + 64, viscosity=0.01))
    sim.init_scene(x)
    pol = D.default_policy(D.POLICY_REFERENCE)
    for _ in range(2):
        sim.step(pol, dt=1e-4)
    em = np.array([[0.3, 0.3, 0.3, 0.05, 0.0, -1.0, 0.0, 5000.0]], np.float32)
    assert sim.emit(em, 64) == 64
    xe, ve = sim.export(D.FIELD_POSITION), sim.export(D.FIELD_VELOCITY)
    ora = O.OracleSim.create(O.make_config(0.01, lo, hi, viscosity=0.01), xe, ve)
    sim.step(pol, dt=1e-4)
    ora.step(1e-4, O.policy("reference"))
    assert np.array_equal(sim.export(D.FIELD_NEIGHBOR_COUNT), ora.get(O.F_NBR))
    sim.close()


def test_droplet_preset_steps(dfsph_mod):
    D = dfsph_mod
    print("case: droplet-impact preset", flush=True)
    p = D.preset(2)
    r = 0.02
    x, v, f, nb = D.preset_fill(p, r, 1 << 20)
    assert x.shape[0] > nb                             # brick and droplet
    sim = D.Sim(D.make_params(r, tuple(p.domain_min), tuple(p.domain_max), capacity=x.shape[0],
                              viscosity=p.viscosity, cohesion=p.cohesion, vorticity=p.vorticity))
    sim.init_scene(x, v, flags=f)
    pol = D.default_policy(D.POLICY_CONVERGED)
    for _ in range(5):
        st = sim.step(pol)
        assert st["converged_density"] == 1
    xs = sim.export(D.FIELD_POSITION)
    assert np.all(np.isfinite(xs))
    # the droplet moves down at ~5 m/s while the pool stays
    assert xs[nb:, 1].mean() < x[nb:, 1].mean() - 4.0 * st["dt"]
    sim.close()


@pytest.mark.parametrize("dom,r,cap", [
    (((0.0, 0.0, 0.0), (2.0, 1.0, 1.0)), 0.01, 0),
    (((-1.5, -1.0, -1.5), (1.5, 2.0, 1.5)), 0.02, 0),
    (((0.1, -0.3, 0.25), (0.537, 0.411, 1.093)), 0.0073, 0),
    (((0.0, 0.0, 0.0), (0.4, 0.03, 0.4)), 0.01, 0),         # thin domain: side faces and y edges skipped
    (((0.0, 0.0, 0.0), (2.0, 1.0, 1.0)), 0.01, 20000),      # capped mid-face
])
def test_device_boundary_sampler(dfsph_mod, dom, r, cap):
    """Samples placed on the device equal the host sampler and the oracle's
    restatement of main.cpp:1332-1431, bit for bit and in sample order."""
    D = dfsph_mod
    print(f"case: boundary sampler {dom} r={r} cap={cap}", flush=True)
    lo, hi = dom
    sim = D.Sim(D.make_params(r, lo, hi, capacity=16))
    sim.init_scene(np.array([[0.5 * (lo[a] + hi[a]) for a in range(3)]], np.float32), boundary_cap=cap)
    got = sim.export(D.FIELD_BOUNDARY_POSITION)
    ref = O.boundary_samples(lo, hi, r, cap)
    assert got.shape == ref.shape and np.array_equal(got, ref)
    assert np.array_equal(got, D.boundary_samples(lo, hi, r, cap))
    sim.close()
