"""Pins the CPU oracle (oracle/dfsph_oracle.cpp) before it is trusted as the
checker of the HIP path.  Every check here is independent of the oracle's own
code: the reference's golden vectors and compiled driver, closed forms of the
cubic kernel on a lattice, an operator identity derived from the reference's
shader algebra, brute force on small inputs, and textbook invariants.
"""
from __future__ import annotations

import json
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden", "cubic_kernel_expected.json")
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "cat3_driver")


# ---- cubic spline ------------------------------------------------------------

def test_cubic_kernel_golden_vectors(oracle_mod):
    """tools/integrity/integrity/cat3_numerical/expected_values.json (h = 1)."""
    O = oracle_mod
    with open(GOLDEN) as f:
        g = json.load(f)
    atol, rtol = g["tolerance"]["atol"], g["tolerance"]["rtol"]
    for p in g["points"]:
        for fn in (O.kernel_f64, O.kernel_f32):
            W, G = fn(p["q"], p["h"])
            assert abs(W - p["W"]) <= atol + rtol * abs(p["W"]), (p, W)
            assert abs(G - p["gradW_magnitude"]) <= atol + rtol * abs(p["gradW_magnitude"]), (p, G)


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="oracle/_ref not built (reference not mounted)")
def test_cubic_kernel_vs_reference_driver(oracle_mod):
    """The reference's own C++ transcription (tools/integrity/drivers/
    integrity_cat3_stack_c/main.cpp), compiled from its source, on a dense q grid."""
    O = oracle_mod
    qs = np.linspace(0.0, 1.2, 97)
    hs = [1.0, 0.04, 0.02, 0.01]
    args = []
    for h in hs:
        for q in qs:
            args += [repr(float(q)), repr(h)]
    out = json.loads(subprocess.run([REF_DRIVER] + args, capture_output=True, text=True,
                                    check=True).stdout)["evaluations"]
    for ev in out:
        W, G = O.kernel_f64(ev["q"], ev["h"])
        assert W == pytest.approx(ev["W"], rel=1e-12, abs=1e-12 * ev["W"] + 1e-300)
        assert G == pytest.approx(ev["gradW_magnitude"], rel=1e-12, abs=1e-300)


def test_cubic_kernel_known_answers(oracle_mod):
    """W(0) = 8/pi, W(0.5) = 2/pi, W(1) = gradW(1) = 0 (tools/integrity/tests/
    test_cat3_cubic_kernel.py:43-59, :99-102)."""
    O = oracle_mod
    assert O.kernel_f64(0.0, 1.0)[0] == pytest.approx(8 / math.pi, abs=1e-12)
    assert O.kernel_f64(0.5, 1.0)[0] == pytest.approx(2 / math.pi, abs=1e-12)
    assert O.kernel_f64(1.0, 1.0) == (0.0, 0.0)


# ---- Morton keys and the counting sort ------------------------------------

def _morton_bitloop(x, y, z):
    k = 0
    for b in range(10):
        k |= ((x >> b) & 1) << (3 * b + 2)
        k |= ((y >> b) & 1) << (3 * b + 1)
        k |= ((z >> b) & 1) << (3 * b)
    return k


def test_morton_known_values_and_bitloop(oracle_mod):
    O = oracle_mod
    assert O.morton3(1, 0, 0) == 4 and O.morton3(0, 1, 0) == 2 and O.morton3(0, 0, 1) == 1
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(0, 1024, size=(2000, 3)):
        assert O.morton3(int(x), int(y), int(z)) == _morton_bitloop(int(x), int(y), int(z))


def test_hash_matches_stable_sort(oracle_mod):
    """keys, sorted order (stable by key then index) and cell starts with the
    sentinel, against numpy's stable argsort (main.cpp:1452-1493 semantics)."""
    O = oracle_mod
    cfg = O.make_config(0.01, (0.0, 0.0, 0.0), (1.0, 0.7, 0.5), boundary=False)
    d = O.derive(cfg)
    rng = np.random.default_rng(1)
    x = (rng.random((5000, 3)) * np.array([1.1, 0.8, 0.6]) - 0.05).astype(np.float32)
    keys, srt, starts = O.hash_reference(cfg, x)
    cell = np.float32(d.cell)
    c = np.clip((x - np.float32(0.0)) / cell, np.float32(0.0),
                np.array(d.axes, np.float32) - np.float32(1.0)).astype(np.uint32)
    ref_keys = np.array([_morton_bitloop(*map(int, cc)) for cc in c], np.uint32)
    assert np.array_equal(keys, ref_keys)
    assert np.array_equal(srt, np.argsort(keys, kind="stable").astype(np.uint32))
    ref_starts = np.searchsorted(np.sort(keys), np.arange(d.total_cells + 1)).astype(np.uint32)
    assert np.array_equal(starts, ref_starts) and starts[-1] == len(x)


# ---- lattice closed forms and operator identities -----------------------------

def _lattice(n, r, origin):
    """Perfect lattice at spacing 2r; r a power of two so positions are exact."""
    i = np.arange(n)
    g = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3)
    return (origin + (g + 0.5) * 2 * r).astype(np.float32), g


R_EXACT = 2.0 ** -7


def _center_index(g, n):
    return int(np.flatnonzero((g == n // 2).all(1))[0])


def test_interior_density_and_alpha_closed_form(oracle_mod):
    """Interior of a perfect lattice, h = 4r, m = rho0 (2r)^3:
    rho = (rho0/pi) [1 + 6/4 + 24 (1 - sqrt2/2)^3 + 16 (1 - sqrt3/2)^3] = 999.972466...;
    by symmetry sum m gradW = 0, so alpha = 1 / sum |m gradW|^2."""
    O = oracle_mod
    n, r = 9, R_EXACT
    x, g = _lattice(n, r, 4 * r)
    cfg = O.make_config(r, (0, 0, 0), (1.0, 1.0, 1.0), boundary=False)
    sim = O.OracleSim.create(cfg, x, fp64=True)
    sim.prepare()
    c = _center_index(g, n)
    q2, q3 = math.sqrt(2) / 2, math.sqrt(3) / 2
    rho_cf = 1000.0 / math.pi * (1 + 6 * 0.25 + 24 * (1 - q2) ** 3 + 16 * (1 - q3) ** 3)
    assert rho_cf == pytest.approx(999.972466, abs=1e-6)
    d = sim.derived()
    rho = sim.get(O.F_RHO)[c]
    assert rho == pytest.approx(rho_cf, rel=1e-6)   # float32 constants of the reference
    m, gn = d.mass, d.gnorm
    s = 6 * (1.5 * gn) ** 2 + 12 * (6 * (1 - q2) ** 2 * gn) ** 2 + 8 * (6 * (1 - q3) ** 2 * gn) ** 2
    assert sim.get(O.F_ALPHA)[c] == pytest.approx(1.0 / (m * m * s), rel=1e-6)
    assert np.abs(sim.get(O.F_GF)[c]).max() < 1e-9 * m * 1.5 * gn


@pytest.mark.parametrize("mode_k,power", [(0, 2), (1, 1)])
def test_jacobi_diagonal_identity(oracle_mod, mode_k, power):
    """(A e_c)_c = -dt^k / (rho_c^3 alpha_c) in the lattice interior, from the
    operator of compute_pressure_accel.comp.glsl:84-157 and
    compute_aij_pj.comp.glsl:89-163 (the diagonal the converged policy inverts)."""
    O = oracle_mod
    n, r, dt = 9, R_EXACT, 1e-3
    x, g = _lattice(n, r, 4 * r)
    sim = O.OracleSim.create(O.make_config(r, (0, 0, 0), (1.0, 1.0, 1.0), boundary=False), x, fp64=True)
    sim.prepare()
    c = _center_index(g, n)
    e = np.zeros(len(x))
    e[c] = 1.0
    ap, _ = sim.apply_A(e, mode_k, dt)
    rho, alpha = sim.get(O.F_RHO)[c], sim.get(O.F_ALPHA)[c]
    assert ap[c] == pytest.approx(-dt ** power / (rho ** 3 * alpha), rel=1e-9)


def test_uniform_pressure_gives_zero_interior_accel_and_A_is_linear(oracle_mod):
    O = oracle_mod
    n, r = 9, R_EXACT
    x, g = _lattice(n, r, 4 * r)
    sim = O.OracleSim.create(O.make_config(r, (0, 0, 0), (1.0, 1.0, 1.0), boundary=False), x, fp64=True)
    sim.prepare()
    c = _center_index(g, n)
    ap, acc = sim.apply_A(np.full(len(x), 1e5), 0, 1e-3)
    scale = np.abs(acc).max()
    assert np.abs(acc[c]).max() <= 1e-12 * scale
    rng = np.random.default_rng(3)
    p1, p2 = rng.random(len(x)) * 1e4, rng.random(len(x)) * 1e4
    a1, _ = sim.apply_A(p1, 0, 1e-3)
    a2, _ = sim.apply_A(p2, 0, 1e-3)
    a12, _ = sim.apply_A(2 * p1 + 3 * p2, 0, 1e-3)
    assert np.abs(a12 - (2 * a1 + 3 * a2)).max() <= 1e-10 * np.abs(a12).max()


# ---- brute force on small inputs -----------------------------------------------

def _jittered(n, r, seed, origin):
    rng = np.random.default_rng(seed)
    x, _ = _lattice(n, r, origin)
    return (x + rng.uniform(-0.3 * r, 0.3 * r, x.shape)).astype(np.float32)


def test_neighbour_counts_brute_force(oracle_mod):
    """j != i with sqrt(r^2)/h < 1 in IEEE float (density_alpha.comp.glsl:125-128), O(N^2)."""
    O = oracle_mod
    r = 0.01
    x = _jittered(10, r, 4, 0.05)
    sim = O.OracleSim.create(O.make_config(r, (0, 0, 0), (0.5, 0.5, 0.5), boundary=False), x)
    sim.step(0.0, O.policy("reference"))
    h = np.float32(4 * r)
    dx = x[:, None, :] - x[None, :, :]
    r2 = (dx[..., 0] * dx[..., 0] + dx[..., 1] * dx[..., 1]) + dx[..., 2] * dx[..., 2]
    nb = (np.sqrt(r2) / h < np.float32(1.0)).sum(1) - 1
    assert np.array_equal(sim.get(O.F_NBR), nb.astype(np.uint32))


def test_density_brute_force_fp64(oracle_mod):
    O = oracle_mod
    r = 0.01
    x = _jittered(10, r, 5, 0.05)
    sim = O.OracleSim.create(O.make_config(r, (0, 0, 0), (0.5, 0.5, 0.5), boundary=False), x, fp64=True)
    sim.prepare()
    d = sim.derived()
    xd = x.astype(np.float64)
    q = np.sqrt(((xd[:, None, :] - xd[None, :, :]) ** 2).sum(-1)) / d.h
    Wv = np.vectorize(lambda qq: O.kernel_f64(qq, d.h)[0])(q)
    assert np.allclose(sim.get(O.F_RHO), d.mass * Wv.sum(1), rtol=1e-6, atol=0)


def test_boundary_volume_brute_force(oracle_mod):
    """V_b = 1 / max(W(0) + sum_{b' != b, q < 1} W, 1e-9) (compute_boundary_volume.comp.glsl:55-92)."""
    O = oracle_mod
    r = 0.01
    lo, hi = (0.0, 0.0, 0.0), (0.2, 0.16, 0.12)
    sim = O.OracleSim.create(O.make_config(r, lo, hi), np.zeros((1, 3), np.float32) + 0.05, fp64=True)
    bx = sim.get(O.F_BX)
    d = sim.derived()
    q = np.sqrt(((bx[:, None, :] - bx[None, :, :]) ** 2).sum(-1)) / d.h
    Wv = np.vectorize(lambda qq: O.kernel_f64(qq, d.h)[0])(q)   # includes W(0) on the diagonal
    assert np.allclose(sim.get(O.F_BV), 1.0 / np.maximum(Wv.sum(1), 1e-9), rtol=1e-6)


def test_boundary_samples_invariants(oracle_mod):
    """Every sample on the AABB surface, no two closer than half a spacing."""
    O = oracle_mod
    lo, hi, r = np.array([0.0, 0.0, 0.0]), np.array([0.3, 0.22, 0.18]), 0.01
    b = O.boundary_samples(lo, hi, r).astype(np.float64)
    on = (np.isclose(b, lo) | np.isclose(b, hi)).any(1)
    assert on.all()
    dd = np.sqrt(((b[:, None] - b[None]) ** 2).sum(-1)) + np.eye(len(b))
    assert dd.min() >= 0.5 * 2 * r - 1e-6


# ---- trajectories and failure handling --------------------------------------------

def test_free_fall_closed_form(oracle_mod):
    """A lone particle: v_n = n g dt, y_n = y_0 + g dt^2 n(n+1)/2 (semi-implicit Euler)."""
    O = oracle_mod
    x = np.array([[0.5, 0.9, 0.5]], np.float32)
    sim = O.OracleSim.create(O.make_config(0.01, (0, 0, 0), (1, 1, 1), boundary=False), x, fp64=True)
    dt, g = 1e-3, np.float64(np.float32(-9.81))
    for k in range(1, 31):
        sim.step(dt, O.policy("reference"))
    y = sim.get(O.F_X)[0, 1]
    v = sim.get(O.F_V)[0, 1]
    dtf = np.float64(np.float32(dt))
    assert v == pytest.approx(30 * g * dtf, rel=1e-12)
    assert y == pytest.approx(np.float64(np.float32(0.9)) + g * dtf ** 2 * 30 * 31 / 2, rel=1e-12)


def test_reference_mode_pressure_is_inert(oracle_mod):
    """The reference's coefficient alpha/rho^2 leaves the column in free fall
    (mean v_y = n g dt): the defect the converged policy exists for."""
    O = oracle_mod
    r = 0.01
    x = _jittered(8, r, 6, 2 * r)
    sim = O.OracleSim.create(O.make_config(r, (0, 0, 0), (0.5, 0.5, 0.5)), x)
    for _ in range(5):
        sim.step(1e-3, O.policy("reference"))
    assert sim.get(O.F_V)[:, 1].mean() == pytest.approx(5 * -9.81e-3, rel=1e-3)


def test_converged_mode_meets_tolerance(oracle_mod):
    O = oracle_mod
    r = 0.01
    x = _jittered(10, r, 7, 2 * r)
    sim = O.OracleSim.create(O.make_config(r, (0, 0, 0), (0.5, 0.5, 0.5)), x)
    for _ in range(4):
        st = sim.step(0.0, O.policy("converged"))
        assert st["converged_density"] == 1 and st["err_density"] <= 1e-3
        assert st["converged_div"] == 1 and st["err_div"] <= 1e-3


def test_clamp_and_nonfinite(oracle_mod):
    """integrate_forces.comp.glsl:81-99: beyond dmax + h/8 -> snapped to dmax with
    outward velocity zeroed; a NaN input is counted and reported (status -5)."""
    O = oracle_mod
    x = np.array([[0.5, 0.5, 0.5], [0.2, 0.2, 0.2]], np.float32)
    v = np.array([[100.0, 0.0, 0.0], [0.0, 0.0, 0.0]], np.float32)
    sim = O.OracleSim.create(O.make_config(0.01, (0, 0, 0), (0.55, 1, 1), boundary=False,
                                           gravity=(0, 0, 0)), x, v)
    st = sim.step(1e-3, O.policy("reference"))
    assert st["status"] == 0
    assert sim.get(O.F_X)[0, 0] == pytest.approx(0.55) and sim.get(O.F_V)[0, 0] == 0.0
    x[1, 2] = np.nan
    sim = O.OracleSim.create(O.make_config(0.01, (0, 0, 0), (1, 1, 1), boundary=False), x)
    st = sim.step(1e-3, O.policy("reference"))
    assert st["status"] == -5 and st["nonfinite"] >= 1


def test_max_iter_reports_not_converged(oracle_mod):
    O = oracle_mod
    r = 0.01
    x = _jittered(10, r, 8, 2 * r)
    sim = O.OracleSim.create(O.make_config(r, (0, 0, 0), (0.5, 0.5, 0.5)), x)
    st = sim.step(0.0, O.policy("converged", max_iter_density=1, min_iter_density=1, tol_density=1e-9))
    assert st["converged_density"] == 0 and st["iters_density"] == 1
