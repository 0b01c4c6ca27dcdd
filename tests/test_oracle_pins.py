"""Independent pins for the oracle's source, force, boundary and error terms.

Each check is fixed by the mathematics, not by retyping the oracle's sums:
closed-form lattice sums of the cubic kernel (itself pinned by the
reference's golden vectors in test_oracle.py), symmetry, and brute force on
inputs of two or three particles.  The lattice is exact in float: r = 2^-7,
spacing 2^-6, h = 2^-5, positions (k + 0.5) 2^-6 + 0.25 and velocity fields
with dyadic coefficients, so every r_ij and v_i - v_j the oracle forms is
exact and the FP64 oracle must meet the closed forms to rounding.

Functions pinned (oracle/dfsph_oracle.cpp):
  density_source(0)   divergence source rho_dot  (compute_density_change.comp.glsl:82-151)
  density_source(1)   density source rho*        (compute_density_adv.comp.glsl:80-153)
  integrate_forces    viscosity and cohesion     (integrate_forces.comp.glsl:102-143)
  boundary terms      rho, G_b, alpha            (density_alpha.comp.glsl:146-172)
                      pressure acceleration      (compute_pressure_accel.comp.glsl:130-154)
                      aij                        (compute_aij_pj.comp.glsl:143-158)
  solve               FP64 error (mean complementarity residual)
"""
from __future__ import annotations

import itertools
import math

import numpy as np
import pytest

R = 2.0 ** -7
S = 2 * R           # lattice spacing
H = 4 * R           # support radius
ORIGIN = 0.25


def lattice(n, origin=(ORIGIN, ORIGIN, ORIGIN)):
    """n^3 lattice points (x fastest) and their integer coordinates."""
    ijk = np.array([(i, j, k) for k in range(n) for j in range(n) for i in range(n)], np.int64)
    x = np.asarray(origin, np.float64) + (ijk + 0.5) * S
    return x.astype(np.float32), ijk


def interior(ijk, n, layers=2):
    return np.all((ijk >= layers) & (ijk <= n - 1 - layers), axis=1)


class Kernel:
    """Cubic spline with the oracle's float-rounded constants, evaluated from
    the pinned oracle_kernel_f64 (W and |grad W| at q for support h)."""

    def __init__(self, O, cfg):
        d = O.derive(cfg)
        self.d = d
        self.h = float(d.h)
        k64 = 8.0 / (math.pi * self.h ** 3)
        g64 = 8.0 / (math.pi * self.h ** 4)
        self.wscale = float(d.knorm) / k64
        self.gscale = float(d.gnorm) / g64
        self.O = O

    def W(self, q):
        return self.O.kernel_f64(q, self.h)[0] * self.wscale

    def dW(self, q):
        """dW/dr (negative inside the support)."""
        return -self.O.kernel_f64(q, self.h)[1] * self.gscale


# Neighbour offsets of a lattice point: |n|^2 < 4 (|n|^2 = 4 sits at q = 1, W = 0).
OFFSETS = [np.array(n) for n in itertools.product((-1, 0, 1), repeat=3) if 0 < sum(c * c for c in n) < 4]


def interior_density(K):
    return float(K.d.mass) * (K.W(0.0) + sum(K.W(math.sqrt(n @ n) * S / K.h) for n in OFFSETS))


def make(O, x, v=None, fp64=True, **kw):
    cfg = O.make_config(R, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), **kw)
    sim = O.OracleSim.create(cfg, x, None if v is None else np.asarray(v, np.float32), fp64=fp64)
    return sim, cfg


# ---------------------------------------------------------------------------
# Divergence source rho_dot and density source rho*


def test_uniform_velocity_has_no_fluid_source(oracle_mod):
    """Galilean invariance: v_i - v_j = 0 exactly, so rho_dot = 0 everywhere
    (no boundary in reach) and rho* = rho / rho0."""
    O = oracle_mod
    x, _ = lattice(8)
    v = np.tile(np.array([0.75, -1.5, 2.0], np.float32), (x.shape[0], 1))
    sim, _ = make(O, x, v, boundary=False)
    sim.prepare(0)
    assert np.all(sim.source(False, 1e-3) == 0.0)
    rho = sim.get(O.F_RHO)
    np.testing.assert_allclose(sim.source(True, 1e-3), rho / 1000.0, rtol=1e-15, atol=0)


@pytest.mark.parametrize("A", [
    [[1.0, 2.0, 0.0], [-0.5, -3.0, 1.0], [0.25, 0.0, 4.0]],     # tr 2
    [[2.0, 0.5, -1.0], [1.0, -1.5, 0.0], [0.0, 4.0, -0.5]],     # traceless
])
def test_linear_field_source_is_trace_times_lattice_sum(oracle_mod, A):
    """v = A x on the lattice interior: sum_j (A r_ij).gradW_ij reduces by
    cubic symmetry to tr(A)/3 sum_n dW(r_n) r_n, so
    rho_dot = (m / rho) tr(A) / 3 sum_n dW(r_n) r_n; off-diagonal entries drop
    out and a contracting field (tr < 0) raises the density."""
    O = oracle_mod
    A = np.array(A)
    n = 10
    x, ijk = lattice(n)
    v = (x.astype(np.float64) @ A.T)
    assert np.array_equal(v.astype(np.float32).astype(np.float64), v)   # exact in float
    sim, cfg = make(O, x, v, boundary=False)
    K = Kernel(O, cfg)
    sim.prepare(0)
    rd = sim.source(False, 1e-3)
    inn = interior(ijk, n)
    rho_in = interior_density(K)
    np.testing.assert_allclose(sim.get(O.F_RHO)[inn], rho_in, rtol=1e-13)
    lat = sum(K.dW(math.sqrt(o @ o) * S / K.h) * math.sqrt(o @ o) * S for o in OFFSETS)
    expect = float(K.d.mass) / rho_in * np.trace(A) / 3.0 * lat
    scale = float(K.d.mass) / rho_in * np.abs(A).sum() * abs(lat)
    assert np.abs(rd[inn] - expect).max() <= 1e-12 * scale, (rd[inn][:4], expect)
    if np.trace(A) > 0:
        assert expect < 0   # expansion lowers the density


def test_density_source_is_rho_over_rho0_plus_dt_rho_dot(oracle_mod):
    """rho* = rho/rho0 + dt rho_dot for the same state, fluid and boundary
    terms included (the two shaders sum the same terms)."""
    O = oracle_mod
    rng = np.random.default_rng(3)
    # A block resting on the floor so boundary samples are in reach.
    x, _ = lattice(8, origin=(0.4, 0.0, 0.4))
    x = x + rng.uniform(-0.1 * R, 0.1 * R, x.shape).astype(np.float32)
    v = rng.uniform(-1, 1, x.shape).astype(np.float32)
    sim, _ = make(O, x, v)
    sim.prepare(0)
    dt = 1e-3
    rd = sim.source(False, dt).copy()
    ra = sim.source(True, dt)
    rho = sim.get(O.F_RHO)
    assert sim.nb > 0 and np.any(sim.get(O.F_NBR_B) > 0)
    np.testing.assert_allclose(ra - rho / 1000.0, dt * rd, rtol=1e-11, atol=1e-13 * np.abs(dt * rd).max())


def test_uniform_velocity_source_near_wall_is_v_dot_boundary_gradient(oracle_mod):
    """With uniform v the fluid sum vanishes exactly and the boundary term of
    rho_dot is sum_b V_b v.gradW_ib = v . G_b (not scaled by V_i)."""
    O = oracle_mod
    x, _ = lattice(6, origin=(0.4, 0.0, 0.4))
    vel = np.array([0.5, -2.0, 0.25])
    v = np.tile(vel.astype(np.float32), (x.shape[0], 1))
    sim, _ = make(O, x, v)
    sim.prepare(0)
    rd = sim.source(False, 1e-3)
    gb = sim.get(O.F_GB)
    assert np.abs(gb).max() > 0
    np.testing.assert_allclose(rd, gb @ vel, rtol=1e-12, atol=1e-12 * np.abs(gb @ vel).max())
    # the floor lies below: moving down (v_y < 0) compresses, rho_dot > 0 there
    near = np.abs(gb[:, 1]) > 0.1 * np.abs(gb[:, 1]).max()
    assert np.all(rd[near] > 0)


# ---------------------------------------------------------------------------
# Viscosity and cohesion (integrate_forces mode 0)


def _force_accel(O, x, v, dt=1e3, **kw):
    sim, cfg = make(O, x, v, boundary=False, gravity=(0.0, 0.0, 0.0), **kw)
    sim.prepare(0)
    vn = sim.forces(dt)
    return (vn - np.asarray(v, np.float64)) / dt, sim, cfg


def test_viscosity_vanishes_for_uniform_and_linear_fields(oracle_mod):
    O = oracle_mod
    n = 10
    x, ijk = lattice(n)
    a, _, _ = _force_accel(O, x, np.tile(np.float32([1.0, -0.5, 2.0]), (x.shape[0], 1)), viscosity=1.0)
    assert np.all(a == 0.0)
    A = np.array([[1.0, 2.0, 0.0], [-0.5, -3.0, 1.0], [0.25, 0.0, 4.0]])
    v = x.astype(np.float64) @ A.T
    a, _, cfg = _force_accel(O, x, v, viscosity=1.0)
    K = Kernel(O, cfg)
    # scale of one neighbour's term
    scale = float(K.d.mass) / interior_density(K) * K.W(0.0) * np.abs(A).sum() * S
    assert np.abs(a[interior(ijk, n)]).max() <= 1e-9 * scale


def test_viscosity_of_quadratic_field_is_lattice_sum(oracle_mod):
    """v_x = c y^2: (v_j - v_i)_x = c (2 y_i d_y + d_y^2); the odd part cancels
    on the lattice, leaving a_x = nu c sum_n (m / rho_n) W(r_n) d_y^2 > 0 for
    c > 0 (momentum diffuses from the faster neighbours)."""
    O = oracle_mod
    n, c, nu = 10, 4.0, 1.0
    x, ijk = lattice(n)
    v = np.zeros((x.shape[0], 3))
    v[:, 0] = c * x[:, 1].astype(np.float64) ** 2
    assert np.array_equal(v.astype(np.float32).astype(np.float64), v)
    a, _, cfg = _force_accel(O, x, v, viscosity=nu)
    K = Kernel(O, cfg)
    rho = interior_density(K)
    expect = nu * c * float(K.d.mass) / rho * sum(K.W(math.sqrt(o @ o) * S / K.h) * (o[1] * S) ** 2
                                                   for o in OFFSETS)
    inn = interior(ijk, n)
    assert expect > 0
    np.testing.assert_allclose(a[inn, 0], expect, rtol=1e-6)
    assert np.abs(a[inn, 1:]).max() <= 1e-9 * expect


def test_viscosity_weights_by_the_neighbours_density(oracle_mod):
    """Three particles on a line at 0, d1, d1 + d2 (brute force): particle 0's
    viscous acceleration is nu sum_j (m / rho_j) (v_j - v_0) W_0j with each
    rho_j the neighbour's own density, which differs between the middle and
    the end particle."""
    O = oracle_mod
    d1, d2 = 0.5 * H, 0.375 * H
    x = np.array([[0.5, 0.5, 0.5], [0.5 + d1, 0.5, 0.5], [0.5 + d1 + d2, 0.5, 0.5]], np.float32)
    v = np.array([[0.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 2.0]], np.float32)
    nu = 1.0
    a, sim, cfg = _force_accel(O, x, v, viscosity=nu)
    K = Kernel(O, cfg)
    m = float(K.d.mass)
    r = lambda i, j: float(np.linalg.norm(x[i].astype(np.float64) - x[j]))
    W = lambda i, j: K.W(r(i, j) / K.h) if r(i, j) < K.h else 0.0
    rho = [m * sum(W(i, j) if i != j else K.W(0.0) for j in range(3)) for i in range(3)]
    np.testing.assert_allclose(sim.get(O.F_RHO), rho, rtol=1e-13)
    assert abs(rho[1] - rho[2]) > 1e-3 * rho[1]
    for i in range(3):
        exp = sum(nu * m / rho[j] * (v[j].astype(np.float64) - v[i]) * W(i, j) for j in range(3) if j != i)
        np.testing.assert_allclose(a[i], exp, rtol=1e-9, atol=1e-12 * np.abs(exp).max())


def test_cohesion_interior_zero_surface_inward_pair_attracts(oracle_mod):
    """Cohesion -c m sum_j W_ij (x_i - x_j)/r: zero in the lattice interior by
    symmetry; at a planar free surface it points into the fluid with
    magnitude c m sum_{n_y = -1} W(r_n) / |n|; two particles attract."""
    O = oracle_mod
    n, coh = 10, 2.0
    x, ijk = lattice(n)
    a, _, cfg = _force_accel(O, x, np.zeros_like(x), cohesion=coh, viscosity=0.0)
    K = Kernel(O, cfg)
    m = float(K.d.mass)
    unit = coh * m * K.W(0.0)
    assert np.abs(a[interior(ijk, n)]).max() <= 1e-12 * unit
    top = (ijk[:, 1] == n - 1) & np.all((ijk[:, [0, 2]] >= 2) & (ijk[:, [0, 2]] <= n - 3), axis=1)
    expect_y = -coh * m * sum(K.W(math.sqrt(o @ o) * S / K.h) / math.sqrt(o @ o) for o in OFFSETS if o[1] == -1)
    assert expect_y < 0
    np.testing.assert_allclose(a[top, 1], expect_y, rtol=1e-9)
    assert np.abs(a[top][:, [0, 2]]).max() <= 1e-9 * abs(expect_y)
    # a pair at distance d: each is pulled toward the other with c m W(d/h)
    d = 0.625 * H
    xp = np.array([[0.5, 0.5, 0.5], [0.5, 0.5 + d, 0.5]], np.float32)
    ap, _, _ = _force_accel(O, xp, np.zeros_like(xp), cohesion=coh, viscosity=0.0)
    w = K.W(d / K.h)
    np.testing.assert_allclose(ap, [[0, coh * m * w, 0], [0, -coh * m * w, 0]], rtol=1e-12, atol=1e-18)


# ---------------------------------------------------------------------------
# Boundary terms


def _lone(O, y, psi=0):
    """One particle above the floor at (0.5, y, 0.5): 0.5 = u_lo + 31 spacings,
    a mirror point of the staggered Akinci face pattern in both tangential
    directions (rows alternate the half-spacing offset)."""
    x = np.array([[0.5, y, 0.5]], np.float32)
    sim, cfg = make(O, x)
    sim.prepare(psi)
    return sim, cfg, x


@pytest.mark.parametrize("y", [0.25 * H, 0.5 * H, 0.8125 * H])
def test_boundary_gradient_is_normal_above_face_centre(oracle_mod, y):
    O = oracle_mod
    sim, _, _ = _lone(O, y)
    gb = sim.get(O.F_GB)[0]
    assert gb[1] < 0   # gradW_ib points from i toward the samples below
    assert max(abs(gb[0]), abs(gb[2])) <= 1e-12 * abs(gb[1])


@pytest.mark.parametrize("y", [0.25 * H, 0.5 * H, 0.8125 * H])
def test_boundary_terms_match_grid_free_brute_force(oracle_mod, y):
    """A lone particle near the floor: density m W(0) + rho0 sum_b V_b W_ib,
    G_b = sum_b V_b gradW_ib and alpha = 1 / max(|psi G_b|^2, 1e-5), summed
    over every sample of the domain with no grid, V_b from the oracle (pinned
    against an O(Nb^2) brute force in test_oracle.py); then the boundary
    pressure force -(p/rho^2) psi G_b and its aij term dt^k a.G_b."""
    O = oracle_mod
    for psi in (0, 1):
        sim, cfg, x = _lone(O, y, psi)
        K = Kernel(O, cfg)
        bx = sim.get(O.F_BX)
        bv = sim.get(O.F_BV)
        d = x[0].astype(np.float64) - bx
        r = np.linalg.norm(d, axis=1)
        sel = r < K.h
        assert sel.sum() >= 4
        w = np.array([K.W(q) for q in r[sel] / K.h])
        dw = np.array([K.dW(q) for q in r[sel] / K.h])
        rho = float(K.d.mass) * K.W(0.0) + 1000.0 * np.sum(bv[sel] * w)
        gb = np.sum((bv[sel] * dw / r[sel])[:, None] * d[sel], axis=0)
        np.testing.assert_allclose(sim.get(O.F_RHO)[0], rho, rtol=1e-12)
        np.testing.assert_allclose(sim.get(O.F_GB)[0], gb, rtol=1e-11, atol=1e-12 * np.abs(gb).max())
        bscale = 1000.0 if psi else 1.0
        alpha = 1.0 / max(float((bscale * gb) @ (bscale * gb)), 1e-5)
        np.testing.assert_allclose(sim.get(O.F_ALPHA)[0], alpha, rtol=1e-11)
        p = np.array([2.0e4])
        for mode_k, dt in ((1, 1e-3), (0, 1e-3)):
            ap, acc = sim.apply_A(p, mode_k, dt, psi)
            a_exp = -(p[0] / rho ** 2) * bscale * gb
            np.testing.assert_allclose(acc[0], a_exp, rtol=1e-11, atol=1e-12 * np.abs(a_exp).max())
            np.testing.assert_allclose(ap[0], dt ** (1 if mode_k else 2) * (a_exp @ gb), rtol=1e-11)


def test_boundary_pressure_force_is_gated_on_small_pressure(oracle_mod):
    """compute_pressure_accel.comp.glsl:130: the wall term applies only when
    |p/rho^2| > 1e-5."""
    O = oracle_mod
    sim, _, _ = _lone(O, 0.5 * H)
    rho = sim.get(O.F_RHO)[0]
    _, acc = sim.apply_A(np.array([0.5e-5 * rho * rho]), 1, 1e-3, 0)
    assert np.all(acc == 0.0)
    _, acc = sim.apply_A(np.array([2.0e-5 * rho * rho]), 1, 1e-3, 0)
    assert np.abs(acc).max() > 0


# ---------------------------------------------------------------------------
# FP64 error of the solves


def _dam(O, n=12):
    x, _ = lattice(n, origin=(0.0, 0.0, 0.0))
    rng = np.random.default_rng(8)
    x = (x + 2 * R + rng.uniform(-0.1 * R, 0.1 * R, x.shape)).astype(np.float32)
    cfg = O.make_config(R, (0.0, 0.0, 0.0), (0.5, 0.3, 0.3))
    return x, cfg


def test_error_at_zero_pressure_is_mean_compression(oracle_mod):
    """One iteration from zero pressure checks the zero iterate, whose
    residual is the source itself: err_density = mean max(rho* - 1, 0) and
    err_div = dt mean max(rho_dot, 0), with rho* and rho_dot as exported."""
    O = oracle_mod
    x, cfg = _dam(O)
    sim = O.OracleSim.create(cfg, x, fp64=True)
    pol = O.policy("converged", min_iter_div=1, max_iter_div=1, min_iter_density=1, max_iter_density=1)
    for _ in range(3):
        st = sim.step(0.0, pol)
    ra = sim.get(O.F_RHO_ADV)
    rd = sim.get(O.F_RHO_CHANGE)
    assert st["iters_density"] == 1 and st["iters_div"] == 1
    assert np.any(ra > 1.0)
    exp_den = math.fsum(np.maximum(ra - 1.0, 0.0)) / ra.size
    exp_div = st["dt"] * math.fsum(np.maximum(rd, 0.0)) / rd.size
    assert st["err_density"] == pytest.approx(exp_den, rel=1e-13)
    assert st["err_div"] == pytest.approx(exp_div, rel=1e-13, abs=1e-300)


def test_error_is_exact_mean_of_residuals(oracle_mod):
    """After several iterations the reported error equals the exactly rounded
    mean (math.fsum) of the per-particle residuals of the checked iterate."""
    O = oracle_mod
    x, cfg = _dam(O)
    for fp64 in (False, True):
        sim = O.OracleSim.create(cfg, x, fp64=fp64)
        pol = O.policy("converged")
        for _ in range(4):
            st = sim.step(0.0, pol)
        assert st["iters_density"] >= 2
        for key, f, k in (("err_density", O.F_RES_DENSITY, 1.0), ("err_div", O.F_RES_DIV, st["dt"])):
            e = sim.get(f)
            assert np.all(e >= 0)
            assert st[key] == pytest.approx(k * math.fsum(e) / e.size, rel=1e-15, abs=1e-300), key
