"""Learned warm start on the GPU: BF16-MFMA forward and backward against the
FP64 oracle (oracle/mlp_oracle.py, evaluated with the device's bfloat16
rounding points), training, and the warm start inside the solver.

Bars: outputs and loss within 1e-3 relative of the oracle; every gradient
block within 2e-2 of its largest magnitude (a hidden activation one bf16 ulp
apart propagates); the solver with the learned start meets the same tolerance.
"""
from __future__ import annotations

import numpy as np
import pytest

import mlp_oracle as M
from conftest import dam_break_small

pytestmark = pytest.mark.gpu


def _sim(D, n=4096):
    x, lo, hi = dam_break_small(16)
    return D.Sim(D.make_params(0.01, lo, hi, capacity=max(n, x.shape[0])))


def _theta(seed=3):
    rng = np.random.default_rng(seed)
    th = M.init(seed)
    th += rng.normal(0, 0.02, M.NPARAM)   # non-zero biases too
    return th.astype(np.float32)


def test_forward_matches_oracle(dfsph_mod):
    D = dfsph_mod
    print("case: mlp forward", flush=True)
    sim = _sim(D)
    th = _theta()
    sim.mlp_set_params(th)
    assert np.array_equal(sim.mlp_get_params(), th)
    rng = np.random.default_rng(1)
    X = rng.normal(size=(1000, 16)).astype(np.float32)
    y = sim.mlp_eval(X)
    ref, _ = M.forward(th.astype(np.float64), X.astype(np.float64), bf16=True)
    scale = np.abs(ref).max()
    err = np.abs(y - ref) / scale
    print(f"  forward max rel err {err.max():.2e}, median {np.median(err):.2e}")
    assert np.median(err) < 1e-4 and err.max() < 1e-3
    sim.close()


@pytest.mark.parametrize("n", [64, 200, 4096])
def test_gradient_matches_oracle(dfsph_mod, n):
    D = dfsph_mod
    print(f"case: mlp gradient n={n}", flush=True)
    sim = _sim(D)
    th = _theta(5)
    sim.mlp_set_params(th)
    rng = np.random.default_rng(n)
    X = rng.normal(size=(n, 16)).astype(np.float32)
    T = rng.normal(size=(n, 2)).astype(np.float32)
    loss, g = sim.mlp_grad(X, T)
    rl, rg = M.loss_and_grad(th.astype(np.float64), X.astype(np.float64), T.astype(np.float64), bf16=True)
    assert loss == pytest.approx(rl, rel=1e-3)
    o = 0
    for name, shp in M.SHAPES:
        k = int(np.prod(shp))
        a, b = g[o:o + k], rg[o:o + k]
        e = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
        print(f"  {name}: max |dg| / max |g| = {e:.2e}")
        assert e < 2e-2, name
        o += k
    sim.close()


def test_training_reduces_heldout_loss(dfsph_mod):
    """A smooth synthetic target: Adam on the device must cut the held-out loss."""
    D = dfsph_mod
    print("case: mlp training", flush=True)
    sim = _sim(D)
    sim.mlp_init(7)
    rng = np.random.default_rng(2)
    X = rng.normal(size=(32768, 16)).astype(np.float32)
    T = np.stack([np.tanh(X[:, 0] - 0.5 * X[:, 3]) + 0.3 * X[:, 5] * X[:, 6],
                  np.maximum(X[:, 1], 0.0) - 0.2 * X[:, 9]], 1).astype(np.float32)
    sim.dataset_import(X, T)
    first = sim.mlp_train(epochs=1, batch=1024, lr=1e-3)
    later = sim.mlp_train(epochs=8, batch=1024, lr=1e-3)
    print(f"  val loss after 1 epoch {first['val_loss']:.4f}, after 9 {later['val_loss']:.4f}")
    assert later["val_loss"] < 0.5 * first["val_loss"]
    assert later["train_rows"] + later["val_rows"] == 32768
    sim.close()


def test_expectile_shifts_density_output_up(dfsph_mod):
    """With expectile tau on the density output, a noisy target is fitted from
    above: the share of rows predicted above their target moves from about 1/2
    (tau = 0.5, MSE) towards tau; the divergence output stays MSE."""
    D = dfsph_mod
    print("case: mlp expectile loss", flush=True)
    rng = np.random.default_rng(5)
    X = rng.normal(size=(32768, 16)).astype(np.float32)
    base = np.tanh(X[:, 0]) + 0.5 * X[:, 2]
    T = np.stack([base + 0.3 * rng.normal(size=base.size), base + 0.3 * rng.normal(size=base.size)],
                 1).astype(np.float32)
    above = {}
    for tau in (0.0, 0.85):
        sim = _sim(D)
        sim.mlp_init(3)
        sim.dataset_import(X, T)
        sim.mlp_train(epochs=6, batch=1024, lr=2e-3, expectile_density=tau)
        y = sim.mlp_eval(X[:8192])
        above[tau] = ((y[:, 0] > T[:8192, 0]).mean(), (y[:, 1] > T[:8192, 1]).mean())
        sim.close()
    print(f"  share above target (div, density): mse {above[0.0]}, tau 0.85 {above[0.85]}")
    assert abs(above[0.0][1] - 0.5) < 0.1 and above[0.85][1] > 0.7
    assert abs(above[0.85][0] - 0.5) < 0.1
    sim = _sim(D)
    sim.mlp_init(3)
    sim.dataset_import(X, T)
    with pytest.raises(D.DfsphError):
        sim.mlp_train(epochs=1, batch=1024, expectile_density=1.0)
    sim.close()


def test_corrupt_parameters_fall_back(dfsph_mod):
    D = dfsph_mod
    print("case: mlp corrupt params", flush=True)
    x, lo, hi = dam_break_small(8)
    sim = D.Sim(D.make_params(0.01, lo, hi, capacity=x.shape[0]))
    sim.init_scene(x)
    th = _theta()
    th[100] = np.nan
    with pytest.raises(D.DfsphError) as e:
        sim.mlp_set_params(th)
    assert e.value.code == D.E_INVALID
    for warm in (D.WARM_MLP, D.WARM_MLP_DENSITY):
        st = sim.step(D.default_policy(D.POLICY_CONVERGED, warm_start=warm))
        assert st["warm_fallback"] == 1 and st["converged_density"] == 1
    sim.close()


def test_learned_warm_start_keeps_tolerance(dfsph_mod):
    """Rows recorded from plain-solver steps, trained on the device; solves
    started from the network meet the same tolerance without falling back."""
    D = dfsph_mod
    print("case: mlp warm start keeps tolerance", flush=True)
    x, lo, hi = dam_break_small(16)
    sim = D.Sim(D.make_params(0.01, lo, hi, capacity=x.shape[0]))
    sim.init_scene(x)
    plain = D.default_policy(D.POLICY_CONVERGED)
    sim.dataset_start(200000, 4)
    for _ in range(40):
        sim.step(plain)
    rows = sim.dataset_stop()
    assert rows == 40 * 2 * (x.shape[0] // 4)
    X, T = sim.dataset_export(rows)
    assert np.isfinite(X).all() and np.isfinite(T).all() and (T >= 0).all()
    sim.mlp_init(11)
    st = sim.mlp_train(epochs=30, batch=1024, lr=2e-3)
    print(f"  train {st['train_loss']:.4f} val {st['val_loss']:.4f} ({st['steps']} steps, {st['ms']:.0f} ms)")
    for warm in (D.WARM_MLP, D.WARM_MLP_DENSITY):   # network for both solves / density only
        learned = D.default_policy(D.POLICY_CONVERGED, warm_start=warm)
        for _ in range(10):
            r = sim.step(learned)
            assert r["converged_density"] == 1 and r["err_density"] <= 1e-3
            assert r["converged_div"] == 1 and r["err_div"] <= 1e-3
            assert r["warm_fallback"] == 0
    sim.close()


def test_learned_warm_start_gain_1m(dfsph_mod):
    """What the trained network gains over the plain solver (zero start) at the
    same 0.1 % tolerance on the 1M dam break: trained on rows of one seeded run,
    evaluated on a scene with another seed from the same state, both policies."""
    D = dfsph_mod
    from dfsph import scenes
    print("case: mlp gain on dam_break_1m", flush=True)
    cfg = scenes.load("c2_dam_break_1m")
    sim = D.Sim(scenes.params(cfg))
    sim.init_scene(scenes.positions(cfg))
    plain = scenes.policy(cfg)
    sim.dataset_start(2_000_000, 512)
    for _ in range(200):
        sim.step(plain)
    sim.dataset_stop()
    sim.mlp_init(1)
    tr = sim.mlp_train(epochs=20, batch=4096, lr=1e-3, expectile_density=0.8)
    theta = sim.mlp_get_params()
    sim.close()
    held = dict(cfg, seed=cfg["seed"] + 101)
    x = scenes.positions(held)
    res = {}
    for name, warm in (("plain", D.WARM_ZERO), ("learned", D.WARM_MLP), ("learned_density", D.WARM_MLP_DENSITY)):
        s2 = D.Sim(scenes.params(cfg))
        s2.init_scene(x)
        s2.mlp_set_params(theta)
        pol = scenes.policy(cfg, warm_start=warm)
        for _ in range(150):                      # same 150-step history for both
            s2.step(scenes.policy(cfg))
        it = 0
        import time
        t0 = time.perf_counter()
        for _ in range(50):
            r = s2.step(pol)
            assert r["converged_density"] == 1 and r["err_density"] <= cfg["tol_density"]
            it += r["iters_density"] + r["iters_div"]
        res[name] = (it / 50, (time.perf_counter() - t0) / 50 * 1e3)
        s2.close()
    print(f"  val loss {tr['val_loss']:.4f}; plain {res['plain'][0]:.2f} iters {res['plain'][1]:.2f} ms/step; "
          f"learned {res['learned'][0]:.2f} iters {res['learned'][1]:.2f} ms/step; "
          f"learned (density solve) {res['learned_density'][0]:.2f} iters {res['learned_density'][1]:.2f} ms/step")
    for name in ("learned", "learned_density"):
        assert res[name][0] <= 0.7 * res["plain"][0]     # >= 30 % fewer iterations
        assert res[name][1] <= 0.8 * res["plain"][1]     # >= 20 % lower step time


def test_bad_guess_is_redone_from_zero(dfsph_mod):
    """A learned start the solver cannot converge from within max_iter is
    discarded and the solve is redone from zero pressure."""
    D = dfsph_mod
    print("case: mlp bad guess fallback", flush=True)
    x, lo, hi = dam_break_small(16)
    sim = D.Sim(D.make_params(0.01, lo, hi, capacity=x.shape[0]))
    sim.init_scene(x)
    th = np.zeros(M.NPARAM, np.float32)
    th[-2:] = 12.0                       # log(1 + kappa) = 12: a huge initial pressure
    sim.mlp_set_params(th)
    pol = D.default_policy(D.POLICY_CONVERGED, warm_start=D.WARM_MLP, max_iter_density=8, max_iter_div=8)
    st = sim.step(pol)
    assert st["warm_fallback"] == 1
    assert st["converged_density"] == 1 and st["err_density"] <= 1e-3
    sim.close()


def test_learned_guess_bounded_for_isolated_particles(dfsph_mod):
    """A learned guess is kappa in units of the inverse Jacobi diagonal, which
    is ~13 orders larger for a particle without neighbours than in the bulk.
    A constant network output of 0.3 (log units, a typical fit error) must not
    turn into a huge pressure on splash particles: after one learned step the
    speeds stay those of free fall plus a small pressure response."""
    D = dfsph_mod
    print("case: mlp guess for isolated particles", flush=True)
    r = 0.005
    x = D.scene_block(8, 8, 8, r, (2 * r, 2 * r, 2 * r), 4)
    lone = np.array([[0.05 + 0.02 * k, 0.30, 0.05] for k in range(5)], np.float32)   # 10h above the block
    x = np.concatenate([x, lone])
    sim = D.Sim(D.make_params(r, (0.0, 0.0, 0.0), (0.2, 0.4, 0.2), capacity=x.shape[0]))
    sim.init_scene(x)
    th = np.zeros(M.NPARAM, np.float32)
    th[-2:] = 0.3
    sim.mlp_set_params(th)
    pol = D.default_policy(D.POLICY_CONVERGED, warm_start=D.WARM_MLP)
    for _ in range(3):
        st = sim.step(pol)
        assert st["converged_density"] == 1
    v = sim.export(D.FIELD_VELOCITY)
    print(f"  max |v| after 3 learned steps: {np.abs(v).max():.3e}, fallback {st['warm_fallback']}")
    assert np.abs(v).max() < 0.5
    sim.close()


def test_recorded_features_match_their_definitions(dfsph_mod):
    """The 16 network inputs recorded for every particle of one step equal a
    numpy evaluation of their written definitions on the exported state:
      0 slog((rho/rho0 - 1) 100)       6-8 v dt/(2r)        12 boundary nbrs / 30
      1 slog(-rho_dot dt 1e3)          9-11 (x - dmin)/ext  13 |v| dt/(2r)
      2 log(alpha rho^3) / 10          5 fluid nbrs / 30    14 slog((1 - rho*) 1e3) (density solve)
      3 log(1 + kappa_v,prev)          4 log(1 + kappa_rho,prev)   15 phase
    with kappa_prev = p_prev (dt_prev/dt)^k dt^k / (alpha rho^3) 1e3 and
    slog(x) = sign(x) log(1 + |x|).  The divergence solve's rows use the
    step's start velocity, the density solve's the velocity after forces,
    v_end - dt a_final (particles not clamped at a wall).  Tolerance 2e-5
    absolute + 2e-5 relative (the device uses the hardware logarithm)."""
    D = dfsph_mod
    print("case: mlp feature definitions", flush=True)
    r = 0.01
    x0, lo, hi = dam_break_small(16)
    sim = D.Sim(D.make_params(r, lo, hi, capacity=x0.shape[0]))
    sim.init_scene(x0)
    pol = D.default_policy(D.POLICY_CONVERGED, warm_start=D.WARM_PREVIOUS)
    for _ in range(30):
        st_prev = sim.step(pol)
    xs, vs = sim.export(D.FIELD_POSITION), sim.export(D.FIELD_VELOCITY)
    pdp, pdn = sim.export(D.FIELD_PRESSURE_DIV), sim.export(D.FIELD_PRESSURE)
    n = xs.shape[0]
    sim.dataset_start(2 * n, 1)
    st = sim.step(pol)
    rows = sim.dataset_stop()
    assert rows == 2 * n
    F, _ = sim.dataset_export(rows)
    f64 = lambda a: np.asarray(a, np.float64)
    rho, alpha = f64(sim.export(D.FIELD_DENSITY)), f64(sim.export(D.FIELD_ALPHA))
    rdot, radv = f64(sim.export(D.FIELD_DENSITY_CHANGE)), f64(sim.export(D.FIELD_DENSITY_ADV))
    nbr, nbrb = f64(sim.export(D.FIELD_NEIGHBOR_COUNT)), f64(sim.export(D.FIELD_BOUNDARY_NEIGHBOR_COUNT))
    xe, ve, ae = f64(sim.export(D.FIELD_POSITION)), f64(sim.export(D.FIELD_VELOCITY)), f64(sim.export(D.FIELD_PRESSURE_ACCEL))
    sim.close()
    dt, dtp = st["dt"], st_prev["dt"]
    ratio = dtp / dt
    cb = alpha * rho ** 3
    slog = lambda v: np.sign(v) * np.log1p(np.abs(v))
    ext = np.asarray(hi, np.float64) - np.asarray(lo, np.float64)
    vsc = dt / (2 * r)
    assert np.abs(f64(pdn)).max() > 0 and np.abs(rdot).max() > 0
    for phase in (0, 1):
        v = f64(vs) if phase == 0 else ve - dt * ae
        ok = np.ones(n, bool) if phase == 0 else np.all((xe > np.asarray(lo)) & (xe < np.asarray(hi)), axis=1)
        E = np.zeros((n, 16))
        E[:, 0] = slog((rho / 1000.0 - 1.0) * 100.0)
        E[:, 1] = slog(-rdot * dt * 1000.0)
        E[:, 2] = np.log(np.maximum(cb, 1e-30)) * 0.1
        E[:, 3] = np.log1p(np.maximum(f64(pdp) * ratio * dt * 1000.0 / cb, 0.0))
        E[:, 4] = np.log1p(np.maximum(f64(pdn) * (ratio * dt) ** 2 * 1000.0 / cb, 0.0))
        E[:, 5] = nbr / 30.0
        E[:, 6:9] = v * vsc
        E[:, 9:12] = (f64(xs) - np.asarray(lo)) / ext
        E[:, 12] = nbrb / 30.0
        E[:, 13] = np.linalg.norm(v, axis=1) * vsc
        E[:, 14] = slog((1.0 - radv) * 1000.0) if phase else 0.0
        E[:, 15] = phase
        G = f64(F[phase::2])
        err = np.abs(G - E) - 2e-5 * (1.0 + np.abs(E))
        worst = err[ok].max(axis=0)
        print(f"  phase {phase}: worst excess per feature {np.array2string(worst, precision=1)}")
        assert np.all(worst <= 0.0), (phase, int(np.argmax(worst)))
