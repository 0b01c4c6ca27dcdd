"""The learned warm start on a tank it was not trained on.

The network is trained as the 1M dam break's configuration says
(configs/c2_dam_break_1m.json `learned`: 520 plain converged steps of the
2:1:1 tank with its 128x64x128 block at the min corner, another jitter seed,
20 epochs).  It is then scored on a different geometry: a shallow 4:1:2 tank
(2.0 x 0.5 x 1.0 m, gravity along -y) with a 50x40x96 block (192,000
particles, r = 0.005) standing against the +x wall, so the column is lower and
narrower and collapses in the opposite direction.  Four runs of 150 steps from
the same seeded scene: the plain solver at the configuration's 0.1 %
tolerance, the classic previous-pressure start and the learned density start
at the same tolerance, and the plain solver at 1e-5 as the converged
solution.  Every 0.1 % run converges every step; the learned run takes fewer
solver iterations than the plain one, and its velocities and positions are as
close to the converged ones as the plain run's (rms within 10 %)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 150


def test_network_trained_on_dam_break_scores_on_a_shallow_tank(dfsph_mod):
    D = dfsph_mod
    from dfsph import scenes

    train = scenes.load("dam_break_1m")
    L = train["learned"]
    xt = scenes.positions(dict(train, seed=L["seed"]))
    n = xt.shape[0]
    print(f"case: learned transfer, training on dam_break_1m ({n} particles)", flush=True)
    sim = D.Sim(scenes.params(train))
    sim.init_scene(xt)
    sim.dataset_start(2 * L["steps"] * (n // L["stride"] + 1), L["stride"])
    plain_train = scenes.policy(train)
    for _ in range(L["steps"]):
        sim.step(plain_train)
    rows = sim.dataset_stop()
    sim.mlp_init(L["init_seed"])
    tr = sim.mlp_train(epochs=L["epochs"], batch=L["batch"], lr=L["lr"], expectile_density=L["expectile_density"])
    theta = sim.mlp_get_params()
    sim.close()
    print(f"  network: {rows} rows, val loss {tr['val_loss']:.4f}", flush=True)

    score = dict(train, name="shallow_tank_4_1_2", domain_min=[0.0, 0.0, 0.0], domain_max=[2.0, 0.5, 1.0],
                 blocks=[{"dims": [50, 40, 96], "origin": [2.0 - 0.01 - 50 * 0.01, 0.01, 0.01]}], seed=7)
    x0 = scenes.positions(score)
    assert x0.shape[0] == 192_000
    assert np.all(x0 > 0.0) and np.all(x0 < np.array(score["domain_max"], np.float32))
    print(f"  scoring on {score['name']}: {x0.shape[0]} particles, {STEPS} steps", flush=True)
    runs = {}
    for run in ("plain", "previous", "learned", "converged"):
        sim = D.Sim(scenes.params(score))
        sim.init_scene(x0)
        if run == "converged":
            pol = scenes.policy(score, tol_div=1e-5, tol_density=1e-5, max_iter_div=5000, max_iter_density=5000,
                                residual_bound=-1.0)
        elif run == "learned":
            pol = scenes.policy(score, warm_start=D.WARM_MLP_DENSITY)
            sim.mlp_set_params(theta)
        elif run == "previous":
            pol = scenes.policy(score, warm_start=D.WARM_PREVIOUS)
        else:
            pol = scenes.policy(score)
        it_div = it_den = fallbacks = 0
        for s in range(STEPS):
            st = sim.step(pol)
            assert st["converged_div"] == 1 and st["converged_density"] == 1, (run, s)
            assert st["nonfinite"] == 0, (run, s)
            it_div += st["iters_div"]
            it_den += st["iters_density"]
            fallbacks += st["warm_fallback"]
        v = sim.export(D.FIELD_VELOCITY).astype(np.float64)
        x = sim.export(D.FIELD_POSITION).astype(np.float64)
        sim.close()
        runs[run] = (v, x, it_div / STEPS, it_den / STEPS)
        sp = np.linalg.norm(v, axis=1)
        print(f"  {run}: iterations/step div {it_div / STEPS:.2f} density {it_den / STEPS:.2f}, fallbacks "
              f"{fallbacks}; speed max {sp.max():.3f} m/s, mean {sp.mean():.4f}", flush=True)
        assert np.all(np.isfinite(v)) and sp.max() < 10.0, run
    vc, xc = runs["converged"][:2]
    rv = {k: float(np.sqrt(np.mean(np.sum((runs[k][0] - vc) ** 2, axis=1)))) for k in ("plain", "previous", "learned")}
    rx = {k: float(np.sqrt(np.mean(np.sum((runs[k][1] - xc) ** 2, axis=1)))) for k in ("plain", "previous", "learned")}
    for k in rv:
        print(f"  {k}: rms |v - v_converged| {rv[k]:.4f} m/s, rms |x - x_converged| {rx[k]:.3e} m", flush=True)
    total = {k: runs[k][2] + runs[k][3] for k in ("plain", "previous", "learned")}
    print(f"  iteration cut vs plain: learned {1 - total['learned'] / total['plain']:.3f}, previous "
          f"{1 - total['previous'] / total['plain']:.3f}", flush=True)
    assert total["learned"] < total["plain"]
    assert rv["learned"] <= 1.1 * rv["plain"] and rx["learned"] <= 1.1 * rx["plain"]
