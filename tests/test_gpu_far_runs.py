"""Neighbour records with far run starts, against the CPU oracle.

A record stores each run's first slot as a byte offset from the first slot of
the wave's window for that run; a run starting more than 255 slots in is
flagged, and every sweep reads its first slot from the cell table instead.
That happens where a wave's particles spread over many cells next to a dense
region.  The scene here makes it happen on purpose: a 16^3 jittered block
(spacing 2r) whose lowest z layer starts one FINE cell above a sheet of 64
particles (one per cell, 8 x 8 under the block's footprint, 2r below it).
Sorted by cell, the sheet is the first wave; its runs toward the block
(dz = +1) share a window over the block's first cell layer under rows 0-7,
about 512 slots, so sheet particles past row 3 start beyond a byte's reach.
test_scene_has_far_runs counts them with a numpy restatement of the grid
order and the window rule (CPU); the GPU tests then hold the step to the
oracle as test_gpu_parity does, with the staged walk (whose far groups fall
back to the memory walk) and with the memory walk throughout.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

R = 0.01
BOX = ((0.0, 0.0, 0.0), (2.0, 1.0, 1.0))


def sheet_under_block(seed=3):
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij"), -1).reshape(-1, 3)
    block = np.array([2 * R, 2 * R, 0.21]) + 2 * R * g + rng.uniform(-0.1 * R, 0.1 * R, g.shape)
    gx, gy = np.meshgrid(0.02 + 0.04 * np.arange(8), 0.02 + 0.04 * np.arange(8), indexing="ij")
    sheet = np.stack([gx.ravel(), gy.ravel(), np.full(gx.size, 0.19)], 1)
    sheet += rng.uniform(-0.2 * R, 0.2 * R, sheet.shape)
    return np.concatenate([sheet, block]).astype(np.float32)


def far_run_count(x, r=R, dmin=BOX[0], dmax=BOX[1]):
    """(runs with a neighbour whose first slot lies over 255 slots past the
    start of the wave's window, runs with a neighbour): FINE cells of
    h (1 + 2^-10) ordered z, y, x; slots by (cell, x, index); waves of 64
    slots; run r = (dz, dy) of a particle spans cells cx-1..cx+1 of row
    (cz+dz, cy+dy); the wave's window for it starts at cell c0 + off - 1, c0 the
    cell of the wave's first particle, off the run's row offset."""
    h = 4 * r
    cell = np.float32(h) * np.float32(1 + 2 ** -10)
    lo, hi = np.asarray(dmin, np.float32), np.asarray(dmax, np.float32)
    axes = np.maximum(np.ceil((hi - lo) / cell).astype(np.int64), 1)
    c = np.clip(np.floor((x - lo) / cell).astype(np.int64), 0, axes - 1)
    key = (c[:, 2] * axes[1] + c[:, 1]) * axes[0] + c[:, 0]
    order = np.lexsort((np.arange(len(x)), x[:, 0], key))
    xs, cs, ks = x[order], c[order], key[order]
    ncell = int(np.prod(axes))
    start = np.zeros(ncell + 1, np.int64)
    np.add.at(start, ks + 1, 1)
    start = np.cumsum(start)
    far = total = 0
    for w0 in range(0, len(x), 64):
        c0 = ks[w0]
        for i in range(w0, min(w0 + 64, len(x))):
            cx, cy, cz = cs[i]
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    zz, yy = cz + dz, cy + dy
                    if not (0 <= zz < axes[2] and 0 <= yy < axes[1]):
                        continue
                    row = (zz * axes[1] + yy) * axes[0]
                    a, b = row + max(cx - 1, 0), row + min(cx + 1, axes[0] - 1)
                    js = np.arange(start[a], start[b + 1])
                    if not np.any((((xs[js] - xs[i]) ** 2).sum(1) < h * h) & (js != i)):
                        continue
                    total += 1
                    off = (dz * axes[1] + dy) * axes[0]
                    if start[a] - start[min(max(c0 + off - 1, 0), ncell)] > 255:
                        far += 1
    return far, total


def test_scene_has_far_runs():
    far, total = far_run_count(sheet_under_block())
    print(f"far runs {far} of {total}", flush=True)
    assert far >= 32


FIELDS = [("density", "FIELD_DENSITY", "F_RHO"), ("density_change", "FIELD_DENSITY_CHANGE", "F_RHO_CHANGE"),
          ("density_adv", "FIELD_DENSITY_ADV", "F_RHO_ADV"), ("pressure_div", "FIELD_PRESSURE_DIV", "F_P_DIV"),
          ("pressure", "FIELD_PRESSURE", "F_P_DENSITY"), ("position", "FIELD_POSITION", "F_X"),
          ("velocity", "FIELD_VELOCITY", "F_V")]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["reference", "converged"])
@pytest.mark.parametrize("walk", ["staged", "memory"])
def test_far_runs_match_oracle(dfsph_mod, oracle_mod, mode, walk):
    D, O = dfsph_mod, oracle_mod
    x = sheet_under_block()
    flags = 0 if walk == "staged" else D.FLAG_NO_STAGED_WINDOWS
    sim = D.Sim(D.make_params(R, *BOX, capacity=x.shape[0], grid=D.GRID_FINE, flags=flags))
    sim.init_scene(x)
    ora = O.OracleSim.create(O.make_config(R, *BOX), x)
    pol = D.default_policy(D.POLICY_REFERENCE if mode == "reference" else D.POLICY_CONVERGED)
    opol = O.policy(mode)
    for step in range(2):
        print(f"case: far runs mode={mode} walk={walk} step={step}", flush=True)
        st = sim.step(pol)
        ost = ora.step(0.0, opol)
        assert (st["iters_div"], st["iters_density"]) == (ost["iters_div"], ost["iters_density"])
        assert np.array_equal(sim.export(D.FIELD_NEIGHBOR_COUNT), ora.get(O.F_NBR))
        for name, f, of in FIELDS:
            e = rel_err(sim.export(getattr(D, f)), ora.get(getattr(O, of)))
            tol = 1e-4 if (mode == "converged" and name.startswith("pressure")) else 1e-5
            assert e <= tol, f"{name}: {e:.3e}"
        ora.set(O.F_X, sim.export(D.FIELD_POSITION))
        ora.set(O.F_V, sim.export(D.FIELD_VELOCITY))
    sim.close()
