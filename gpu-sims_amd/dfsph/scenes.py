"""Workload configurations (gpu-sims_amd/configs/*.json) and their seeded scenes.

A configuration names the particle radius, the tank (domain AABB), one or more
jittered lattice blocks and the solver tolerances.  Blocks are generated by the
library's own lattice generator (dfsph_scene_block) so the oracle and the GPU
receive identical bits.
"""
from __future__ import annotations

import glob
import json
import os

import numpy as np

from . import scene_block, make_params, default_policy, POLICY_CONVERGED, GRID_FINE

CONFIG_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "configs")


def names() -> list[str]:
    return sorted(os.path.splitext(os.path.basename(p))[0]
                  for p in glob.glob(os.path.join(CONFIG_DIR, "*.json")))


def load(name: str) -> dict:
    """Load a configuration by file stem (e.g. 'c2_dam_break_1m') or by its 'name'."""
    for stem in names():
        with open(os.path.join(CONFIG_DIR, stem + ".json")) as f:
            cfg = json.load(f)
        if name in (stem, cfg["name"]):
            return cfg
    raise KeyError(f"unknown configuration {name!r}; have {names()}")


def positions(cfg: dict) -> np.ndarray:
    """Fluid positions of all blocks, ids consecutive across blocks."""
    out, id0 = [], 0
    for blk in cfg["blocks"]:
        nx, ny, nz = blk["dims"]
        out.append(scene_block(nx, ny, nz, cfg["radius"], blk["origin"], cfg["seed"],
                               cfg.get("jitter", 0.1), id0))
        id0 += nx * ny * nz
    return np.concatenate(out, axis=0)


def count(cfg: dict) -> int:
    if "batch" in cfg:
        return int(sum(int(np.prod(d)) for d, _, _ in batch_layout(cfg)))
    return int(sum(int(np.prod(b["dims"])) for b in cfg["blocks"]))


def _splitmix64(state: int):
    """SplitMix64 step: (new state, 64-bit output)."""
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def batch_layout(cfg: dict) -> list:
    """Per scene of a batched configuration: (block dims, tank max corner, seed).

    Scene s draws u0, u1 in [0, 1) from one SplitMix64 stream started at the
    configuration's seed: N_s = round(exp(ln n_min + u0 (ln n_max - ln n_min))),
    aspect a = a_min + u1 (a_max - a_min); the block is ny = nz =
    max(2, round((N_s / a)^(1/3))) and nx = max(2, round(N_s / (ny nz))), so
    nx : ny ~ a.  The tank is the block extent times cfg.batch.tank plus 4r.
    """
    B = cfg["batch"]
    r = cfg["radius"]
    state = B["seed"]
    out = []
    for s in range(B["scenes"]):
        state, z0 = _splitmix64(state)
        state, z1 = _splitmix64(state)
        u0, u1 = (z0 >> 11) * 2.0 ** -53, (z1 >> 11) * 2.0 ** -53
        n = int(round(np.exp(np.log(B["n_min"]) + u0 * (np.log(B["n_max"]) - np.log(B["n_min"])))))
        a = B["aspect_min"] + u1 * (B["aspect_max"] - B["aspect_min"])
        ny = max(2, int(round((n / a) ** (1.0 / 3.0))))
        nx = max(2, int(round(n / (ny * ny))))
        dims = (nx, ny, ny)
        ext = [d * 2.0 * r for d in dims]
        tank = tuple(float(e * f + 4.0 * r) for e, f in zip(ext, B["tank"]))
        out.append((dims, tank, B["seed"] * 1000 + s))
    return out


def batch_scenes(cfg: dict, **param_kw) -> tuple[list, list]:
    """(params list, positions list) of a batched configuration."""
    r = cfg["radius"]
    plist, xs = [], []
    for dims, tank, seed in batch_layout(cfg):
        x = scene_block(dims[0], dims[1], dims[2], r, (2 * r, 2 * r, 2 * r), seed, cfg.get("jitter", 0.1), 0)
        plist.append(make_params(r, (0.0, 0.0, 0.0), tank, x.shape[0], viscosity=cfg.get("viscosity", 0.01),
                                 grid=GRID_FINE, **param_kw))
        xs.append(x)
    return plist, xs


def params(cfg: dict, capacity: int | None = None, grid=GRID_FINE, **kw):
    return make_params(cfg["radius"], cfg["domain_min"], cfg["domain_max"],
                       capacity or count(cfg), viscosity=cfg.get("viscosity", 0.01), grid=grid, **kw)


def policy(cfg: dict, mode=POLICY_CONVERGED, **kw):
    pol = default_policy(mode, tol_div=cfg["tol_div"], tol_density=cfg["tol_density"],
                         cfl=cfg["cfl"], dt_max=cfg["dt_max"])
    for k, v in kw.items():
        setattr(pol, k, v)
    return pol
