#!/usr/bin/env python3
"""Benchmark: DFSPH water substeps on one MI355X, particle-steps/s.

Workload (default): the 1,048,576-particle dam break (gpu-sims_amd/configs/
c2_dam_break_1m.json) with the converged solver policy at 0.1 % density and
divergence tolerance.  W untimed warm-up steps, then K timed steps of the same
simulation; inputs are resident in HBM throughout.  Per-kernel durations come
from the library's HIP-event profiler on its own stream over the timed region;
`roofline.achieved` is the algorithmic bytes of the dominant neighbour-loop
kernel divided by its mean launch duration.  The CPU baseline runs the oracle
(the reference step restated in C++, OpenMP) for one substep of the same
workload, started from the GPU state after the timed steps.

Prints one JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-sims_amd"))

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E

# Algorithmic bytes per particle per launch (FP32 SoA, compulsory reads and
# writes; neighbour reuse free; reorder traffic and pair lists not counted).
# Fused kernels carry the sum of the stages they fuse.
ALG_BYTES = {
    "hash_keys": 20,            # x in; key, id out
    "hash_scatter": 16,         # key, id in; key, id out
    "density_alpha": 20 + 32,   # rho, alpha (x in) + divergence source (x, v, rho in; rho_dot out)
    "pressure_accel": 32,       # x, rho, p in; a out
    "jacobi": 44,               # x, rho, a, p, alpha, s in; p out
    "accel_apply_div": 44,      # x, rho, p, v in; v out
    "forces": 40,               # x, v, rho in; v out
    "density_adv": 32,          # x, v, rho in; rho* out
    "accel_apply_final": 44 + 48,   # solve finish + position update (x, v in; x, v out)
}
NEIGHBOUR_LOOPS = ["density_alpha", "pressure_accel", "jacobi", "accel_apply_div", "forces",
                   "density_adv", "accel_apply_final"]


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2_dam_break_1m")
    ap.add_argument("--warm", default="zero", choices=["zero", "previous"])
    ap.add_argument("--block-size", type=int, default=0)
    ap.add_argument("--flags", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "decomposition of the dam break is a design, not a build.", file=sys.stderr)
        return 2

    import numpy as np

    import dfsph as D
    from dfsph import scenes

    cfg = scenes.load(args.config)
    x = scenes.positions(cfg)
    n = x.shape[0]
    sim = D.Sim(scenes.params(cfg, block_size=args.block_size, flags=args.flags))
    sim.init_scene(x)
    nb = sim.field_size(D.FIELD_BOUNDARY_VOLUME) // 4
    warm = {"zero": D.WARM_ZERO, "previous": D.WARM_PREVIOUS}[args.warm]
    pol = scenes.policy(cfg, warm_start=warm)
    log = (lambda *a: print(*a, file=sys.stderr, flush=True)) if args.verbose else (lambda *a: None)
    log(f"bench: {cfg['name']} n={n} boundary={nb}")

    for s in range(args.warmup):
        sim.step(pol)
    sim.profile(True)
    it_v = it_d = 0
    err_max = 0.0
    t0 = time.perf_counter()
    for s in range(args.steps):
        st = sim.step(pol)
        it_v += st["iters_div"]
        it_d += st["iters_density"]
        err_max = max(err_max, st["err_density"])
        if not (st["converged_div"] and st["converged_density"]):
            log(f"step {s}: solver stopped at max_iter")
    t1 = time.perf_counter()
    wall = t1 - t0
    kt = sim.kernel_times()
    sim.profile(False)

    value = n * args.steps / wall
    kernels = {}
    for name, t in kt.items():
        rec = {"launches": t["launches"], "avg_us": round(t["avg_ms"] * 1e3, 2),
               "total_ms": round(t["total_ms"], 3)}
        if name in ALG_BYTES:
            b = ALG_BYTES[name] * n + (16 * nb if name == "density_alpha" else 0)
            rec["alg_bytes"] = b
            rec["achieved_gbs"] = round(b / (t["avg_ms"] * 1e-3) / 1e9, 1)
        kernels[name] = rec
    dom = max((k for k in NEIGHBOUR_LOOPS if k in kernels), key=lambda k: kernels[k]["total_ms"])
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{cfg['name']}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get(dom)
    ach = kernels[dom]["achieved_gbs"]
    roofline = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(cfg, sim, pol, log)

    out = {
        "metric": "particle-steps/s @1M DFSPH, rho-err<=0.1%; learned warm start vs plain solver",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": cfg["name"], "particles": n, "boundary_particles": nb,
                   "solver": "converged", "warm_start": args.warm, "tol_density": cfg["tol_density"],
                   "tol_div": cfg["tol_div"], "grid": "fine"},
        "solver": {"mean_iters_div": round(it_v / args.steps, 3),
                   "mean_iters_density": round(it_d / args.steps, 3),
                   "max_err_density": err_max},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline(cfg, sim, pol, log):
    """One converged substep of the same workload on the host cores, run by the
    oracle from the GPU's current state."""
    import dfsph as D

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.build()
    x = sim.export(D.FIELD_POSITION)
    v = sim.export(D.FIELD_VELOCITY)
    ocfg = O.make_config(cfg["radius"], cfg["domain_min"], cfg["domain_max"],
                         viscosity=cfg.get("viscosity", 0.01))
    ora = O.OracleSim.create(ocfg, x, v)
    opol = O.policy("converged", tol_div=cfg["tol_div"], tol_density=cfg["tol_density"],
                    cfl=cfg["cfl"], dt_max=cfg["dt_max"])
    t0 = time.perf_counter()
    st = ora.step(0.0, opol)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    log(f"cpu baseline: {dt:.2f} s for one substep, iters {st['iters_div']}/{st['iters_density']}")
    return {"value": round(x.shape[0] / dt, 1), "unit": "particle-steps/s", "cores": cores,
            "kind": "port",
            "sample": (f"1 converged substep of {cfg['name']} ({x.shape[0]} particles) from the GPU "
                       f"state after the timed steps; oracle (C++ restatement of the reference "
                       f"step, OpenMP, reference 2h Morton grid); {dt:.2f} s wall; iterations "
                       f"{st['iters_div']}/{st['iters_density']}")}


if __name__ == "__main__":
    sys.exit(main())
