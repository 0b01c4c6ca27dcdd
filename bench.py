#!/usr/bin/env python3
"""Benchmark: DFSPH water substeps on one MI355X, particle-steps/s.

Workload (default): the 1,048,576-particle dam break (gpu-sims_amd/configs/
c2_dam_break_1m.json), converged solver at 0.1 % density and divergence
tolerance (every solve also keeps each particle's residual within 25 x the
tolerance), learned warm start.  Sequence:

0. autotuner (untimed, reported in config.tile): every launch shape and cell
   size candidate on the workload's own scene (dfsph_autotune); all runs below
   use its choice;
1. training run (untimed, reported): the same tank and block with another
   jitter seed, stepped by the plain solver while every stride-th particle's
   features and converged stiffness are recorded (about 4M rows), then the
   network is trained on the GPU by the library's own kernels;
2. headline run: W untimed warm-up steps, then K timed steps of the seeded
   scene with the learned warm start -> `value` (by default the network starts
   the density solve and the previous pressure the divergence solve; --warm mlp uses
   the network for both);
3. comparison runs, the same W + K steps from the same scene: plain (zero
   start), previous (the previous step's pressure, classic warm start), and
   the other network mode -> `plain`, `previous`, ..., `learned_vs_plain`;
4. horizon block (when K is not the configuration's step count): every policy
   over the configuration's whole run from rest (500 steps), unprofiled.

All timed runs of steps 2-3 are profiled the same way: HIP events around
every launch on the library's own stream, on every PROFILE_EVERY-th timed
step, read once they have completed (no extra host wait).  State is
resident in HBM throughout; the step's own host waits (the tolerance loop
reads one status word per chunk) are inside the timed region.
`roofline.achieved` is the algorithmic bytes of the dominant neighbour-loop
kernel divided by its mean launch duration from those events; the whole
substep and a measured device copy are reported beside it, and each sweep's
FP32 rate at 30 FLOP per true neighbour pair.  The CPU baseline runs the
oracle (the reference step restated in C++, OpenMP) for a few substeps of the
same workload from the GPU state after the timed steps, and the
32,768-particle dam break on all host threads and on one.

Prints one JSON line on stdout.

Helper modes (measurement plumbing, no JSON line):
  --copy-calibration     1 GiB device copies, the known-size kernel that scales
                         rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes
  --summarize-trace DIR  write a kernel-trace summary (kernel_stats.csv +
                         kernel_trace.csv under DIR) to --out
  --summarize-pmc F W CF CW  write the traffic summary of FETCH_SIZE / WRITE_SIZE
                         runs (bench, copy calibration) to --out and the
                         per-kernel HBM bytes per launch to --out-json
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-sims_amd"))

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3   # MI355X vector FP32
FLOP_PER_PAIR = 30      # per true neighbour pair and sweep (the FP32 fraction's count)
PROFILE_EVERY = 4       # per-kernel HIP events on every 4th timed step (host cost of events)

# Algorithmic bytes per particle per launch (FP32 SoA, compulsory reads and
# writes; neighbour reuse free; reorder traffic and neighbour masks not counted).
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
    "mlp_warm": 72,             # 16 features in (64 B), kappa out (8 B)
}
BOUNDARY_BYTES = {"density_alpha": 16}   # per boundary sample: position + volume, read once


def kernel_alg_bytes(name, kt):
    """ALG_BYTES of a profiled kernel; a sweep that took over a neighbouring
    stage (its launch then absent from kt) carries that stage's own bytes too:
    the forces sweep the divergence velocity update (a, p in; p_div out), the
    final velocity update the next cell build's keys (key out)."""
    b = ALG_BYTES[name]
    if name == "forces" and "accel_apply_div" not in kt:
        b += 24
    if name == "accel_apply_final" and "hash_keys" not in kt:
        b += 4
    return b
PAIR_SWEEPS = ["density_alpha", "pressure_accel", "jacobi", "forces", "density_adv"]
NEIGHBOUR_LOOPS = ["density_alpha", "pressure_accel", "jacobi", "accel_apply_div", "forces",
                   "density_adv", "accel_apply_final"]
# Profiler names -> kernel symbol in a rocprofv3 trace
KERNEL_SYMBOL = {"density_alpha": "k_density_alpha", "pressure_accel": "k_pressure_accel",
                 "jacobi": "k_jacobi", "accel_apply_div": "k_accel_apply",
                 "accel_apply_final": "k_accel_apply", "forces": "k_forces",
                 "density_adv": "k_density_adv", "hash_keys": "k_keys",
                 "hash_scatter": "k_scatter", "mlp_warm": "k_mlp_warm"}


def log_fn(verbose):
    return (lambda *a: print(*a, file=sys.stderr, flush=True)) if verbose else (lambda *a: None)


def run_timed(D, sim, pol, warmup, steps, log, label, profile=True):
    """W untimed steps, then K timed steps; returns solver statistics and, with
    profile, the per-kernel HIP-event times of every PROFILE_EVERY-th timed
    step (the library reads them once they complete: no extra host wait)."""
    for _ in range(warmup):
        sim.step(pol)
    if profile:
        sim.profile(PROFILE_EVERY)
    it_v = it_d = fallbacks = unconverged = syncs = 0
    err_max = 0.0
    D.lib()  # loaded before the clock starts
    sim.trace_mark(1)   # k_trace_mark pair: the timed steps, for trace summaries
    t0 = time.perf_counter()
    for s in range(steps):
        st = sim.step(pol)
        it_v += st["iters_div"]
        it_d += st["iters_density"]
        fallbacks += st["warm_fallback"]
        syncs += st["host_syncs"]
        err_max = max(err_max, st["err_density"])
        if not (st["converged_div"] and st["converged_density"]):
            unconverged += 1
    wall = time.perf_counter() - t0
    sim.trace_mark(2)
    kt = sim.kernel_times() if profile else None
    if profile:
        sim.profile(False)
    res = {"wall_s": wall, "ms_per_step": wall / steps * 1e3,
           "mean_iters_div": it_v / steps, "mean_iters_density": it_d / steps,
           "max_err_density": err_max, "warm_fallbacks": fallbacks, "steps_at_max_iter": unconverged,
           "host_syncs_per_step": syncs / steps, "profiled": f"every {PROFILE_EVERY}th step" if profile else "off"}
    log(f"{label}: {res['ms_per_step']:.3f} ms/step, iters {res['mean_iters_div']:.2f}/"
        f"{res['mean_iters_density']:.2f}, fallbacks {fallbacks}, max err {err_max:.2e}")
    return res, kt


def train_network(D, scenes, cfg, params_kw, log, steps=0):
    """Record (features, converged stiffness) from plain-solver steps of the
    training scene and train the network on the GPU; returns the parameters."""
    L = dict(cfg["learned"])
    if steps:
        L["steps"] = steps
    tcfg = dict(cfg, seed=L["seed"])
    x = scenes.positions(tcfg)
    n = x.shape[0]
    sim = D.Sim(scenes.params(cfg, **params_kw))
    sim.init_scene(x)
    plain = scenes.policy(cfg)
    cap = 2 * L["steps"] * (n // L["stride"] + 1)
    sim.dataset_start(cap, L["stride"])
    t0 = time.perf_counter()
    for _ in range(L["steps"]):
        sim.step(plain)
    rows = sim.dataset_stop()
    t_rec = time.perf_counter() - t0
    sim.mlp_init(L["init_seed"])
    t0 = time.perf_counter()
    tr = sim.mlp_train(epochs=L["epochs"], batch=L["batch"], lr=L["lr"],
                       expectile_density=L.get("expectile_density", 0.0))
    t_train = time.perf_counter() - t0
    theta = sim.mlp_get_params()
    sim.close()
    log(f"training: {rows} rows from {L['steps']} plain steps (seed {L['seed']}) in {t_rec:.2f} s; "
        f"{L['epochs']} epochs in {t_train:.2f} s, val loss {tr['val_loss']:.4f}")
    info = {"seed": L["seed"], "record_steps": L["steps"], "stride": L["stride"], "rows": rows,
            "train_rows": tr["train_rows"], "val_rows": tr["val_rows"], "epochs": L["epochs"],
            "expectile_density": L.get("expectile_density", 0.0),
            "adam_steps": tr["steps"], "train_loss": round(tr["train_loss"], 5),
            "val_loss": round(tr["val_loss"], 5), "record_s": round(t_rec, 3),
            "train_s": round(t_train, 3)}
    return theta, info


def measure_copy_gbs(nbytes=1 << 30, reps=5):
    """Device-to-device copy of nbytes through the HIP runtime the library
    already loaded (read + write bytes / time): the achievable-bandwidth line
    beside the vendor peak."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        return None
    e0, e1 = C.c_void_p(), C.c_void_p()
    hip.hipEventCreate(C.byref(e0))
    hip.hipEventCreate(C.byref(e1))
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    for _ in range(2):
        hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)
    hip.hipEventRecord(e0, None)
    for _ in range(reps):
        hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None)
    hip.hipEventRecord(e1, None)
    hip.hipEventSynchronize(e1)
    ms = C.c_float()
    hip.hipEventElapsedTime(C.byref(ms), e0, e1)
    for e in (e0, e1):
        hip.hipEventDestroy(e)
    hip.hipFree(a)
    hip.hipFree(b)
    return round(2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9, 1)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2_dam_break_1m")
    ap.add_argument("--warm", default="mlp-density", choices=["mlp-density", "mlp", "zero", "previous"],
                    help="warm start of the headline run (mlp-density = learned for the density solve, "
                         "previous pressure for the divergence solve; mlp = learned for both)")
    ap.add_argument("--no-plain", action="store_true", help="skip the comparison runs")
    ap.add_argument("--horizon", type=int, default=-1,
                    help="steps of the horizon block (-1: the configuration's step count; 0: none)")
    ap.add_argument("--train-steps", type=int, default=0,
                    help="recorded plain steps of the training run (0 = the configuration's)")
    ap.add_argument("--block-size", type=int, default=0,
                    help="threads per workgroup (with --flags and --cell-scale: skips the autotuner)")
    ap.add_argument("--flags", type=int, default=0)
    ap.add_argument("--cell-scale", type=float, default=0.0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="defaults (256 threads, masks, XCD remap, cell h) instead of the autotuner's choice")
    ap.add_argument("--cpu-steps", type=int, default=3, help="converged substeps of the workload on all host threads")
    ap.add_argument("--cpu-c1-steps", type=int, default=20,
                    help="steps of the 32,768-particle dam break on all host threads (200: the whole run)")
    ap.add_argument("--cpu-c1-steps-1t", type=int, default=4, help="the same on one thread")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-render", action="store_true", help="skip the render-pass timing")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--summarize-trace", metavar="DIR")
    ap.add_argument("--summarize-pmc", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"))
    ap.add_argument("--cmd", default="", help="command line recorded in a summary")
    ap.add_argument("--out")
    ap.add_argument("--out-json")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "decomposition of the dam break is a design, not a build.", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration()
    if args.summarize_trace:
        return summarize_trace(args.summarize_trace, args.out, args.cmd)
    if args.summarize_pmc:
        return summarize_pmc(*args.summarize_pmc, args.out, args.out_json, args.cmd, args.config)

    import dfsph as D
    from dfsph import scenes

    log = log_fn(args.verbose)
    cfg = scenes.load(args.config)
    if "batch" in cfg:
        return bench_batch(D, scenes, cfg, args, log)
    x = scenes.positions(cfg)
    n = x.shape[0]
    params_kw = dict(block_size=args.block_size, flags=args.flags, fine_cell_scale=args.cell_scale)
    tile = {"block_size": args.block_size or 256, "flags": args.flags, "fine_cell_scale": args.cell_scale,
            "autotuned": False}
    if not args.no_autotune and not (args.block_size or args.flags or args.cell_scale):
        # Untimed: every candidate launch shape and cell size on the workload's
        # own scene, plain solver, 2 timed substeps each.
        t0 = time.perf_counter()
        tr = D.autotune(scenes.params(cfg), x, scenes.policy(cfg), steps=2)
        params_kw = dict(block_size=tr["block_size"], flags=tr["flags"], fine_cell_scale=tr["fine_cell_scale"])
        tile = {"block_size": tr["block_size"], "flags": tr["flags"], "fine_cell_scale": tr["fine_cell_scale"],
                "autotuned": True, "candidates": tr["candidates"], "tune_s": round(time.perf_counter() - t0, 2),
                "ms_per_step_choice": round(tr["ms_per_step"], 4),
                "ms_per_step_slowest": round(tr["worst_ms_per_step"], 4)}
        log(f"autotune: {tile}")
    names = {"zero": D.WARM_ZERO, "previous": D.WARM_PREVIOUS, "mlp": D.WARM_MLP, "mlp-density": D.WARM_MLP_DENSITY}
    warm = names[args.warm]

    theta, train_info = (None, None)
    needs_net = warm in (D.WARM_MLP, D.WARM_MLP_DENSITY) or not args.no_plain
    if needs_net:
        theta, train_info = train_network(D, scenes, cfg, params_kw, log, args.train_steps)

    def new_sim():
        sm = D.Sim(scenes.params(cfg, **params_kw))
        sm.init_scene(x)
        if theta is not None:
            sm.mlp_set_params(theta)
        return sm

    # Headline and comparison runs: the same W + K steps from the same scene,
    # each profiled the same way (HIP events on every PROFILE_EVERY-th step).
    sim = new_sim()
    nb = sim.field_size(D.FIELD_BOUNDARY_VOLUME) // 4
    dv = D.derive(scenes.params(cfg, **params_kw))
    cells = int(dv.fine_axes[0]) * int(dv.fine_axes[1]) * int(dv.fine_axes[2])
    log(f"bench: {cfg['name']} n={n} boundary={nb} cells={cells} warm={args.warm}")
    head, kt = run_timed(D, sim, scenes.policy(cfg, warm_start=warm), args.warmup, args.steps, log,
                         f"headline ({args.warm})")
    mean_nbr = float(sim.export(D.FIELD_NEIGHBOR_COUNT).astype(np.float64).mean())
    render = None if args.no_render else render_frame(D, sim, cfg, log)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(cfg, sim, args, log)
    sim.close()

    comps = {}
    if not args.no_plain:
        for label, w in (("plain", "zero"), ("previous", "previous"), ("learned_density", "mlp-density"),
                         ("learned_both_solves", "mlp")):
            if w == args.warm:
                continue
            sm = new_sim()
            r, _ = run_timed(D, sm, scenes.policy(cfg, warm_start=names[w]), args.warmup, args.steps, log, label)
            sm.close()
            r["value"] = round(n * args.steps / r["wall_s"], 1)
            comps[label] = r

    # Horizon block: every policy over the configuration's whole run from the
    # scene at rest (no warm-up), unprofiled; reported, not the headline.
    horizon = cfg.get("steps", 0) if args.horizon < 0 else args.horizon
    hz = None
    if horizon and horizon != args.steps and not args.no_plain:
        hz = {"steps": horizon, "profiled": "off"}
        for label, w in (("plain", "zero"), ("previous", "previous"), ("learned_density", "mlp-density"),
                         ("learned_both_solves", "mlp")):
            sm = new_sim()
            r, _ = run_timed(D, sm, scenes.policy(cfg, warm_start=names[w]), 0, horizon, log,
                             f"horizon {label}", profile=False)
            sm.close()
            hz[label] = {"ms_per_step": round(r["ms_per_step"], 4), "mean_iters_div": round(r["mean_iters_div"], 3),
                         "mean_iters_density": round(r["mean_iters_density"], 3),
                         "warm_fallbacks": r["warm_fallbacks"], "steps_at_max_iter": r["steps_at_max_iter"]}
        p = hz["plain"]
        for label in ("previous", "learned_density", "learned_both_solves"):
            q = hz[label]
            q["iteration_cut_vs_plain"] = round(1 - (q["mean_iters_div"] + q["mean_iters_density"]) /
                                                (p["mean_iters_div"] + p["mean_iters_density"]), 4)
            q["step_time_cut_vs_plain"] = round(1 - q["ms_per_step"] / p["ms_per_step"], 4)

    value = n * args.steps / head["wall_s"]
    kernels = {}
    for name, t in kt.items():
        rec = {"launches": t["launches"], "avg_us": round(t["avg_ms"] * 1e3, 2),
               "total_ms": round(t["total_ms"], 3)}
        if name in ALG_BYTES and t["avg_ms"] > 0:
            b = kernel_alg_bytes(name, kt) * n + BOUNDARY_BYTES.get(name, 0) * nb
            rec["alg_bytes"] = b
            rec["achieved_gbs"] = round(b / (t["avg_ms"] * 1e-3) / 1e9, 1)
            rec["hbm_frac"] = round(rec["achieved_gbs"] / HBM_PEAK_GBS, 4)
        if name in PAIR_SWEEPS and t["avg_ms"] > 0:
            tf = FLOP_PER_PAIR * mean_nbr * n / (t["avg_ms"] * 1e-3) / 1e12
            rec["fp32_tflops"] = round(tf, 2)
            rec["fp32_frac"] = round(tf / FP32_PEAK_TFLOPS, 4)
        kernels[name] = rec
    dom = max((k for k in NEIGHBOUR_LOOPS if k in kernels), key=lambda k: kernels[k]["total_ms"])
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{cfg['name']}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            per_launch = json.load(f).get("bytes_per_launch", {})
            sym = KERNEL_SYMBOL.get(dom, dom)
            # the forces sweep that applies the divergence update is k_forces_div
            traffic = per_launch.get(sym, per_launch.get(sym + "_div"))
    ach = kernels[dom]["achieved_gbs"]
    roofline = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "alg_bytes_per_particle": kernel_alg_bytes(dom, kt), "particles_per_launch": n,
                "avg_us": kernels[dom]["avg_us"], "launches_timed": kernels[dom]["launches"],
                "window": (f"HIP events on the library stream, every {PROFILE_EVERY}th of timed steps "
                           f"{args.warmup + 1}-{args.warmup + args.steps} of the headline run")}
    # Whole substep against the same roof: algorithmic bytes of all stages at
    # the run's mean iteration counts over the measured step time.
    mlp_calls = {D.WARM_MLP: 2, D.WARM_MLP_DENSITY: 1}.get(warm, 0)
    bpp = alg_bytes_substep(head["mean_iters_div"], head["mean_iters_density"], cells / n, nb / n, mlp_calls)
    sub_gbs = bpp * n / (head["ms_per_step"] * 1e-3) / 1e9
    copy_gbs = measure_copy_gbs()
    roofline["substep"] = {"alg_bytes_per_particle": round(bpp, 1), "achieved": round(sub_gbs, 1),
                           "frac": round(sub_gbs / HBM_PEAK_GBS, 4), "ms_per_step": round(head["ms_per_step"], 4)}
    roofline["copy_gbs"] = copy_gbs
    roofline["copy_frac"] = round(copy_gbs / HBM_PEAK_GBS, 4) if copy_gbs else None

    def solver(r):
        return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items() if k != "wall_s"}

    plain = comps.get("plain")
    out = {
        "metric": f"particle-steps/s @{n / 1e6:.1f}M DFSPH, rho-err<=0.1%; learned warm start vs plain solver"
                  .replace(".0M", "M"),
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": cfg["name"], "particles": n, "boundary_particles": nb, "cells": cells,
                   "solver": "converged", "warm_start": args.warm, "tol_density": cfg["tol_density"],
                   "tol_div": cfg["tol_div"], "residual_bound": "25 x tolerance (every solve)",
                   "grid": "fine", "tile": tile},
        "solver": solver(head),
        **{k: solver(v) for k, v in comps.items()},
        "learned_vs_plain": ({
            "iteration_cut": round(1 - (head["mean_iters_div"] + head["mean_iters_density"])
                                   / (plain["mean_iters_div"] + plain["mean_iters_density"]), 4),
            "density_iteration_cut": round(1 - head["mean_iters_density"] / plain["mean_iters_density"], 4),
            "step_time_cut": round(1 - head["ms_per_step"] / plain["ms_per_step"], 4),
            "speedup": round(plain["ms_per_step"] / head["ms_per_step"], 3),
            "window": f"timed steps {args.warmup + 1}-{args.warmup + args.steps}, both profiled the same way"}
            if plain else None),
        "horizon": hz,
        "training": train_info,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "render": render,
    }
    print(json.dumps(out), flush=True)
    return 0


def render_frame(D, sim, cfg, log, width=1920, height=1080, reps=5):
    """The screen-space render passes over the final state of the headline
    run (untimed for the metric): one 1920x1080 frame with the reference's
    settings, the camera framing the fluid; ms per frame including the colour
    readback to the host."""
    x = sim.export(D.FIELD_POSITION).reshape(-1, 3).astype(np.float64)
    centre = x.mean(axis=0)
    ext = float(np.max(x.max(axis=0) - x.min(axis=0)))
    eye = centre + np.array([0.3, 0.5, 1.2]) * ext
    p = D.render_defaults(eye.tolist(), centre.tolist(), 60.0, width, height)
    img = sim.render(p, rgba=True, depth=True, smoothed=False, thickness=False)
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.render(p, rgba=True, depth=False, smoothed=False, thickness=False)
    ms = (time.perf_counter() - t0) / reps * 1e3
    covered = float((img["depth"] < 1.0).mean())
    log(f"render: {width}x{height}, {ms:.2f} ms per frame, {100 * covered:.1f} % of pixels covered")
    return {"width": width, "height": height, "ms_per_frame": round(ms, 3), "covered_fraction": round(covered, 4),
            "passes": "depth sprites, 4 bilateral iterations, thickness, composite; rgba read back"}


def alg_bytes_substep(it_div, it_den, cells_per_particle=0.0, boundary_per_particle=0.0, mlp_calls=0):
    """Algorithmic bytes per particle of one substep with the given iteration
    counts: fixed stages 296 B + 76 B per Jacobi iteration (both sweeps), the
    cell-range table (8 B a cell), the boundary samples (16 B each, read by
    the density pass, the sources and twice per iteration) and 72 B per
    network call (16 features in, 2 outputs out)."""
    k = it_div + it_den
    return (296.0 + 76.0 * k + 8.0 * cells_per_particle + 16.0 * boundary_per_particle * (5 + 2 * k)
            + 72.0 * mlp_calls)


def bench_batch(D, scenes, cfg, args, log):
    """Batched scenes: autotuned persistent grid, K timed substeps of every
    scene, beside the same scenes stepped one after another by single-scene
    contexts (per-stage launches)."""
    warm_name = "previous" if args.warm in ("mlp", "mlp-density") else args.warm
    warm = {"zero": D.WARM_ZERO, "previous": D.WARM_PREVIOUS}[warm_name]
    plist, xs = scenes.batch_scenes(cfg)
    n = int(sum(x.shape[0] for x in xs))
    pol = scenes.policy(cfg, warm_start=warm)
    t0 = time.perf_counter()
    tune = D.batch_autotune(plist, [(x,) for x in xs], pol, steps=3)
    t_tune = time.perf_counter() - t0
    log(f"batch autotune ({t_tune:.1f} s): {tune}")
    bt = D.Batch(plist, [(x,) for x in xs], workgroups=tune["workgroups"], flags=tune["batch_flags"])
    info = bt.info()
    for _ in range(args.warmup):
        bt.step(pol)
    it_v = it_d = 0.0
    err_max = 0.0
    unconv = 0
    gpu_ms = 0.0
    bt.trace_mark(1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        per, tot = bt.step(pol)
        gpu_ms += tot["gpu_ms"]
        it_v += sum(q["iters_div"] * c for q, c in zip(per, bt.counts)) / n
        it_d += sum(q["iters_density"] * c for q, c in zip(per, bt.counts)) / n
        err_max = max(err_max, tot["err_density"])
        unconv += int(not (tot["converged_div"] and tot["converged_density"]))
    wall = time.perf_counter() - t0
    bt.trace_mark(2)
    bt.close()
    value = n * args.steps / wall
    log(f"batched: {wall / args.steps * 1e3:.3f} ms/substep, {value:.3e} particle-steps/s")

    # Same scenes, one single-scene context each, stepped in turn (fewer steps).
    seq_steps = max(1, min(args.steps, 20))
    sims = []
    for p, x in zip(plist, xs):
        p.block_size = 256
        sm = D.Sim(p)
        sm.init_scene(x)
        for _ in range(min(args.warmup, 5)):
            sm.step(pol)
        sims.append(sm)
    t0 = time.perf_counter()
    for _ in range(seq_steps):
        for sm in sims:
            sm.step(pol)
    seq_wall = time.perf_counter() - t0
    for sm in sims:
        sm.close()
    seq_value = n * seq_steps / seq_wall
    log(f"one context per scene: {seq_wall / seq_steps * 1e3:.3f} ms per substep of all scenes")

    mean_it_v, mean_it_d = it_v / args.steps, it_d / args.steps
    bpp = alg_bytes_substep(mean_it_v, mean_it_d, info["cells"] / n, info["boundary_particles"] / n)
    launch_s = gpu_ms / args.steps * 1e-3   # HIP events around each graph replay
    ach = bpp * n / launch_s / 1e9
    # HBM bytes per launch of the persistent kernel, from the committed counter
    # summary of this workload (rocprofv3 FETCH_SIZE / WRITE_SIZE).
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{cfg['name']}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch", {}).get("k_batch_substep")
    out = {
        "metric": "particle-steps/s, batched DFSPH scenes (64 scenes, 4k-128k particles), rho-err<=0.1%",
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
        "config": {"workload": cfg["name"], "scenes": len(xs), "particles": n,
                   "boundary_particles": info["boundary_particles"], "solver": "converged",
                   "warm_start": warm_name, "tol_density": cfg["tol_density"], "grid": "fine",
                   "launch": "one persistent kernel per substep, HIP graph replay",
                   "workgroups": info["workgroups"], "batch_flags": tune["batch_flags"]},
        "solver": {"mean_iters_div": round(mean_it_v, 3), "mean_iters_density": round(mean_it_d, 3),
                   "max_err_density": err_max, "steps_at_max_iter": unconv},
        "autotune": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in tune.items()},
        "one_context_per_scene": {"value": round(seq_value, 1), "steps": seq_steps,
                                  "ms_per_step": round(seq_wall / seq_steps * 1e3, 4),
                                  "speedup_of_batch": round(value / seq_value, 3)},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "batch_substep",
                     "alg_bytes_per_particle": round(bpp, 1), "particles_per_launch": n,
                     "avg_us": round(launch_s * 1e6, 1)},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, sim, args, log):
    """The oracle (the reference's step restated in C++, OpenMP, the reference's
    2h Morton grid) on the host: a few converged substeps of the workload from
    the GPU's state (all threads) -> value; and the 32,768-particle dam break
    from rest, on all threads and on one."""
    import dfsph as D
    from dfsph import scenes

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.build()
    threads = O.max_threads()

    def steps_of(ocfg, x, v, steps, nthreads):
        O.set_threads(nthreads)
        ora = O.OracleSim.create(ocfg, x, v)
        opol = O.policy("converged", tol_div=cfg["tol_div"], tol_density=cfg["tol_density"], cfl=cfg["cfl"],
                        dt_max=cfg["dt_max"])
        its = []
        t0 = time.perf_counter()
        for _ in range(steps):
            st = ora.step(0.0, opol)
            its.append(f"{st['iters_div']}/{st['iters_density']}")
        dt = time.perf_counter() - t0
        del ora
        O.set_threads(threads)
        return dt, its

    x = sim.export(D.FIELD_POSITION)
    v = sim.export(D.FIELD_VELOCITY)
    ocfg = O.make_config(cfg["radius"], cfg["domain_min"], cfg["domain_max"], viscosity=cfg.get("viscosity", 0.01))
    dt, its = steps_of(ocfg, x, v, args.cpu_steps, threads)
    log(f"cpu baseline: {dt:.2f} s for {args.cpu_steps} substeps, iters {its}")
    out = {"value": round(x.shape[0] * args.cpu_steps / dt, 1), "unit": "particle-steps/s", "cores": threads,
           "kind": "port",
           "sample": (f"{args.cpu_steps} converged zero-start substeps of {cfg['name']} ({x.shape[0]} particles) "
                      f"from the GPU state after the timed steps; oracle (C++ restatement of the "
                      f"reference step, OpenMP, reference 2h Morton grid); {dt:.2f} s wall; "
                      f"iterations div/density {', '.join(its)}"),
           "cpu_model": _cpu_model()}
    c1 = scenes.load("dam_break_32k")
    x1 = scenes.positions(c1)
    c1cfg = O.make_config(c1["radius"], c1["domain_min"], c1["domain_max"], viscosity=c1.get("viscosity", 0.01))
    for key, steps, nthreads in (("c1_all_threads", args.cpu_c1_steps, threads),
                                 ("c1_one_thread", args.cpu_c1_steps_1t, 1)):
        if steps <= 0:
            continue
        dt, its = steps_of(c1cfg, x1, None, steps, nthreads)
        it = np.array([[int(a) for a in s.split("/")] for s in its], float).mean(axis=0)
        out[key] = {"value": round(x1.shape[0] * steps / dt, 1), "unit": "particle-steps/s", "threads": nthreads,
                    "ms_per_step": round(dt / steps * 1e3, 2), "steps": steps,
                    "sample": (f"dam_break_32k (32,768 particles), converged policy from rest, steps 1-{steps}; "
                               f"mean iterations div/density {it[0]:.2f}/{it[1]:.2f}")}
        log(f"cpu baseline {key}: {dt / steps * 1e3:.1f} ms/step over {steps} steps")
    return out


# ---------------------------------------------------------------------------
# Measurement plumbing: copy calibration and rocprofv3 summaries

def copy_calibration() -> int:
    """Device-to-device copies of 1 GiB: the kernel of known traffic that scales
    FETCH_SIZE / WRITE_SIZE (KiB per dispatch) to bytes."""
    import torch
    n = (1 << 30) // 4
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    t = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t[0].record()
    for _ in range(5):
        b.copy_(a)
    t[1].record()
    torch.cuda.synchronize()
    ms = t[0].elapsed_time(t[1]) / 5
    print(f"copy 1 GiB: {ms:.3f} ms, {2 * (1 << 30) / ms / 1e6:.1f} GB/s (read + write)", flush=True)
    return 0


def _short(name: str) -> str:
    m = re.search(r"\b(k_\w+)(<(\d)>)?", name)
    if m:
        return m.group(1) + ({"0": "<reference>", "1": "<fine>"}.get(m.group(3), "") if m.group(2) else "")
    return name.split("(")[0][:48]


def _csv(directory: str, suffix: str):
    hits = sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {directory}")
    with open(hits[0]) as f:
        return list(csv.DictReader(f))


def _marked(rows):
    """Rows (one or more per dispatch) between the first two k_trace_mark
    dispatches, in dispatch order: the timed steps of the first measured run
    (bench marks each run's timed region; autotuning and training come
    before it).  Returns (rows, number of marks seen)."""
    marks = sorted({int(r["Dispatch_Id"]) for r in rows if "k_trace_mark" in r["Kernel_Name"]})
    if len(marks) < 2:
        return [], len(marks)
    a, b = marks[0], marks[1]
    return [r for r in rows if a < int(r["Dispatch_Id"]) < b], len(marks)


def summarize_trace(directory: str, out: str, cmd: str, skip_us: float = 5.0) -> int:
    stats = _csv(directory, "kernel_stats.csv")
    trace = _csv(directory, "kernel_trace.csv")
    win, nmarks = _marked(trace)
    lines = ["# rocprofv3 --kernel-trace --stats summary", f"# command: {cmd}", ""]

    def table(rows, title):
        d = collections.defaultdict(list)
        for r in rows:
            d[_short(r["Kernel_Name"])].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
        tot = sum(sum(v) for v in d.values()) or 1.0
        lines.extend(title)
        lines.append(f"{'kernel':36s} {'calls':>7s} {'avg_us':>9s} {'>=5us':>7s} {'avg_us':>9s} {'pct':>6s}")
        for k, v in sorted(d.items(), key=lambda kv: -sum(kv[1])):
            if not k.startswith("k_") or k == "k_trace_mark":
                continue
            w = [t for t in v if t >= skip_us]
            lines.append(f"{k:36s} {len(v):7d} {sum(v) / len(v):9.2f} {len(w):7d} "
                         f"{(sum(w) / len(w) if w else 0.0):9.2f} {100 * sum(v) / tot:6.2f}")
        lines.append("")

    table(win, [f"# timed steps of the headline run: the {len(win)} dispatches between the first two",
                "# k_trace_mark launches (bench marks each run's timed steps; autotuning and",
                f"# training come before); {nmarks} marks in the trace.  '>=5us' keeps launches of",
                "# at least 5 us: shorter ones are solver iterations past convergence (the kernel",
                "# reads the done flag and returns), which the library's HIP-event profiler skips"])
    lines += ["# whole command (autotune candidates, training, every run), rocprofv3 --stats",
              f"{'kernel':36s} {'calls':>7s} {'total_ms':>10s} {'avg_us':>9s} {'min_us':>8s} {'max_us':>8s} {'pct':>6s}"]
    for r in stats:
        lines.append(f"{_short(r['Name']):36s} {int(r['Calls']):7d} {float(r['TotalDurationNs']) / 1e6:10.3f} "
                     f"{float(r['AverageNs']) / 1e3:9.2f} {float(r['MinNs']) / 1e3:8.2f} "
                     f"{float(r['MaxNs']) / 1e3:8.2f} {float(r['Percentage']):6.2f}")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))
    return 0


def summarize_pmc(fetch_dir, write_dir, cfetch_dir, cwrite_dir, out, out_json, cmd, config) -> int:
    def agg(directory, counter, marked=True):
        rows = [r for r in _csv(directory, "counter_collection.csv") if r["Counter_Name"] == counter]
        if marked:
            rows, nm = _marked(rows)
            if not rows:
                raise RuntimeError(f"{directory}: {nm} k_trace_mark dispatches, need two around the timed steps")
        d = collections.defaultdict(list)
        for r in rows:
            d[_short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        return d

    def copy_kib(directory, counter):
        d = agg(directory, counter, marked=False)
        vals = [v for k, vs in d.items() if not k.startswith("k_") for v in vs if v > 0]
        return max(vals)   # the 1 GiB copies dominate every other dispatch in that run

    known_kib = (1 << 30) / 1024.0
    cf, cw = copy_kib(cfetch_dir, "FETCH_SIZE"), copy_kib(cwrite_dir, "WRITE_SIZE")
    sf, sw = known_kib / cf, known_kib / cw
    f, w = agg(fetch_dir, "FETCH_SIZE"), agg(write_dir, "WRITE_SIZE")
    lines = ["# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (separate runs), per dispatch",
             f"# command: {cmd}",
             f"# calibration: a 1 GiB device copy reads FETCH_SIZE {cf:.0f} KiB -> scale {sf:.3f};"
             f" WRITE_SIZE {cw:.0f} KiB -> scale {sw:.3f}", "",
             f"{'kernel':36s} {'dispatches':>10s} {'fetch_MB':>9s} {'write_MB':>9s} {'traffic_MB':>10s}"]
    lines.insert(3, "# timed steps of the headline run only (dispatches between the first two k_trace_mark")
    lines.insert(4, "# launches); dispatches reading under 2 % of that kernel's largest fetch are dropped:")
    lines.insert(5, "# solver iterations past convergence, which read the done flag and return")
    res = {}
    for k in sorted(f):
        if not k.startswith("k_"):
            continue
        top = max(f[k])
        fl = [v for v in f[k] if v >= 0.02 * top] or [0.0]
        wl = w.get(k, [0.0])
        wl = [v for v in wl if v >= 0.02 * max(wl)] or [0.0]
        fm = sum(fl) / len(fl) * sf * 1024 / 1e6
        wm = sum(wl) / len(wl) * sw * 1024 / 1e6
        lines.append(f"{k:36s} {len(fl):10d} {fm:9.2f} {wm:9.2f} {fm + wm:10.2f}")
        res[k.split("<")[0]] = round((fm + wm) * 1e6)
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    if out_json:
        with open(out_json, "w") as fh:
            json.dump({"command": cmd, "config": config, "fetch_scale": round(sf, 4),
                       "write_scale": round(sw, 4), "bytes_per_launch": res}, fh, indent=1)
    print("\n".join(lines))
    return 0


if __name__ == "__main__":
    sys.exit(main())
