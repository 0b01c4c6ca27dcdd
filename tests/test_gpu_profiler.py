"""The per-kernel profiler (dfsph_profile_enable / dfsph_profile_read, the
counterpart of the reference's GpuProfiler scopes, gpu_profiler.cpp:55-132).

Every n-th step is timed, launch by launch: each timed launch's interval
starts at the kernel's own dispatch (the start event is bound to the launch)
and ends after it, so the intervals of one step are disjoint pieces of that
step's stream time.  Checked on a 32k-particle dam break: the sampled steps
are the ones asked for, the per-pass launch counts follow the step structure
(one density pass per step, at least one Jacobi launch per reported solver
iteration), and the timed kernels add up to no more than the steps' own GPU
time (cell builds aside, which run ahead of the step they serve)."""
from __future__ import annotations

import pytest

from conftest import dam_break_small

pytestmark = pytest.mark.gpu


def test_profiler_samples_and_accounts_for_the_step(dfsph_mod):
    D = dfsph_mod
    print("case: profiler on a 32k dam break, every 2nd of 12 steps", flush=True)
    x, lo, hi = dam_break_small(32)
    sim = D.Sim(D.make_params(0.01, lo, hi, capacity=x.shape[0]))
    sim.init_scene(x)
    pol = D.default_policy(D.POLICY_CONVERGED)
    for _ in range(4):
        sim.step(pol)
    sim.profile(2)
    timed, iters, gpu_ms = 0, 0, 0.0
    for s in range(12):
        st = sim.step(pol)
        # which steps are timed depends only on their parity: half of 12
        iters_s = st["iters_div"] + st["iters_density"]
        gpu_ms_s = st["gpu_ms"]
        kt = sim.kernel_times()
        n_now = kt.get("density_alpha", {}).get("launches", 0)
        if n_now > timed:
            timed = n_now
            iters += iters_s
            gpu_ms += gpu_ms_s
    kt = sim.kernel_times()
    sim.close()
    assert timed == 6
    for name in ("reorder", "density_alpha", "forces", "density_adv"):
        assert kt[name]["launches"] == 6, (name, kt[name])
    # the substep's finish rides on the next cell build's first scan launch
    assert "step_finish" not in kt
    # a step's cell build is launched at the end of the step before it (and
    # timed with the step it serves, once profiling was on when it launched);
    # its keys pass rides on the velocity update after the density solve
    for name in ("hash_scan", "hash_scatter"):
        assert kt[name]["launches"] in (5, 6), (name, kt[name])
    assert "hash_keys" not in kt
    assert kt["accel_apply_final"]["launches"] == 6
    assert kt["jacobi"]["launches"] >= iters
    assert kt["jacobi"]["launches"] <= iters + 12
    for name, t in kt.items():
        assert 0.0 < t["min_ms"] <= t["avg_ms"] <= t["max_ms"], (name, t)
    # gpu_ms covers a step's launches from its first to its last, without the
    # next step's cell build
    total = sum(t["total_ms"] for k, t in kt.items() if not k.startswith("hash_"))
    print(f"  timed kernels {total:.3f} ms of {gpu_ms:.3f} ms step GPU time; jacobi "
          f"{kt['jacobi']['launches']} launches for {iters} iterations, avg {kt['jacobi']['avg_ms'] * 1e3:.1f} us",
          flush=True)
    assert total <= gpu_ms * 1.001
