"""Screen-space render passes on the GPU (dfsph_render) against the render
oracle on the same exported positions.

Bars: the depth image bit-exact (same coverage, same winning particle -- the
highest index, as the reference's depth pass without a depth test leaves it --
same impostor depth); thickness within 1e-5 relative (the GPU sums fixed-point
contributions, the oracle blends floats in draw order); the smoothed depth
within 2e-6 (the GPU's expf against the host library's); the composite stage,
run by the oracle on the GPU's own smoothed depth and thickness, within 1e-5;
end to end, at most 1 % of pixels' colours more than 1e-3 apart (normals are
differences of NDC depth across a pixel quad, which magnify a last-place
difference of the smoothed depth at silhouettes).  Cases: the reference's settings over a stepped dam break, the
non-reference switches (alternating smoothing, thickness in the composite,
roughness), a close-up with large sprites, a 262k-particle frame at 1280x720,
argument errors, and that rendering leaves the simulation untouched."""
from __future__ import annotations

import time

import numpy as np
import pytest

from conftest import dam_break_small

pytestmark = pytest.mark.gpu


def _scene(D, n_side, steps=3, box=(2.0, 1.0, 1.0)):
    x, lo, hi = dam_break_small(n_side, box=box)
    sim = D.Sim(D.make_params(0.01, lo, hi, capacity=x.shape[0]))
    sim.init_scene(x)
    pol = D.default_policy(D.POLICY_CONVERGED)
    for _ in range(steps):
        sim.step(pol)
    return sim, lo, hi


def _compare(D, O, sim, params, label):
    got = sim.render(params)
    x = sim.export(D.FIELD_POSITION).reshape(-1, 3)
    want = O.render_from(params, x)
    covered = int((got["depth"] < 1.0).sum())
    assert covered > 100
    assert np.array_equal(got["depth"], want["depth"]), label
    assert np.allclose(got["thickness"], want["thickness"], rtol=1e-5, atol=1e-9), label
    assert np.allclose(got["smoothed"], want["smoothed"], rtol=0, atol=2e-6), label
    # The composite stage on the GPU's own inputs: expf / powf rounding only.
    p = params
    stage = O.render_composite(got["smoothed"], got["thickness"], list(p.view), list(p.proj), list(p.camera_pos),
                               list(p.water_tint), p.roughness, list(p.sky_zenith), list(p.sky_horizon), p.exposure,
                               p.fresnel_f0, p.absorption, p.thickness_scale)
    assert np.allclose(got["rgba"], stage, rtol=0, atol=1e-5), label
    # End to end, normals come from differences of NDC depth between
    # neighbouring pixels, which turn a last-place difference of the smoothed
    # depth into a visible one at silhouettes: most pixels agree closely.
    off = np.abs(got["rgba"] - want["rgba"]).max(axis=-1)
    frac = float((off > 1e-3).mean())
    print(f"  {label}: {covered} covered pixels of {p.width * p.height}; colour off by > 1e-3 at "
          f"{100 * frac:.3f} % of pixels, median {np.median(off):.1e}", flush=True)
    assert frac <= 0.01 and float(np.median(off)) <= 1e-5, label
    return got


def test_render_matches_the_oracle(dfsph_mod, oracle_mod):
    D, O = dfsph_mod, oracle_mod
    print("case: render passes, 4096-particle dam break after 3 steps", flush=True)
    sim, lo, hi = _scene(D, 16)
    pr = float(sim.params.particle_radius)
    center = [(lo[a] + hi[a]) / 2 for a in range(3)]
    eye = [center[0] + 0.4, center[1] + 0.6, center[2] + 2.2]
    p = D.render_defaults(eye, center, 60.0, 320, 180, particle_radius=pr)
    _compare(D, O, sim, p, "reference settings, 320x180")
    p = D.render_defaults(eye, center, 60.0, 320, 180, particle_radius=pr, bilateral_alternate=1,
                          bilateral_iterations=5, thickness_scale=40.0, roughness=0.3, exposure=1.3)
    _compare(D, O, sim, p, "alternating smoothing, thickness, roughness")
    mid = sim.export(D.FIELD_POSITION).reshape(-1, 3).mean(axis=0).tolist()
    near = [mid[0] + 0.05, mid[1] + 0.15, mid[2] + 0.4]
    p = D.render_defaults(near, mid, 50.0, 256, 256, particle_radius=pr, sigma_spatial=5.0)
    _compare(D, O, sim, p, "close-up, large sprites")
    sim.close()


def test_render_262k_frame(dfsph_mod, oracle_mod):
    D, O = dfsph_mod, oracle_mod
    print("case: render passes, 262144 particles at 1280x720", flush=True)
    sim, lo, hi = _scene(D, 64, steps=2, box=(3.0, 2.0, 2.0))
    center = [(lo[a] + hi[a]) / 2 for a in range(3)]
    p = D.render_defaults([center[0] + 1.0, center[1] + 1.2, center[2] + 3.0], center, 60.0, 1280, 720,
                          particle_radius=float(sim.params.particle_radius))
    sim.render(p)   # buffers sized, kernels loaded
    t0 = time.perf_counter()
    for _ in range(5):
        sim.render(p, depth=False, smoothed=False, thickness=False)
    ms = (time.perf_counter() - t0) / 5 * 1e3
    print(f"  {ms:.2f} ms per frame including the colour readback", flush=True)
    _compare(D, O, sim, p, "1280x720")
    sim.close()


def test_render_arguments_and_no_side_effects(dfsph_mod):
    D = dfsph_mod
    print("case: render argument errors; a render between steps changes nothing", flush=True)
    a, lo, hi = _scene(D, 12, steps=2)
    b, _, _ = _scene(D, 12, steps=2)
    center = [(lo[k] + hi[k]) / 2 for k in range(3)]
    p = D.render_defaults([center[0], center[1] + 0.5, center[2] + 2.0], center, 60.0, 64, 48)
    for field, bad in (("width", 0), ("bilateral_iterations", 65), ("sigma_spatial", 0.0),
                       ("point_scale", float("nan")), ("thickness_per_particle", -1.0),
                       ("exposure", float("inf"))):
        q = D.render_defaults([center[0], center[1] + 0.5, center[2] + 2.0], center, 60.0, 64, 48)
        setattr(q, field, bad)
        with pytest.raises(D.DfsphError) as e:
            a.render(q)
        assert e.value.code == D.E_INVALID, field
    q = D.render_defaults([center[0], center[1] + 0.5, center[2] + 2.0], center, 60.0, 64, 48)
    q.reserved[0] = 1
    with pytest.raises(D.DfsphError):
        a.render(q)
    q = D.render_defaults([center[0], center[1] + 0.5, center[2] + 2.0], center, 60.0, 64, 48)
    q.proj[:] = [0.0] * 16
    with pytest.raises(D.DfsphError):
        a.render(q)
    a.render(p)
    pol = D.default_policy(D.POLICY_CONVERGED)
    sa, sb = a.step(pol), b.step(pol)
    assert sa["iters_density"] == sb["iters_density"] and sa["dt"] == sb["dt"]
    for f in (D.FIELD_POSITION, D.FIELD_VELOCITY, D.FIELD_PRESSURE):
        assert np.array_equal(a.export(f), b.export(f))
    a.close()
    b.close()
