"""Capture format on the GPU: save -> load -> step round trip, the written
files against the library's own exports and the nlohmann dump(2) layout, and
loading a capture laid out as the reference writes it.

Bars: state.json byte-identical to a dump(2) of its parsed content; every
buffer equal to the matching export (records: slot 0 = position and radius,
slot 1 = velocity, slot 5 = original index and flags); a context loaded from
the capture steps bit-identically to the one that wrote it (zero warm start).
"""
from __future__ import annotations

import importlib.util
import json
import os

import numpy as np
import pytest

from conftest import dam_break_small

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture_mod():
    spec = importlib.util.spec_from_file_location("mcf", os.path.join(GOLD, "make_capture_fixture.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _sim(D, x, cap=None):
    lo, hi = (0.0, 0.0, 0.0), (2.0, 1.0, 1.0)
    return D.Sim(D.make_params(0.01, lo, hi, capacity=cap or x.shape[0]))


def test_save_load_round_trip(dfsph_mod, tmp_path):
    D = dfsph_mod
    print("case: capture save/load round trip", flush=True)
    x, lo, hi = dam_break_small(12)
    a = D.Sim(D.make_params(0.01, lo, hi, capacity=x.shape[0]))
    flags = np.zeros(x.shape[0], np.uint32)
    flags[::7] = 2
    a.init_scene(x, flags=flags)
    pol = D.default_policy(D.POLICY_CONVERGED)
    for _ in range(4):
        a.step(pol)
    d = a.capture_save(str(tmp_path), 42, iteration=4, tier_index=1, preset_index=0, substeps=2,
                       camera_json='{"yaw": 0.5, "target": [0, 1, 2]}')
    assert d.endswith("capture_0042")
    n = x.shape[0]
    # state.json: the dump(2) layout, the reference's meta and buffer list
    text = open(os.path.join(d, "state.json"), encoding="utf-8").read()
    st = json.loads(text)
    assert text == _fixture_mod().dump2(st)
    meta = st["meta"]["sphWater"]
    assert meta == {"iteration": 4, "particle_count": n, "tier_index": 1, "preset_index": 0, "substeps": 2,
                    "camera": {"yaw": 0.5, "target": [0, 1, 2]}}
    assert st["frame"] == 42
    names = [b["name"] for b in st["buffers"]]
    assert names == ["particles", "density_alpha", "pressure", "morton_codes", "sorted_index", "cell_starts"]
    for b in st["buffers"]:
        assert b["format"] == "raw" and b["file"] == b["name"] + ".bin"
        assert os.path.getsize(os.path.join(d, b["file"])) == b["bytes"] == b["count"] * b["stride"]

    def raw(name, dt):
        return np.fromfile(os.path.join(d, name + ".bin"), dt)

    rec = raw("particles", np.uint8).reshape(n, 128)
    recf, recu = rec.view(np.float32), rec.view(np.uint32)
    assert np.array_equal(recf[:, 0:3], a.export(D.FIELD_POSITION))
    assert np.all(recf[:, 3] == np.float32(0.01))
    assert np.array_equal(recf[:, 4:7], a.export(D.FIELD_VELOCITY))
    assert np.array_equal(recu[:, 20], np.arange(n, dtype=np.uint32))
    assert np.array_equal(recu[:, 22], flags)
    da = raw("density_alpha", np.float32).reshape(n, 4)
    for c, f in enumerate((D.FIELD_DENSITY, D.FIELD_ALPHA_OVER_RHO2, D.FIELD_DENSITY_ADV, D.FIELD_DENSITY_CHANGE)):
        assert np.array_equal(da[:, c], a.export(f))
    assert np.array_equal(raw("pressure", np.float32), a.export(D.FIELD_PRESSURE))
    assert np.array_equal(raw("morton_codes", np.uint32), a.export(D.FIELD_MORTON_KEY))
    assert np.array_equal(raw("sorted_index", np.uint32), a.export(D.FIELD_SORTED_INDEX))
    assert np.array_equal(raw("cell_starts", np.uint32), a.export(D.FIELD_CELL_STARTS))
    # load into a fresh context and step both
    b = D.Sim(D.make_params(0.01, lo, hi, capacity=n))
    info = b.capture_load(d)
    assert info["particle_count"] == n and info["iteration"] == 4 and info["frame"] == 42
    assert np.array_equal(b.export(D.FIELD_POSITION), a.export(D.FIELD_POSITION))
    for _ in range(3):
        sa, sb = a.step(pol), b.step(pol)
        assert sa["iters_density"] == sb["iters_density"]
    assert np.array_equal(b.export(D.FIELD_POSITION), a.export(D.FIELD_POSITION))
    assert np.array_equal(b.export(D.FIELD_VELOCITY), a.export(D.FIELD_VELOCITY))
    # flags survive a second save
    d2 = b.capture_save(str(tmp_path / "again"), 0)
    rec2 = np.fromfile(os.path.join(d2, "particles.bin"), np.uint32).reshape(n, 32)
    assert np.array_equal(rec2[:, 22], flags)
    a.close()
    b.close()


def test_load_reference_layout(dfsph_mod, tmp_path):
    D = dfsph_mod
    print("case: load a reference-layout capture", flush=True)
    d = os.path.join(GOLD, "captures", "capture_0007")
    rec = np.fromfile(os.path.join(d, "particles.bin"), np.uint8).reshape(-1, 128)
    s = D.Sim(D.make_params(0.01, (0.0, 0.0, 0.0), (2.0, 1.0, 1.0), capacity=8))
    info = s.capture_load(d)
    assert info == {"iteration": 345, "frame": 7, "particle_count": 3, "tier_index": 1, "preset_index": 2,
                    "substeps": 4}
    assert np.array_equal(s.export(D.FIELD_POSITION), rec.view(np.float32)[:, 0:3])
    assert np.array_equal(s.export(D.FIELD_VELOCITY), rec.view(np.float32)[:, 4:7])
    out = s.capture_save(str(tmp_path), 7, iteration=info["iteration"])
    rec2 = np.fromfile(os.path.join(out, "particles.bin"), np.uint8).reshape(-1, 128)
    assert np.array_equal(rec2.view(np.uint32)[:, 22], rec.view(np.uint32)[:, 22])   # droplet flag kept
    s.close()


def test_capture_errors(dfsph_mod, tmp_path):
    D = dfsph_mod
    print("case: capture errors", flush=True)
    s = D.Sim(D.make_params(0.01, (0.0, 0.0, 0.0), (2.0, 1.0, 1.0), capacity=2))
    with pytest.raises(D.DfsphError) as e:           # three particles, capacity two
        s.capture_load(os.path.join(GOLD, "captures", "capture_0007"))
    assert e.value.code == D.E_CAPACITY
    s.init_scene(np.zeros((0, 3), np.float32))
    with pytest.raises(D.DfsphError) as e:           # nothing to capture (main.cpp:2441-2444)
        s.capture_save(str(tmp_path), 0)
    assert e.value.code == D.E_INVALID
    s.init_scene(np.array([[0.5, 0.5, 0.5]], np.float32))
    with pytest.raises(D.DfsphError):
        s.capture_save(str(tmp_path), 0, camera_json="[1, 2]")   # camera must be an object
    s.close()


def test_alembic_frames_from_the_simulation(dfsph_mod, tmp_path):
    """dfsph_abc_write_ctx (main.cpp:2875-2897): each frame holds the
    context's positions and velocities in original-index order and ids
    0..N-1, read back with the independent Ogawa reader."""
    from abc_read import Archive
    D = dfsph_mod
    print("case: Alembic export of 4 stepped frames", flush=True)
    x, lo, hi = dam_break_small(10)
    sim = D.Sim(D.make_params(0.01, lo, hi, capacity=x.shape[0]))
    sim.init_scene(x)
    pol = D.default_policy(D.POLICY_CONVERGED)
    path = str(tmp_path / "sph_water.abc")
    want = []
    with D.AbcWriter(path, fps=30.0) as w:
        for _ in range(4):
            sim.step(pol)
            w.write_sim(sim)
            want.append((sim.export(D.FIELD_POSITION), sim.export(D.FIELD_VELOCITY)))
    sim.close()
    a = Archive(path)
    grp = a.children(a.top)[0][2]
    geom = a.properties(a.object_properties(grp)[".geom"][1])
    n = x.shape[0]
    for q, (xq, vq) in enumerate(want):
        assert np.array_equal(a.sample(geom["P"], q)[0], xq.reshape(n, 3))
        assert np.array_equal(a.sample(geom[".velocities"], q)[0], vq.reshape(n, 3))
        assert np.array_equal(a.sample(geom[".pointIds"], q)[0], np.arange(n, dtype=np.uint64))
