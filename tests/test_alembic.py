"""Alembic particle export (dfsph_abc_*, replacing the reference's
abc::ParticleWriter, common/common-cpp/src/alembic_writer.cpp:38-116, fed at
main.cpp:2875-2897).  Archives written by the library are read back with an
independent reader of the published Ogawa/AbcCoreOgawa layout
(tests/abc_read.py): the root layout and versions, the uniform 1/fps time
sampling, the OPoints object and its schema metadata, every frame's P,
.pointIds, .velocities and .selfBnds, the MurmurHash3 sample keys, and the
first/last changed indices of repeated samples.  The Alembic library is not
available to check the bytes against (parity of the container unpinned
there).  CPU only: the writer is host code."""
from __future__ import annotations

import numpy as np
import pytest

from abc_read import Archive, murmur3_x64_128


def _frames(rng, counts):
    out = []
    for n in counts:
        x = rng.uniform(-1, 2, (n, 3)).astype(np.float32)
        v = rng.normal(0, 1, (n, 3)).astype(np.float32)
        out.append((x, v))
    return out


def _open_points(path):
    a = Archive(path)
    kids = a.children(a.top)
    assert [k[0] for k in kids] == ["particles"]
    name, meta, grp = kids[0]
    assert meta == {"schema": "AbcGeom_Points_v1", "schemaBaseType": "AbcGeom_GeomBase_v1",
                    "schemaObjTitle": "AbcGeom_Points_v1:.geom"}
    top = a.object_properties(grp)
    assert list(top) == [".geom"] and top[".geom"][0]["type"] == "compound"
    assert top[".geom"][0]["metadata"]["schema"] == "AbcGeom_Points_v1"
    return a, a.properties(top[".geom"][1]), grp


def test_murmur3_published_vector():
    assert murmur3_x64_128(b"The quick brown fox jumps over the lazy dog").hex() == \
        "6c1b07bc7bbc4be347939ac4a93c437a"


def test_frames_round_trip(dfsph_mod, tmp_path):
    D = dfsph_mod
    rng = np.random.default_rng(7)
    frames = _frames(rng, [5, 1000, 1000, 17, 70001])
    path = str(tmp_path / "sph_water.abc")
    with D.AbcWriter(path, fps=30.0) as w:
        for x, v in frames:
            w.write(x, v)
    a, geom, _ = _open_points(path)
    assert a.o.frozen and a.o.version == 1
    assert a.file_version == 0 and a.library_version == 10810
    assert a.time_samplings[0] == {"max_samples": 0, "time_per_cycle": 1.0, "times": [0.0]}
    assert a.time_samplings[1] == {"max_samples": 5, "time_per_cycle": 1.0 / 30.0, "times": [0.0]}
    assert list(geom) == [".selfBnds", "P", ".pointIds", ".velocities"]
    hp, hi, hv, hb = geom["P"][0], geom[".pointIds"][0], geom[".velocities"][0], geom[".selfBnds"][0]
    assert (hp["type"], hp["pod"], hp["extent"], hp["time_sampling"]) == ("array", 10, 3, 1)
    assert (hi["type"], hi["pod"], hi["extent"]) == ("array", 7, 1)
    assert (hv["type"], hv["pod"], hv["extent"]) == ("array", 10, 3)
    assert (hb["type"], hb["pod"], hb["extent"]) == ("scalar", 11, 6)
    assert hp["metadata"] == {"geoScope": "var", "interpretation": "point"}
    assert hv["metadata"] == {"interpretation": "vector"}
    assert hb["metadata"] == {"interpretation": "box"}
    for h in (hp, hi, hv, hb):
        assert h["num_samples"] == 5
        assert not h["homogenous"] or h is hb
    for q, (x, v) in enumerate(frames):
        px, key = a.sample(geom["P"], q)
        assert np.array_equal(px, x)
        assert key == murmur3_x64_128(x.tobytes())
        assert np.array_equal(a.sample(geom[".velocities"], q)[0], v)
        assert np.array_equal(a.sample(geom[".pointIds"], q)[0], np.arange(x.shape[0], dtype=np.uint64))
        box = a.sample(geom[".selfBnds"], q)[0].reshape(-1)
        assert np.array_equal(box, np.concatenate([x.min(0), x.max(0)]).astype(np.float64))


def test_repeated_samples_are_stored_once(dfsph_mod, tmp_path):
    """Alembic's sample bookkeeping: a property whose samples repeat stores each
    run once; the header's first/last changed indices map every frame back."""
    D = dfsph_mod
    rng = np.random.default_rng(3)
    x0, v0 = _frames(rng, [64])[0]
    x1 = x0 + 0.5
    ids = np.arange(100, 164, dtype=np.uint64)
    seq = [x0, x0, x0, x1, x1, x0, x0]   # P: first change at 3, last at 5
    path = str(tmp_path / "rep.abc")
    with D.AbcWriter(path, fps=24.0) as w:
        for x in seq:
            w.write(x, v0, ids)
    a, geom, _ = _open_points(path)
    hp, hi, hv = geom["P"][0], geom[".pointIds"][0], geom[".velocities"][0]
    assert (hp["num_samples"], hp["first_changed"], hp["last_changed"]) == (7, 3, 5)
    assert (hi["first_changed"], hi["last_changed"]) == (0, 0)      # constant
    assert (hv["first_changed"], hv["last_changed"]) == (0, 0)
    assert hp["homogenous"] and hi["homogenous"]
    assert len(a.o.group(geom[".pointIds"][1])) == 2                   # one stored sample + dims
    blocks = {c for c in a.o.group(geom["P"][1])[::2]}
    assert len(blocks) == 2                                            # x0 and x1 stored once each
    for q, x in enumerate(seq):
        assert np.array_equal(a.sample(geom["P"], q)[0], x)
        assert np.array_equal(a.sample(geom[".pointIds"], q)[0], ids)
    assert a.time_samplings[1]["time_per_cycle"] == pytest.approx(1 / 24.0, rel=0, abs=0)


def test_velocities_from_a_later_frame(dfsph_mod, tmp_path):
    """Frames before the first one with velocities get empty samples, as
    OPointsSchema back-fills a property created late."""
    D = dfsph_mod
    rng = np.random.default_rng(5)
    (xa, _), (xb, vb) = _frames(rng, [10, 12])
    path = str(tmp_path / "late.abc")
    with D.AbcWriter(path) as w:
        w.write(xa)
        w.write(xb, vb)
        w.write(xa)
    a, geom, _ = _open_points(path)
    hv = geom[".velocities"][0]
    assert hv["num_samples"] == 3
    assert a.sample(geom[".velocities"], 0)[0].size == 0
    assert np.array_equal(a.sample(geom[".velocities"], 1)[0], vb)
    assert a.sample(geom[".velocities"], 2)[0].size == 0
    assert np.array_equal(a.sample(geom["P"], 2)[0], xa)


def test_empty_archive_and_errors(dfsph_mod, tmp_path):
    D = dfsph_mod
    path = str(tmp_path / "empty.abc")
    D.AbcWriter(path).close()
    a, geom, _ = _open_points(path)
    assert a.time_samplings[1]["max_samples"] == 0
    assert all(h["num_samples"] == 0 for h, _ in geom.values())
    with pytest.raises(D.DfsphError) as e:
        D.AbcWriter(str(tmp_path / "no" / "such" / "dir.abc"))
    assert e.value.code == D.E_IO
    with pytest.raises(D.DfsphError) as e:
        D.AbcWriter(path, fps=0.0)
    assert e.value.code == D.E_INVALID
    with D.AbcWriter(str(tmp_path / "z.abc")) as w:
        with pytest.raises(D.DfsphError) as e:
            w.write(np.zeros((0, 3), np.float32))
        assert e.value.code == D.E_INVALID
