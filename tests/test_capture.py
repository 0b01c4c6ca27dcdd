"""Capture format on the host (no GPU): the reader accepts captures laid out
as the reference's StateWriter writes them (tests/golden/captures, made by
tests/golden/make_capture_fixture.py) and rejects what the reference's loader
rejects (main.cpp:2493-2527): no sphWater meta, a missing buffer, a buffer
whose size disagrees with particle_count; plus malformed JSON and file names
that leave the capture directory.  StateReader::findLatest picks the last
capture_* directory in name order (state_reader.cpp:12-28)."""
from __future__ import annotations

import json
import os
import shutil

import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "captures")


def test_reads_reference_layout(dfsph_mod):
    info = dfsph_mod.capture_read_info(os.path.join(GOLD, "capture_0007"))
    assert info == {"iteration": 345, "frame": 7, "particle_count": 3, "tier_index": 1,
                    "preset_index": 2, "substeps": 4}
    info = dfsph_mod.capture_read_info(os.path.join(GOLD, "capture_0003"))
    assert info["iteration"] == 120 and info["frame"] == 3


def test_find_latest(dfsph_mod, tmp_path):
    assert dfsph_mod.capture_find_latest(GOLD).endswith("capture_0007")
    (tmp_path / "capture_0010.txt").write_text("not a directory")
    (tmp_path / "other").mkdir()
    with pytest.raises(dfsph_mod.DfsphError) as e:
        dfsph_mod.capture_find_latest(str(tmp_path))
    assert e.value.code == dfsph_mod.E_IO


def _copy(tmp_path):
    d = tmp_path / "capture_0007"
    shutil.copytree(os.path.join(GOLD, "capture_0007"), d)
    return d


def _state(d):
    return json.loads((d / "state.json").read_text(encoding="utf-8"))


def _write(d, st):
    (d / "state.json").write_text(json.dumps(st), encoding="utf-8")


def _rejects(D, d, what):
    with pytest.raises(D.DfsphError) as e:
        D.capture_read_info(str(d))
    assert e.value.code == D.E_IO, e.value
    assert what in str(e.value), str(e.value)


def test_rejects_size_mismatch(dfsph_mod, tmp_path):
    d = _copy(tmp_path)
    data = (d / "particles.bin").read_bytes()
    (d / "particles.bin").write_bytes(data[:-128])
    _rejects(dfsph_mod, d, "buffer-size mismatch")


def test_rejects_missing_meta_and_buffers(dfsph_mod, tmp_path):
    d = _copy(tmp_path)
    st = _state(d)
    del st["meta"]["sphWater"]
    _write(d, st)
    _rejects(dfsph_mod, d, "sphWater")
    st = _state(_copy(tmp_path / "b"))
    d = tmp_path / "b" / "capture_0007"
    st["buffers"] = [b for b in st["buffers"] if b["name"] != "pressure"]
    _write(d, st)
    _rejects(dfsph_mod, d, "pressure")
    # the cell-start table is optional, as in the reference's loader
    st = _state(d)
    st["buffers"] = [b for b in _state(_copy(tmp_path / "c"))["buffers"] if b["name"] != "cell_starts"]
    _write(d, st)
    assert dfsph_mod.capture_read_info(str(d))["particle_count"] == 3


def test_rejects_bad_json_and_paths(dfsph_mod, tmp_path):
    d = _copy(tmp_path)
    (d / "state.json").write_text('{"frame": 7, "meta": {"sphWater": {"particle_count": 3}}, "buffers": [')
    _rejects(dfsph_mod, d, "invalid JSON")
    d = _copy(tmp_path / "b")
    st = _state(d)
    st["buffers"][0]["file"] = "../particles.bin"
    _write(d, st)
    _rejects(dfsph_mod, d, "leaves the capture directory")
    d = _copy(tmp_path / "c")
    st = _state(d)
    st["meta"]["sphWater"]["particle_count"] = -3
    _write(d, st)
    _rejects(dfsph_mod, d, "particle_count")


def test_json_escapes_and_unicode(dfsph_mod, tmp_path):
    d = _copy(tmp_path)
    text = (d / "state.json").read_text(encoding="utf-8")
    cam = r'{"note": "tab\there \u00e9 \ud83d\ude00 \"q\" \/", "v": [1e-3, -2.5E+2, 0, true, null]}'
    text = text.replace('"camera": {', '"camera": ' + cam + ', "camera_old": {', 1)
    json.loads(text)                       # still valid JSON
    (d / "state.json").write_text(text, encoding="utf-8")
    assert dfsph_mod.capture_read_info(str(d))["substeps"] == 4
