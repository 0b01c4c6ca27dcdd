"""The library's host code -- the capture reader and writer
(gpu-sims_amd/csrc/capture.cpp), the domain checks, derivation, boundary
sampler and preset fill (gpu-sims_amd/csrc/host_util.cpp) and the Alembic
writer (gpu-sims_amd/csrc/alembic.cpp) -- built for the
host with AddressSanitizer and UndefinedBehaviorSanitizer (float-to-integer
overflow included); tests/sanitize/capture_harness.cpp stands in for the
device-backed calls with a context held in host vectors.  Runs the reader over
the reference-layout captures, the rejections of tests/test_capture.py, edge
cases (zero particles, deep nesting, long strings, huge numbers, ill-formed
UTF-8) and a seeded fuzz of state.json, save -> read_info -> load round trips,
and the sampler, derivation and preset fill over ordinary and extreme
arguments; any sanitizer report fails the test.  CPU only."""
from __future__ import annotations

import json
import os
import random
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden", "captures")
SRC = [os.path.join(HERE, "sanitize", "capture_harness.cpp"),
       os.path.join(REPO, "gpu-sims_amd", "csrc", "capture.cpp"),
       os.path.join(REPO, "gpu-sims_amd", "csrc", "host_util.cpp"),
       os.path.join(REPO, "gpu-sims_amd", "csrc", "alembic.cpp")]
ENV = dict(os.environ,
           ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=86",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=87")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++ on this host")
    out = str(tmp_path_factory.mktemp("asan") / "capture_harness")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined,float-cast-overflow", "-fno-sanitize-recover=all",
           "-Wall", "-Wextra", "-Werror", "-o", out] + SRC
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return out


def run(harness, *args):
    r = subprocess.run([harness, *map(str, args)], capture_output=True, text=True, env=ENV, timeout=120)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-6000:]
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout.splitlines()


def _copy(tmp_path, name="capture_0007"):
    d = tmp_path / name
    shutil.copytree(os.path.join(GOLD, "capture_0007"), d)
    return d


def _state(d):
    return json.loads((d / "state.json").read_text(encoding="utf-8"))


def _write(d, st):
    (d / "state.json").write_text(json.dumps(st), encoding="utf-8")


def test_reference_layout_and_latest(harness):
    out = run(harness, "info", os.path.join(GOLD, "capture_0007"), os.path.join(GOLD, "capture_0003"))
    assert out == ["OK 345 7 3 1 2 4", "OK 120 3 3 1 2 4"]
    assert run(harness, "latest", GOLD)[0].endswith("capture_0007")
    need = len(os.path.join(GOLD, "capture_0007")) + 1
    assert run(harness, "latest", GOLD, need)[0].startswith("OK ")
    assert run(harness, "latest", GOLD, need - 1)[0].startswith("ERR -1 output path buffer too small")
    assert run(harness, "latest", "/nonexistent-dir")[0].startswith("ERR -7")
    assert run(harness, "load", os.path.join(GOLD, "capture_0007"))[0].startswith("OK 345 7 3 1 2 4 ")


def test_rejections(harness, tmp_path):
    cases = []
    d = _copy(tmp_path / "a")
    (d / "particles.bin").write_bytes((d / "particles.bin").read_bytes()[:-128])
    cases.append((d, "buffer-size mismatch"))
    d = _copy(tmp_path / "b")
    st = _state(d)
    del st["meta"]["sphWater"]
    _write(d, st)
    cases.append((d, "sphWater"))
    d = _copy(tmp_path / "c")
    st = _state(d)
    st["buffers"] = [b for b in st["buffers"] if b["name"] != "pressure"]
    _write(d, st)
    cases.append((d, "pressure"))
    d = _copy(tmp_path / "d")
    (d / "state.json").write_text('{"frame": 7, "meta": {"sphWater": {"particle_count": 3}}, "buffers": [')
    cases.append((d, "invalid JSON"))
    for fname in ("../particles.bin", "..", "/etc/passwd"):
        d = _copy(tmp_path / ("e" + str(len(cases))))
        st = _state(d)
        st["buffers"][0]["file"] = fname
        _write(d, st)
        cases.append((d, "leaves the capture directory"))
    d = _copy(tmp_path / "f")
    st = _state(d)
    st["meta"]["sphWater"]["particle_count"] = -3
    _write(d, st)
    cases.append((d, "particle_count"))
    out = run(harness, "info", *[str(c[0]) for c in cases])
    for line, (_, what) in zip(out, cases):
        assert line.startswith("ERR -7 ") and what in line, line
    for d, _ in cases:
        assert run(harness, "load", d)[0].startswith("ERR -7 ")


def test_edge_cases(harness, tmp_path):
    dirs, expect = [], []

    def case(name, text, ok):
        d = _copy(tmp_path, name)
        (d / "state.json").write_bytes(text if isinstance(text, bytes) else text.encode("utf-8"))
        dirs.append(str(d))
        expect.append(ok)

    base = (_copy(tmp_path, "base") / "state.json").read_text(encoding="utf-8")
    case("deep_array", "[" * 100000 + "]" * 100000, False)
    case("deep_object", '{"a":' * 5000 + "1" + "}" * 5000, False)
    case("long_string", base.replace('"camera": {', '"pad": "' + "x" * 200000 + '", "camera": {', 1), True)
    case("huge_count", base.replace('"particle_count": 3', '"particle_count": 184467440737095516160'), False)
    case("u32_overflow", base.replace('"particle_count": 3', '"particle_count": 4294967296'), False)
    case("float_count", base.replace('"particle_count": 3', '"particle_count": 3.0'), False)
    case("huge_ints", base.replace('"substeps": 4', '"substeps": 99999999999999999999999'), True)
    case("lone_surrogate", base.replace('"camera": {', '"s": "\\ud800", "camera": {', 1), False)
    case("cut_escape", base.replace('"camera": {', '"s": "\\u12', 1), False)
    case("nul_byte", base.encode("utf-8").replace(b'"camera": {', b'"s": "a\x00b", "camera": {', 1), False)
    case("lone_low_surrogate", base.replace('"camera": {', '"s": "\\udc00", "camera": {', 1), False)
    # UTF-8 inside strings must be well formed, as nlohmann::json requires
    case("utf8_ok", base.replace('"camera": {', '"s": "é€😀", "camera": {', 1), True)
    for name, seq in (("bad_lead", b"\xff\xfe"), ("overlong", b"\xc0\xaf"), ("surrogate", b"\xed\xa0\x80"),
                      ("past_max", b"\xf4\x90\x80\x80"), ("cut_seq", b"\xe2\x82"), ("stray_cont", b"\x80")):
        case("utf8_" + name, base.encode("utf-8").replace(b'"camera": {', b'"s": "' + seq + b'", "camera": {', 1),
             False)
    case("empty", "", False)
    case("only_ws", " \n\t ", False)
    out = run(harness, "info", *dirs)
    assert len(out) == len(dirs)
    for line, ok, d in zip(out, expect, dirs):
        assert line.startswith("OK ") == ok, (os.path.basename(d), line)
    # zero particles with empty buffers: accepted by the reader, refused by load
    d = _copy(tmp_path, "zero")
    st = _state(d)
    st["meta"]["sphWater"]["particle_count"] = 0
    _write(d, st)
    for b in st["buffers"]:
        (d / b["file"]).write_bytes(b"")
    assert run(harness, "info", d)[0] == "OK 345 7 0 1 2 4"
    assert run(harness, "load", d)[0].startswith("ERR -1 ")


def _mutate(rng, text):
    b = bytearray(text.encode("utf-8"))
    op = rng.randrange(6)
    if op == 0:                                   # truncate
        return bytes(b[:rng.randrange(len(b))])
    if op == 1:                                   # flip bytes
        for _ in range(rng.randint(1, 8)):
            b[rng.randrange(len(b))] = rng.randrange(256)
        return bytes(b)
    if op == 2:                                   # insert JSON punctuation
        for _ in range(rng.randint(1, 6)):
            b.insert(rng.randrange(len(b) + 1), ord(rng.choice('{}[]:,"\\-+.eE0u')))
        return bytes(b)
    if op == 3:                                   # delete a span
        a = rng.randrange(len(b))
        del b[a:a + rng.randint(1, 40)]
        return bytes(b)
    st = json.loads(text)
    if op == 4:                                   # retype a meta or buffer field
        vals = [None, True, -1, 0, 2**64, 1e308, "3", [], {}, "../x", "a" * 300]
        if rng.random() < 0.5:
            k = rng.choice(sorted(st["meta"]["sphWater"]))
            st["meta"]["sphWater"][k] = rng.choice(vals)
        else:
            buf = rng.choice(st["buffers"])
            buf[rng.choice(sorted(buf))] = rng.choice(vals)
        return json.dumps(st).encode("utf-8")
    st["buffers"] = rng.choice([None, 7, [], [1, "x", None], st["buffers"][::-1], st["buffers"] * 3])
    return json.dumps(st).encode("utf-8")


def test_fuzzed_state_json(harness, tmp_path):
    rng = random.Random(20261204)
    src = _copy(tmp_path, "src")
    text = (src / "state.json").read_text(encoding="utf-8")
    dirs = []
    for q in range(400):
        d = tmp_path / ("f%03d" % q)
        d.mkdir()
        for f in os.listdir(src):
            if f != "state.json":
                os.symlink(src / f, d / f)
        (d / "state.json").write_bytes(_mutate(rng, text))
        dirs.append(str(d))
    out = []
    for q in range(0, len(dirs), 100):
        out += run(harness, "info", *dirs[q:q + 100])
    assert len(out) == len(dirs)
    assert all(line.startswith(("OK ", "ERR ")) for line in out)
    assert 0 < sum(line.startswith("ERR ") for line in out) < len(out)
    for d in dirs[:60]:
        assert run(harness, "load", d)[0].startswith(("OK ", "ERR "))


@pytest.mark.parametrize("n,stepped", [(1, 0), (1000, 1), (70001, 1), (12, 0)])
def test_save_load_roundtrip(harness, tmp_path, n, stepped):
    out = run(harness, "roundtrip", tmp_path, n, stepped)
    assert out == ["OK %s/capture_0012" % tmp_path]
    st = json.loads((tmp_path / "capture_0012" / "state.json").read_text(encoding="utf-8"))
    assert st["meta"]["sphWater"]["camera"] == {"pos": [1, 2.5, -0.03], "name": "cam é\n"}
    sizes = {b["name"]: b["bytes"] for b in st["buffers"]}
    assert sizes["particles"] == 128 * n and sizes["density_alpha"] == 16 * n
    assert sizes["cell_starts"] == (65 * 4 if stepped else 0)


def test_domain_checks_and_derivation(harness):
    ok = run(harness, "domain", -2, -1, -1, 2, 2, 1, 0.005, 0)[0].split()
    assert ok[0] == "OK" and ok[1] == "128" and ok[2] == str(128 ** 3)
    for args, what in [((0, 0, 0, 41, 1, 1, 0.005, 0), "1024 reference cells"),   # 41/0.04 > 1024
                       ((0, 0, 0, 1, 1, 1, 0, 0), "particle_radius"),
                       ((0, 0, 0, 1, 1, 1, "nan", 0), "particle_radius"),
                       ((0, 0, 0, "inf", 1, 1, 0.005, 0), "finite"),
                       ((0, 0, 0, 1, "nan", 1, 0.005, 0), "finite"),
                       ((0, 1, 0, 1, 1, 1, 0.005, 0), "domain_max > domain_min"),
                       ((0, 0, 0, 40, 40, 40, 0.005, 0), "2^31 FINE grid cells")]:
        line = run(harness, "domain", *args)[0]
        assert line.startswith("ERR ") and what in line, (args, line)
    # the largest accepted Morton table: 1024 cells per axis, 2^30 in all
    ok = run(harness, "domain", 0, 0, 0, 40.96, 1, 1, 0.005, 4)[0].split()
    assert ok[:3] == ["OK", "1024", str(1024 ** 3)]


@pytest.mark.parametrize("box,r,cap", [
    ((-2, -1, -1, 2, 2, 1), 0.005, 512000),      # Dam-Break domain, the reference cap
    ((-2, -1, -1, 2, 2, 1), 0.005, 0),
    ((-1, -1, -1, 1, 2, 1), 0.01, 123457),       # cap in the middle of a face row
    ((0, 0, 0, 0.05, 0.03, 0.04), 0.005, 0),     # a few rows per face
    ((0, 0, 0, 0.01, 0.01, 0.01), 0.005, 0),     # faces and edges empty: corners only
    ((0, 0, 0, 0, 1, 1), 0.005, 0),              # flat box
    ((0, 0, 0, 1, 1, 1), 0.005, 7),              # cap inside the first row
])
def test_boundary_sampler(harness, box, r, cap):
    line = run(harness, "boundary", *box, r, cap)[0].split()
    assert line[0] == "OK"
    if cap:
        assert int(line[1]) <= cap


def test_boundary_sampler_rejects(harness):
    for args in [(0, 0, 0, 1, 1, 1, 0, 0), (0, 0, 0, "inf", 1, 1, 0.005, 0), (0, 0, 0, 1, 1, 1, 1e-9, 10),
                 (1, 0, 0, 0, 1, 1, 0.005, 0), (0, 0, 0, 1, 1, 1, "nan", 0)]:
        assert run(harness, "boundary", *args) == ["ERR extent"], args


def test_preset_fill(harness):
    for idx in range(4):
        for r, cap in [(0.005, 2000000), (0.02, 1000), (0.5, 10), (0.005, 0), (3.0, 100)]:
            line = run(harness, "preset", idx, r, cap)[0]
            assert line.startswith("OK "), (idx, r, cap, line)
    assert run(harness, "preset", 2, 0.005, 100000)[0] == "OK 92116 87616"
    for r in (1e-9, 0, -1, "nan", "inf"):
        assert run(harness, "preset", 0, r, 100)[0].startswith("ERR "), r


def test_library_rejects_bad_extents(dfsph_mod):
    """The same checks through the C ABI (host-only entry points)."""
    D = dfsph_mod
    with pytest.raises(D.DfsphError) as e:
        D.derive(D.make_params(0.005, (0, 0, 0), (41, 1, 1), 1000))
    assert e.value.code == D.E_INVALID and "1024 reference cells" in str(e.value)
    with pytest.raises(D.DfsphError) as e:
        D.derive(D.make_params(0.005, (0, 0, 0), (float("inf"), 1, 1), 1000))
    assert e.value.code == D.E_INVALID
    with pytest.raises(D.DfsphError) as e:
        D.boundary_samples((0, 0, 0), (1, 1, 1), 1e-9, 10)
    assert e.value.code == D.E_INVALID
    assert D.boundary_samples((0, 0, 0), (0.01, 0.01, 0.01), 0.005).shape == (8, 3)
    p = D.preset(0)
    with pytest.raises(D.DfsphError) as e:
        D.preset_fill(p, 1e-9, 100)
    assert e.value.code == D.E_INVALID and "2^20" in str(e.value)


@pytest.mark.parametrize("n,frames", [(1, 1), (3, 7), (5000, 9)])
def test_alembic_writer(harness, tmp_path, n, frames):
    from abc_read import Archive
    path = tmp_path / "h.abc"
    assert run(harness, "abc", path, n, frames) == [f"OK {frames}"]
    a = Archive(str(path))
    grp = a.children(a.top)[0][2]
    geom = a.properties(a.object_properties(grp)[".geom"][1])
    assert all(h["num_samples"] == frames for h, _ in geom.values())
    assert a.sample(geom["P"], frames - 1)[0].shape == (n, 3)
