"""Writes tests/golden/captures/: two small sphWater captures laid out as the
reference's StateWriter writes them (common/common-cpp/src/state_writer.cpp:
26-87 -- capture_%04u/, state.json from nlohmann::json::dump(2): keys sorted,
two-space indent, no trailing newline; one raw <name>.bin per buffer with
count/stride/format meta), with the sphWater meta and buffers of
particle-fluids/sph-water/src/main.cpp:2438-2487 (camera meta included, as
camera.toJson writes a JSON object).

The content is synthetic (3 particles, a 4-entry cell table) and fixed; the
reference cannot be built here (Vulkan), so the layout is restated from its
writer's source, not recorded from a run.  Run from the repository root:
    python tests/golden/make_capture_fixture.py
"""
from __future__ import annotations

import json
import os
import struct

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "captures")


def f32(x: float) -> float:
    """A float32 value as the double nlohmann stores (camera floats)."""
    return struct.unpack("<f", struct.pack("<f", x))[0]


def dump2(v, level=0) -> str:
    """nlohmann::json::dump(2) layout (keys sorted; Python floats print the
    shortest round-trip form, as nlohmann's grisu2 does)."""
    pad, close = "  " * (level + 1), "  " * level
    if isinstance(v, dict):
        if not v:
            return "{}"
        items = [f'{pad}{json.dumps(k, ensure_ascii=False)}: {dump2(v[k], level + 1)}' for k in sorted(v)]
        return "{\n" + ",\n".join(items) + "\n" + close + "}"
    if isinstance(v, list):
        if not v:
            return "[]"
        return "[\n" + ",\n".join(pad + dump2(x, level + 1) for x in v) + "\n" + close + "]"
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, float):
        return repr(v)
    return json.dumps(v, ensure_ascii=False)


def record(pos, radius, vel, pid, flags) -> bytes:
    """128-byte particle record (shaders/_struct_layouts.txt:11-42)."""
    rec = bytearray(128)
    struct.pack_into("<4f", rec, 0, *pos, radius)
    struct.pack_into("<4f", rec, 16, *vel, 0.0)
    struct.pack_into("<4I", rec, 80, pid, 0, flags, 0)
    return bytes(rec)


def write_capture(frame: int, iteration: int, n: int, shift: float) -> None:
    d = os.path.join(ROOT, f"capture_{frame:04d}")
    os.makedirs(d, exist_ok=True)
    r = 0.01
    parts = b"".join(record((0.1 + 0.02 * i + shift, 0.2, 0.3), r, (0.0, -0.5 * i, 0.25), i, 2 if i == n - 1 else 0)
                     for i in range(n))
    da = b"".join(struct.pack("<4f", 1000.0 + i, 1e-9 * (i + 1), 1.0 + 1e-4 * i, -0.5 * i) for i in range(n))
    pressure = b"".join(struct.pack("<f", 10.0 * i) for i in range(n))
    morton = b"".join(struct.pack("<I", 4 * i) for i in range(n))
    sorted_index = b"".join(struct.pack("<I", i) for i in range(n))
    cell_starts = b"".join(struct.pack("<I", v) for v in (0, 1, 2, n))
    bufs = []
    for name, data, count, stride in (("particles", parts, n, 128), ("density_alpha", da, n, 16),
                                      ("pressure", pressure, n, 4), ("morton_codes", morton, n, 4),
                                      ("sorted_index", sorted_index, n, 4), ("cell_starts", cell_starts, 4, 4)):
        with open(os.path.join(d, name + ".bin"), "wb") as fh:
            fh.write(data)
        bufs.append({"count": count, "stride": stride, "format": "raw", "name": name, "file": name + ".bin",
                     "bytes": len(data)})
    camera = {"distance": f32(3.5), "pitch": f32(-0.35), "target": [0.0, f32(0.1), 0.0], "yaw": f32(0.785),
              "name": "orbit é \"main\""}
    meta = {"iteration": iteration, "particle_count": n, "tier_index": 1, "preset_index": 2, "substeps": 4,
            "camera": camera}
    state = {"frame": frame, "meta": {"sphWater": meta}, "buffers": bufs}
    with open(os.path.join(d, "state.json"), "w", encoding="utf-8") as fh:
        fh.write(dump2(state))


if __name__ == "__main__":
    write_capture(3, 120, 3, 0.0)
    write_capture(7, 345, 3, 0.5)
    print("wrote", ROOT)
