"""A small reader of Alembic Ogawa archives, enough to check the particle
export: the Ogawa container (header, groups, data blocks), the AbcCoreOgawa
root layout (file and library version, top object, archive metadata, time
samplings, indexed metadata), object and property headers, and array/scalar
samples with their first/last changed indices.  Written from the published
format for the tests; the Alembic library itself is not available here."""
from __future__ import annotations

import struct

import numpy as np

DATA_BIT = 1 << 63
POD_DTYPE = {0: np.bool_, 1: np.uint8, 2: np.int8, 3: np.uint16, 4: np.int16, 5: np.uint32, 6: np.int32,
             7: np.uint64, 8: np.int64, 9: np.float16, 10: np.float32, 11: np.float64}


class Ogawa:
    def __init__(self, path):
        with open(path, "rb") as f:
            self.buf = f.read()
        b = self.buf
        if len(b) < 16 or b[:5] != b"Ogawa":
            raise ValueError("not an Ogawa file")
        self.frozen = b[5] == 0xFF
        self.version = (b[6] << 8) | b[7]
        self.root = struct.unpack_from("<Q", b, 8)[0]

    def group(self, pos):
        """Child offsets of the group at pos (0: the empty group)."""
        if pos == 0:
            return []
        n = struct.unpack_from("<Q", self.buf, pos)[0]
        if pos + 8 + 8 * n > len(self.buf):
            raise ValueError("group table past the end of the file")
        return list(struct.unpack_from(f"<{n}Q", self.buf, pos + 8))

    def data(self, child):
        if not child & DATA_BIT:
            raise ValueError("child is a group, not data")
        pos = child & ~DATA_BIT
        if pos == 0:
            return b""
        n = struct.unpack_from("<Q", self.buf, pos)[0]
        if pos + 8 + n > len(self.buf):
            raise ValueError("data block past the end of the file")
        return self.buf[pos + 8: pos + 8 + n]


def _hint(buf, pos, hint):
    size = (1, 2, 4)[hint]
    return int.from_bytes(buf[pos:pos + size], "little"), pos + size


def _metadata(s):
    return dict(kv.split("=", 1) for kv in s.split(";") if kv) if s else {}


def property_headers(buf, indexed):
    out, pos = [], 0
    while pos < len(buf):
        info = struct.unpack_from("<I", buf, pos)[0]
        pos += 4
        h = {"type": ("compound", "scalar", "array", "array")[info & 3], "scalar_like": (info & 3) == 3}
        hint = (info >> 2) & 3
        if h["type"] != "compound":
            h["pod"] = (info >> 4) & 0xF
            h["extent"] = (info >> 12) & 0xFF
            h["homogenous"] = bool(info & 0x400)
            h["num_samples"], pos = _hint(buf, pos, hint)
            if info & 0x200:
                h["first_changed"], pos = _hint(buf, pos, hint)
                h["last_changed"], pos = _hint(buf, pos, hint)
            elif info & 0x800:
                h["first_changed"] = h["last_changed"] = 0
            else:
                h["first_changed"], h["last_changed"] = 1, h["num_samples"] - 1
            h["time_sampling"] = 0
            if info & 0x100:
                h["time_sampling"], pos = _hint(buf, pos, hint)
        n, pos = _hint(buf, pos, hint)
        h["name"] = buf[pos:pos + n].decode()
        pos += n
        md = (info >> 20) & 0xFF
        if md == 0xFF:
            m, pos = _hint(buf, pos, hint)
            h["metadata"] = _metadata(buf[pos:pos + m].decode())
            pos += m
        else:
            h["metadata"] = _metadata(indexed[md])
        out.append(h)
    return out


def object_headers(buf, indexed):
    body, out, pos = buf[:-32], [], 0
    while pos < len(body):
        n = struct.unpack_from("<I", body, pos)[0]
        pos += 4
        name = body[pos:pos + n].decode()
        pos += n
        md = body[pos]
        pos += 1
        if md == 0xFF:
            m = struct.unpack_from("<I", body, pos)[0]
            pos += 4
            meta = body[pos:pos + m].decode()
            pos += m
        else:
            meta = indexed[md]
        out.append({"name": name, "metadata": _metadata(meta)})
    return out


class Archive:
    def __init__(self, path):
        self.o = Ogawa(path)
        root = self.o.group(self.o.root)
        if len(root) < 6 or not all(root[i] & DATA_BIT for i in (0, 1, 3, 4, 5)) or root[2] & DATA_BIT:
            raise ValueError("root group is not [data, data, group, data, data, data]")
        self.file_version = struct.unpack("<i", self.o.data(root[0]))[0]
        self.library_version = struct.unpack("<i", self.o.data(root[1]))[0]
        self.metadata = _metadata(self.o.data(root[3]).decode())
        ts, self.time_samplings, pos = self.o.data(root[4]), [], 0
        while pos < len(ts):
            max_samples, tpc, n = struct.unpack_from("<IdI", ts, pos)
            pos += 16
            times = list(struct.unpack_from(f"<{n}d", ts, pos))
            pos += 8 * n
            self.time_samplings.append({"max_samples": max_samples, "time_per_cycle": tpc, "times": times})
        md, self.indexed, pos = self.o.data(root[5]), [""], 0
        while pos < len(md):
            n = md[pos]
            self.indexed.append(md[pos + 1:pos + 1 + n].decode())
            pos += 1 + n
        self.top = root[2]

    def children(self, obj_group):
        """(name, metadata, group offset) of an object's child objects."""
        ch = self.o.group(obj_group)
        if not ch:
            return []
        hdr = object_headers(self.o.data(ch[-1]), self.indexed) if ch[-1] & DATA_BIT else []
        return [(h["name"], h["metadata"], ch[1 + i]) for i, h in enumerate(hdr)]

    def properties(self, compound_group):
        """{name: (header, group offset)} of a compound property."""
        ch = self.o.group(compound_group)
        if not ch or not ch[-1] & DATA_BIT:
            return {}
        hdrs = property_headers(self.o.data(ch[-1]), self.indexed)
        return {h["name"]: (h, ch[i]) for i, h in enumerate(hdrs)}

    def object_properties(self, obj_group):
        ch = self.o.group(obj_group)
        return self.properties(ch[0]) if ch else {}

    def sample(self, prop, index):
        """(values, key) of sample `index` of a scalar or array property."""
        h, g = prop
        if not 0 <= index < h["num_samples"]:
            raise IndexError(index)
        first, last = h["first_changed"], h["last_changed"]
        if index < first or (first == 0 and last == 0):
            index = 0
        elif index >= last:
            index = last
        ch = self.o.group(g)
        per = 2 if h["type"] == "array" else 1
        raw = self.o.data(ch[per * index])
        key, body = raw[:16], raw[16:]
        vals = np.frombuffer(body, POD_DTYPE[h["pod"]])
        if h["type"] == "array":
            dims = self.o.data(ch[per * index + 1])
            if dims:
                n = int(np.prod(np.frombuffer(dims, np.uint64)))
                assert n * h["extent"] == vals.size
        return vals.reshape(-1, h["extent"]) if h["extent"] > 1 else vals, key


def murmur3_x64_128(data: bytes, seed: int = 0) -> bytes:
    """MurmurHash3_x64_128 (Austin Appleby, public domain), pure Python."""
    M = (1 << 64) - 1
    c1, c2 = 0x87C37B91114253D5, 0x4CF5AD432745937F

    def rotl(x, r):
        return ((x << r) | (x >> (64 - r))) & M

    def fmix(k):
        k ^= k >> 33
        k = (k * 0xFF51AFD7ED558CCD) & M
        k ^= k >> 33
        k = (k * 0xC4CEB9FE1A85EC53) & M
        return k ^ (k >> 33)

    h1 = h2 = seed
    n = len(data) // 16
    for i in range(n):
        k1, k2 = struct.unpack_from("<QQ", data, 16 * i)
        k1 = (rotl((k1 * c1) & M, 31) * c2) & M
        h1 ^= k1
        h1 = (rotl(h1, 27) + h2) & M
        h1 = (h1 * 5 + 0x52DCE729) & M
        k2 = (rotl((k2 * c2) & M, 33) * c1) & M
        h2 ^= k2
        h2 = (rotl(h2, 31) + h1) & M
        h2 = (h2 * 5 + 0x38495AB5) & M
    tail = data[16 * n:]
    k1 = int.from_bytes(tail[:8], "little")
    k2 = int.from_bytes(tail[8:], "little")
    if len(tail) > 8:
        h2 ^= (rotl((k2 * c2) & M, 33) * c1) & M
    if tail:
        h1 ^= (rotl((k1 * c1) & M, 31) * c2) & M
    h1 ^= len(data)
    h2 ^= len(data)
    h1 = (h1 + h2) & M
    h2 = (h2 + h1) & M
    h1, h2 = fmix(h1), fmix(h2)
    h1 = (h1 + h2) & M
    h2 = (h2 + h1) & M
    return struct.pack("<QQ", h1, h2)
