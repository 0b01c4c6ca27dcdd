// alembic.cpp -- particle export as an Alembic archive, replacing the
// reference's abc::ParticleWriter (common/common-cpp/src/alembic_writer.cpp:38-86,
// created at main.cpp:1567-1580 with fps 30 and fed at main.cpp:2875-2897).
// The reference builds it on the Alembic library (1.8.10, Ogawa back end),
// which is absent here, so this file writes the same archive layout directly:
// one OPoints object "particles" under the top object, its schema compound
// ".geom" holding ".selfBnds" (Box3d), "P" (V3f positions), ".pointIds"
// (uint64) and, from the first frame that has them, ".velocities" (V3f), all
// on a uniform time sampling of 1/fps starting at 0.
//
// Ogawa container (the published Alembic format): a 16-byte header -- "Ogawa",
// a frozen byte (0xff once the archive is complete), version 0x0001 big
// endian, the root group's offset (u64) -- then groups (u64 child count, u64
// child offsets; bit 63 marks a data child, offset 0 an empty group,
// 2^63 an empty data) and data blocks (u64 size, bytes).  The root group holds
// [file version i32 = 0, library version i32, top object group, archive
// metadata, time samplings, indexed metadata].  An object group is [property
// compound, child objects..., child headers + 32 hash bytes]; a compound
// group is [property groups..., property headers]; an array property holds
// two children per stored sample (16-byte key + values, dimensions: empty for
// rank 1), a scalar property one.  As in Alembic, a sample equal to an
// earlier one (same key and size) refers to the stored block, and trailing or
// leading repeats are described by first/last changed indices.  Frames stream to
// the file as they are written; the group tables and headers go out at close.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "dfsph_internal.h"

using namespace dfsph;

namespace {

constexpr uint64_t kDataBit = 0x8000000000000000ull;   // child offset of a data block
constexpr uint64_t kEmptyData = kDataBit;
constexpr uint64_t kEmptyGroup = 0;
constexpr int32_t kLibraryVersion = 10810;               // Alembic 1.8.10, the reference's pin
// Alembic PlainOldDataType values of the PODs used here.
constexpr uint32_t kPodUint64 = 7, kPodFloat32 = 10, kPodFloat64 = 11;

// MurmurHash3_x64_128 with seed 0 (Austin Appleby's public-domain hash, which
// Alembic uses for sample keys).
inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
inline uint64_t fmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}
void murmur3_128(const void* data, size_t len, uint8_t out[16]) {
    const uint8_t* p = static_cast<const uint8_t*>(data);
    const size_t nblocks = len / 16;
    uint64_t h1 = 0, h2 = 0;
    const uint64_t c1 = 0x87c37b91114253d5ull, c2 = 0x4cf5ad432745937full;
    for (size_t i = 0; i < nblocks; ++i) {
        uint64_t k1, k2;
        std::memcpy(&k1, p + 16 * i, 8);
        std::memcpy(&k2, p + 16 * i + 8, 8);
        k1 *= c1; k1 = rotl64(k1, 31); k1 *= c2; h1 ^= k1;
        h1 = rotl64(h1, 27); h1 += h2; h1 = h1 * 5 + 0x52dce729;
        k2 *= c2; k2 = rotl64(k2, 33); k2 *= c1; h2 ^= k2;
        h2 = rotl64(h2, 31); h2 += h1; h2 = h2 * 5 + 0x38495ab5;
    }
    const uint8_t* tail = p + nblocks * 16;
    uint64_t k1 = 0, k2 = 0;
    const size_t rem = len & 15;
    for (size_t q = rem; q > 8; --q) k2 ^= uint64_t(tail[q - 1]) << (8 * (q - 9));
    if (rem > 8) { k2 *= c2; k2 = rotl64(k2, 33); k2 *= c1; h2 ^= k2; }
    for (size_t q = rem < 8 ? rem : 8; q > 0; --q) k1 ^= uint64_t(tail[q - 1]) << (8 * (q - 1));
    if (rem) { k1 *= c1; k1 = rotl64(k1, 31); k1 *= c2; h1 ^= k1; }
    h1 ^= uint64_t(len); h2 ^= uint64_t(len);
    h1 += h2; h2 += h1;
    h1 = fmix64(h1); h2 = fmix64(h2);
    h1 += h2; h2 += h1;
    std::memcpy(out, &h1, 8);
    std::memcpy(out + 8, &h2, 8);
}

struct Key {
    uint8_t d[16];
    bool operator==(const Key& o) const { return std::memcmp(d, o.d, 16) == 0; }
};

// One scalar or array property: the offsets of its stored samples and the
// bookkeeping of Alembic's first/last changed indices.
struct Prop {
    std::string name, meta;
    bool array = true;
    uint32_t pod = 0, extent = 1;
    std::vector<uint64_t> children;   // per stored sample: data (and dims for arrays)
    uint32_t next = 0, first_changed = 0, last_changed = 0;
    Key prev{};
    uint64_t prev_data = kEmptyData;
    uint64_t first_count = 0;
    bool homogenous = true, scalar_like = true;
};

void put_u32(std::vector<uint8_t>& o, uint32_t v) { o.insert(o.end(), reinterpret_cast<uint8_t*>(&v), reinterpret_cast<uint8_t*>(&v) + 4); }
void put_hint(std::vector<uint8_t>& o, uint32_t v, uint32_t hint) {
    if (hint == 0) o.push_back(uint8_t(v));
    else if (hint == 1) { const uint16_t s = uint16_t(v); o.insert(o.end(), reinterpret_cast<const uint8_t*>(&s), reinterpret_cast<const uint8_t*>(&s) + 2); }
    else put_u32(o, v);
}
void put_str(std::vector<uint8_t>& o, const std::string& s) { o.insert(o.end(), s.begin(), s.end()); }

// Indexed metadata (AbcCoreOgawa): strings under 256 bytes get an index
// 1..254 on first use; 0 is the empty string, 0xff means "stored inline".
struct MetaIndex {
    std::vector<std::string> list;   // index i + 1
    uint32_t get(const std::string& s) {
        if (s.empty()) return 0;
        for (size_t i = 0; i < list.size(); ++i)
            if (list[i] == s) return uint32_t(i + 1);
        if (s.size() < 256 && list.size() < 254) {
            list.push_back(s);
            return uint32_t(list.size());
        }
        return 0xff;
    }
};

}  // namespace

struct dfsph_abc {
    std::FILE* f = nullptr;
    std::string path;
    uint64_t end = 0;             // file size so far
    double fps = 30.0;
    uint32_t frames = 0;
    uint64_t root_children[2] = {kEmptyData, kEmptyData};   // the two version blocks
    Prop bounds, pos, ids, vel;
    bool has_vel = false;
    bool failed = false;
    // Every stored sample by (key, size): a sample equal to any earlier one
    // refers to its block (Alembic's written-sample map).
    std::map<std::pair<std::string, uint64_t>, uint64_t> written;

    bool write(const void* p, size_t n) {
        if (failed) return false;
        if (n && std::fwrite(p, 1, n, f) != n) { failed = true; return false; }
        end += n;
        return true;
    }
    // A data block; returns its child offset (bit 63 set), or kEmptyData.
    uint64_t data(const std::vector<std::pair<const void*, size_t>>& parts) {
        uint64_t size = 0;
        for (const auto& q : parts) size += q.second;
        if (size == 0) return kEmptyData;
        const uint64_t at = end;
        write(&size, 8);
        for (const auto& q : parts) write(q.first, q.second);
        return at | kDataBit;
    }
    uint64_t data(const std::vector<uint8_t>& v) { return data({{v.data(), v.size()}}); }
    uint64_t group(const std::vector<uint64_t>& ch) {
        if (ch.empty()) return kEmptyGroup;
        const uint64_t at = end;
        const uint64_t n = ch.size();
        write(&n, 8);
        write(ch.data(), ch.size() * 8);
        return at;
    }

    // Alembic's sample bookkeeping (ApwImpl/SpwImpl::setSample): a sample equal
    // to the previous one is not written; when a different one follows, the
    // repeats are recorded as further references to the stored block.
    void set(Prop& p, const void* bytes, size_t nbytes, uint64_t count) {
        Key k;
        murmur3_128(bytes, nbytes, k.d);
        if (p.next == 0) p.first_count = count;
        else if (count != p.first_count) p.homogenous = false;
        if (count != 1) p.scalar_like = false;
        if (p.next == 0 || !(k == p.prev)) {
            for (uint32_t i = p.last_changed + 1; i < p.next; ++i) {
                p.children.push_back(p.prev_data);
                if (p.array) p.children.push_back(kEmptyData);
            }
            const auto id = std::make_pair(std::string(reinterpret_cast<const char*>(k.d), 16), uint64_t(nbytes));
            const auto hit = written.find(id);
            if (hit != written.end()) {
                p.prev_data = hit->second;
            } else {
                p.prev_data = data({{k.d, 16}, {bytes, nbytes}});
                written.emplace(id, p.prev_data);
            }
            p.children.push_back(p.prev_data);
            if (p.array) p.children.push_back(kEmptyData);   // rank-1 dimensions follow from the size
            if (p.next > 0) {
                if (p.first_changed == 0) p.first_changed = p.next;
                p.last_changed = p.next;
            }
            p.prev = k;
        }
        ++p.next;
    }

    // Property header (AbcCoreOgawa WritePropertyInfo layout): an info word
    // (bits 0-1 type: 0 compound, 1 scalar, 2 array, 3 scalar-like array;
    // 2-3 size hint; 4-7 POD; 8 has time sampling index; 9 first/last changed
    // follow; 10 homogenous; 11 constant; 12-19 extent; 20-27 metadata index),
    // then sample count, [first, last], [time sampling], name, [metadata].
    static void header(std::vector<uint8_t>& o, const Prop* p, const std::string& name, const std::string& meta,
                       uint32_t tsidx, MetaIndex& mi) {
        const uint32_t md = mi.get(meta);
        uint32_t maxv = uint32_t(name.size());
        if (md == 0xff) maxv = std::max<uint32_t>(maxv, uint32_t(meta.size()));
        uint32_t info = 0;
        bool first_last = false;
        if (p) {
            info |= p->array ? (p->scalar_like && p->next ? 3u : 2u) : 1u;
            info |= (p->pod & 0xf) << 4;
            if (tsidx) info |= 0x100;
            if (p->first_changed == 0 && p->last_changed == 0) info |= 0x800;
            else if (!(p->first_changed == 1 && p->last_changed + 1 == p->next)) { info |= 0x200; first_last = true; }
            if (p->array && p->homogenous) info |= 0x400;
            info |= (p->extent & 0xff) << 12;
            maxv = std::max({maxv, p->next, tsidx, p->first_changed, p->last_changed});
        }
        info |= (md & 0xff) << 20;
        const uint32_t hint = maxv < 256 ? 0 : (maxv < 65536 ? 1 : 2);
        info |= hint << 2;
        put_u32(o, info);
        if (p) {
            put_hint(o, p->next, hint);
            if (first_last) { put_hint(o, p->first_changed, hint); put_hint(o, p->last_changed, hint); }
            if (tsidx) put_hint(o, tsidx, hint);
        }
        put_hint(o, uint32_t(name.size()), hint);
        put_str(o, name);
        if (md == 0xff) { put_hint(o, uint32_t(meta.size()), hint); put_str(o, meta); }
    }

    // Object headers of a group's children followed by 32 hash bytes (Alembic
    // stores SpookyHash digests of the properties and children there; readers
    // only compare them, so a MurmurHash3 of the stored sample keys stands in).
    static std::vector<uint8_t> object_headers(const std::vector<std::pair<std::string, std::string>>& kids,
                                               MetaIndex& mi, const std::vector<uint8_t>& hashed) {
        std::vector<uint8_t> o;
        for (const auto& k : kids) {
            put_u32(o, uint32_t(k.first.size()));
            put_str(o, k.first);
            const uint32_t md = mi.get(k.second);
            o.push_back(uint8_t(md));
            if (md == 0xff) { put_u32(o, uint32_t(k.second.size())); put_str(o, k.second); }
        }
        uint8_t h[16];
        murmur3_128(hashed.data(), hashed.size(), h);
        o.insert(o.end(), h, h + 16);
        o.insert(o.end(), 16, uint8_t(0));
        return o;
    }

    int finish() {
        MetaIndex mi;
        // .geom: its properties in creation order (OGeomBaseSchema makes
        // .selfBnds, OPointsSchema P and .pointIds, the first velocities .velocities).
        std::vector<const Prop*> props = {&bounds, &pos, &ids};
        if (has_vel) props.push_back(&vel);
        std::vector<uint64_t> geom_children;
        std::vector<uint8_t> geom_headers, keys;
        for (const Prop* p : props) {
            geom_children.push_back(group(p->children));
            header(geom_headers, p, p->name, p->meta, 1, mi);
            keys.insert(keys.end(), p->prev.d, p->prev.d + 16);
        }
        geom_children.push_back(data(geom_headers));
        const uint64_t geom = group(geom_children);
        const std::string schema = "schema=AbcGeom_Points_v1;schemaBaseType=AbcGeom_GeomBase_v1";
        std::vector<uint8_t> top_props_headers;
        header(top_props_headers, nullptr, ".geom", schema, 0, mi);
        const uint64_t particle_props = group({geom, data(top_props_headers)});
        const uint64_t particles = group({particle_props, data(object_headers({}, mi, keys))});
        const std::string obj_meta = schema + ";schemaObjTitle=AbcGeom_Points_v1:.geom";
        const uint64_t top = group({kEmptyGroup, particles, data(object_headers({{"particles", obj_meta}}, mi, keys))});
        const std::string archive_meta = "_ai_Application=dfsph particle export;_ai_Description=DFSPH water";
        std::vector<uint8_t> ts;
        auto sampling = [&](uint32_t max_samples, double per_cycle) {
            put_u32(ts, max_samples);
            ts.insert(ts.end(), reinterpret_cast<uint8_t*>(&per_cycle), reinterpret_cast<uint8_t*>(&per_cycle) + 8);
            put_u32(ts, 1u);
            const double start = 0.0;
            ts.insert(ts.end(), reinterpret_cast<const uint8_t*>(&start), reinterpret_cast<const uint8_t*>(&start) + 8);
        };
        sampling(0u, 1.0);               // index 0: the identity sampling every archive has
        sampling(frames, 1.0 / fps);     // index 1: TimeSampling(1 / fps, 0)
        std::vector<uint8_t> md;
        for (const std::string& s : mi.list) {
            md.push_back(uint8_t(s.size()));
            put_str(md, s);
        }
        std::vector<uint8_t> am(archive_meta.begin(), archive_meta.end());
        const uint64_t root = group({root_children[0], root_children[1], top, data(am), data(ts), data(md)});
        if (failed) return set_error(DFSPH_E_IO, "write failed: " + path);
        const uint8_t frozen = 0xff;
        if (std::fseek(f, 8, SEEK_SET) != 0 || std::fwrite(&root, 8, 1, f) != 1 || std::fseek(f, 5, SEEK_SET) != 0 ||
            std::fwrite(&frozen, 1, 1, f) != 1 || std::fflush(f) != 0)
            return set_error(DFSPH_E_IO, "cannot finish " + path);
        return DFSPH_OK;
    }
};

extern "C" {

int dfsph_abc_open(const char* path, double fps, dfsph_abc** out) {
    if (!path || !out) return set_error(DFSPH_E_INVALID, "NULL argument");
    *out = nullptr;
    if (!(fps > 0.0) || !std::isfinite(fps)) return set_error(DFSPH_E_INVALID, "fps must be finite and > 0");
    std::unique_ptr<dfsph_abc> w(new dfsph_abc);
    w->path = path;
    w->fps = fps;
    w->f = std::fopen(path, "wb");
    if (!w->f) return set_error(DFSPH_E_IO, std::string("cannot create ") + path);
    const uint8_t head[16] = {'O', 'g', 'a', 'w', 'a', 0x00, 0x00, 0x01, 0, 0, 0, 0, 0, 0, 0, 0};
    w->write(head, 16);
    const int32_t file_version = 0, lib = kLibraryVersion;
    w->root_children[0] = w->data({{&file_version, 4}});
    w->root_children[1] = w->data({{&lib, 4}});
    w->bounds.name = ".selfBnds"; w->bounds.meta = "interpretation=box";
    w->bounds.array = false; w->bounds.pod = kPodFloat64; w->bounds.extent = 6;
    w->pos.name = "P"; w->pos.meta = "geoScope=var;interpretation=point";
    w->pos.pod = kPodFloat32; w->pos.extent = 3;
    w->ids.name = ".pointIds"; w->ids.meta = "geoScope=var";
    w->ids.pod = kPodUint64; w->ids.extent = 1;
    w->vel.name = ".velocities"; w->vel.meta = "interpretation=vector";
    w->vel.pod = kPodFloat32; w->vel.extent = 3;
    if (w->failed) {
        std::fclose(w->f);
        return set_error(DFSPH_E_IO, std::string("cannot write ") + path);
    }
    *out = w.release();
    return DFSPH_OK;
}

// ParticleWriter::writeFrame (alembic_writer.cpp:51-80): positions, ids (the
// caller's, else 0..count-1) and, when given, velocities; the bounds are
// those of the positions, as OPointsSchema computes them.
int dfsph_abc_write(dfsph_abc* w, const float* positions, const float* velocities, const uint64_t* ids,
                    uint32_t count) {
    if (!w || !positions) return set_error(DFSPH_E_INVALID, "NULL argument");
    if (count == 0) return set_error(DFSPH_E_INVALID, "a frame needs at least one particle");
    if (w->failed) return set_error(DFSPH_E_IO, "an earlier write to " + w->path + " failed");
    double box[6] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
                     std::numeric_limits<double>::max(), -std::numeric_limits<double>::max(),
                     -std::numeric_limits<double>::max(), -std::numeric_limits<double>::max()};
    for (uint32_t i = 0; i < count; ++i)
        for (int a = 0; a < 3; ++a) {
            const double x = positions[3 * size_t(i) + a];
            box[a] = std::min(box[a], x);
            box[3 + a] = std::max(box[3 + a], x);
        }
    std::vector<uint64_t> seq;
    if (!ids) {
        seq.resize(count);
        for (uint32_t i = 0; i < count; ++i) seq[i] = i;
        ids = seq.data();
    }
    w->set(w->bounds, box, sizeof(box), 1);
    w->set(w->pos, positions, size_t(count) * 12, count);
    w->set(w->ids, ids, size_t(count) * 8, count);
    if (velocities && !w->has_vel) {
        // Created at the first frame with velocities; earlier frames get
        // empty samples, as OPointsSchema back-fills them.
        w->has_vel = true;
        for (uint32_t q = 0; q < w->frames; ++q) w->set(w->vel, nullptr, 0, 0);
    }
    if (w->has_vel) w->set(w->vel, velocities, velocities ? size_t(count) * 12 : 0, velocities ? count : 0);
    ++w->frames;
    if (w->failed) return set_error(DFSPH_E_IO, "write failed: " + w->path);
    return DFSPH_OK;
}

// The reference's per-frame export (main.cpp:2875-2897): read back the
// particles, unpack position and velocity, write a frame with ids 0..N-1.
int dfsph_abc_write_ctx(dfsph_abc* w, dfsph_ctx* c) {
    if (!w || !c) return set_error(DFSPH_E_INVALID, "NULL argument");
    size_t bytes = 0;
    if (int rc = dfsph_field_size(c, DFSPH_FIELD_POSITION, &bytes)) return rc;
    const size_t n = bytes / 12;
    if (n == 0) return set_error(DFSPH_E_INVALID, "no particles to export");
    std::vector<float> x(n * 3), v(n * 3);
    if (int rc = dfsph_export(c, DFSPH_FIELD_POSITION, x.data(), n * 12)) return rc;
    if (int rc = dfsph_export(c, DFSPH_FIELD_VELOCITY, v.data(), n * 12)) return rc;
    return dfsph_abc_write(w, x.data(), v.data(), nullptr, uint32_t(n));
}

int dfsph_abc_close(dfsph_abc* w) {
    if (!w) return set_error(DFSPH_E_INVALID, "NULL argument");
    const int rc = w->finish();
    const bool closed = std::fclose(w->f) == 0;
    const std::string path = w->path;
    delete w;
    if (rc) return rc;
    if (!closed) return set_error(DFSPH_E_IO, "cannot close " + path);
    return DFSPH_OK;
}

}  // extern "C"
