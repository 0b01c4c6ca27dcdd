// capture_harness.cpp -- drives the library's host code (the capture reader
// and writer of gpu-sims_amd/csrc/capture.cpp, and the domain checks,
// derivation, boundary sampler and preset fill of host_util.cpp) in a
// host-only build under AddressSanitizer and UndefinedBehaviorSanitizer, with
// the Alembic writer of alembic.cpp.  The device-backed calls that capture.cpp
// and alembic.cpp make (dfsph_field_size, dfsph_export,
// dfsph_init_scene and the ctx_* helpers) are replaced here by a context held
// in host vectors, so the JSON parser, the buffer checks and the save/load
// paths run exactly as they do in the library while every byte they touch is
// checked.
//
//   capture_harness info DIR...          read_info on each DIR
//   capture_harness latest ROOT [BYTES]  find_latest with a BYTES-long output buffer
//   capture_harness load DIR             capture_load into a host context
//   capture_harness roundtrip ROOT N S   save N particles (S=1: after a step),
//                                        read back, load, compare bit for bit
//   capture_harness domain X0 Y0 Z0 X1 Y1 Z1 R SCALE
//                                        domain_problem, then derive and make_grid
//   capture_harness boundary X0 Y0 Z0 X1 Y1 Z1 R CAP
//                                        samples and their segment description,
//                                        which must rebuild every sample exactly
//   capture_harness preset INDEX R CAPACITY
//                                        preset_problem, then preset_fill
//   capture_harness abc PATH N FRAMES    an Alembic archive of FRAMES frames
//                                        of N particles (repeats, late
//                                        velocities, a write_ctx frame)
//
// Each call prints one line: "OK ..." or "ERR <code> <message>".  The exit
// status is 0 unless a comparison fails (1) or a sanitizer stops the run.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../gpu-sims_amd/csrc/dfsph_internal.h"

struct dfsph_ctx {
    uint32_t n = 0;
    std::vector<float> x, v, p;
    std::vector<uint32_t> flags, starts;
    bool stepped = false;
    uint64_t step = 0;
    float radius = 0.005f;
};

namespace {
std::string g_error;

size_t field_bytes(const dfsph_ctx* c, int32_t field) {
    switch (field) {
        case DFSPH_FIELD_POSITION:
        case DFSPH_FIELD_VELOCITY: return size_t(c->n) * 12;
        case DFSPH_FIELD_CELL_STARTS: return c->starts.size() * 4;
        default: return size_t(c->n) * 4;
    }
}

// Per-step values of a stepped context: a fixed function of (field, i).
float synth(int32_t field, uint32_t i) { return float(field) + 1e-3f * float(i % 977) + 0.25f; }
}  // namespace

namespace dfsph {
int set_error(int code, const std::string& msg) {
    g_error = msg;
    return code;
}
float ctx_particle_radius(const dfsph_ctx* c) { return c->radius; }
uint32_t ctx_particle_flags(const dfsph_ctx* c, uint32_t i) { return c->flags.at(i); }
int ctx_resume(dfsph_ctx* c, const float* p_density, const uint32_t* flags, uint64_t step) {
    c->p.assign(p_density, p_density + c->n);
    c->flags.assign(flags, flags + c->n);
    c->step = step;
    return DFSPH_OK;
}
}  // namespace dfsph

extern "C" {
int dfsph_field_size(dfsph_ctx* c, int32_t field, size_t* bytes) {
    if (!c || !bytes) return dfsph::set_error(DFSPH_E_INVALID, "NULL argument");
    *bytes = field_bytes(c, field);
    return DFSPH_OK;
}

int dfsph_export(dfsph_ctx* c, int32_t field, void* dst, size_t bytes) {
    if (!c || !dst) return dfsph::set_error(DFSPH_E_INVALID, "NULL argument");
    if (bytes != field_bytes(c, field)) return dfsph::set_error(DFSPH_E_INVALID, "size mismatch");
    if (field == DFSPH_FIELD_POSITION) { std::memcpy(dst, c->x.data(), bytes); return DFSPH_OK; }
    if (field == DFSPH_FIELD_VELOCITY) { std::memcpy(dst, c->v.data(), bytes); return DFSPH_OK; }
    if (!c->stepped) return dfsph::set_error(DFSPH_E_INVALID, "no step taken yet");
    if (field == DFSPH_FIELD_CELL_STARTS) { std::memcpy(dst, c->starts.data(), bytes); return DFSPH_OK; }
    if (field == DFSPH_FIELD_PRESSURE) { std::memcpy(dst, c->p.data(), bytes); return DFSPH_OK; }
    for (uint32_t i = 0; i < c->n; ++i) {
        if (field == DFSPH_FIELD_MORTON_KEY || field == DFSPH_FIELD_SORTED_INDEX) {
            const uint32_t u = field == DFSPH_FIELD_SORTED_INDEX ? c->n - 1 - i : i * 2654435761u;
            std::memcpy(static_cast<char*>(dst) + 4 * size_t(i), &u, 4);
        } else {
            static_cast<float*>(dst)[i] = synth(field, i);
        }
    }
    return DFSPH_OK;
}

int dfsph_init_scene(dfsph_ctx* c, const dfsph_scene* s) {
    if (!c || !s) return dfsph::set_error(DFSPH_E_INVALID, "NULL argument");
    if (s->count == 0 || !s->positions) return dfsph::set_error(DFSPH_E_INVALID, "empty scene");
    c->n = s->count;
    c->x.assign(s->positions, s->positions + size_t(s->count) * 3);
    if (s->velocities) c->v.assign(s->velocities, s->velocities + size_t(s->count) * 3);
    else c->v.assign(size_t(s->count) * 3, 0.0f);
    if (s->flags) c->flags.assign(s->flags, s->flags + s->count);
    else c->flags.assign(s->count, 0u);
    c->p.assign(s->count, 0.0f);
    c->stepped = false;
    return DFSPH_OK;
}
}  // extern "C"

namespace {
int report(int rc, const std::string& ok) {
    if (rc == DFSPH_OK) std::printf("OK %s\n", ok.c_str());
    else std::printf("ERR %d %s\n", rc, g_error.c_str());
    return rc;
}

std::string info_text(const dfsph_capture_info& in) {
    char buf[160];
    std::snprintf(buf, sizeof(buf), "%" PRIu64 " %u %u %d %d %d", in.iteration, in.frame, in.particle_count,
                  in.tier_index, in.preset_index, in.substeps);
    return buf;
}

void fill_scene(dfsph_ctx& c, uint32_t n, bool stepped) {
    c.n = n;
    c.x.resize(size_t(n) * 3);
    c.v.resize(size_t(n) * 3);
    c.p.resize(n);
    c.flags.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            c.x[3 * size_t(i) + a] = 0.01f * float((i >> (5 * a)) & 31u) + 1e-4f * float(a);
            c.v[3 * size_t(i) + a] = 0.5f - 1e-3f * float((i * 7u + a) % 1000u);
        }
        c.p[i] = stepped ? 10.0f + float(i % 13u) : 0.0f;
        c.flags[i] = i % 4u;
    }
    c.starts.assign(stepped ? 65 : 0, 0u);
    for (size_t q = 0; q < c.starts.size(); ++q) c.starts[q] = uint32_t(q * n / 64);
    c.stepped = stepped;
}

int roundtrip(const char* root, uint32_t n, bool stepped) {
    dfsph_ctx a;
    fill_scene(a, n, stepped);
    dfsph_capture_info info{};
    info.iteration = 4242;
    info.tier_index = 2;
    info.preset_index = 1;
    info.substeps = 3;
    info.camera_json = "{\"pos\": [1, 2.5, -3e-2], \"name\": \"cam \\u00e9\\n\"}";
    char dir[512];
    int rc = dfsph_capture_save(&a, root, 12, &info, dir, sizeof(dir));
    if (rc) return report(rc, "");
    dfsph_capture_info back{};
    if ((rc = dfsph_capture_read_info(dir, &back))) return report(rc, "");
    if (back.iteration != 4242 || back.frame != 12 || back.particle_count != n || back.tier_index != 2 ||
        back.preset_index != 1 || back.substeps != 3) {
        std::printf("FAIL info %s\n", info_text(back).c_str());
        return 1;
    }
    dfsph_ctx b;
    if ((rc = dfsph_capture_load(&b, dir, 1, 0, &back))) return report(rc, "");
    const bool same = b.n == n && std::memcmp(b.x.data(), a.x.data(), a.x.size() * 4) == 0 &&
                      std::memcmp(b.v.data(), a.v.data(), a.v.size() * 4) == 0 &&
                      std::memcmp(b.p.data(), a.p.data(), a.p.size() * 4) == 0 && b.flags == a.flags &&
                      b.step == 4242;
    if (!same) {
        std::printf("FAIL load differs\n");
        return 1;
    }
    std::printf("OK %s\n", dir);
    return 0;
}
}  // namespace

float arg_f(char** argv, int q) { return std::strtof(argv[q], nullptr); }

int domain_cmd(char** argv) {
    dfsph_params p{};
    for (int a = 0; a < 3; ++a) { p.domain_min[a] = arg_f(argv, 2 + a); p.domain_max[a] = arg_f(argv, 5 + a); }
    p.particle_radius = arg_f(argv, 8);
    p.fine_cell_scale = arg_f(argv, 9);
    p.density0 = 1000.0f;
    if (const char* why = dfsph::domain_problem(p.domain_min, p.domain_max, p.particle_radius, p.fine_cell_scale)) {
        std::printf("ERR %s\n", why);
        return 0;
    }
    const dfsph::Derived d = dfsph::derive(p);
    const dfsph::Grid f = dfsph::make_grid(p, d, DFSPH_GRID_FINE);
    const dfsph::Grid r = dfsph::make_grid(p, d, DFSPH_GRID_REFERENCE);
    std::printf("OK %u %u %u %u %u\n", d.max_axis, r.cells, f.axes[0] * f.axes[1], f.axes[2], f.cells);
    return 0;
}

int boundary_cmd(char** argv) {
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) { lo[a] = arg_f(argv, 2 + a); hi[a] = arg_f(argv, 5 + a); }
    const float r = arg_f(argv, 8);
    const uint32_t cap = uint32_t(std::strtoul(argv[9], nullptr, 10));
    if (!dfsph::sampler_extent_ok(lo, hi, r)) {
        std::printf("ERR extent\n");
        return 0;
    }
    const std::vector<float> pts = dfsph::boundary_samples(lo, hi, r, cap);
    uint32_t total = 0;
    const std::vector<dfsph::BoundarySeg> segs = dfsph::boundary_segments(lo, hi, r, cap, &total);
    if (size_t(total) * 3 != pts.size()) {
        std::printf("FAIL segments hold %u samples, sampler %zu\n", total, pts.size() / 3);
        return 1;
    }
    const float sp = 2.0f * r;
    uint32_t next = 0;
    for (const dfsph::BoundarySeg& g : segs) {
        if (g.start != next) { std::printf("FAIL segment gap at %u\n", next); return 1; }
        for (uint32_t i = 0; i < g.count; ++i) {
            float q[3];
            q[g.ax_a] = g.kind == 2 ? g.base : g.base + float(i) * sp;
            q[g.ax_b] = g.c1;
            q[g.ax_c] = g.c2;
            if (std::memcmp(q, &pts[3 * size_t(g.start + i)], 12) != 0) {
                std::printf("FAIL sample %u differs\n", g.start + i);
                return 1;
            }
        }
        next = g.start + g.count;
    }
    std::printf("OK %u\n", total);
    return 0;
}

int preset_cmd(char** argv) {
    const int idx = std::atoi(argv[2]);
    if (idx < 0 || idx >= DFSPH_PRESET_COUNT) { std::printf("ERR index\n"); return 0; }
    const dfsph_preset& p = dfsph::preset_table()[idx];
    const float r = arg_f(argv, 3);
    const uint32_t capacity = uint32_t(std::strtoul(argv[4], nullptr, 10));
    if (const char* why = dfsph::preset_problem(p, r)) { std::printf("ERR %s\n", why); return 0; }
    uint32_t n = 0, brick = 0;
    dfsph::preset_fill(p, r, capacity, nullptr, nullptr, nullptr, &n, &brick);
    std::vector<float> x(size_t(n) * 3), v(size_t(n) * 3);
    std::vector<uint32_t> flags(n);
    uint32_t n2 = 0, brick2 = 0;
    dfsph::preset_fill(p, r, capacity, x.data(), v.data(), flags.data(), &n2, &brick2);
    for (size_t q = 0; q < x.size(); ++q)
        if (!std::isfinite(x[q]) || !std::isfinite(v[q])) { std::printf("FAIL non-finite at %zu\n", q / 3); return 1; }
    if (n2 != n || brick2 != brick || n > capacity) { std::printf("FAIL counts %u %u\n", n, n2); return 1; }
    std::printf("OK %u %u\n", n, brick);
    return 0;
}

int abc_cmd(char** argv) {
    const uint32_t n = uint32_t(std::strtoul(argv[3], nullptr, 10));
    const uint32_t frames = uint32_t(std::strtoul(argv[4], nullptr, 10));
    dfsph_abc* w = nullptr;
    if (int rc = dfsph_abc_open(argv[2], 30.0, &w)) return report(rc, "");
    dfsph_ctx c;
    fill_scene(c, n, false);
    std::vector<uint64_t> ids(n);
    for (uint32_t i = 0; i < n; ++i) ids[i] = 7u * i;
    int rc = DFSPH_OK;
    for (uint32_t f = 0; f < frames && !rc; ++f) {
        if (f % 3 == 2) for (float& x : c.x) x += 0.001f;           // two equal frames, then a change
        if (f == frames - 1) rc = dfsph_abc_write_ctx(w, &c);
        else rc = dfsph_abc_write(w, c.x.data(), f >= 2 ? c.v.data() : nullptr, f % 2 ? ids.data() : nullptr, n);
    }
    const int rc2 = dfsph_abc_close(w);
    return report(rc ? rc : rc2, std::to_string(frames));
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s info DIR... | latest ROOT [BYTES] | load DIR | roundtrip ROOT N S\n",
                     argv[0]);
        return 2;
    }
    const std::string cmd = argv[1];
    if (cmd == "info") {
        for (int q = 2; q < argc; ++q) {
            dfsph_capture_info in{};
            const int rc = dfsph_capture_read_info(argv[q], &in);
            report(rc, rc ? "" : info_text(in));
        }
        return 0;
    }
    if (cmd == "latest") {
        const size_t bytes = argc > 3 ? size_t(std::strtoul(argv[3], nullptr, 10)) : 512;
        std::vector<char> out(bytes ? bytes : 1);
        const int rc = dfsph_capture_find_latest(argv[2], out.data(), bytes);
        report(rc, rc ? "" : std::string(out.data()));
        return 0;
    }
    if (cmd == "load") {
        dfsph_ctx c;
        dfsph_capture_info in{};
        const int rc = dfsph_capture_load(&c, argv[2], 0, 0, &in);
        double sum = 0.0;
        for (float f : c.x) sum += f;
        for (float f : c.p) sum += f;
        char buf[64];
        std::snprintf(buf, sizeof(buf), " %.9g", sum);
        report(rc, rc ? "" : info_text(in) + buf);
        return 0;
    }
    if (cmd == "domain" && argc >= 10) return domain_cmd(argv);
    if (cmd == "boundary" && argc >= 10) return boundary_cmd(argv);
    if (cmd == "preset" && argc >= 5) return preset_cmd(argv);
    if (cmd == "abc" && argc >= 5) return abc_cmd(argv) ? 1 : 0;
    if (cmd == "roundtrip" && argc >= 5) {
        return roundtrip(argv[2], uint32_t(std::strtoul(argv[3], nullptr, 10)), std::atoi(argv[4]) != 0);
    }
    std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
    return 2;
}
