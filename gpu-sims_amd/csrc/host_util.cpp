// host_util.cpp -- host-side parts of the DFSPH library that need no device:
// parameter derivation, the per-pair test thresholds, the Akinci boundary
// sampler and the seeded lattice generator.  Compiled into libdfsph.so.
#include "dfsph_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace dfsph {

namespace {

uint32_t pow2_at_least(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

inline float bits_to_float(uint32_t b) { float f; std::memcpy(&f, &b, 4); return f; }

// Smallest non-negative float x with pred(x) true, for a predicate that is
// false below some point and true from it on (the IEEE sqrt and division are
// correctly rounded, hence monotone, so every test here has that shape).
template <class P>
float first_true(P pred) {
    uint32_t lo = 0, hi = 0x7f800000u;  // +inf is always true for these predicates
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (pred(bits_to_float(mid))) hi = mid; else lo = mid + 1;
    }
    return bits_to_float(lo);
}

}  // namespace

const char* domain_problem(const float dmin[3], const float dmax[3], float radius, float fine_cell_scale) {
    if (!(radius > 0.0f) || !std::isfinite(radius)) return "particle_radius must be finite and > 0";
    double fine_cells = 1.0;
    const double fine = 4.0 * double(radius) * std::max(double(fine_cell_scale), 1.0009765625);
    for (int a = 0; a < 3; ++a) {
        if (!std::isfinite(dmin[a]) || !std::isfinite(dmax[a]) || !(dmax[a] > dmin[a]))
            return "domain bounds must be finite with domain_max > domain_min on every axis";
        const double ext = double(dmax[a]) - double(dmin[a]);
        // Morton table: max_axis^3 cells held in 32 bits (max_axis <= 1024).
        if (std::ceil(ext / (8.0 * double(radius))) > 1024.0)
            return "domain spans more than 1024 reference cells (8 particle radii) on an axis";
        fine_cells *= std::max(1.0, std::ceil(ext / fine));
    }
    if (fine_cells > 2147483648.0) return "domain holds more than 2^31 FINE grid cells";
    return nullptr;
}

bool sampler_extent_ok(const float dmin[3], const float dmax[3], float radius) {
    if (!(radius > 0.0f) || !std::isfinite(radius)) return false;
    for (int a = 0; a < 3; ++a) {
        if (!std::isfinite(dmin[a]) || !std::isfinite(dmax[a]) || !(dmax[a] >= dmin[a])) return false;
        if ((double(dmax[a]) - double(dmin[a])) / (2.0 * double(radius)) > 16777216.0) return false;
    }
    return true;
}

Derived derive(const dfsph_params& p) {
    Derived d{};
    const float r = p.particle_radius;
    // main.cpp:268-270: h = r*4, Morton cell = r*4*2 (both exact power-of-two scalings).
    d.h = r * 4.0f;
    d.cell = r * 4.0f * 2.0f;
    const float spacing = 2.0f * r;                    // main.cpp:2008-2011
    d.mass = p.density0 * spacing * spacing * spacing;
    const float pi = float(M_PI);                      // main.cpp:2012-2019
    d.knorm = 8.0f / (pi * d.h * d.h * d.h);
    d.gnorm = 8.0f / (pi * d.h * d.h * d.h * d.h);
    d.relax = 0.5f;                                    // main.cpp:122
    for (int a = 0; a < 3; ++a) {                      // main.cpp:1994-2005
        const float ext = p.domain_max[a] - p.domain_min[a];
        const uint32_t cells = std::max(1u, uint32_t(std::ceil(ext / d.cell)));
        d.axes[a] = pow2_at_least(cells);
    }
    d.max_axis = std::max(d.axes[0], std::max(d.axes[1], d.axes[2]));
    d.total_cells = d.max_axis * d.max_axis * d.max_axis;   // main.cpp:2024 (no 2^18 cap)
    // FINE grid: edge slightly above h so that floating-point rounding of
    // (x - dmin)/cell can never put two particles closer than h three cells
    // apart; a caller may ask for larger cells (the autotuner's cell tiles).
    d.fine_cell = d.h * (p.fine_cell_scale > 0.0f ? std::max(p.fine_cell_scale, 1.0009765625f) : 1.0009765625f);
    for (int a = 0; a < 3; ++a) {
        const float ext = p.domain_max[a] - p.domain_min[a];
        d.fine_axes[a] = std::max(1u, uint32_t(std::ceil(ext / d.fine_cell)));
    }
    const float h = d.h;
    d.r2_q1 = first_true([h](float r2) { return !(std::sqrt(r2) / h < 1.0f); });
    d.r2_eps_gt = first_true([](float r2) { return std::sqrt(r2) > 1e-7f; });
    d.r2_eps_lt = first_true([](float r2) { return !(std::sqrt(r2) < 1e-7f); });
    return d;
}

// Akinci boundary samples on the AABB surface (main.cpp:1332-1431): hex rows
// on the six face interiors trimmed one spacing from each rim, one line per
// edge, one sample per corner.  Float arithmetic in the reference's order.
std::vector<float> boundary_samples(const float dmin[3], const float dmax[3], float radius,
                                    uint32_t cap) {
    std::vector<float> pts;
    const float sp = 2.0f * radius;
    if (!(sp > 0.0f)) return pts;
    bool full = false;
    auto emit = [&](float x, float y, float z) {
        if (full || (cap && pts.size() / 3 >= cap)) { full = true; return false; }
        pts.push_back(x); pts.push_back(y); pts.push_back(z);
        return true;
    };
    auto put = [&](int axis_n, float plane, int axis_u, float u, int axis_v, float v) {
        float q[3];
        q[axis_n] = plane; q[axis_u] = u; q[axis_v] = v;
        return emit(q[0], q[1], q[2]);
    };
    struct Face { int n; bool hi; int u, v; };
    const Face faces[6] = {{1, false, 0, 2}, {1, true, 0, 2}, {0, false, 1, 2},
                           {0, true, 1, 2},  {2, false, 0, 1}, {2, true, 0, 1}};
    for (const Face& f : faces) {
        const float plane = f.hi ? dmax[f.n] : dmin[f.n];
        const float u0 = dmin[f.u] + sp, u1 = dmax[f.u] - sp;
        const float v0 = dmin[f.v] + sp, v1 = dmax[f.v] - sp;
        if (u0 > u1 || v0 > v1) continue;
        const int rows = int(std::floor((v1 - v0) / sp));
        bool stop = false;
        for (int j = 0; j <= rows && !stop; ++j) {
            const float v = v0 + float(j) * sp;
            if (v > v1 + 1e-6f) break;
            const float shift = (j & 1) ? 0.5f * sp : 0.0f;
            const int cols = int(std::floor((u1 - u0 - shift) / sp));
            for (int i = 0; i <= cols; ++i) {
                const float u = u0 + shift + float(i) * sp;
                if (u > u1 + 1e-6f) break;
                if (!put(f.n, plane, f.u, u, f.v, v)) { stop = true; break; }
            }
        }
    }
    // Edges: free axis, then the two pinned coordinates (lo/hi) in the
    // reference's order.
    const int edge_free[12] = {0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2};
    const bool edge_hi[12][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}, {0, 0}, {0, 1},
                                 {1, 0}, {1, 1}, {0, 0}, {0, 1}, {1, 0}, {1, 1}};
    for (int e = 0; e < 12; ++e) {
        const int fa = edge_free[e];
        const int pa = fa == 0 ? 1 : 0;
        const int pb = fa == 2 ? 1 : 2;
        const float t0 = dmin[fa] + sp, t1 = dmax[fa] - sp;
        if (t0 > t1) continue;
        const int cnt = int(std::floor((t1 - t0) / sp));
        for (int i = 0; i <= cnt; ++i) {
            const float t = t0 + float(i) * sp;
            if (t > t1 + 1e-6f) break;
            float q[3];
            q[fa] = t;
            q[pa] = edge_hi[e][0] ? dmax[pa] : dmin[pa];
            q[pb] = edge_hi[e][1] ? dmax[pb] : dmin[pb];
            if (!emit(q[0], q[1], q[2])) break;
        }
    }
    for (int c = 0; c < 8; ++c) {   // x fastest, then y, then z (main.cpp:1423-1430)
        emit((c & 1) ? dmax[0] : dmin[0], (c & 2) ? dmax[1] : dmin[1], (c & 4) ? dmax[2] : dmin[2]);
    }
    return pts;
}

// The emission order of boundary_samples as segments, with the same loop
// bounds and float expressions; a row or edge stops at its first coordinate
// beyond the rim, as the sampler's loops do.
std::vector<BoundarySeg> boundary_segments(const float dmin[3], const float dmax[3], float radius, uint32_t cap,
                                           uint32_t* total) {
    std::vector<BoundarySeg> segs;
    uint32_t n = 0;
    const float sp = 2.0f * radius;
    bool full = false;
    auto add = [&](BoundarySeg g) {
        if (full || g.count == 0) return;
        if (cap && n + uint64_t(g.count) >= cap) {
            g.count = cap - n;
            full = true;
            if (g.count == 0) return;
        }
        g.start = n;
        n += g.count;
        segs.push_back(g);
    };
    // Prefix of i = 0..last with base + i*sp <= hi + 1e-6 (non-decreasing in i).
    auto run_len = [&](float base, int last, float hi) -> uint32_t {
        if (last < 0) return 0u;
        int k = last;
        while (k >= 0 && base + float(k) * sp > hi + 1e-6f) --k;
        return uint32_t(k + 1);
    };
    if (sp > 0.0f) {
        struct Face { int n; bool hi; int u, v; };
        const Face faces[6] = {{1, false, 0, 2}, {1, true, 0, 2}, {0, false, 1, 2},
                               {0, true, 1, 2},  {2, false, 0, 1}, {2, true, 0, 1}};
        for (const Face& f : faces) {
            const float plane = f.hi ? dmax[f.n] : dmin[f.n];
            const float u0 = dmin[f.u] + sp, u1 = dmax[f.u] - sp;
            const float v0 = dmin[f.v] + sp, v1 = dmax[f.v] - sp;
            if (u0 > u1 || v0 > v1) continue;
            const int rows = int(std::floor((v1 - v0) / sp));
            for (int j = 0; j <= rows; ++j) {
                const float v = v0 + float(j) * sp;
                if (v > v1 + 1e-6f) break;
                const float shift = (j & 1) ? 0.5f * sp : 0.0f;
                const int cols = int(std::floor((u1 - u0 - shift) / sp));
                const float base = u0 + shift;
                add(BoundarySeg{0, run_len(base, cols, u1), 0, f.u, f.n, f.v, base, plane, v});
            }
        }
        const int edge_free[12] = {0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2};
        const bool edge_hi[12][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}, {0, 0}, {0, 1},
                                     {1, 0}, {1, 1}, {0, 0}, {0, 1}, {1, 0}, {1, 1}};
        for (int e = 0; e < 12; ++e) {
            const int fa = edge_free[e];
            const int pa = fa == 0 ? 1 : 0;
            const int pb = fa == 2 ? 1 : 2;
            const float t0 = dmin[fa] + sp, t1 = dmax[fa] - sp;
            if (t0 > t1) continue;
            const int cnt = int(std::floor((t1 - t0) / sp));
            add(BoundarySeg{0, run_len(t0, cnt, t1), 1, fa, pa, pb, t0, edge_hi[e][0] ? dmax[pa] : dmin[pa],
                            edge_hi[e][1] ? dmax[pb] : dmin[pb]});
        }
    }
    if (sp > 0.0f)
        for (int c = 0; c < 8; ++c)
            add(BoundarySeg{0, 1, 2, 0, 1, 2, (c & 1) ? dmax[0] : dmin[0], (c & 2) ? dmax[1] : dmin[1],
                            (c & 4) ? dmax[2] : dmin[2]});
    if (total) *total = n;
    return segs;
}

namespace {
inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
}  // namespace

void scene_block(uint32_t nx, uint32_t ny, uint32_t nz, float radius, const float origin[3],
                 uint64_t seed, float jitter, uint32_t id0, float* out) {
    const float sp = 2.0f * radius;
    const float amp = jitter * radius;
    const uint64_t n = uint64_t(nx) * ny * nz;
    for (uint64_t id = 0; id < n; ++id) {
        const uint32_t idx[3] = {uint32_t(id % nx), uint32_t((id / nx) % ny), uint32_t(id / (uint64_t(nx) * ny))};
        for (int a = 0; a < 3; ++a) {
            const uint64_t key = (seed << 40) ^ ((uint64_t(id0) + id) * 3u + uint64_t(a));
            const float u = float(splitmix64(key) >> 40) * 0x1p-24f;
            const float jit = (2.0f * u - 1.0f) * amp;
            out[3 * id + a] = (origin[a] + (float(idx[a]) + 0.5f) * sp) + jit;
        }
    }
}

// Scene presets, main.cpp:191-232 (values as listed there).
const dfsph_preset* preset_table() {
    static const dfsph_preset t[DFSPH_PRESET_COUNT] = {
        {"Dam-Break", {+0.52f, -0.98f, -0.98f}, {+1.98f, +0.5f, +0.98f}, 0, {0.0f, 0.0f, 0.0f}, 0.0f,
         {0.0f, 0.0f, 0.0f}, 0, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, 0.0f,
         {-2.0f, -1.0f, -1.0f}, {+2.0f, +2.0f, +1.0f}, 0.005f, 0.0f, 0.5f, 0.0f},
        {"Central-Fountain", {-0.98f, -0.98f, -0.98f}, {+0.98f, -0.95f, +0.98f}, 0, {0.0f, 0.0f, 0.0f}, 0.0f,
         {0.0f, 0.0f, 0.0f}, 1, {0.0f, -0.9f, 0.0f}, {0.1f, 0.05f, 0.0f}, {0.0f, +5.0f, 0.0f}, 5000.0f,
         {-1.0f, -1.0f, -1.0f}, {+1.0f, +2.0f, +1.0f}, 0.01f, 0.05f, 0.8f, 0.0f},
        {"Droplet-Impact", {-1.48f, -0.98f, -1.48f}, {+1.48f, -0.97f, +1.48f}, 1, {0.0f, +1.0f, 0.0f}, 0.15f,
         {0.0f, -5.0f, 0.0f}, 0, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, 0.0f,
         {-1.5f, -1.0f, -1.5f}, {+1.5f, +2.0f, +1.5f}, 0.02f, 0.1f, 1.2f, 0.02f},
        {"Pour-from-Source", {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, 0, {0.0f, 0.0f, 0.0f}, 0.0f,
         {0.0f, 0.0f, 0.0f}, 2, {0.0f, +1.95f, 0.0f}, {0.3f, 0.0f, 0.3f}, {0.0f, -2.0f, 0.0f}, 8000.0f,
         {-1.5f, -1.0f, -1.5f}, {+1.5f, +2.0f, +1.5f}, 0.01f, 0.05f, 0.6f, 0.0f},
    };
    return t;
}

const char* preset_problem(const dfsph_preset& p, float radius) {
    if (!(radius > 0.0f) || !std::isfinite(radius)) return "radius must be finite and > 0";
    const double sp = 2.0 * double(radius);
    for (int a = 0; a < 3; ++a) {
        if (!std::isfinite(p.brick_min[a]) || !std::isfinite(p.brick_max[a]) || !(p.brick_max[a] >= p.brick_min[a]))
            return "brick_max must be >= brick_min (finite)";
        if ((double(p.brick_max[a]) - double(p.brick_min[a])) / sp > 1048576.0)
            return "the brick spans more than 2^20 particle spacings on an axis";
    }
    if (p.add_droplet) {
        if (!(p.droplet_radius >= 0.0f) || !std::isfinite(p.droplet_radius))
            return "droplet_radius must be finite and >= 0";
        if (2.0 * double(p.droplet_radius) / sp > 1048576.0)
            return "the droplet spans more than 2^20 particle spacings";
        for (int a = 0; a < 3; ++a)
            if (!std::isfinite(p.droplet_center[a]) || !std::isfinite(p.droplet_velocity[a]))
                return "droplet centre and velocity must be finite";
    }
    return nullptr;
}

// Counts of seedInitialFill (main.cpp:2266-2294) and the positions of
// initial_fill (shaders/initial_fill.comp.glsl:22-65), float operations in the
// shader's order without contraction.
void preset_fill(const dfsph_preset& p, float radius, uint32_t capacity, float* x, float* v, uint32_t* flags,
                 uint32_t* count, uint32_t* brick_count) {
    const float spacing = 2.0f * radius;
    float ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = p.brick_max[a] - p.brick_min[a];
    // 64-bit products: equal to the reference's 32-bit ones wherever those do
    // not wrap (preset_problem bounds every factor by 2^20).
    const uint64_t dx = std::max(1u, uint32_t(ext[0] / spacing));
    const uint64_t dy = std::max(1u, uint32_t(ext[1] / spacing));
    const uint64_t dz = std::max(1u, uint32_t(ext[2] / spacing));
    const uint32_t brick = uint32_t(std::min<uint64_t>(capacity, dx * dy * dz));
    uint32_t droplet = 0;
    if (p.add_droplet) {
        const uint64_t dr = std::max(1u, uint32_t(p.droplet_radius / spacing));
        droplet = uint32_t(std::min<uint64_t>((4u * dr * dr * dr) / 3u, capacity - brick));
    }
    const uint32_t total = brick + droplet;
    if (count) *count = total;
    if (brick_count) *brick_count = brick;
    if (!x && !v && !flags) return;
    uint32_t dims[3], ddims[3];
    for (int a = 0; a < 3; ++a) {
        dims[a] = uint32_t(std::max(ext[a] / spacing, 1.0f));
        ddims[a] = uint32_t(std::max(2.0f * p.droplet_radius / spacing, 1.0f));
    }
    for (uint32_t gid = 0; gid < total; ++gid) {
        float pos[3], vel[3] = {0.0f, 0.0f, 0.0f};
        uint32_t f = 0;
        if (gid < brick) {
            const uint32_t i = gid % std::max(dims[0], 1u);
            const uint32_t j = (gid / std::max(dims[0], 1u)) % std::max(dims[1], 1u);
            const uint32_t k = uint32_t(gid / std::max<uint64_t>(uint64_t(dims[0]) * dims[1], 1u));
            const uint32_t ijk[3] = {i, j, k};
            for (int a = 0; a < 3; ++a) pos[a] = p.brick_min[a] + (float(ijk[a]) + 0.5f) * spacing;
        } else {
            const uint32_t dg = gid - brick;
            const uint32_t i = dg % std::max(ddims[0], 1u);
            const uint32_t j = (dg / std::max(ddims[0], 1u)) % std::max(ddims[1], 1u);
            const uint32_t k = uint32_t(dg / std::max<uint64_t>(uint64_t(ddims[0]) * ddims[1], 1u));
            const uint32_t ijk[3] = {i, j, k};
            float l[3];
            for (int a = 0; a < 3; ++a) l[a] = (float(ijk[a]) - float(ddims[a]) * 0.5f + 0.5f) * spacing;
            const float ll = std::sqrt(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
            if (ll > p.droplet_radius)
                for (int a = 0; a < 3; ++a) l[a] = l[a] / std::max(ll, 1e-7f) * p.droplet_radius * 0.95f;
            for (int a = 0; a < 3; ++a) {
                pos[a] = p.droplet_center[a] + l[a];
                vel[a] = p.droplet_velocity[a];
            }
            f = 2u;
        }
        for (int a = 0; a < 3; ++a) {
            if (x) x[3 * size_t(gid) + a] = pos[a];
            if (v) v[3 * size_t(gid) + a] = vel[a];
        }
        if (flags) flags[gid] = f;
    }
}

Grid make_grid(const dfsph_params& p, const Derived& d, int kind) {
    Grid g{};
    g.kind = kind;
    for (int a = 0; a < 3; ++a) g.dmin[a] = p.domain_min[a];
    if (kind == DFSPH_GRID_REFERENCE) {
        g.cell = d.cell;
        for (int a = 0; a < 3; ++a) g.axes[a] = d.axes[a];
        g.cells = d.total_cells;
    } else {
        g.cell = d.fine_cell;
        for (int a = 0; a < 3; ++a) g.axes[a] = d.fine_axes[a];
        g.cells = d.fine_axes[0] * d.fine_axes[1] * d.fine_axes[2];
    }
    return g;
}

// alpha of an interior particle of the rest lattice (spacing 2r, h = 4r): the
// 26 neighbours at q = 1/2, sqrt(2)/2, sqrt(3)/2 (6, 12, 8 of them), all in the
// outer branch |gradW| = 6 l (1 - q)^2, and sum_j m gradW_ij = 0 by symmetry,
// so alpha = 1 / sum_j (m |gradW_ij|)^2.  About 2.354e-10 at r = 0.005.
static double alpha_lattice(const Derived& d) {
    double s = 0.0;
    for (int x = -1; x <= 1; ++x)
        for (int y = -1; y <= 1; ++y)
            for (int z = -1; z <= 1; ++z) {
                const int n2 = x * x + y * y + z * z;
                if (n2 == 0) continue;
                const double q = 0.5 * std::sqrt(double(n2));
                const double g = 6.0 * double(d.gnorm) * (1.0 - q) * (1.0 - q);
                s += (double(d.mass) * g) * (double(d.mass) * g);
            }
    return 1.0 / s;
}

SimConst make_sim_const(const dfsph_params& p, const Derived& d, float bscale, uint32_t n, uint32_t nb) {
    SimConst k{};
    k.h = d.h;
    k.inv_h = 1.0f / d.h;
    k.mass = d.mass;
    k.knorm = d.knorm;
    k.gnorm = d.gnorm;
    k.gh = d.gnorm / d.h;
    k.m6g = -6.0f * d.gnorm;
    k.rho0 = p.density0;
    k.relax = d.relax;
    k.visc = p.viscosity;
    k.coh = p.cohesion;
    k.vort = p.vorticity;
    for (int a = 0; a < 3; ++a) {
        k.g[a] = p.gravity[a];
        k.dmin[a] = p.domain_min[a];
        k.dmax[a] = p.domain_max[a];
    }
    k.r2_q1 = d.r2_q1;
    k.r2_eps_gt = d.r2_eps_gt;
    k.r2_eps_lt = d.r2_eps_lt;
    k.bscale = bscale;
    k.alpha_cap = float(10.0 * alpha_lattice(d));
    const double h = d.h;
    k.gfac = float(-6.0 * double(d.gnorm) / (h * h));
    k.wfac = float(2.0 * double(d.knorm) / (h * h * h));
    k.h3half = float(0.5 * h * h * h);
    k.n = n;
    k.nb = nb;
    k.slot0 = 0;
    return k;
}

}  // namespace dfsph

namespace dfsph {

void mul4(const float a[16], const float b[16], float out[16]) {
    float r[16];
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row)
            r[c * 4 + row] = ((a[row] * b[c * 4] + a[4 + row] * b[c * 4 + 1]) + a[8 + row] * b[c * 4 + 2]) +
                             a[12 + row] * b[c * 4 + 3];
    std::memcpy(out, r, sizeof(r));
}

// Cofactor inverse through the 2x2 minors of the top two and bottom two rows
// (Laplace expansion by complementary minors).
bool invert4(const float m[16], float out[16]) {
    auto A = [&](int r, int c) { return m[c * 4 + r]; };
    const float s0 = A(0, 0) * A(1, 1) - A(1, 0) * A(0, 1), s1 = A(0, 0) * A(1, 2) - A(1, 0) * A(0, 2);
    const float s2 = A(0, 0) * A(1, 3) - A(1, 0) * A(0, 3), s3 = A(0, 1) * A(1, 2) - A(1, 1) * A(0, 2);
    const float s4 = A(0, 1) * A(1, 3) - A(1, 1) * A(0, 3), s5 = A(0, 2) * A(1, 3) - A(1, 2) * A(0, 3);
    const float c5 = A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3), c4 = A(2, 1) * A(3, 3) - A(3, 1) * A(2, 3);
    const float c3 = A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2), c2 = A(2, 0) * A(3, 3) - A(3, 0) * A(2, 3);
    const float c1 = A(2, 0) * A(3, 2) - A(3, 0) * A(2, 2), c0 = A(2, 0) * A(3, 1) - A(3, 0) * A(2, 1);
    const float det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    if (!(det != 0.0f) || !std::isfinite(det)) return false;
    const float id = 1.0f / det;
    float r[16];
    auto I = [&](int row, int col, float v) { r[col * 4 + row] = v * id; };
    I(0, 0, A(1, 1) * c5 - A(1, 2) * c4 + A(1, 3) * c3);
    I(0, 1, -A(0, 1) * c5 + A(0, 2) * c4 - A(0, 3) * c3);
    I(0, 2, A(3, 1) * s5 - A(3, 2) * s4 + A(3, 3) * s3);
    I(0, 3, -A(2, 1) * s5 + A(2, 2) * s4 - A(2, 3) * s3);
    I(1, 0, -A(1, 0) * c5 + A(1, 2) * c2 - A(1, 3) * c1);
    I(1, 1, A(0, 0) * c5 - A(0, 2) * c2 + A(0, 3) * c1);
    I(1, 2, -A(3, 0) * s5 + A(3, 2) * s2 - A(3, 3) * s1);
    I(1, 3, A(2, 0) * s5 - A(2, 2) * s2 + A(2, 3) * s1);
    I(2, 0, A(1, 0) * c4 - A(1, 1) * c2 + A(1, 3) * c0);
    I(2, 1, -A(0, 0) * c4 + A(0, 1) * c2 - A(0, 3) * c0);
    I(2, 2, A(3, 0) * s4 - A(3, 1) * s2 + A(3, 3) * s0);
    I(2, 3, -A(2, 0) * s4 + A(2, 1) * s2 - A(2, 3) * s0);
    I(3, 0, -A(1, 0) * c3 + A(1, 1) * c1 - A(1, 2) * c0);
    I(3, 1, A(0, 0) * c3 - A(0, 1) * c1 + A(0, 2) * c0);
    I(3, 2, -A(3, 0) * s3 + A(3, 1) * s1 - A(3, 2) * s0);
    I(3, 3, A(2, 0) * s3 - A(2, 1) * s1 + A(2, 2) * s0);
    for (float v : r)
        if (!std::isfinite(v)) return false;
    std::memcpy(out, r, sizeof(r));
    return true;
}

}  // namespace dfsph
