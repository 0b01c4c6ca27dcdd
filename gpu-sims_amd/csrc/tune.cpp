// tune.cpp -- autotuner for the launch shape of the substep (host code on top
// of the library's own C ABI).  Each candidate configuration is created on
// the caller's scene(s), stepped once untimed, then `steps` substeps are timed
// with the host clock (every step ends with a device wait, so wall time is
// device time plus launch overhead).  Results may be cached in a text file,
// one line per choice:
//   single <arch> <n> <hash> : <block> <tuned flag bits> <ms> <cell scale>
//   batch <arch> <scenes> <n> <hash> : <workgroups> <tuned flag bits> <ms> 0
// where <hash> is a 64-bit FNV-1a hash of every input the choice depends on
// (each scene's count, constants, domain and boundary, the grid kind, the
// caller's untuned flags, the whole solver policy).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "dfsph_internal.h"

using namespace dfsph;

namespace {

std::string arch_of(int device) {
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "unknown";
    std::string a(prop.gcnArchName);
    const size_t colon = a.find(':');
    return colon == std::string::npos ? a : a.substr(0, colon);
}

int cu_count(int device) {
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 0;
    return prop.multiProcessorCount;
}

// Flag bits the tuners choose; every other bit is the caller's and part of the key.
constexpr uint32_t kTunedFlags = DFSPH_FLAG_NEIGHBOR_LIST | DFSPH_FLAG_NO_XCD_REMAP | DFSPH_FLAG_NO_NEIGHBOR_MASKS |
                                  DFSPH_FLAG_NO_STAGED_WINDOWS;
constexpr uint32_t kTunedBatchFlags = DFSPH_BATCH_NO_XCD_MAP;

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    template <class T>
    Fnv& add(const T& v) {
        const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
        for (size_t q = 0; q < sizeof(T); ++q) {
            h ^= p[q];
            h *= 1099511628211ull;
        }
        return *this;
    }
    template <class T, size_t N>
    Fnv& add_n(const T (&v)[N]) {
        for (size_t q = 0; q < N; ++q) add(v[q]);
        return *this;
    }
};

void hash_scene(Fnv& f, const dfsph_params& p, const dfsph_scene& sc) {
    f.add(sc.count).add(sc.boundary).add(sc.boundary_cap).add(p.particle_radius).add(p.density0);
    f.add_n(p.gravity).add(p.viscosity).add(p.cohesion).add(p.vorticity).add_n(p.domain_min).add_n(p.domain_max);
    f.add(p.grid);
}

void hash_policy(Fnv& f, const dfsph_solver_policy& pol) {
    f.add(pol.mode).add(pol.warm_start).add(pol.div_enabled).add(pol.min_iter_div).add(pol.max_iter_div);
    f.add(pol.min_iter_density).add(pol.max_iter_density).add(pol.tol_div).add(pol.tol_density).add(pol.cfl);
    f.add(pol.dt_max).add(pol.boundary_psi).add(pol.strict).add(pol.residual_bound);
}

std::string fmt_key(const char* kind, const std::string& arch, uint64_t scenes, uint64_t n, uint64_t hash) {
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s %s %llu %llu %016llx", kind, arch.c_str(), (unsigned long long)scenes,
                  (unsigned long long)n, (unsigned long long)hash);
    return std::string(buf);
}

// Looks the key up; returns true with (a, b, ms, c) on a hit.
bool cache_find(const char* path, const std::string& key, uint32_t* a, uint32_t* b, double* ms, float* c) {
    if (!path) return false;
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    while (std::getline(in, line)) {
        const size_t sep = line.find(" : ");
        if (sep == std::string::npos || line.compare(0, sep, key) != 0 || sep != key.size()) continue;
        std::istringstream rest(line.substr(sep + 3));
        unsigned long long x = 0, y = 0;
        double t = 0.0, z = 0.0;
        if (rest >> x >> y >> t >> z) {
            *a = uint32_t(x);
            *b = uint32_t(y);
            *ms = t;
            *c = float(z);
            return true;
        }
    }
    return false;
}

void cache_add(const char* path, const std::string& key, uint32_t a, uint32_t b, double ms, float c) {
    if (!path) return;
    std::ofstream out(path, std::ios::app);
    if (out) out << key << " : " << a << " " << b << " " << ms << " " << c << "\n";
}

using Clock = std::chrono::steady_clock;

}  // namespace

extern "C" {

int dfsph_autotune(const dfsph_params* params, const dfsph_scene* scene, const dfsph_solver_policy* pol,
                   uint32_t steps, const char* cache_path, dfsph_tune_result* out) {
    if (!params || !scene || !pol || !out) return set_error(DFSPH_E_INVALID, "NULL argument");
    if (int rc = check_params(params)) return rc;
    if (int rc = check_policy(pol)) return rc;
    if (steps == 0) return set_error(DFSPH_E_INVALID, "steps must be >= 1");
    *out = dfsph_tune_result{};
    Fnv f;
    hash_scene(f, *params, *scene);
    hash_policy(f, *pol);
    f.add(params->flags & ~kTunedFlags);
    const std::string key = fmt_key("single", arch_of(params->device), 1, scene->count, f.h);
    uint32_t a = 0, b = 0;
    double ms = 0.0;
    float cs = 0.0f;
    if (cache_find(cache_path, key, &a, &b, &ms, &cs)) {
        out->block_size = a;
        out->flags = (params->flags & ~kTunedFlags) | (b & kTunedFlags);
        out->fine_cell_scale = cs;
        out->ms_per_step = ms;
        out->cached = 1;
        return DFSPH_OK;
    }
    const uint32_t blocks[3] = {64, 128, 256};
    std::vector<uint32_t> formats = {0u};
    std::vector<float> cells = {0.0f};
    if (params->grid == DFSPH_GRID_FINE) {
        formats.push_back(DFSPH_FLAG_NO_STAGED_WINDOWS);
        formats.push_back(DFSPH_FLAG_NEIGHBOR_LIST);
        cells.push_back(1.25f);
        cells.push_back(1.5f);
    }
    const uint32_t maps[2] = {0u, DFSPH_FLAG_NO_XCD_REMAP};
    double best = 1e300, worst = 0.0;
    for (float cell : cells)
        for (uint32_t blk : blocks)
            for (uint32_t fm : formats)
                for (uint32_t m : maps) {
                    dfsph_params p = *params;
                    p.block_size = blk;
                    p.flags = (params->flags & ~kTunedFlags) | fm | m;
                    p.fine_cell_scale = cell;
                    dfsph_ctx* c = nullptr;
                    if (int rc = dfsph_create(&p, &c)) return rc;
                    int rc = dfsph_init_scene(c, scene);
                    if (!rc) rc = dfsph_step(c, 0.0f, 1, pol, nullptr);
                    double t = 0.0;
                    if (!rc) {
                        const auto t0 = Clock::now();
                        for (uint32_t s = 0; s < steps && !rc; ++s) rc = dfsph_step(c, 0.0f, 1, pol, nullptr);
                        t = std::chrono::duration<double, std::milli>(Clock::now() - t0).count() / steps;
                    }
                    dfsph_destroy(c);
                    if (rc) return rc;
                    ++out->candidates;
                    worst = std::max(worst, t);
                    if (t < best) {
                        best = t;
                        out->block_size = blk;
                        out->flags = p.flags;
                        out->fine_cell_scale = cell;
                    }
                }
    out->ms_per_step = best;
    out->worst_ms_per_step = worst;
    cache_add(cache_path, key, out->block_size, out->flags & kTunedFlags, best, out->fine_cell_scale);
    return DFSPH_OK;
}

int dfsph_batch_autotune(const dfsph_batch_config* cfg, const dfsph_params* params, const dfsph_scene* scenes,
                         uint32_t count, const dfsph_solver_policy* pol, uint32_t steps, const char* cache_path,
                         dfsph_tune_result* out) {
    if (!cfg || !params || !scenes || !pol || !out || count == 0)
        return set_error(DFSPH_E_INVALID, "NULL argument or empty batch");
    if (int rc = check_policy(pol)) return rc;
    if (steps == 0) return set_error(DFSPH_E_INVALID, "steps must be >= 1");
    *out = dfsph_tune_result{};
    uint64_t n = 0;
    Fnv f;
    for (uint32_t s = 0; s < count; ++s) {
        n += scenes[s].count;
        hash_scene(f, params[s], scenes[s]);
        f.add(params[s].fine_cell_scale);
    }
    hash_policy(f, *pol);
    f.add(cfg->flags & ~kTunedBatchFlags);
    const std::string key = fmt_key("batch", arch_of(cfg->device), count, n, f.h);
    uint32_t a = 0, b = 0;
    double ms = 0.0;
    float cs = 0.0f;
    if (cache_find(cache_path, key, &a, &b, &ms, &cs)) {
        out->workgroups = a;
        out->batch_flags = (cfg->flags & ~kTunedBatchFlags) | (b & kTunedBatchFlags);
        out->ms_per_step = ms;
        out->cached = 1;
        return DFSPH_OK;
    }
    const uint32_t cus = uint32_t(std::max(cu_count(cfg->device), 1));
    // The co-resident limit comes from a probe batch of the caller's scenes.
    dfsph_batch* probe = nullptr;
    if (int rc = dfsph_batch_create(cfg, params, scenes, count, &probe)) return rc;
    dfsph_batch_info_t info{};
    dfsph_batch_info(probe, &info);
    dfsph_batch_destroy(probe);
    std::vector<uint32_t> grids;
    for (uint32_t per : {1u, 2u, 4u, 8u}) {
        const uint32_t g = std::min(per * cus, info.max_workgroups) & ~7u;
        if (g >= 8 && std::find(grids.begin(), grids.end(), g) == grids.end()) grids.push_back(g);
    }
    double best = 1e300, worst = 0.0;
    for (uint32_t g : grids)
        for (uint32_t m : {0u, uint32_t(DFSPH_BATCH_NO_XCD_MAP)}) {
            dfsph_batch_config c2 = *cfg;
            c2.workgroups = g;
            c2.flags = (cfg->flags & ~kTunedBatchFlags) | m;
            dfsph_batch* bt = nullptr;
            if (int rc = dfsph_batch_create(&c2, params, scenes, count, &bt)) return rc;
            int rc = dfsph_batch_step(bt, 0.0f, pol, nullptr, nullptr);
            double t = 0.0;
            if (!rc) {
                const auto t0 = Clock::now();
                for (uint32_t s = 0; s < steps && !rc; ++s) rc = dfsph_batch_step(bt, 0.0f, pol, nullptr, nullptr);
                t = std::chrono::duration<double, std::milli>(Clock::now() - t0).count() / steps;
            }
            dfsph_batch_destroy(bt);
            if (rc) return rc;
            ++out->candidates;
            worst = std::max(worst, t);
            if (t < best) {
                best = t;
                out->workgroups = g;
                out->batch_flags = c2.flags;
            }
        }
    out->ms_per_step = best;
    out->worst_ms_per_step = worst;
    cache_add(cache_path, key, out->workgroups, out->batch_flags & kTunedBatchFlags, best, 0.0f);
    return DFSPH_OK;
}

}  // extern "C"
