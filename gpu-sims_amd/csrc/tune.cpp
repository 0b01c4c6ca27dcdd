// tune.cpp -- autotuner for the launch shape of the substep (host code on top
// of the library's own C ABI).  Each candidate configuration is created on
// the caller's scene(s), stepped once untimed, then `steps` substeps are timed
// with the host clock (every step ends with a device wait, so wall time is
// device time plus launch overhead).  Results may be cached in a text file,
// one line per choice:
//   single <arch> <n> <radius> <dmin xyz> <dmax xyz> <mode> <warm> : <block> <flags> <ms>
//   batch <arch> <scenes> <n> <radius0> <dmin0 xyz> <dmax0 xyz> <mode> <warm> : <workgroups> <flags> <ms>
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

std::string fmt_key(const char* kind, const std::string& arch, uint64_t scenes, uint64_t n,
                    const dfsph_params& p, const dfsph_solver_policy& pol) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s %s %llu %llu %.9g %.9g %.9g %.9g %.9g %.9g %.9g %d %d", kind, arch.c_str(),
                  (unsigned long long)scenes, (unsigned long long)n, double(p.particle_radius),
                  double(p.domain_min[0]), double(p.domain_min[1]), double(p.domain_min[2]),
                  double(p.domain_max[0]), double(p.domain_max[1]), double(p.domain_max[2]), pol.mode,
                  pol.warm_start);
    return std::string(buf);
}

// Looks the key up; returns true with (a, b, ms) on a hit.
bool cache_find(const char* path, const std::string& key, uint32_t* a, uint32_t* b, double* ms) {
    if (!path) return false;
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    while (std::getline(in, line)) {
        const size_t sep = line.find(" : ");
        if (sep == std::string::npos || line.compare(0, sep, key) != 0 || sep != key.size()) continue;
        std::istringstream rest(line.substr(sep + 3));
        unsigned long long x = 0, y = 0;
        double t = 0.0;
        if (rest >> x >> y >> t) {
            *a = uint32_t(x);
            *b = uint32_t(y);
            *ms = t;
            return true;
        }
    }
    return false;
}

void cache_add(const char* path, const std::string& key, uint32_t a, uint32_t b, double ms) {
    if (!path) return;
    std::ofstream out(path, std::ios::app);
    if (out) out << key << " : " << a << " " << b << " " << ms << "\n";
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
    const std::string key = fmt_key("single", arch_of(params->device), 1, scene->count, *params, *pol);
    uint32_t a = 0, b = 0;
    double ms = 0.0;
    if (cache_find(cache_path, key, &a, &b, &ms)) {
        out->block_size = a;
        out->flags = b;
        out->ms_per_step = ms;
        out->cached = 1;
        return DFSPH_OK;
    }
    const uint32_t blocks[3] = {64, 128, 256};
    std::vector<uint32_t> formats = {0u};
    if (params->grid == DFSPH_GRID_FINE) formats.push_back(DFSPH_FLAG_NEIGHBOR_LIST);
    const uint32_t maps[2] = {0u, DFSPH_FLAG_NO_XCD_REMAP};
    double best = 1e300, worst = 0.0;
    for (uint32_t blk : blocks)
        for (uint32_t f : formats)
            for (uint32_t m : maps) {
                dfsph_params p = *params;
                p.block_size = blk;
                p.flags = (params->flags & ~(DFSPH_FLAG_NEIGHBOR_LIST | DFSPH_FLAG_NO_XCD_REMAP |
                                              DFSPH_FLAG_NO_NEIGHBOR_MASKS)) | f | m;
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
                }
            }
    out->ms_per_step = best;
    out->worst_ms_per_step = worst;
    cache_add(cache_path, key, out->block_size, out->flags, best);
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
    for (uint32_t s = 0; s < count; ++s) n += scenes[s].count;
    const std::string key = fmt_key("batch", arch_of(cfg->device), count, n, params[0], *pol);
    uint32_t a = 0, b = 0;
    double ms = 0.0;
    if (cache_find(cache_path, key, &a, &b, &ms)) {
        out->workgroups = a;
        out->batch_flags = b;
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
            c2.flags = (cfg->flags & ~uint32_t(DFSPH_BATCH_NO_XCD_MAP)) | m;
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
    cache_add(cache_path, key, out->workgroups, out->batch_flags, best);
    return DFSPH_OK;
}

}  // extern "C"
