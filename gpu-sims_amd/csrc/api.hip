// api.hip -- the C ABI (include/dfsph.h): context, scene set-up, substep
// orchestration and export.  Host code only; kernels live in hash.hip and
// sweeps.hip and are reached through their launch wrappers.
//
// Substep order follows the reference's dispatch chain (main.cpp:2683-2867):
// hash -> density/alpha (+ divergence source) -> divergence solve -> forces ->
// density source -> density solve -> position update.  The reference keeps
// particles in a fixed AoS record and walks them through sorted_index; this
// library physically reorders SoA state into cell order each substep, keeps the
// original indices alongside, and converts back on export.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dfsph_internal.h"

using namespace dfsph;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define DFSPH_HIP(call)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(DFSPH_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline uint32_t blocks_for(uint64_t n, uint32_t t) { return uint32_t((n + t - 1) / t); }

// Launch sites timed by the profiler (one name per kernel or kernel group).
enum KernelId {
    K_BEGIN_STEP, K_KEYS, K_SCAN, K_SCATTER, K_REORDER, K_DENSITY_ALPHA, K_BEGIN_SOLVE,
    K_PRESSURE_ACCEL, K_JACOBI, K_SOLVE_FINISH, K_ACCEL_APPLY_DIV, K_FORCES, K_DENSITY_ADV,
    K_ACCEL_APPLY_FINAL, K_STEP_FINISH, K_COUNT
};
const char* const kKernelNames[K_COUNT] = {
    "begin_step", "hash_keys", "hash_scan", "hash_scatter", "reorder", "density_alpha",
    "begin_solve", "pressure_accel", "jacobi", "solve_finish", "accel_apply_div", "forces",
    "density_adv", "accel_apply_final", "step_finish"};

struct ProfTotal {
    uint64_t launches = 0;
    double total_ms = 0.0, min_ms = 1e300, max_ms = 0.0;
};
struct ProfRec {
    int kid, solver, iter;
};

}  // namespace

struct dfsph_ctx {
    dfsph_params params{};
    Derived d{};
    Grid grid{};       // grid of the step
    Grid ref_grid{};   // reference Morton grid (exports)
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint32_t n = 0, nb = 0;
    bool stepped = false;
    Buffers b{};
    DeviceStatus* h_status = nullptr;   // pinned
    int predicted[2] = {4, 4};          // iterations launched before the first convergence check
    // reference-hash export scratch (lazy)
    uint32_t *ref_keys = nullptr, *ref_count = nullptr, *ref_start = nullptr, *ref_perm = nullptr,
             *ref_scan = nullptr;
    void* export_tmp = nullptr;
    size_t export_tmp_bytes = 0;
    std::vector<void*> allocs;
    uint32_t launches = 0;
    int iters_done[2] = {0, 0};
    // profiler
    bool prof_on = false;
    ProfTotal prof_tot[K_COUNT];
    std::vector<hipEvent_t> prof_ev;   // pairs
    std::vector<ProfRec> prof_rec;
    size_t prof_used = 0;

    int prof_begin(int kid, int solver, int iter) {
        if (!prof_on) return DFSPH_OK;
        if (2 * (prof_used + 1) > prof_ev.size()) {
            for (int q = 0; q < 2; ++q) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return fail(DFSPH_E_DEVICE, "hipEventCreate failed");
                prof_ev.push_back(e);
            }
            prof_rec.emplace_back();
        }
        prof_rec[prof_used] = ProfRec{kid, solver, iter};
        if (hipEventRecord(prof_ev[2 * prof_used], stream) != hipSuccess)
            return fail(DFSPH_E_DEVICE, "hipEventRecord failed");
        return DFSPH_OK;
    }
    int prof_end() {
        if (!prof_on) return DFSPH_OK;
        if (hipEventRecord(prof_ev[2 * prof_used + 1], stream) != hipSuccess)
            return fail(DFSPH_E_DEVICE, "hipEventRecord failed");
        ++prof_used;
        return DFSPH_OK;
    }
    // After the stream is idle: accumulate the recorded pairs, skipping solver
    // iterations beyond the converged count (their kernels exit at once).
    int prof_collect() {
        for (size_t q = 0; q < prof_used; ++q) {
            const ProfRec& r = prof_rec[q];
            if (r.iter >= 0 && r.iter >= iters_done[r.solver]) continue;
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, prof_ev[2 * q], prof_ev[2 * q + 1]) != hipSuccess)
                return fail(DFSPH_E_DEVICE, "hipEventElapsedTime failed");
            ProfTotal& t = prof_tot[r.kid];
            ++t.launches;
            t.total_ms += ms;
            t.min_ms = std::min(t.min_ms, double(ms));
            t.max_ms = std::max(t.max_ms, double(ms));
        }
        prof_used = 0;
        return DFSPH_OK;
    }

    int alloc(void** p, size_t bytes) {
        if (hipMalloc(p, bytes ? bytes : 16) != hipSuccess)
            return fail(DFSPH_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        allocs.push_back(*p);
        return DFSPH_OK;
    }
    template <class T>
    int alloc_n(T** p, size_t count) {
        return alloc(reinterpret_cast<void**>(p), count * sizeof(T));
    }
    void release(void* p) {
        if (!p) return;
        auto it = std::find(allocs.begin(), allocs.end(), p);
        if (it != allocs.end()) { (void)hipFree(p); allocs.erase(it); }
    }
    ~dfsph_ctx() {
        for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
        for (void* p : allocs) (void)hipFree(p);
        if (h_status) (void)hipHostFree(h_status);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    SimConst sim_const(float bscale) const {
        SimConst k{};
        k.h = d.h;
        k.inv_h = 1.0f / d.h;
        k.mass = d.mass;
        k.knorm = d.knorm;
        k.gnorm = d.gnorm;
        k.gh = d.gnorm / d.h;
        k.m6g = -6.0f * d.gnorm;
        k.rho0 = params.density0;
        k.relax = d.relax;
        k.visc = params.viscosity;
        k.coh = params.cohesion;
        k.vort = params.vorticity;
        for (int a = 0; a < 3; ++a) {
            k.g[a] = params.gravity[a];
            k.dmin[a] = params.domain_min[a];
            k.dmax[a] = params.domain_max[a];
        }
        k.r2_q1 = d.r2_q1;
        k.r2_eps_gt = d.r2_eps_gt;
        k.r2_eps_lt = d.r2_eps_lt;
        k.bscale = bscale;
        k.n = n;
        k.nb = nb;
        return k;
    }
    SweepArgs args(float bscale) const {
        SweepArgs a{};
        a.k = sim_const(bscale);
        a.grid = grid;
        a.b = b;
        a.block = params.block_size ? params.block_size : 256u;
        a.xcd_remap = (params.flags & DFSPH_FLAG_NO_XCD_REMAP) ? 0 : 1;
        return a;
    }
};

namespace {

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

int validate_params(const dfsph_params* p) {
    if (!p) return fail(DFSPH_E_INVALID, "params is NULL");
    if (p->abi_version != DFSPH_ABI_VERSION)
        return fail(DFSPH_E_INVALID, "abi_version mismatch: expected " +
                                         std::to_string(DFSPH_ABI_VERSION));
    if (p->capacity == 0) return fail(DFSPH_E_INVALID, "capacity must be > 0");
    if (!(p->particle_radius > 0.0f) || !std::isfinite(p->particle_radius))
        return fail(DFSPH_E_INVALID, "particle_radius must be finite and > 0");
    if (!(p->density0 > 0.0f)) return fail(DFSPH_E_INVALID, "density0 must be > 0");
    for (int a = 0; a < 3; ++a)
        if (!(p->domain_max[a] > p->domain_min[a]))
            return fail(DFSPH_E_INVALID, "domain_max must exceed domain_min on every axis");
    if (p->grid != DFSPH_GRID_REFERENCE && p->grid != DFSPH_GRID_FINE)
        return fail(DFSPH_E_INVALID, "grid must be DFSPH_GRID_REFERENCE or DFSPH_GRID_FINE");
    if (p->block_size && (p->block_size % 64 || p->block_size > 256))
        return fail(DFSPH_E_INVALID, "block_size must be a multiple of 64 up to 256");
    return DFSPH_OK;
}

int validate_policy(const dfsph_solver_policy* pol) {
    if (!pol) return fail(DFSPH_E_INVALID, "policy is NULL");
    if (pol->mode != DFSPH_POLICY_REFERENCE && pol->mode != DFSPH_POLICY_CONVERGED)
        return fail(DFSPH_E_INVALID, "policy.mode must be REFERENCE or CONVERGED");
    if (pol->warm_start < DFSPH_WARM_ZERO || pol->warm_start > DFSPH_WARM_MLP)
        return fail(DFSPH_E_INVALID, "policy.warm_start out of range");
    if (pol->max_iter_div < 1 || pol->max_iter_density < 1 || pol->min_iter_div < 0 ||
        pol->min_iter_density < 0 || pol->max_iter_div > 100000 || pol->max_iter_density > 100000)
        return fail(DFSPH_E_INVALID, "policy iteration limits out of range");
    if (!(pol->dt_max > 0.0f) || !(pol->cfl > 0.0f))
        return fail(DFSPH_E_INVALID, "policy.cfl and policy.dt_max must be > 0");
    return DFSPH_OK;
}

}  // namespace

extern "C" {

const char* dfsph_last_error(void) { return g_last_error.c_str(); }

int dfsph_default_policy(int32_t mode, dfsph_solver_policy* out) {
    if (!out) return fail(DFSPH_E_INVALID, "out is NULL");
    if (mode != DFSPH_POLICY_REFERENCE && mode != DFSPH_POLICY_CONVERGED)
        return fail(DFSPH_E_INVALID, "unknown policy mode");
    dfsph_solver_policy p{};
    p.mode = mode;
    p.warm_start = DFSPH_WARM_ZERO;
    p.div_enabled = 1;                 // main.cpp:120
    p.min_iter_div = 1;                // main.cpp:261
    p.max_iter_div = 100;              // main.cpp:119
    p.min_iter_density = 2;            // main.cpp:115
    p.max_iter_density = 100;          // main.cpp:116
    p.tol_div = 1e-3f;
    p.tol_density = 1e-3f;
    p.cfl = 0.4f;
    p.dt_max = 1e-3f;
    p.boundary_psi = -1;
    p.strict = 0;
    *out = p;
    return DFSPH_OK;
}

int dfsph_derive(const dfsph_params* params, dfsph_derived* out) {
    if (int rc = validate_params(params)) return rc;
    if (!out) return fail(DFSPH_E_INVALID, "out is NULL");
    *out = derive(*params);
    return DFSPH_OK;
}

int dfsph_boundary_samples(const float dmin[3], const float dmax[3], float radius, uint32_t cap,
                           float* out, uint32_t cap_out, uint32_t* count) {
    if (!dmin || !dmax || !count) return fail(DFSPH_E_INVALID, "NULL argument");
    std::vector<float> s = boundary_samples(dmin, dmax, radius, cap);
    *count = uint32_t(s.size() / 3);
    if (out) std::memcpy(out, s.data(), sizeof(float) * 3 * std::min<size_t>(*count, cap_out));
    return DFSPH_OK;
}

int dfsph_scene_block(uint32_t nx, uint32_t ny, uint32_t nz, float radius, const float origin[3],
                      uint64_t seed, float jitter, uint32_t id0, float* out) {
    if (!origin || !out) return fail(DFSPH_E_INVALID, "NULL argument");
    if (!(radius > 0.0f)) return fail(DFSPH_E_INVALID, "radius must be > 0");
    scene_block(nx, ny, nz, radius, origin, seed, jitter, id0, out);
    return DFSPH_OK;
}

int dfsph_create(const dfsph_params* params, dfsph_ctx** out) {
    if (!out) return fail(DFSPH_E_INVALID, "out is NULL");
    *out = nullptr;
    if (int rc = validate_params(params)) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(DFSPH_E_DEVICE, "no HIP device available");
    if (params->device < 0 || params->device >= ndev)
        return fail(DFSPH_E_DEVICE, "device ordinal out of range");
    hipDeviceProp_t prop{};
    DFSPH_HIP(hipGetDeviceProperties(&prop, params->device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(DFSPH_E_DEVICE, std::string("device is ") + prop.gcnArchName +
                                        "; this library is built for gfx950 only");
    DFSPH_HIP(hipSetDevice(params->device));

    auto* c = new (std::nothrow) dfsph_ctx();
    if (!c) return fail(DFSPH_E_NOMEM, "context allocation failed");
    c->params = *params;
    c->device = params->device;
    c->d = derive(*params);
    c->grid = make_grid(*params, c->d, params->grid);
    c->ref_grid = make_grid(*params, c->d, DFSPH_GRID_REFERENCE);
    auto bail = [&](int rc) { delete c; return rc; };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return bail(fail(DFSPH_E_DEVICE, "stream/event creation failed"));
    if (hipHostMalloc(reinterpret_cast<void**>(&c->h_status), sizeof(DeviceStatus),
                      hipHostMallocDefault) != hipSuccess)
        return bail(fail(DFSPH_E_NOMEM, "pinned status allocation failed"));

    const size_t N = params->capacity;
    const size_t cells = c->grid.cells;
    Buffers& b = c->b;
    int rc = DFSPH_OK;
    auto A4 = [&](void** p) { if (!rc) rc = c->alloc(p, N * 16); };
    auto AF = [&](float** p) { if (!rc) rc = c->alloc_n(p, N); };
    auto AU = [&](uint32_t** p, size_t cnt) { if (!rc) rc = c->alloc_n(p, cnt); };
    A4(&b.pos4); A4(&b.vel4); A4(&b.vel4b); A4(&b.acc4); A4(&b.gf4); A4(&b.gb4);
    A4(&b.x_stage); A4(&b.v_stage);
    AF(&b.rho); AF(&b.alpha); AF(&b.aor2); AF(&b.rho_dot); AF(&b.rho_adv); AF(&b.p); AF(&b.pr);
    AF(&b.p_div); AF(&b.p_den); AF(&b.pdiv_prev); AF(&b.pden_prev); AF(&b.pdiv_stage);
    AF(&b.pden_stage);
    AU(&b.nbr, N); AU(&b.nbrb, N); AU(&b.ids, N); AU(&b.keys, N); AU(&b.id_stage, N);
    AU(&b.perm, N); AU(&b.cell_count, cells); AU(&b.cell_start, cells + 1);
    AU(&b.scan_blocks, scan_scratch_elems(uint32_t(cells)));
    if (params->grid == DFSPH_GRID_FINE && !(params->flags & DFSPH_FLAG_NO_NEIGHBOR_MASKS)) {
        AU(&b.nmask, 9 * N);
        AU(&b.novf, N);
        b.mask_stride = uint32_t(N);
    }
    if (!rc) rc = c->alloc_n(&b.partials, blocks_for(N, 64) + 64);
    if (!rc) rc = c->alloc_n(&b.step_part, 2 * (blocks_for(N, 64) + 64));
    if (!rc) rc = c->alloc(reinterpret_cast<void**>(&b.status), sizeof(DeviceStatus));
    if (rc) return bail(rc);
    if (hipMemsetAsync(b.cell_count, 0, cells * 4, c->stream) != hipSuccess ||
        hipMemsetAsync(b.status, 0, sizeof(DeviceStatus), c->stream) != hipSuccess ||
        hipMemsetAsync(b.pdiv_stage, 0, N * 4, c->stream) != hipSuccess ||
        hipMemsetAsync(b.pden_stage, 0, N * 4, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return bail(fail(DFSPH_E_DEVICE, "initial clears failed"));
    *out = c;
    return DFSPH_OK;
}

int dfsph_destroy(dfsph_ctx* ctx) {
    if (!ctx) return DFSPH_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return DFSPH_OK;
}

int dfsph_init_scene(dfsph_ctx* c, const dfsph_scene* sc) {
    if (!c || !sc) return fail(DFSPH_E_INVALID, "NULL argument");
    if (sc->count > c->params.capacity)
        return fail(DFSPH_E_CAPACITY, "scene count " + std::to_string(sc->count) +
                                          " exceeds capacity " + std::to_string(c->params.capacity));
    if (sc->count && !sc->positions) return fail(DFSPH_E_INVALID, "positions is NULL");
    const uint32_t n = sc->count;
    DFSPH_HIP(hipSetDevice(c->device));
    std::vector<float> x4(size_t(n) * 4, 0.0f), v4(size_t(n) * 4, 0.0f);
    std::vector<uint32_t> ids(n);
    float vmax2 = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            x4[4 * i + a] = sc->positions[3 * i + a];
            v4[4 * i + a] = sc->velocities ? sc->velocities[3 * i + a] : 0.0f;
        }
        const float vx = v4[4 * i], vy = v4[4 * i + 1], vz = v4[4 * i + 2];
        const float s2 = (vx * vx + vy * vy) + vz * vz;
        if (s2 > vmax2) vmax2 = s2;
        ids[i] = sc->ids ? sc->ids[i] : i;
    }
    if (sc->ids) {   // must be a permutation of 0..n-1 (export scatters by it)
        std::vector<uint8_t> seen(n, 0);
        for (uint32_t i = 0; i < n; ++i) {
            if (ids[i] >= n || seen[ids[i]])
                return fail(DFSPH_E_INVALID, "ids must be a permutation of 0..count-1");
            seen[ids[i]] = 1;
        }
    }
    Buffers& b = c->b;
    c->n = n;
    c->stepped = false;
    DFSPH_HIP(hipMemcpyAsync(b.x_stage, x4.data(), x4.size() * 4, hipMemcpyHostToDevice, c->stream));
    DFSPH_HIP(hipMemcpyAsync(b.v_stage, v4.data(), v4.size() * 4, hipMemcpyHostToDevice, c->stream));
    DFSPH_HIP(hipMemcpyAsync(b.id_stage, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, c->stream));
    DFSPH_HIP(hipMemsetAsync(b.pdiv_stage, 0, size_t(c->params.capacity) * 4, c->stream));
    DFSPH_HIP(hipMemsetAsync(b.pden_stage, 0, size_t(c->params.capacity) * 4, c->stream));
    DeviceStatus st{};
    std::memcpy(&st.vmax2_bits, &vmax2, 4);
    DFSPH_HIP(hipMemcpyAsync(b.status, &st, sizeof(st), hipMemcpyHostToDevice, c->stream));

    // Boundary: samples on the AABB, sorted into the step grid, volumes on device.
    c->release(b.bpos4); c->release(b.bvol); c->release(b.bcell_start); c->release(b.bids);
    b.bpos4 = nullptr; b.bvol = nullptr; b.bcell_start = nullptr; b.bids = nullptr;
    c->nb = 0;
    std::vector<float> bs;
    if (sc->boundary) bs = boundary_samples(c->params.domain_min, c->params.domain_max,
                                            c->params.particle_radius, sc->boundary_cap);
    const uint32_t nb = uint32_t(bs.size() / 3);
    int rc = DFSPH_OK;
    rc = c->alloc_n(&b.bcell_start, size_t(c->grid.cells) + 1);
    if (rc) return rc;
    if (nb) {
        void* bstage = nullptr;
        uint32_t *bkeys = nullptr, *bperm = nullptr;
        if ((rc = c->alloc(&b.bpos4, size_t(nb) * 16)) || (rc = c->alloc_n(&b.bvol, nb)) ||
            (rc = c->alloc_n(&b.bids, nb)) || (rc = c->alloc(&bstage, size_t(nb) * 16)) ||
            (rc = c->alloc_n(&bkeys, nb)) || (rc = c->alloc_n(&bperm, nb)))
            return rc;
        std::vector<float> b4(size_t(nb) * 4, 0.0f);
        for (uint32_t i = 0; i < nb; ++i)
            for (int a = 0; a < 3; ++a) b4[4 * i + a] = bs[3 * i + a];
        DFSPH_HIP(hipMemcpyAsync(bstage, b4.data(), b4.size() * 4, hipMemcpyHostToDevice, c->stream));
        launch_keys(c->grid, bstage, nb, bkeys, b.cell_count, c->stream);
        launch_scan(c->grid.cells, b.cell_count, b.bcell_start, b.scan_blocks, c->stream);
        launch_scatter(nb, bkeys, b.bcell_start, b.cell_count, bperm, nullptr, c->grid.cells, c->stream);
        c->nb = nb;
        SweepArgs a = c->args(1.0f);
        launch_reorder_boundary(a, bstage, bperm, c->stream);
        launch_boundary_volume(a, c->stream);
        DFSPH_HIP(hipGetLastError());
        DFSPH_HIP(hipStreamSynchronize(c->stream));
        c->release(bstage); c->release(bkeys); c->release(bperm);
    } else {
        DFSPH_HIP(hipMemsetAsync(b.bcell_start, 0, (size_t(c->grid.cells) + 1) * 4, c->stream));
    }
    DFSPH_HIP(hipStreamSynchronize(c->stream));
    return DFSPH_OK;
}

int dfsph_set_warm_start(dfsph_ctx* c, const float* p_div, const float* p_density) {
    if (!c) return fail(DFSPH_E_INVALID, "ctx is NULL");
    DFSPH_HIP(hipSetDevice(c->device));
    Buffers& b = c->b;
    const size_t N = c->params.capacity;
    int rc = DFSPH_OK;
    if (!b.warm_ext_div && (rc = c->alloc_n(&b.warm_ext_div, N))) return rc;
    if (!b.warm_ext_den && (rc = c->alloc_n(&b.warm_ext_den, N))) return rc;
    if (p_div) DFSPH_HIP(hipMemcpyAsync(b.warm_ext_div, p_div, size_t(c->n) * 4, hipMemcpyHostToDevice, c->stream));
    else DFSPH_HIP(hipMemsetAsync(b.warm_ext_div, 0, N * 4, c->stream));
    if (p_density) DFSPH_HIP(hipMemcpyAsync(b.warm_ext_den, p_density, size_t(c->n) * 4, hipMemcpyHostToDevice, c->stream));
    else DFSPH_HIP(hipMemsetAsync(b.warm_ext_den, 0, N * 4, c->stream));
    DFSPH_HIP(hipStreamSynchronize(c->stream));
    return DFSPH_OK;
}

int dfsph_step(dfsph_ctx* c, float dt, int32_t substeps, const dfsph_solver_policy* pol,
               dfsph_step_stats* stats) {
    if (!c) return fail(DFSPH_E_INVALID, "ctx is NULL");
    if (int rc = validate_policy(pol)) return rc;
    if (substeps < 1) return fail(DFSPH_E_INVALID, "substeps must be >= 1");
    if (!(dt >= 0.0f) || !std::isfinite(dt)) return fail(DFSPH_E_INVALID, "dt must be finite and >= 0");
    DFSPH_HIP(hipSetDevice(c->device));
    if (c->n == 0) {   // nothing to advance
        if (stats) *stats = dfsph_step_stats{};
        return DFSPH_OK;
    }
    const bool reference = pol->mode == DFSPH_POLICY_REFERENCE;
    const int psi = pol->boundary_psi < 0 ? (reference ? 0 : 1) : pol->boundary_psi;
    const float bscale = psi ? c->params.density0 : 1.0f;
    int warm = reference ? DFSPH_WARM_ZERO : pol->warm_start;
    int warm_fallback = 0;
    if (warm == DFSPH_WARM_EXTERNAL && !c->b.warm_ext_div) warm = DFSPH_WARM_ZERO;
    if (warm == DFSPH_WARM_MLP && !c->b.warm_mlp) { warm = DFSPH_WARM_ZERO; warm_fallback = 1; }
    const float cfl_2r = pol->cfl * (2.0f * c->params.particle_radius);
    uint32_t syncs = 0;
    c->launches = 0;
    int it_tot[2] = {0, 0}, conv_all[2] = {1, 1};
    double err_max[2] = {0.0, 0.0};
    DFSPH_HIP(hipEventRecord(c->ev0, c->stream));
    for (int sub = 0; sub < substeps; ++sub) {
        Buffers& b = c->b;
        SweepArgs a = c->args(bscale);
        // Launch one kernel (or kernel group) under the profiler.
        auto run = [&](int kid, int solver, int iter, auto&& launch) -> int {
            if (int rc = c->prof_begin(kid, solver, iter)) return rc;
            launch();
            ++c->launches;
            return c->prof_end();
        };
#define DFSPH_RUN(kid, solver, iter, ...)                                   \
        do {                                                                \
            if (int rc_ = run(kid, solver, iter, [&] { __VA_ARGS__; })) return rc_; \
        } while (0)
        DFSPH_RUN(K_BEGIN_STEP, -1, -1, launch_begin_step(a, dt, cfl_2r, pol->dt_max, c->stream));
        DFSPH_RUN(K_KEYS, -1, -1, launch_keys(c->grid, b.x_stage, c->n, b.keys, b.cell_count, c->stream));
        DFSPH_RUN(K_SCAN, -1, -1, launch_scan(c->grid.cells, b.cell_count, b.cell_start, b.scan_blocks, c->stream));
        DFSPH_RUN(K_SCATTER, -1, -1, launch_scatter(c->n, b.keys, b.cell_start, b.cell_count, b.perm,
                                                   b.id_stage, c->grid.cells, c->stream));
        DFSPH_RUN(K_REORDER, -1, -1, launch_reorder(a, c->stream));
        DFSPH_RUN(K_DENSITY_ALPHA, -1, -1, launch_density_alpha(a, c->stream));

        auto run_solve = [&](int s) -> int {
            const int min_it = s == 0 ? pol->min_iter_div : pol->min_iter_density;
            const int max_it = s == 0 ? pol->max_iter_div : pol->max_iter_density;
            const float tol = s == 0 ? pol->tol_div : pol->tol_density;
            DFSPH_RUN(K_BEGIN_SOLVE, -1, -1, launch_begin_solve(a, s, warm, c->stream));
            auto iterate = [&](int it, int ref, int mn, int mx, float tl) -> int {
                DFSPH_RUN(K_PRESSURE_ACCEL, s, it, launch_pressure_accel(a, s, it, c->stream));
                DFSPH_RUN(K_JACOBI, s, it, launch_jacobi(a, s, it, ref, mn, mx, tl, c->stream));
                DFSPH_RUN(K_SOLVE_FINISH, s, it, launch_solve_finish(a, s, it, ref, mn, mx, tl, c->stream));
                return DFSPH_OK;
            };
            if (reference) {
                const int fixed = std::max(min_it, 1);   // main.cpp:2774, :2830
                for (int it = 0; it < fixed; ++it)
                    if (int rc = iterate(it, 1, fixed, fixed, 0.0f)) return rc;
                c->iters_done[s] = fixed;
            } else {
                int it = 0;
                int chunk = std::max(std::max(min_it, 1), c->predicted[s]);
                while (it < max_it) {
                    const int cnt = std::min(chunk, max_it - it);
                    for (int q = 0; q < cnt; ++q, ++it)
                        if (int rc = iterate(it, 0, min_it, max_it, tol)) return rc;
                    DFSPH_HIP(hipMemcpyAsync(&c->h_status->solve[s], &b.status->solve[s],
                                             sizeof(SolveStatus), hipMemcpyDeviceToHost, c->stream));
                    DFSPH_HIP(hipStreamSynchronize(c->stream));
                    ++syncs;
                    if (c->h_status->solve[s].done) break;
                    chunk = std::max(1, std::min(8, it / 4 + 1));
                }
                c->iters_done[s] = c->h_status->solve[s].iters;
                c->predicted[s] = std::max(1, c->h_status->solve[s].iters);
            }
            DFSPH_RUN(s == 0 ? K_ACCEL_APPLY_DIV : K_ACCEL_APPLY_FINAL, -1, -1,
                      launch_accel_apply(a, s, c->stream));
            return DFSPH_OK;
        };
        c->iters_done[0] = c->iters_done[1] = 0;
        if (pol->div_enabled) {
            if (int rc = run_solve(0)) return rc;
        } else {
            DFSPH_HIP(hipMemsetAsync(b.p_div, 0, size_t(c->n) * 4, c->stream));
        }
        DFSPH_RUN(K_FORCES, -1, -1, launch_forces(a, c->stream));
        DFSPH_RUN(K_DENSITY_ADV, -1, -1, launch_density_adv(a, c->stream));
        if (int rc = run_solve(1)) return rc;
        DFSPH_RUN(K_STEP_FINISH, -1, -1, launch_step_finish(a, c->stream));
#undef DFSPH_RUN
        DFSPH_HIP(hipGetLastError());
        // The step's sorted order becomes the stage order of the next step.
        std::swap(b.ids, b.id_stage);
        std::swap(b.p_div, b.pdiv_stage);
        std::swap(b.p_den, b.pden_stage);
        c->stepped = true;
        // Per-substep statistics (one status read per substep).
        DFSPH_HIP(hipMemcpyAsync(c->h_status, b.status, sizeof(DeviceStatus), hipMemcpyDeviceToHost,
                                 c->stream));
        DFSPH_HIP(hipStreamSynchronize(c->stream));
        ++syncs;
        if (int rc = c->prof_collect()) return rc;
        for (int s = 0; s < 2; ++s) {
            if (s == 0 && !pol->div_enabled) continue;
            it_tot[s] += c->h_status->solve[s].iters;
            conv_all[s] &= c->h_status->solve[s].converged;
            err_max[s] = std::max(err_max[s], c->h_status->solve[s].err);
        }
    }
    DFSPH_HIP(hipEventRecord(c->ev1, c->stream));
    DFSPH_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    DFSPH_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    const DeviceStatus& h = *c->h_status;
    if (stats) {
        dfsph_step_stats s{};
        s.dt = h.dt;
        s.iters_div = it_tot[0];
        s.iters_density = it_tot[1];
        s.converged_div = conv_all[0];
        s.converged_density = conv_all[1];
        s.err_div = err_max[0];
        s.err_density = err_max[1];
        float v2;
        std::memcpy(&v2, &h.vmax2_start, 4);
        s.max_speed = std::sqrt(double(v2));
        s.nonfinite = h.nonfinite;
        s.gpu_ms = ms;
        s.launches = c->launches;
        s.host_syncs = syncs;
        s.warm_fallback = warm_fallback;
        *stats = s;
    }
    if (h.nonfinite) return fail(DFSPH_E_NONFINITE, std::to_string(h.nonfinite) +
                                                        " non-finite position/velocity components");
    if (pol->strict && (!conv_all[0] || !conv_all[1]))
        return fail(DFSPH_E_NOT_CONVERGED, "a pressure solve stopped at max_iter above tolerance");
    return DFSPH_OK;
}

int dfsph_profile_enable(dfsph_ctx* c, int32_t enable) {
    if (!c) return fail(DFSPH_E_INVALID, "ctx is NULL");
    c->prof_on = enable != 0;
    for (auto& t : c->prof_tot) t = ProfTotal{};
    c->prof_used = 0;
    return DFSPH_OK;
}

int dfsph_profile_read(dfsph_ctx* c, dfsph_kernel_time* out, uint32_t cap, uint32_t* count) {
    if (!c || !count) return fail(DFSPH_E_INVALID, "NULL argument");
    uint32_t k = 0;
    for (int id = 0; id < K_COUNT; ++id) {
        const ProfTotal& t = c->prof_tot[id];
        if (!t.launches) continue;
        if (out && k < cap) {
            dfsph_kernel_time& o = out[k];
            std::memset(&o, 0, sizeof(o));
            std::snprintf(o.name, sizeof(o.name), "%s", kKernelNames[id]);
            o.launches = t.launches;
            o.total_ms = t.total_ms;
            o.min_ms = t.min_ms;
            o.max_ms = t.max_ms;
        }
        ++k;
    }
    *count = k;
    return DFSPH_OK;
}

int dfsph_field_size(dfsph_ctx* c, int32_t field, size_t* bytes) {
    if (!c || !bytes) return fail(DFSPH_E_INVALID, "NULL argument");
    const size_t n = c->n;
    switch (field) {
        case DFSPH_FIELD_POSITION: case DFSPH_FIELD_VELOCITY: case DFSPH_FIELD_PRESSURE_ACCEL:
            *bytes = n * 12; return DFSPH_OK;
        case DFSPH_FIELD_DENSITY: case DFSPH_FIELD_ALPHA_OVER_RHO2: case DFSPH_FIELD_DENSITY_ADV:
        case DFSPH_FIELD_DENSITY_CHANGE: case DFSPH_FIELD_PRESSURE_DIV: case DFSPH_FIELD_PRESSURE:
        case DFSPH_FIELD_ALPHA: case DFSPH_FIELD_NEIGHBOR_COUNT:
        case DFSPH_FIELD_BOUNDARY_NEIGHBOR_COUNT: case DFSPH_FIELD_MORTON_KEY:
        case DFSPH_FIELD_SORTED_INDEX: case DFSPH_FIELD_ID:
            *bytes = n * 4; return DFSPH_OK;
        case DFSPH_FIELD_CELL_STARTS:
            *bytes = (size_t(c->d.total_cells) + 1) * 4; return DFSPH_OK;
        case DFSPH_FIELD_BOUNDARY_POSITION: *bytes = size_t(c->nb) * 12; return DFSPH_OK;
        case DFSPH_FIELD_BOUNDARY_VOLUME: *bytes = size_t(c->nb) * 4; return DFSPH_OK;
        default: return fail(DFSPH_E_INVALID, "unknown field " + std::to_string(field));
    }
}

namespace {
int ensure_tmp(dfsph_ctx* c, size_t bytes) {
    if (c->export_tmp_bytes >= bytes) return DFSPH_OK;
    c->release(c->export_tmp);
    c->export_tmp = nullptr;
    c->export_tmp_bytes = 0;
    if (int rc = c->alloc(&c->export_tmp, bytes)) return rc;
    c->export_tmp_bytes = bytes;
    return DFSPH_OK;
}

// Reference Morton hash of the step's start positions (pos4, ids in id_stage).
int reference_hash(dfsph_ctx* c) {
    const size_t cells = c->ref_grid.cells;
    int rc = DFSPH_OK;
    if (!c->ref_start) {
        if ((rc = c->alloc_n(&c->ref_keys, c->params.capacity)) ||
            (rc = c->alloc_n(&c->ref_perm, c->params.capacity)) ||
            (rc = c->alloc_n(&c->ref_count, cells)) || (rc = c->alloc_n(&c->ref_start, cells + 1)) ||
            (rc = c->alloc_n(&c->ref_scan, scan_scratch_elems(uint32_t(cells)))))
            return rc;
        DFSPH_HIP(hipMemsetAsync(c->ref_count, 0, cells * 4, c->stream));
    }
    launch_keys(c->ref_grid, c->b.pos4, c->n, c->ref_keys, c->ref_count, c->stream);
    launch_scan(uint32_t(cells), c->ref_count, c->ref_start, c->ref_scan, c->stream);
    launch_scatter(c->n, c->ref_keys, c->ref_start, c->ref_count, c->ref_perm, c->b.id_stage,
                   uint32_t(cells), c->stream);
    DFSPH_HIP(hipGetLastError());
    return DFSPH_OK;
}
}  // namespace

int dfsph_export(dfsph_ctx* c, int32_t field, void* dst, size_t bytes) {
    size_t need = 0;
    if (int rc = dfsph_field_size(c, field, &need)) return rc;
    if (!dst && need) return fail(DFSPH_E_INVALID, "dst is NULL");
    if (bytes != need)
        return fail(DFSPH_E_INVALID, "bytes must be " + std::to_string(need) + " for field " +
                                         std::to_string(field));
    DFSPH_HIP(hipSetDevice(c->device));
    const bool step_field = field != DFSPH_FIELD_POSITION && field != DFSPH_FIELD_VELOCITY &&
                            field != DFSPH_FIELD_BOUNDARY_POSITION &&
                            field != DFSPH_FIELD_BOUNDARY_VOLUME && field != DFSPH_FIELD_ID;
    if (step_field && !c->stepped)
        return fail(DFSPH_E_INVALID, "field " + std::to_string(field) + " exists only after a step");
    if (!need) return DFSPH_OK;
    if (int rc = ensure_tmp(c, need)) return rc;
    Buffers& b = c->b;
    const uint32_t n = c->n;
    float* tf = static_cast<float*>(c->export_tmp);
    uint32_t* tu = static_cast<uint32_t*>(c->export_tmp);
    const uint32_t* ids = b.id_stage;   // order of the last step (or of the scene before any step)
    switch (field) {
        case DFSPH_FIELD_POSITION: launch_to_original_vec(b.x_stage, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_VELOCITY: launch_to_original_vec(b.v_stage, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_DENSITY: launch_to_original_f32(b.rho, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_ALPHA_OVER_RHO2: launch_to_original_f32(b.aor2, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_DENSITY_ADV: launch_to_original_f32(b.rho_adv, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_DENSITY_CHANGE: launch_to_original_f32(b.rho_dot, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_PRESSURE_DIV: launch_to_original_f32(b.pdiv_stage, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_PRESSURE: launch_to_original_f32(b.pden_stage, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_PRESSURE_ACCEL: launch_to_original_vec(b.acc4, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_ALPHA: launch_to_original_f32(b.alpha, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_NEIGHBOR_COUNT: launch_to_original_u32(b.nbr, ids, n, tu, c->stream); break;
        case DFSPH_FIELD_BOUNDARY_NEIGHBOR_COUNT: launch_to_original_u32(b.nbrb, ids, n, tu, c->stream); break;
        case DFSPH_FIELD_MORTON_KEY:
            if (int rc = reference_hash(c)) return rc;
            launch_to_original_u32(c->ref_keys, ids, n, tu, c->stream);
            break;
        case DFSPH_FIELD_SORTED_INDEX:
            if (int rc = reference_hash(c)) return rc;
            launch_permute_u32(ids, c->ref_perm, n, tu, c->stream);
            break;
        case DFSPH_FIELD_CELL_STARTS:
            if (int rc = reference_hash(c)) return rc;
            DFSPH_HIP(hipMemcpyAsync(tu, c->ref_start, need, hipMemcpyDeviceToDevice, c->stream));
            break;
        case DFSPH_FIELD_BOUNDARY_POSITION: launch_to_original_vec(b.bpos4, b.bids, c->nb, tf, c->stream); break;
        case DFSPH_FIELD_BOUNDARY_VOLUME: launch_to_original_w(b.bpos4, b.bids, c->nb, tf, c->stream); break;
        case DFSPH_FIELD_ID:
            DFSPH_HIP(hipMemcpyAsync(tu, ids, need, hipMemcpyDeviceToDevice, c->stream));
            break;
        default: return fail(DFSPH_E_INVALID, "unknown field");
    }
    DFSPH_HIP(hipGetLastError());
    DFSPH_HIP(hipMemcpyAsync(dst, c->export_tmp, need, hipMemcpyDeviceToHost, c->stream));
    DFSPH_HIP(hipStreamSynchronize(c->stream));
    return DFSPH_OK;
}

}  // extern "C"
