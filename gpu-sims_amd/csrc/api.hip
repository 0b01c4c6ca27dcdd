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
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "dfsph_internal.h"
#include "launch.h"
#include "mlp.h"

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
    K_ACCEL_APPLY_FINAL, K_STEP_FINISH, K_MLP_WARM, K_RECORD, K_COUNT
};
const char* const kKernelNames[K_COUNT] = {
    "begin_step", "hash_keys", "hash_scan", "hash_scatter", "reorder", "density_alpha",
    "begin_solve", "pressure_accel", "jacobi", "solve_finish", "accel_apply_div", "forces",
    "density_adv", "accel_apply_final", "step_finish", "mlp_warm", "record"};

struct ProfTotal {
    uint64_t launches = 0;
    double total_ms = 0.0, min_ms = 1e300, max_ms = 0.0;
};
struct ProfRec {
    int kid, solver, iter;
    bool skip;   // a solver iteration past the solve's converged count (exits at once)
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
    // The next step's cell build (keys, scan, scatter + order) is already on
    // the stream: enqueued at the end of the last step so it runs while the
    // host returns; cleared by anything that changes the particles.
    bool prefix_ready = false;
    // The velocity update counted the next build's keys into cell_count but
    // the build was not enqueued (an error between): cleared before the next.
    bool counts_dirty = false;
    Buffers b{};
    DeviceStatus* h_status = nullptr;   // pinned
    SolveMirror* h_mirror = nullptr;    // pinned, coherent: decisions of the fused solve path
    StepMirror* h_step = nullptr;       // pinned, coherent: each substep's final status
    int predicted[2] = {4, 4};          // iterations launched before the first convergence check
    // reference-hash export scratch (lazy)
    uint32_t *ref_keys = nullptr, *ref_count = nullptr, *ref_start = nullptr, *ref_perm = nullptr,
             *ref_scan = nullptr;
    void* export_tmp = nullptr;
    size_t export_tmp_bytes = 0;
    // render passes: per-particle arrays for render_n particles, images for render_px pixels
    RenderBuffers rb{};
    uint32_t render_n = 0;
    size_t render_px = 0;
    std::vector<void*> allocs;
    uint32_t launches = 0;
    int iters_done[2] = {0, 0};
    // learned warm start
    MlpDev mlp{};
    bool mlp_ready = false;
    uint64_t step_count = 0;
    std::vector<uint32_t> flags;        // per particle, original index (bit0 emitter, bit1 droplet)
    uint32_t ds_cap = 0, ds_stride = 0, ds_rows = 0;
    bool ds_recording = false;
    uint32_t* train_rows = nullptr;
    size_t train_rows_cap = 0;
    float* train_part = nullptr;
    float* train_loss = nullptr;
    float* grad_tmp = nullptr;
    float* eval_tmp = nullptr;
    size_t eval_tmp_bytes = 0;
    uint64_t adam_t = 0;
    // profiler
    bool prof_on = false;      // timing the current step
    int32_t prof_every = 0;    // 0: off; n: every n-th step
    ProfTotal prof_tot[K_COUNT];
    std::vector<hipEvent_t> prof_ev;   // pairs
    std::vector<ProfRec> prof_rec;
    size_t prof_head = 0, prof_used = 0;   // pairs [prof_head, prof_used) not yet accumulated

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
        prof_rec[prof_used] = ProfRec{kid, solver, iter, false};
        // The group's launches take the events (launch.h).
        pending_timing() = PendingTiming{prof_ev[2 * prof_used], prof_ev[2 * prof_used + 1], false};
        return DFSPH_OK;
    }
    int prof_end() {
        if (!prof_on) return DFSPH_OK;
        PendingTiming& pt = pending_timing();
        const bool launched = pt.used;
        pt = PendingTiming{};
        if (!launched) {   // nothing was launched: an empty interval
            if (hipEventRecord(prof_ev[2 * prof_used], stream) != hipSuccess ||
                hipEventRecord(prof_ev[2 * prof_used + 1], stream) != hipSuccess)
                return fail(DFSPH_E_DEVICE, "hipEventRecord failed");
        }
        ++prof_used;
        return DFSPH_OK;
    }
    // Once solve `solver` has finished with `iters` iterations: its recorded
    // iterations from pair `first` on that lie past the count are not timed
    // (their kernels read the done flag and exit).
    void prof_mark(size_t first, int solver, int iters) {
        for (size_t q = first; q < prof_used; ++q) {
            ProfRec& r = prof_rec[q];
            if (r.solver == solver && r.iter >= iters) r.skip = true;
        }
    }
    // Accumulate the recorded pairs in order.  wait = false takes only pairs
    // whose end event has completed and leaves the rest for a later call, so
    // a profiled step needs no stream wait of its own; wait = true (stream
    // drained by the caller) takes them all.
    int prof_collect(bool wait) {
        for (; prof_head < prof_used; ++prof_head) {
            hipEvent_t e1 = prof_ev[2 * prof_head + 1];
            if (wait) {
                if (hipEventSynchronize(e1) != hipSuccess) return fail(DFSPH_E_DEVICE, "hipEventSynchronize failed");
            } else {
                const hipError_t q = hipEventQuery(e1);
                if (q == hipErrorNotReady) break;
                if (q != hipSuccess) return fail(DFSPH_E_DEVICE, std::string("HIP: ") + hipGetErrorString(q));
            }
            const ProfRec& r = prof_rec[prof_head];
            if (r.skip) continue;
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, prof_ev[2 * prof_head], e1) != hipSuccess)
                return fail(DFSPH_E_DEVICE, "hipEventElapsedTime failed");
            ProfTotal& t = prof_tot[r.kid];
            ++t.launches;
            t.total_ms += ms;
            t.min_ms = std::min(t.min_ms, double(ms));
            t.max_ms = std::max(t.max_ms, double(ms));
        }
        if (prof_head == prof_used) prof_head = prof_used = 0;
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
        if (h_mirror) (void)hipHostFree(h_mirror);
        if (h_step) (void)hipHostFree(h_step);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    SimConst sim_const(float bscale) const { return make_sim_const(params, d, bscale, n, nb); }
    SweepArgs args(float bscale) const {
        SweepArgs a{};
        a.k = sim_const(bscale);
        a.grid = grid;
        a.b = b;
        a.block = params.block_size ? params.block_size : 256u;
        a.xcd_remap = (params.flags & DFSPH_FLAG_NO_XCD_REMAP) ? 0 : 1;
        a.staged = (params.flags & DFSPH_FLAG_NO_STAGED_WINDOWS) ? 0 : 1;
        return a;
    }
};

// ---- scratch of dfsph_mlp_grad_noalloc ---------------------------------------
// One file-scope device array holds everything that entry point stages or
// accumulates: the caller's rows and targets, the row list, the per-workgroup
// gradient partials, the per-workgroup loss partials and the summed gradient.
// Each device holds its own copy until the process ends.  Host code never
// passes the array's name to a kernel (in host code it is a stand-in, not a
// device address) but the per-device address that hipGetSymbolAddress
// returned in dfsph_create.
namespace {
constexpr uint32_t kNoAllocMaxRows = 4096;   // DFSPH_MLP_GRAD_NOALLOC_MAX_ROWS
constexpr size_t align256(size_t v) { return (v + 255) & ~size_t(255); }
constexpr size_t kNaWorkgroups = kNoAllocMaxRows / 64;
constexpr size_t kNaFeatures = 0;
constexpr size_t kNaTargets = kNaFeatures + align256(size_t(kNoAllocMaxRows) * kNF * 4);
constexpr size_t kNaRows = kNaTargets + align256(size_t(kNoAllocMaxRows) * kNO * 4);
constexpr size_t kNaPart = kNaRows + align256(size_t(kNoAllocMaxRows) * 4);
constexpr size_t kNaLoss = kNaPart + align256(kNaWorkgroups * kNParam * 4);
constexpr size_t kNaGrad = kNaLoss + align256(kNaWorkgroups * 4);
constexpr size_t kNaBytes = kNaGrad + align256(size_t(kNParam) * 4);   // 2.79 MB
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
constexpr int kNaMaxDevices = 64;
std::mutex g_na_mutex;
unsigned char* g_na_ptr[kNaMaxDevices] = {};   // device address of `scratch`, per device

// Called with `dev` current (dfsph_create), once per device.
int noalloc_register(int dev) {
    if (dev < 0 || dev >= kNaMaxDevices) return DFSPH_OK;   // the entry point then reports it
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (g_na_ptr[dev]) return DFSPH_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(DFSPH_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_na_ptr[dev] = static_cast<unsigned char*>(p);
    return DFSPH_OK;
}

unsigned char* noalloc_ptr(int dev) {
    if (dev < 0 || dev >= kNaMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_na_mutex);
    return g_na_ptr[dev];
}

__global__ void k_iota(uint32_t* __restrict__ out, uint32_t n) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = q;
}
}  // namespace

namespace {

int validate_params(const dfsph_params* p) {
    if (!p) return fail(DFSPH_E_INVALID, "params is NULL");
    if (p->abi_version != DFSPH_ABI_VERSION)
        return fail(DFSPH_E_INVALID, "abi_version mismatch: expected " +
                                         std::to_string(DFSPH_ABI_VERSION));
    if (p->capacity == 0) return fail(DFSPH_E_INVALID, "capacity must be > 0");
    if (!(p->particle_radius > 0.0f) || !std::isfinite(p->particle_radius))
        return fail(DFSPH_E_INVALID, "particle_radius must be finite and > 0");
    if (!(p->density0 > 0.0f)) return fail(DFSPH_E_INVALID, "density0 must be > 0");
    if (p->grid != DFSPH_GRID_REFERENCE && p->grid != DFSPH_GRID_FINE)
        return fail(DFSPH_E_INVALID, "grid must be DFSPH_GRID_REFERENCE or DFSPH_GRID_FINE");
    if (p->block_size && (p->block_size % 64 || p->block_size > 256))
        return fail(DFSPH_E_INVALID, "block_size must be a multiple of 64 up to 256");
    if (p->fine_cell_scale != 0.0f && !(p->fine_cell_scale >= 1.0009765625f && p->fine_cell_scale <= 4.0f))
        return fail(DFSPH_E_INVALID, "fine_cell_scale must be 0 or in [1 + 2^-10, 4]");
    if (const char* why = domain_problem(p->domain_min, p->domain_max, p->particle_radius, p->fine_cell_scale))
        return fail(DFSPH_E_INVALID, why);
    return DFSPH_OK;
}

int validate_policy(const dfsph_solver_policy* pol) {
    if (!pol) return fail(DFSPH_E_INVALID, "policy is NULL");
    if (pol->mode != DFSPH_POLICY_REFERENCE && pol->mode != DFSPH_POLICY_CONVERGED)
        return fail(DFSPH_E_INVALID, "policy.mode must be REFERENCE or CONVERGED");
    if (pol->warm_start < DFSPH_WARM_ZERO || pol->warm_start > DFSPH_WARM_MLP_DENSITY)
        return fail(DFSPH_E_INVALID, "policy.warm_start out of range");
    if (pol->max_iter_div < 1 || pol->max_iter_density < 1 || pol->min_iter_div < 0 ||
        pol->min_iter_density < 0 || pol->max_iter_div > 100000 || pol->max_iter_density > 100000)
        return fail(DFSPH_E_INVALID, "policy iteration limits out of range");
    if (!(pol->dt_max > 0.0f) || !(pol->cfl > 0.0f))
        return fail(DFSPH_E_INVALID, "policy.cfl and policy.dt_max must be > 0");
    if (!std::isfinite(pol->residual_bound))
        return fail(DFSPH_E_INVALID, "policy.residual_bound must be finite");
    if (pol->div_start < DFSPH_DIV_START_DEFAULT || pol->div_start > DFSPH_DIV_START_MLP)
        return fail(DFSPH_E_INVALID, "policy.div_start out of range");
    return DFSPH_OK;
}

}  // namespace

// An empty kernel whose only purpose is its name in a kernel trace or a
// counter run: a caller brackets a region with two of them.
__global__ void k_trace_mark(uint32_t) {}

namespace dfsph {
PendingTiming& pending_timing() {
    thread_local PendingTiming t;
    return t;
}
void launch_trace_mark(uint32_t tag, void* stream) {
    DFSPH_LAUNCH(k_trace_mark, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), tag);
}
int set_error(int code, const std::string& msg) { return fail(code, msg); }
int check_params(const dfsph_params* p) { return validate_params(p); }
int check_policy(const dfsph_solver_policy* pol) { return validate_policy(pol); }

float ctx_particle_radius(const dfsph_ctx* c) { return c->params.particle_radius; }

uint32_t ctx_particle_flags(const dfsph_ctx* c, uint32_t i) { return i < c->flags.size() ? c->flags[i] : 0u; }

int ctx_resume(dfsph_ctx* c, const float* p_density, const uint32_t* flags, uint64_t step) {
    DFSPH_HIP(hipSetDevice(c->device));
    if (c->stepped) return fail(DFSPH_E_INVALID, "resume applies to a freshly initialised scene");
    // After dfsph_init_scene without ids the stage order is the original order.
    if (p_density && c->n)
        DFSPH_HIP(hipMemcpyAsync(c->b.pden_stage, p_density, size_t(c->n) * 4, hipMemcpyHostToDevice, c->stream));
    if (flags) c->flags.assign(flags, flags + c->n);
    c->step_count = step;
    DFSPH_HIP(hipStreamSynchronize(c->stream));
    return DFSPH_OK;
}

}  // namespace dfsph

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
    if (!sampler_extent_ok(dmin, dmax, radius))
        return fail(DFSPH_E_INVALID, "radius must be finite and > 0 and the box finite, with dmax >= dmin and "
                                     "at most 2^24 sample spacings per axis");
    std::vector<float> s;
    try {
        s = boundary_samples(dmin, dmax, radius, cap);
    } catch (const std::bad_alloc&) {
        return fail(DFSPH_E_NOMEM, "boundary samples do not fit in host memory; pass a cap");
    }
    *count = uint32_t(s.size() / 3);
    const size_t copy = std::min<size_t>(*count, cap_out);
    if (out && copy) std::memcpy(out, s.data(), sizeof(float) * 3 * copy);
    return DFSPH_OK;
}

int dfsph_preset_get(int32_t index, dfsph_preset* out) {
    if (!out) return fail(DFSPH_E_INVALID, "out is NULL");
    if (index < 0 || index >= DFSPH_PRESET_COUNT)
        return fail(DFSPH_E_INVALID, "preset index must be 0.." + std::to_string(DFSPH_PRESET_COUNT - 1));
    *out = preset_table()[index];
    return DFSPH_OK;
}

int dfsph_preset_fill(const dfsph_preset* preset, float radius, uint32_t capacity, float* positions,
                      float* velocities, uint32_t* flags, uint32_t* count, uint32_t* brick_count) {
    if (!preset) return fail(DFSPH_E_INVALID, "preset is NULL");
    if (const char* why = preset_problem(*preset, radius)) return fail(DFSPH_E_INVALID, why);
    preset_fill(*preset, radius, capacity, positions, velocities, flags, count, brick_count);
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
    if (int rc = noalloc_register(params->device)) return rc;

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
    // The decision mirror is optional: without it the host reads decisions
    // with a copy and a stream wait.
    if (hipHostMalloc(reinterpret_cast<void**>(&c->h_mirror), sizeof(SolveMirror), hipHostMallocCoherent) ==
        hipSuccess) {
        std::memset(c->h_mirror, 0, sizeof(SolveMirror));
        void* dm = nullptr;
        if (hipHostGetDevicePointer(&dm, c->h_mirror, 0) == hipSuccess) c->b.mirror = static_cast<SolveMirror*>(dm);
    } else {
        c->h_mirror = nullptr;
        (void)hipGetLastError();
    }
    // Likewise the substep mirror: without it the end of a substep is a copy
    // and a stream wait.
    if (hipHostMalloc(reinterpret_cast<void**>(&c->h_step), sizeof(StepMirror), hipHostMallocCoherent) ==
        hipSuccess) {
        std::memset(c->h_step, 0, sizeof(StepMirror));
        void* dm = nullptr;
        if (hipHostGetDevicePointer(&dm, c->h_step, 0) == hipSuccess) c->b.step_mirror = static_cast<StepMirror*>(dm);
    } else {
        c->h_step = nullptr;
        (void)hipGetLastError();
    }

    const size_t N = params->capacity;
    const size_t cells = c->grid.cells;
    Buffers& b = c->b;
    int rc = DFSPH_OK;
    auto A4 = [&](void** p) { if (!rc) rc = c->alloc(p, N * 16); };
    auto AF = [&](float** p) { if (!rc) rc = c->alloc_n(p, N); };
    auto AU = [&](uint32_t** p, size_t cnt) { if (!rc) rc = c->alloc_n(p, cnt); };
    A4(&b.pos4); A4(&b.vel4); A4(&b.vel4b); A4(&b.acc4); A4(&b.gb4);
    A4(&b.x_stage); A4(&b.v_stage);
    AF(&b.rho); AF(&b.alpha); AF(&b.aor2); AF(&b.rho_dot); AF(&b.rho_adv); AF(&b.p);
    AF(&b.p_div); AF(&b.p_den); AF(&b.pdiv_prev); AF(&b.pden_prev); AF(&b.pdiv_stage);
    AF(&b.pden_stage);
    AU(&b.nbr, N); AU(&b.nbrb, N); AU(&b.ids, N); AU(&b.keys, N); AU(&b.id_stage, N);
    AU(&b.perm, N); AU(&b.cell_count, cells); AU(&b.cell_start, cells + 1);
    if (!rc) rc = c->alloc_n(&b.skey, N);
    AU(&b.scan_blocks, scan_scratch_elems(uint32_t(cells)));
    if (params->grid == DFSPH_GRID_FINE && (params->flags & DFSPH_FLAG_NEIGHBOR_LIST)) {
        AU(&b.nlist, size_t(kListSlots) * N);
        AU(&b.ncnt, N);
        b.mask_stride = uint32_t(N);
    } else if (params->grid == DFSPH_GRID_FINE && !(params->flags & DFSPH_FLAG_NO_NEIGHBOR_MASKS)) {
        AU(&b.nrec, 12 * N);
        b.mask_stride = uint32_t(N);
        if (!rc) rc = c->alloc_n(&b.wave_cells, 2 * (N / 64 + 4));   // waves of the largest block (256)
    }
    if (!rc) rc = c->alloc_n(&b.partials, blocks_for(N, 64) + 64);
    if (!rc) rc = c->alloc_n(&b.res_max_part, blocks_for(N, 64) + 64);
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
    c->prefix_ready = false;
    c->flags.assign(n, 0u);
    if (sc->flags)
        for (uint32_t i = 0; i < n; ++i) c->flags[ids[i]] = sc->flags[i];
    DFSPH_HIP(hipMemcpyAsync(b.x_stage, x4.data(), x4.size() * 4, hipMemcpyHostToDevice, c->stream));
    DFSPH_HIP(hipMemcpyAsync(b.v_stage, v4.data(), v4.size() * 4, hipMemcpyHostToDevice, c->stream));
    DFSPH_HIP(hipMemcpyAsync(b.id_stage, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, c->stream));
    DFSPH_HIP(hipMemsetAsync(b.pdiv_stage, 0, size_t(c->params.capacity) * 4, c->stream));
    DFSPH_HIP(hipMemsetAsync(b.pden_stage, 0, size_t(c->params.capacity) * 4, c->stream));
    DeviceStatus st{};
    std::memcpy(&st.vmax2_bits, &vmax2, 4);
    DFSPH_HIP(hipMemcpyAsync(b.status, &st, sizeof(st), hipMemcpyHostToDevice, c->stream));

    // Boundary: samples on the AABB, sorted into the step grid, volumes on device.
    c->release(b.bpos4); c->release(b.bvol); c->release(b.bcell_start); c->release(b.bids); c->release(b.bnear);
    b.bpos4 = nullptr; b.bvol = nullptr; b.bcell_start = nullptr; b.bids = nullptr; b.bnear = nullptr;
    c->nb = 0;
    // Samples are placed on the device from a table of face rows, edges and
    // corners (a few thousand entries), not uploaded one by one.
    uint32_t nb = 0;
    std::vector<BoundarySeg> segs;
    if (sc->boundary) segs = boundary_segments(c->params.domain_min, c->params.domain_max,
                                               c->params.particle_radius, sc->boundary_cap, &nb);
    int rc = DFSPH_OK;
    rc = c->alloc_n(&b.bcell_start, size_t(c->grid.cells) + 1);
    if (rc) return rc;
    if (nb) {
        void *bstage = nullptr, *dsegs = nullptr;
        uint32_t *bkeys = nullptr, *bperm = nullptr;
        if ((rc = c->alloc(&b.bpos4, size_t(nb) * 16)) || (rc = c->alloc_n(&b.bvol, nb)) ||
            (rc = c->alloc_n(&b.bids, nb)) || (rc = c->alloc(&bstage, size_t(nb) * 16)) ||
            (rc = c->alloc_n(&bkeys, nb)) || (rc = c->alloc_n(&bperm, nb)) ||
            (rc = c->alloc(&dsegs, segs.size() * sizeof(BoundarySeg))))
            return rc;
        DFSPH_HIP(hipMemcpyAsync(dsegs, segs.data(), segs.size() * sizeof(BoundarySeg), hipMemcpyHostToDevice,
                                 c->stream));
        launch_boundary_sample(static_cast<const BoundarySeg*>(dsegs), uint32_t(segs.size()), nb,
                               2.0f * c->params.particle_radius, bstage, c->stream);
        launch_keys(c->grid, bstage, nb, bkeys, b.cell_count, c->stream);
        launch_scan(c->grid.cells, b.cell_count, b.bcell_start, b.scan_blocks, c->stream);
        launch_scatter(nb, bkeys, b.bcell_start, b.cell_count, bperm, nullptr, nullptr, nullptr, c->grid.cells, c->stream);
        c->nb = nb;
        SweepArgs a = c->args(1.0f);
        launch_reorder_boundary(a, bstage, bperm, c->stream);
        launch_boundary_volume(a, c->stream);
        if (c->grid.kind == DFSPH_GRID_FINE) {
            if ((rc = c->alloc(reinterpret_cast<void**>(&b.bnear), c->grid.cells))) return rc;
            launch_boundary_near(c->grid, b.bcell_start, b.bnear, c->stream);
        }
        DFSPH_HIP(hipGetLastError());
        DFSPH_HIP(hipStreamSynchronize(c->stream));
        c->release(bstage); c->release(bkeys); c->release(bperm); c->release(dsegs);
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

int dfsph_emit(dfsph_ctx* c, const dfsph_emitter* emitters, uint32_t count, uint32_t n_new,
               uint32_t* emitted) {
    if (!c || !emitters) return fail(DFSPH_E_INVALID, "NULL argument");
    if (count < 1 || count > DFSPH_MAX_EMITTERS)
        return fail(DFSPH_E_INVALID, "emitter count must be 1.." + std::to_string(DFSPH_MAX_EMITTERS));
    for (uint32_t e = 0; e < count; ++e) {
        const dfsph_emitter& m = emitters[e];
        bool ok = std::isfinite(m.radius) && m.radius >= 0.0f;
        for (int a = 0; a < 3; ++a) ok = ok && std::isfinite(m.pos[a]) && std::isfinite(m.velocity[a]);
        if (!ok) return fail(DFSPH_E_INVALID, "emitter " + std::to_string(e) + " has a non-finite field");
    }
    DFSPH_HIP(hipSetDevice(c->device));
    const uint32_t room = c->params.capacity - c->n;
    const uint32_t k = std::min(n_new, room);
    if (emitted) *emitted = k;
    if (k == 0) return DFSPH_OK;
    float e8[DFSPH_MAX_EMITTERS * 8] = {};
    for (uint32_t e = 0; e < count; ++e) {
        for (int a = 0; a < 3; ++a) {
            e8[8 * e + a] = emitters[e].pos[a];
            e8[8 * e + 4 + a] = emitters[e].velocity[a];
        }
        e8[8 * e + 3] = emitters[e].radius;
        e8[8 * e + 7] = emitters[e].rate;
    }
    launch_emit(c->b, e8, count, c->n, k, c->stream);
    DFSPH_HIP(hipGetLastError());
    DFSPH_HIP(hipStreamSynchronize(c->stream));
    c->flags.resize(size_t(c->n) + k, 1u);   // bit0: emitter (shaders/apply_emitter.comp.glsl:63)
    c->n += k;
    c->stepped = false;                     // per-step fields cover the old count until the next step
    c->prefix_ready = false;
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
    // The divergence solve's start: as warm_start says unless div_start names one.
    int warm_div = warm == DFSPH_WARM_MLP_DENSITY ? DFSPH_WARM_PREVIOUS : warm;
    if (!reference && pol->div_start != DFSPH_DIV_START_DEFAULT) {
        const int starts[] = {DFSPH_WARM_ZERO, DFSPH_WARM_PREVIOUS, DFSPH_WARM_MLP};
        warm_div = starts[pol->div_start - 1];
    }
    if ((warm == DFSPH_WARM_MLP || warm == DFSPH_WARM_MLP_DENSITY || warm_div == DFSPH_WARM_MLP) && !c->mlp_ready) {
        warm = warm_div = DFSPH_WARM_ZERO;
        warm_fallback = 1;
    }
    const float cfl_2r = pol->cfl * (2.0f * c->params.particle_radius);
    uint32_t syncs = 0;
    c->launches = 0;
    int it_tot[2] = {0, 0}, conv_all[2] = {1, 1}, it_redo[2] = {0, 0};
    double err_max[2] = {0.0, 0.0};
    DFSPH_HIP(hipEventRecord(c->ev0, c->stream));
    for (int sub = 0; sub < substeps; ++sub) {
        Buffers& b = c->b;
        if (int rc = c->prof_collect(false)) return rc;
        c->prof_on = c->prof_every > 0 && c->step_count % uint64_t(c->prof_every) == 0;
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
        // keys_done: the keys and histogram came with the velocity update;
        // fin: the substep's finish, carried by the first scan launch.
        auto cell_build = [&](bool keys_done, const StepFinish* fin) -> int {
            if (!keys_done)
                DFSPH_RUN(K_KEYS, -1, -1, launch_keys(c->grid, b.x_stage, c->n, b.keys, b.cell_count, c->stream));
            DFSPH_RUN(K_SCAN, -1, -1,
                      launch_scan(c->grid.cells, b.cell_count, b.cell_start, b.scan_blocks, c->stream, fin));
            // FINE grid: cells ordered by x, so each run of cells is one x-sorted range.
            DFSPH_RUN(K_SCATTER, -1, -1, launch_scatter(c->n, b.keys, b.cell_start, b.cell_count, b.perm,
                                                       b.id_stage,
                                                       c->grid.kind == DFSPH_GRID_FINE ? b.x_stage : nullptr,
                                                       c->grid.kind == DFSPH_GRID_FINE ? b.skey : nullptr,
                                                       c->grid.cells, c->stream));
            return DFSPH_OK;
        };
        if (c->counts_dirty) {
            DFSPH_HIP(hipMemsetAsync(b.cell_count, 0, size_t(c->grid.cells) * 4, c->stream));
            c->counts_dirty = false;
        }
        if (!c->prefix_ready)
            if (int rc = cell_build(false, nullptr)) return rc;
        c->prefix_ready = false;
        DFSPH_RUN(K_REORDER, -1, -1, launch_reorder(a, dt, cfl_2r, pol->dt_max, c->stream));
        // The divergence solve's start rides on the density pass unless the
        // network provides it.
        const bool div_start_fused = pol->div_enabled && warm_div != DFSPH_WARM_MLP;
        // The density solve's start (DFSPH_WARM_MLP_DENSITY: the network) rides
        // on the density-source sweep unless the network provides it.
        const int warm_den = warm == DFSPH_WARM_MLP_DENSITY ? DFSPH_WARM_MLP : warm;
        const bool den_start_fused = warm_den != DFSPH_WARM_MLP;
        DFSPH_RUN(K_DENSITY_ALPHA, -1, -1, launch_density_alpha(a, div_start_fused ? warm_div : -1, c->stream));
        // Dataset rows of this step (one particle in ds_stride, two rows each).
        const uint32_t ds_step = uint32_t(c->step_count);
        const uint32_t ds_base = c->ds_rows;
        const uint32_t ds_end = c->ds_recording
            ? uint32_t(std::min<uint64_t>(c->ds_cap, uint64_t(ds_base) + 2ull * (c->n / c->ds_stride)))
            : 0u;
        const bool rec = c->ds_recording && ds_end > ds_base;
        // The substep's status comes back through the pinned mirror unless
        // launches follow the publishing one (dataset rows); then the next
        // step's cell build is enqueued at the end of this one, and the
        // velocity update after the density solve takes its keys pass.
        const bool poll_end = c->h_step && b.step_mirror && !rec;
        if (rec)
            DFSPH_RUN(K_RECORD, -1, -1, launch_record_features(a.k, a.b, c->mlp, 0, ds_step, c->ds_stride,
                                                              ds_base, ds_end, c->stream));

        // The decision after `launched` iterations of solve s into h_status: from
        // the pinned mirror as soon as the deciding launch publishes it (that
        // launch is still running), else by a copy and a stream wait.
        auto wait_decision = [&](int s, int launched, bool fused) -> int {
            if (fused && c->h_mirror) {
                uint64_t* pw = c->h_mirror->solve[s].w;
                for (uint32_t spin = 1;; ++spin) {
                    uint64_t w[4];
                    for (int u = 0; u < 4; ++u) w[u] = __atomic_load_n(&pw[u], __ATOMIC_ACQUIRE);
                    const uint32_t tag = uint32_t(w[0] >> 32);
                    if ((tag & kPubTagBit) && uint32_t(w[1] >> 32) == tag && uint32_t(w[2] >> 32) == tag &&
                        uint32_t(w[3] >> 32) == tag) {
                        SolveStatus st{};
                        st.iters = int32_t(tag & ~kPubTagBit);
                        st.done = int32_t(uint32_t(w[0]) & 1u);
                        st.converged = int32_t((uint32_t(w[0]) >> 1) & 1u);
                        const uint32_t rb = uint32_t(w[1]);
                        std::memcpy(&st.res_max, &rb, 4);
                        const uint64_t eb = uint64_t(uint32_t(w[2])) | (uint64_t(uint32_t(w[3])) << 32);
                        std::memcpy(&st.err, &eb, 8);
                        st.done_at = st.done ? st.iters : 0;
                        if (st.done || st.iters >= launched) {
                            c->h_status->solve[s] = st;
                            return DFSPH_OK;
                        }
                    }
                    if ((spin & 1023u) == 0u) {
                        const hipError_t e = hipStreamQuery(c->stream);
                        if (e == hipSuccess) break;   // idle without the decision: read it directly
                        if (e != hipErrorNotReady) return fail(DFSPH_E_DEVICE, std::string("HIP: ") +
                                                                                  hipGetErrorString(e));
                    }
                }
            }
            DFSPH_HIP(hipMemcpyAsync(&c->h_status->solve[s], &b.status->solve[s], sizeof(SolveStatus),
                                     hipMemcpyDeviceToHost, c->stream));
            DFSPH_HIP(hipStreamSynchronize(c->stream));
            return DFSPH_OK;
        };
        auto run_solve = [&](int s) -> int {
            const int min_it = s == 0 ? pol->min_iter_div : pol->min_iter_density;
            const int max_it = s == 0 ? pol->max_iter_div : pol->max_iter_density;
            const float tol = s == 0 ? pol->tol_div : pol->tol_density;
            // This solve's start: DFSPH_WARM_MLP_DENSITY is the network for the
            // density solve and the previous pressure for the divergence solve.
            const int warm_s = s == 0 ? warm_div : warm_den;
            if (warm_s == DFSPH_WARM_MLP)
                DFSPH_RUN(K_MLP_WARM, -1, -1, launch_mlp_warm(a.k, a.b, c->mlp, s, c->stream));
            else if (!(s == 0 ? div_start_fused : den_start_fused))
                DFSPH_RUN(K_BEGIN_SOLVE, -1, -1, launch_begin_solve(a, s, warm_s, c->stream));
            // With 256-thread workgroups the decision of iteration it rides on the
            // pressure-acceleration launch of iteration it + 1 (workgroup 0), and
            // the solve's last such launch leaves the final acceleration in acc4,
            // so the velocity update needs no neighbour walk of its own.
            // Otherwise each iteration ends with its own one-workgroup decision.
            const bool fused = fused_decision(a);
            auto decide = [&](int ref, int mn, int mx, float tl, float rb) {
                return SolveDecide{fused ? 1 : 0, ref, mn, mx, tl, rb};
            };
            auto iterate = [&](int it, int ref, int mn, int mx, float tl, float rb) -> int {
                const SolveDecide d = decide(ref, mn, mx, tl, rb);
                if (!fused) DFSPH_RUN(K_PRESSURE_ACCEL, s, it, launch_pressure_accel(a, s, it, d, c->stream));
                DFSPH_RUN(K_JACOBI, s, it, launch_jacobi(a, s, ref, c->stream));
                if (fused)   // recorded as iteration it: it runs whenever iteration it did
                    DFSPH_RUN(K_PRESSURE_ACCEL, s, it, launch_pressure_accel(a, s, it + 1, d, c->stream));
                else
                    DFSPH_RUN(K_SOLVE_FINISH, s, it, launch_solve_finish(a, s, it, ref, mn, mx, tl, rb, c->stream));
                return DFSPH_OK;
            };
            auto first_accel = [&](int ref, int mn, int mx, float tl, float rb) -> int {
                if (fused)
                    DFSPH_RUN(K_PRESSURE_ACCEL, s, -1, launch_pressure_accel(a, s, 0, decide(ref, mn, mx, tl, rb),
                                                                             c->stream));
                return DFSPH_OK;
            };
            if (reference) {
                const int fixed = std::max(min_it, 1);   // main.cpp:2774, :2830
                if (int rc = first_accel(1, fixed, fixed, 0.0f, 0.0f)) return rc;
                for (int it = 0; it < fixed; ++it)
                    if (int rc = iterate(it, 1, fixed, fixed, 0.0f, 0.0f)) return rc;
                c->iters_done[s] = fixed;
            } else {
                // Converged policy: iterate to tolerance.  The average criterion
                // alone accepts a few over-pressured particles, so every solve
                // also bounds the largest per-particle residual (the zero start
                // stays within the default bound by itself; a warm start is held
                // to the same accuracy).  A warm-started solve that still misses
                // either criterion at max_iter is redone once from zero (the
                // plain solver is the fallback and sets the accuracy).
                const float bound = pol->residual_bound == 0.0f ? 25.0f * tol : pol->residual_bound;
                for (int attempt = 0; attempt < 2; ++attempt) {
                    const float rb = std::max(bound, 0.0f);
                    const size_t prof_first = c->prof_used;
                    if (fused && c->h_mirror)   // no decision of this solve is in flight
                        for (int u = 0; u < 4; ++u) __atomic_store_n(&c->h_mirror->solve[s].w[u], 0ull, __ATOMIC_RELEASE);
                    if (int rc = first_accel(0, min_it, max_it, tol, rb)) return rc;
                    int it = 0;
                    // With the decision published at the start of the next
                    // pressure launch, the host learns it while that launch
                    // still runs and enqueues the next iteration (or stage)
                    // before the queue drains: iterations are launched one at a
                    // time past min_iter, and no launch past convergence is
                    // queued ahead (each costs a few us).  A separate decision
                    // launch ends its iteration: there one more is queued than
                    // the last solve needed, as falling short costs a host
                    // round trip.
                    int chunk = fused ? std::max(min_it, 1) : std::max(std::max(min_it, 1), c->predicted[s] + 1);
                    while (it < max_it) {
                        const int cnt = std::min(chunk, max_it - it);
                        for (int q = 0; q < cnt; ++q, ++it)
                            if (int rc = iterate(it, 0, min_it, max_it, tol, rb)) return rc;
                        if (int rc = wait_decision(s, it, fused)) return rc;
                        ++syncs;
                        if (c->h_status->solve[s].done) break;
                        chunk = fused ? 1 : std::max(2, std::min(8, it / 4 + 1));
                    }
                    const SolveStatus& hs = c->h_status->solve[s];
                    c->iters_done[s] = hs.iters;
                    c->predicted[s] = std::max(1, hs.iters);
                    c->prof_mark(prof_first, s, hs.iters);
                    if (hs.converged || warm_s == DFSPH_WARM_ZERO || attempt) break;
                    warm_fallback = 1;
                    DFSPH_HIP(hipStreamSynchronize(c->stream));  // its deciding launch may still run
                    DFSPH_RUN(K_BEGIN_SOLVE, -1, -1, launch_reset_solve(a, s, c->stream));
                    it_redo[s] += c->h_status->solve[s].iters;
                }
            }
            // The divergence solve's velocity update rides on the forces sweep
            // when its final acceleration is in acc4.
            if (s == 0 && fused) return DFSPH_OK;
            uint32_t* keys = s == 1 && poll_end ? b.keys : nullptr;
            uint32_t* count = keys ? b.cell_count : nullptr;
            if (keys) c->counts_dirty = true;
            DFSPH_RUN(s == 0 ? K_ACCEL_APPLY_DIV : K_ACCEL_APPLY_FINAL, -1, -1,
                      fused ? launch_accel_apply_acc(a, s, keys, count, c->stream)
                            : launch_accel_apply(a, s, keys, count, c->stream));
            return DFSPH_OK;
        };
        c->iters_done[0] = c->iters_done[1] = 0;
        if (pol->div_enabled) {
            if (int rc = run_solve(0)) return rc;
        } else {
            DFSPH_HIP(hipMemsetAsync(b.p_div, 0, size_t(c->n) * 4, c->stream));
        }
        DFSPH_RUN(K_FORCES, -1, -1, launch_forces(a, pol->div_enabled && fused_decision(a), c->stream));
        DFSPH_RUN(K_DENSITY_ADV, -1, -1, launch_density_adv(a, den_start_fused ? warm_den : -1, c->stream));
        if (rec)
            DFSPH_RUN(K_RECORD, -1, -1, launch_record_features(a.k, a.b, c->mlp, 1, ds_step, c->ds_stride,
                                                              ds_base, ds_end, c->stream));
        if (int rc = run_solve(1)) return rc;
        // A profiled step takes the same path: its events are read once they
        // have completed.
        const uint32_t ticket = uint32_t(c->step_count % 0x7FFFFFFFull) + 1u;
        // With the next step's cell build enqueued at the end of this one, the
        // substep's finish (and its publication) rides on the build's first
        // scan launch; otherwise it has a launch of its own, unpublished.
        const StepFinish fin{b.status, b.step_part, (c->n + a.block - 1u) / a.block, b.step_mirror, ticket};
        if (poll_end) {
            // The last substep published has already been read: clear its
            // sequence so a ticket seen again (state reloaded) is not stale.
            __atomic_store_n(&c->h_step->seq, 0u, __ATOMIC_RELEASE);
        } else {
            SweepArgs af = a;
            af.b.step_mirror = nullptr;
            DFSPH_RUN(K_STEP_FINISH, -1, -1, launch_step_finish(af, ticket, c->stream));
        }
        if (rec) {
            DFSPH_RUN(K_RECORD, -1, -1, launch_record_targets(a.k, a.b, c->mlp, ds_step, c->ds_stride,
                                                             ds_base, ds_end, c->stream));
            c->ds_rows = ds_end;
        }
        ++c->step_count;
        // The step's sorted order becomes the stage order of the next step.
        std::swap(b.ids, b.id_stage);
        std::swap(b.p_div, b.pdiv_stage);
        std::swap(b.p_den, b.pden_stage);
        c->stepped = true;
        if (sub + 1 == substeps) DFSPH_HIP(hipEventRecord(c->ev1, c->stream));
        // The next step's cell build depends only on the stage arrays just
        // written: enqueued now, it runs while the host reads this step's
        // status and returns.  When the next step is profiled, these launches
        // are timed as part of it.  Not ahead of a stream wait.
        if (poll_end) {
            c->prof_on = c->prof_every > 0 && c->step_count % uint64_t(c->prof_every) == 0;
            if (int rc = cell_build(true, &fin)) return rc;
            c->prefix_ready = true;
            c->counts_dirty = false;
        }
#undef DFSPH_RUN
        DFSPH_HIP(hipGetLastError());
        // Per-substep statistics (one status read per substep).
        bool polled = false;
        if (poll_end) {
            StepMirror* mr = c->h_step;
            for (uint32_t spin = 1;; ++spin) {
                const uint32_t q1 = __atomic_load_n(&mr->seq, __ATOMIC_ACQUIRE);
                if (q1 == 2u * ticket) {
                    std::memcpy(c->h_status, const_cast<const DeviceStatus*>(&mr->status), sizeof(DeviceStatus));
                    std::atomic_thread_fence(std::memory_order_acquire);
                    if (__atomic_load_n(&mr->seq, __ATOMIC_ACQUIRE) == q1) { polled = true; break; }
                }
                if ((spin & 1023u) == 0u) {
                    const hipError_t e = hipStreamQuery(c->stream);
                    if (e == hipSuccess && __atomic_load_n(&mr->seq, __ATOMIC_ACQUIRE) != 2u * ticket) break;
                    if (e != hipSuccess && e != hipErrorNotReady)
                        return fail(DFSPH_E_DEVICE, std::string("HIP: ") + hipGetErrorString(e));
                }
            }
        }
        if (!polled) {
            DFSPH_HIP(hipMemcpyAsync(c->h_status, b.status, sizeof(DeviceStatus), hipMemcpyDeviceToHost,
                                     c->stream));
            DFSPH_HIP(hipStreamSynchronize(c->stream));
        }
        ++syncs;
        if (int rc = c->prof_collect(false)) return rc;
        for (int s = 0; s < 2; ++s) {
            if (s == 0 && !pol->div_enabled) continue;
            it_tot[s] += c->h_status->solve[s].iters + it_redo[s];
            it_redo[s] = 0;
            conv_all[s] &= c->h_status->solve[s].converged;
            err_max[s] = std::max(err_max[s], c->h_status->solve[s].err);
        }
    }
    for (;;) {   // polled: recorded after the last substep's final launch
        const hipError_t e = hipEventQuery(c->ev1);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) return fail(DFSPH_E_DEVICE, std::string("HIP: ") + hipGetErrorString(e));
    }
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
        s.max_residual_div = pol->div_enabled ? h.solve[0].res_max : 0.0f;
        s.max_residual_density = h.solve[1].res_max;
        *stats = s;
    }
    if (h.nonfinite) return fail(DFSPH_E_NONFINITE, std::to_string(h.nonfinite) +
                                                        " non-finite position/velocity components");
    if (pol->strict && (!conv_all[0] || !conv_all[1]))
        return fail(DFSPH_E_NOT_CONVERGED, "a pressure solve stopped at max_iter above tolerance");
    return DFSPH_OK;
}

// Render buffers for n particles and px pixels, grown as needed.
static int ensure_render(dfsph_ctx* c, uint32_t n, size_t px) {
    RenderBuffers& r = c->rb;
    if (n > c->render_n) {
        c->release(r.proj4); c->release(r.view4);
        r.proj4 = r.view4 = nullptr;
        c->render_n = 0;
        if (int rc = c->alloc(&r.proj4, size_t(n) * 16)) return rc;
        if (int rc = c->alloc(&r.view4, size_t(n) * 16)) return rc;
        c->render_n = n;
    }
    if (px > c->render_px) {
        void** imgs[] = {reinterpret_cast<void**>(&r.winner), reinterpret_cast<void**>(&r.thick_acc),
                         reinterpret_cast<void**>(&r.depth), reinterpret_cast<void**>(&r.smooth_a),
                         reinterpret_cast<void**>(&r.smooth_b), reinterpret_cast<void**>(&r.thick), &r.rgba};
        const size_t bytes[] = {4, 8, 4, 4, 4, 4, 16};
        for (void** q : imgs) { c->release(*q); *q = nullptr; }
        c->render_px = 0;
        for (int i = 0; i < 7; ++i)
            if (int rc = c->alloc(imgs[i], px * bytes[i])) return rc;
        c->render_px = px;
    }
    return DFSPH_OK;
}

int dfsph_render_defaults(const float eye[3], const float target[3], float fov_deg, uint32_t width,
                          uint32_t height, dfsph_render_params* out) {
    if (!eye || !target || !out) return fail(DFSPH_E_INVALID, "NULL argument");
    if (!(fov_deg > 0.0f && fov_deg < 180.0f)) return fail(DFSPH_E_INVALID, "fov_deg must be in (0, 180)");
    if (width == 0 || height == 0) return fail(DFSPH_E_INVALID, "width and height must be >= 1");
    dfsph_render_params r{};
    // glm::lookAt (right-handed, +y up) and glm::perspective with depth in
    // [-1, 1], y flipped for Vulkan (common/common-cpp/src/camera.cpp:107-118).
    float f[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
    const float fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (!(fl > 0.0f) || !std::isfinite(fl)) return fail(DFSPH_E_INVALID, "eye and target must differ (finite)");
    for (float& q : f) q *= 1.0f / fl;   // glm::normalize: v * inversesqrt(dot(v, v))
    float sx[3] = {f[1] * 0.0f - f[2] * 1.0f, f[2] * 0.0f - f[0] * 0.0f, f[0] * 1.0f - f[1] * 0.0f};   // f x up
    const float sl = std::sqrt(sx[0] * sx[0] + sx[1] * sx[1] + sx[2] * sx[2]);
    if (!(sl > 0.0f)) return fail(DFSPH_E_INVALID, "the view direction must not be vertical");
    for (float& q : sx) q *= 1.0f / sl;
    const float u[3] = {sx[1] * f[2] - sx[2] * f[1], sx[2] * f[0] - sx[0] * f[2], sx[0] * f[1] - sx[1] * f[0]};
    float* v = r.view;
    for (int a = 0; a < 3; ++a) { v[a * 4 + 0] = sx[a]; v[a * 4 + 1] = u[a]; v[a * 4 + 2] = -f[a]; v[a * 4 + 3] = 0.0f; }
    v[12] = -(sx[0] * eye[0] + sx[1] * eye[1] + sx[2] * eye[2]);
    v[13] = -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]);
    v[14] = f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2];
    v[15] = 1.0f;
    const float near = 0.05f, far = 100.0f;                  // main.cpp:155-156
    const float aspect = float(width) / float(height);
    const float th = std::tan(fov_deg * 0.01745329251994329576923690768489f * 0.5f);   // glm::radians
    r.proj[0] = 1.0f / (aspect * th);
    r.proj[5] = -(1.0f / th);
    r.proj[10] = -(far + near) / (far - near);
    r.proj[11] = -1.0f;
    r.proj[14] = -(2.0f * far * near) / (far - near);
    for (int a = 0; a < 3; ++a) r.camera_pos[a] = eye[a];
    r.width = width;
    r.height = height;
    r.point_scale = float(height) / (2.0f * th);             // main.cpp:2135
    r.thickness_per_particle = 0.005f;                       // main.cpp:140
    r.bilateral_iterations = 4;                              // main.cpp:137-139
    r.sigma_spatial = 3.0f;
    r.sigma_depth = 0.05f;
    const float tint[3] = {0.10f, 0.30f, 0.50f}, zen[3] = {0.40f, 0.55f, 0.85f}, hor[3] = {0.85f, 0.90f, 0.95f};
    for (int a = 0; a < 3; ++a) { r.water_tint[a] = tint[a]; r.sky_zenith[a] = zen[a]; r.sky_horizon[a] = hor[a]; }
    r.exposure = 1.0f;
    r.fresnel_f0 = 0.02f;                                    // main.cpp:2191
    r.absorption = r.thickness_per_particle;                 // main.cpp:2186
    r.thickness_scale = 0.0f;
    *out = r;
    return DFSPH_OK;
}

int dfsph_render(dfsph_ctx* c, const dfsph_render_params* p, float* rgba, float* depth, float* smoothed,
                 float* thickness) {
    if (!c || !p) return fail(DFSPH_E_INVALID, "NULL argument");
    if (p->width == 0 || p->height == 0 || p->width > 16384 || p->height > 16384)
        return fail(DFSPH_E_INVALID, "width and height must be in 1..16384");
    if (p->bilateral_iterations < 0 || p->bilateral_iterations > 64)
        return fail(DFSPH_E_INVALID, "bilateral_iterations must be in 0..64");
    if (!(p->sigma_spatial > 0.0f && p->sigma_spatial <= 256.0f) || !(p->sigma_depth > 0.0f) ||
        !std::isfinite(p->sigma_depth))
        return fail(DFSPH_E_INVALID, "sigma_spatial must be in (0, 256] and sigma_depth finite and > 0");
    if (!(p->point_scale > 0.0f) || !std::isfinite(p->point_scale) || !(p->particle_radius >= 0.0f) ||
        !std::isfinite(p->particle_radius))
        return fail(DFSPH_E_INVALID, "point_scale must be finite and > 0, particle_radius finite and >= 0");
    // Thickness sums are 64-bit fixed point in 2^-32 units: per-fragment
    // values up to thickness_per_particle / 0.1 must leave room for every
    // particle landing on one pixel.
    if (!(p->thickness_per_particle >= 0.0f && p->thickness_per_particle <= 1.0f))
        return fail(DFSPH_E_INVALID, "thickness_per_particle must be in [0, 1]");
    const float shade[] = {p->roughness, p->exposure, p->fresnel_f0, p->absorption, p->thickness_scale};
    for (float v : shade)
        if (!std::isfinite(v)) return fail(DFSPH_E_INVALID, "render shading parameters must be finite");
    const float* scalars[] = {p->view, p->proj, p->camera_pos, p->water_tint, p->sky_zenith, p->sky_horizon};
    const int counts[] = {16, 16, 3, 3, 3, 3};
    for (int q = 0; q < 6; ++q)
        for (int i = 0; i < counts[q]; ++i)
            if (!std::isfinite(scalars[q][i])) return fail(DFSPH_E_INVALID, "render matrices and colours must be finite");
    for (uint32_t v : p->reserved)
        if (v) return fail(DFSPH_E_INVALID, "render_params.reserved must be zero");
    RenderConst k{};
    std::memcpy(k.view, p->view, sizeof(k.view));
    std::memcpy(k.proj, p->proj, sizeof(k.proj));
    float vp[16];
    mul4(p->proj, p->view, vp);
    if (!invert4(vp, k.inv_vp)) return fail(DFSPH_E_INVALID, "proj * view is singular");
    for (int a = 0; a < 3; ++a) {
        k.cam[a] = p->camera_pos[a];
        k.tint[a] = p->water_tint[a];
        k.zenith[a] = p->sky_zenith[a];
        k.horizon[a] = p->sky_horizon[a];
    }
    k.w = p->width;
    k.h = p->height;
    k.n = c->n;
    k.radius = p->particle_radius > 0.0f ? p->particle_radius : c->params.particle_radius;
    k.point_scale = p->point_scale;
    k.tpp = p->thickness_per_particle;
    k.sigma_s = p->sigma_spatial;
    k.sigma_d = p->sigma_depth;
    k.rough = p->roughness;
    k.exposure = p->exposure;
    k.f0 = p->fresnel_f0;
    k.absorption = p->absorption;
    k.thick_scale = p->thickness_scale;
    const size_t px = size_t(k.w) * k.h;
    DFSPH_HIP(hipSetDevice(c->device));
    if (int rc = ensure_render(c, std::max(k.n, 1u), px)) return rc;
    RenderBuffers b = c->rb;
    b.x4 = c->b.x_stage;
    b.ids = c->b.id_stage;
    const float* final_depth = launch_render(k, b, p->bilateral_iterations, p->bilateral_alternate, c->stream);
    DFSPH_HIP(hipGetLastError());
    if (rgba) DFSPH_HIP(hipMemcpyAsync(rgba, b.rgba, px * 16, hipMemcpyDeviceToHost, c->stream));
    if (depth) DFSPH_HIP(hipMemcpyAsync(depth, b.depth, px * 4, hipMemcpyDeviceToHost, c->stream));
    if (smoothed) DFSPH_HIP(hipMemcpyAsync(smoothed, final_depth, px * 4, hipMemcpyDeviceToHost, c->stream));
    if (thickness) DFSPH_HIP(hipMemcpyAsync(thickness, b.thick, px * 4, hipMemcpyDeviceToHost, c->stream));
    DFSPH_HIP(hipStreamSynchronize(c->stream));
    return DFSPH_OK;
}

int dfsph_trace_mark(dfsph_ctx* c, uint32_t tag) {
    if (!c) return fail(DFSPH_E_INVALID, "ctx is NULL");
    DFSPH_HIP(hipSetDevice(c->device));
    launch_trace_mark(tag, c->stream);
    DFSPH_HIP(hipGetLastError());
    return DFSPH_OK;
}

int dfsph_profile_enable(dfsph_ctx* c, int32_t enable) {
    if (!c) return fail(DFSPH_E_INVALID, "ctx is NULL");
    if (enable < 0) return fail(DFSPH_E_INVALID, "enable must be >= 0");
    c->prof_every = enable;
    c->prof_on = false;
    for (auto& t : c->prof_tot) t = ProfTotal{};
    c->prof_head = c->prof_used = 0;   // pairs still in flight are dropped
    return DFSPH_OK;
}

int dfsph_profile_read(dfsph_ctx* c, dfsph_kernel_time* out, uint32_t cap, uint32_t* count) {
    if (!c || !count) return fail(DFSPH_E_INVALID, "NULL argument");
    DFSPH_HIP(hipSetDevice(c->device));
    DFSPH_HIP(hipStreamSynchronize(c->stream));
    if (int rc = c->prof_collect(true)) return rc;
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
    launch_scatter(c->n, c->ref_keys, c->ref_start, c->ref_count, c->ref_perm, c->b.id_stage, nullptr, nullptr,
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


namespace {

inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int ensure_mlp(dfsph_ctx* c) {
    if (c->mlp.theta) return DFSPH_OK;
    if (!mlp_smem_ok()) return fail(DFSPH_E_DEVICE, "MLP LDS layout exceeds the device limit");
    if (!mlp_set_smem_limits()) return fail(DFSPH_E_DEVICE, "hipFuncSetAttribute for the MLP kernels failed");
    int cus = 0;
    DFSPH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    c->mlp.cus = uint32_t(std::max(cus, 1));
    int rc;
    if ((rc = c->alloc_n(&c->mlp.theta, kNParam)) || (rc = c->alloc(&c->mlp.wpack, mlp_pack_bytes())) ||
        (rc = c->alloc_n(&c->mlp.adam_m, kNParam)) || (rc = c->alloc_n(&c->mlp.adam_v, kNParam)))
        return rc;
    return DFSPH_OK;
}

int ensure_dataset(dfsph_ctx* c, uint32_t cap) {
    if (c->mlp.ds_x && c->ds_cap >= cap) return DFSPH_OK;
    c->release(c->mlp.ds_x);
    c->release(c->mlp.ds_t);
    c->mlp.ds_x = c->mlp.ds_t = nullptr;
    int rc;
    if ((rc = c->alloc_n(&c->mlp.ds_x, size_t(cap) * kNF)) || (rc = c->alloc_n(&c->mlp.ds_t, size_t(cap) * kNO)))
        return rc;
    c->ds_cap = cap;
    return DFSPH_OK;
}

// Upload a parameter vector; a vector with a non-finite entry is rejected and
// leaves the learned warm start disabled (steps then fall back to zero init).
int upload_theta(dfsph_ctx* c, const float* theta) {
    for (int q = 0; q < kNParam; ++q)
        if (!std::isfinite(theta[q])) {
            c->mlp_ready = false;
            return fail(DFSPH_E_INVALID, "network parameter " + std::to_string(q) + " is not finite");
        }
    if (int rc = ensure_mlp(c)) return rc;
    DFSPH_HIP(hipMemcpyAsync(c->mlp.theta, theta, kNParam * 4, hipMemcpyHostToDevice, c->stream));
    DFSPH_HIP(hipMemsetAsync(c->mlp.adam_m, 0, kNParam * 4, c->stream));
    DFSPH_HIP(hipMemsetAsync(c->mlp.adam_v, 0, kNParam * 4, c->stream));
    launch_mlp_pack(c->mlp, c->stream);
    DFSPH_HIP(hipGetLastError());
    DFSPH_HIP(hipStreamSynchronize(c->stream));
    c->adam_t = 0;
    c->mlp_ready = true;
    return DFSPH_OK;
}

int ensure_train_scratch(dfsph_ctx* c, size_t rows, size_t nparts, size_t nloss) {
    if (c->train_rows_cap < std::max({rows, nparts * kNParam, nloss})) {
        c->release(c->train_rows); c->release(c->train_part); c->release(c->train_loss);
        c->train_rows = nullptr; c->train_part = nullptr; c->train_loss = nullptr;
        const size_t cap = std::max({rows, nparts * kNParam, nloss});
        int rc;
        if ((rc = c->alloc_n(&c->train_rows, cap)) || (rc = c->alloc_n(&c->train_part, cap)) ||
            (rc = c->alloc_n(&c->train_loss, cap)))
            return rc;
        c->train_rows_cap = cap;
    }
    return DFSPH_OK;
}

}  // namespace

int dfsph_mlp_init(dfsph_ctx* c, uint64_t seed) {
    if (!c) return fail(DFSPH_E_INVALID, "ctx is NULL");
    DFSPH_HIP(hipSetDevice(c->device));
    std::vector<float> th(kNParam, 0.0f);
    uint64_t ctr = 0;
    auto fill = [&](int off, int fan_in, int count) {
        const double lim = std::sqrt(6.0 / fan_in);
        for (int q = 0; q < count; ++q) {
            const double u = double(mix64(seed * 0x100000001B3ull + ctr++) >> 11) * 0x1p-53;
            th[off + q] = float((2.0 * u - 1.0) * lim);
        }
    };
    fill(kOW1, kNF, kNF * kNH);
    fill(kOW2, kNH, kNH * kNH);
    fill(kOW3, kNH, kNH * kNH);
    fill(kOW4, kNH, kNH * kNO);
    return upload_theta(c, th.data());
}

int dfsph_mlp_set_params(dfsph_ctx* c, const float* theta, uint32_t count) {
    if (!c || !theta) return fail(DFSPH_E_INVALID, "NULL argument");
    if (count != uint32_t(kNParam)) return fail(DFSPH_E_INVALID, "count must be DFSPH_MLP_PARAMS");
    DFSPH_HIP(hipSetDevice(c->device));
    return upload_theta(c, theta);
}

int dfsph_mlp_get_params(dfsph_ctx* c, float* theta, uint32_t count) {
    if (!c || !theta) return fail(DFSPH_E_INVALID, "NULL argument");
    if (count != uint32_t(kNParam)) return fail(DFSPH_E_INVALID, "count must be DFSPH_MLP_PARAMS");
    if (!c->mlp_ready) return fail(DFSPH_E_INVALID, "no network parameters set");
    DFSPH_HIP(hipSetDevice(c->device));
    DFSPH_HIP(hipMemcpyAsync(theta, c->mlp.theta, kNParam * 4, hipMemcpyDeviceToHost, c->stream));
    DFSPH_HIP(hipStreamSynchronize(c->stream));
    return DFSPH_OK;
}

int dfsph_dataset_start(dfsph_ctx* c, uint32_t capacity_rows, uint32_t stride) {
    if (!c) return fail(DFSPH_E_INVALID, "ctx is NULL");
    if (!capacity_rows || !stride) return fail(DFSPH_E_INVALID, "capacity_rows and stride must be > 0");
    DFSPH_HIP(hipSetDevice(c->device));
    if (int rc = ensure_mlp(c)) return rc;
    if (int rc = ensure_dataset(c, capacity_rows)) return rc;
    c->ds_cap = capacity_rows;
    c->ds_stride = stride;
    c->ds_rows = 0;
    c->ds_recording = true;
    return DFSPH_OK;
}

int dfsph_dataset_stop(dfsph_ctx* c, uint32_t* rows) {
    if (!c) return fail(DFSPH_E_INVALID, "ctx is NULL");
    c->ds_recording = false;
    if (rows) *rows = c->ds_rows;
    return DFSPH_OK;
}

int dfsph_dataset_export(dfsph_ctx* c, float* features, float* targets, uint32_t rows) {
    if (!c || (!features && !targets)) return fail(DFSPH_E_INVALID, "NULL argument");
    if (rows > c->ds_rows) return fail(DFSPH_E_INVALID, "rows exceeds the recorded count");
    DFSPH_HIP(hipSetDevice(c->device));
    if (features) DFSPH_HIP(hipMemcpyAsync(features, c->mlp.ds_x, size_t(rows) * kNF * 4, hipMemcpyDeviceToHost, c->stream));
    if (targets) DFSPH_HIP(hipMemcpyAsync(targets, c->mlp.ds_t, size_t(rows) * kNO * 4, hipMemcpyDeviceToHost, c->stream));
    DFSPH_HIP(hipStreamSynchronize(c->stream));
    return DFSPH_OK;
}

int dfsph_dataset_import(dfsph_ctx* c, const float* features, const float* targets, uint32_t rows) {
    if (!c || !features || !targets || !rows) return fail(DFSPH_E_INVALID, "bad argument");
    DFSPH_HIP(hipSetDevice(c->device));
    if (int rc = ensure_mlp(c)) return rc;
    if (int rc = ensure_dataset(c, rows)) return rc;
    DFSPH_HIP(hipMemcpyAsync(c->mlp.ds_x, features, size_t(rows) * kNF * 4, hipMemcpyHostToDevice, c->stream));
    DFSPH_HIP(hipMemcpyAsync(c->mlp.ds_t, targets, size_t(rows) * kNO * 4, hipMemcpyHostToDevice, c->stream));
    DFSPH_HIP(hipStreamSynchronize(c->stream));
    c->ds_rows = rows;
    c->ds_recording = false;
    return DFSPH_OK;
}

int dfsph_mlp_train(dfsph_ctx* c, const dfsph_train_config* cfg, dfsph_train_stats* stats) {
    if (!c || !cfg) return fail(DFSPH_E_INVALID, "NULL argument");
    if (!c->mlp_ready) return fail(DFSPH_E_INVALID, "initialise or set the network parameters first");
    if (!c->ds_rows) return fail(DFSPH_E_INVALID, "no dataset rows");
    if (cfg->batch < 64 || cfg->batch % 64 || cfg->batch > 65536)
        return fail(DFSPH_E_INVALID, "batch must be a multiple of 64 in [64, 65536]");
    if (!(cfg->lr > 0.0f) || !(cfg->val_fraction >= 0.0f && cfg->val_fraction < 1.0f))
        return fail(DFSPH_E_INVALID, "lr must be > 0 and val_fraction in [0, 1)");
    if (!(cfg->expectile_density == 0.0f || (cfg->expectile_density > 0.0f && cfg->expectile_density < 1.0f)))
        return fail(DFSPH_E_INVALID, "expectile_density must be 0 (MSE) or in (0, 1)");
    const float tau = cfg->expectile_density == 0.0f ? 0.5f : cfg->expectile_density;
    DFSPH_HIP(hipSetDevice(c->device));
    std::vector<uint32_t> tr, va;
    const uint64_t vthr = uint64_t(double(cfg->val_fraction) * 10000.0);
    for (uint32_t r = 0; r < c->ds_rows; ++r)
        (mix64(uint64_t(r) ^ 0xD1B54A32D192ED03ull) % 10000ull < vthr ? va : tr).push_back(r);
    const uint32_t batch = cfg->batch, wg = batch / 64;
    const size_t nb = tr.size() / batch;
    if (!nb) return fail(DFSPH_E_INVALID, "fewer training rows than one batch");
    if (int rc = ensure_train_scratch(c, std::max(tr.size(), va.size()), wg, nb * wg + va.size() / 64 + 1)) return rc;
    DFSPH_HIP(hipEventRecord(c->ev0, c->stream));
    uint64_t rng = cfg->seed;
    std::vector<float> lp;
    double train_loss = 0.0;
    uint32_t steps = 0;
    for (uint32_t ep = 0; ep < cfg->epochs; ++ep) {
        for (size_t q = tr.size(); q > 1; --q) {   // Fisher-Yates
            rng = mix64(rng);
            std::swap(tr[q - 1], tr[rng % q]);
        }
        DFSPH_HIP(hipMemcpyAsync(c->train_rows, tr.data(), tr.size() * 4, hipMemcpyHostToDevice, c->stream));
        for (size_t bq = 0; bq < nb; ++bq) {
            launch_mlp_train(c->mlp, c->train_rows + bq * batch, batch, c->train_part,
                             c->train_loss + bq * wg, tau, c->stream);
            ++c->adam_t;
            const float c1 = 1.0f - std::pow(cfg->beta1, float(c->adam_t));
            const float c2 = 1.0f - std::pow(cfg->beta2, float(c->adam_t));
            launch_mlp_adam(c->mlp, c->train_part, wg, cfg->lr, cfg->beta1, cfg->beta2, cfg->eps, c1, c2,
                            nullptr, c->stream);
            launch_mlp_pack(c->mlp, c->stream);
            ++steps;
        }
        DFSPH_HIP(hipGetLastError());
        lp.resize(nb * wg);
        DFSPH_HIP(hipMemcpyAsync(lp.data(), c->train_loss, lp.size() * 4, hipMemcpyDeviceToHost, c->stream));
        DFSPH_HIP(hipStreamSynchronize(c->stream));
        double sum = 0.0;
        for (float v : lp) sum += v;
        train_loss = sum / double(nb * batch);
    }
    // Held-out loss (forward only; the gradient partials are discarded).
    double val_loss = 0.0;
    if (!va.empty()) {
        DFSPH_HIP(hipMemcpyAsync(c->train_rows, va.data(), va.size() * 4, hipMemcpyHostToDevice, c->stream));
        size_t done = 0, nl = 0;
        while (done < va.size()) {
            const uint32_t cnt = uint32_t(std::min<size_t>(batch, va.size() - done));
            launch_mlp_train(c->mlp, c->train_rows + done, cnt, c->train_part, c->train_loss + nl, tau, c->stream);
            nl += (cnt + 63) / 64;
            done += cnt;
        }
        lp.resize(nl);
        DFSPH_HIP(hipMemcpyAsync(lp.data(), c->train_loss, nl * 4, hipMemcpyDeviceToHost, c->stream));
        DFSPH_HIP(hipStreamSynchronize(c->stream));
        double sum = 0.0;
        for (float v : lp) sum += v;
        val_loss = sum / double(va.size());
    }
    DFSPH_HIP(hipEventRecord(c->ev1, c->stream));
    for (;;) {   // polled: the event is at most one launch away
        const hipError_t e = hipEventQuery(c->ev1);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) return fail(DFSPH_E_DEVICE, std::string("HIP: ") + hipGetErrorString(e));
    }
    float ms = 0.0f;
    DFSPH_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (stats) {
        dfsph_train_stats st{};
        st.train_loss = train_loss;
        st.val_loss = val_loss;
        st.steps = steps;
        st.train_rows = uint32_t(tr.size());
        st.val_rows = uint32_t(va.size());
        st.ms = ms;
        *stats = st;
    }
    return DFSPH_OK;
}

int dfsph_mlp_eval(dfsph_ctx* c, const float* features, uint32_t n, float* out) {
    if (!c || !features || !out) return fail(DFSPH_E_INVALID, "NULL argument");
    if (!c->mlp_ready) return fail(DFSPH_E_INVALID, "no network parameters set");
    DFSPH_HIP(hipSetDevice(c->device));
    const size_t need = size_t(n) * (kNF + kNO) * 4;
    if (c->eval_tmp_bytes < need) {
        c->release(c->eval_tmp);
        c->eval_tmp = nullptr;
        if (int rc = c->alloc(reinterpret_cast<void**>(&c->eval_tmp), need)) return rc;
        c->eval_tmp_bytes = need;
    }
    float* xin = c->eval_tmp;
    float* yout = c->eval_tmp + size_t(n) * kNF;
    DFSPH_HIP(hipMemcpyAsync(xin, features, size_t(n) * kNF * 4, hipMemcpyHostToDevice, c->stream));
    launch_mlp_eval(c->mlp, xin, n, yout, c->stream);
    DFSPH_HIP(hipGetLastError());
    DFSPH_HIP(hipMemcpyAsync(out, yout, size_t(n) * kNO * 4, hipMemcpyDeviceToHost, c->stream));
    DFSPH_HIP(hipStreamSynchronize(c->stream));
    return DFSPH_OK;
}

int dfsph_mlp_grad(dfsph_ctx* c, const float* features, const float* targets, uint32_t n, float* grad,
                   double* loss) {
    if (!c || !features || !targets || !grad || !loss) return fail(DFSPH_E_INVALID, "NULL argument");
    if (!n || n > 65536) return fail(DFSPH_E_INVALID, "n must be in [1, 65536]");
    if (!c->mlp_ready) return fail(DFSPH_E_INVALID, "no network parameters set");
    DFSPH_HIP(hipSetDevice(c->device));
    const uint32_t wg = (n + 63) / 64;
    if (int rc = ensure_train_scratch(c, n, wg, wg)) return rc;
    if (!c->grad_tmp)
        if (int rc = c->alloc_n(&c->grad_tmp, kNParam)) return rc;
    float *gx = nullptr, *gt = nullptr;
    if (int rc = c->alloc_n(&gx, size_t(n) * kNF)) return rc;
    if (int rc = c->alloc_n(&gt, size_t(n) * kNO)) { c->release(gx); return rc; }
    std::vector<uint32_t> rows(n);
    for (uint32_t q = 0; q < n; ++q) rows[q] = q;
    MlpDev m = c->mlp;
    m.ds_x = gx;
    m.ds_t = gt;
    int rc = DFSPH_OK;
    std::vector<float> lp(wg);
    if (hipMemcpyAsync(gx, features, size_t(n) * kNF * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(gt, targets, size_t(n) * kNO * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(c->train_rows, rows.data(), size_t(n) * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = fail(DFSPH_E_DEVICE, "upload failed");
    if (!rc) {
        launch_mlp_train(m, c->train_rows, n, c->train_part, c->train_loss, 0.5f, c->stream);
        launch_mlp_adam(m, c->train_part, wg, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f, c->grad_tmp, c->stream);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(grad, c->grad_tmp, kNParam * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipMemcpyAsync(lp.data(), c->train_loss, wg * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail(DFSPH_E_DEVICE, "gradient kernels failed");
    }
    c->release(gx);
    c->release(gt);
    if (rc) return rc;
    double sum = 0.0;
    for (float v : lp) sum += v;
    *loss = sum / double(n);
    return DFSPH_OK;
}

int dfsph_mlp_grad_noalloc(dfsph_ctx* c, const float* features, const float* targets, uint32_t n, float* grad,
                           double* loss) {
    if (!c || !features || !targets || !grad || !loss) return fail(DFSPH_E_INVALID, "NULL argument");
    if (!n || n > kNoAllocMaxRows)
        return fail(DFSPH_E_INVALID, "n must be in [1, " + std::to_string(kNoAllocMaxRows) +
                                         "] (the scratch array's size)");
    if (!c->mlp_ready) return fail(DFSPH_E_INVALID, "no network parameters set");
    DFSPH_HIP(hipSetDevice(c->device));
    unsigned char* s = noalloc_ptr(c->device);
    if (!s) return fail(DFSPH_E_DEVICE, "no scratch address for device " + std::to_string(c->device));
    const uint32_t wg = (n + 63) / 64;
    float* gx = reinterpret_cast<float*>(s + kNaFeatures);
    float* gt = reinterpret_cast<float*>(s + kNaTargets);
    uint32_t* rows = reinterpret_cast<uint32_t*>(s + kNaRows);
    float* part = reinterpret_cast<float*>(s + kNaPart);
    float* part_loss = reinterpret_cast<float*>(s + kNaLoss);
    float* gsum = reinterpret_cast<float*>(s + kNaGrad);
    MlpDev m = c->mlp;
    m.ds_x = gx;
    m.ds_t = gt;
    // Every word the kernels read is written first in this call (rows by
    // k_iota, one partial per workgroup and parameter by k_mlp_train), so
    // nothing is cleared.
    float lp[kNaWorkgroups];
    DFSPH_HIP(hipMemcpyAsync(gx, features, size_t(n) * kNF * 4, hipMemcpyHostToDevice, c->stream));
    DFSPH_HIP(hipMemcpyAsync(gt, targets, size_t(n) * kNO * 4, hipMemcpyHostToDevice, c->stream));
    DFSPH_LAUNCH(k_iota, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, rows, n);
    launch_mlp_train(m, rows, n, part, part_loss, 0.5f, c->stream);
    launch_mlp_adam(m, part, wg, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f, gsum, c->stream);
    DFSPH_HIP(hipGetLastError());
    DFSPH_HIP(hipMemcpyAsync(grad, gsum, kNParam * 4, hipMemcpyDeviceToHost, c->stream));
    DFSPH_HIP(hipMemcpyAsync(lp, part_loss, wg * 4, hipMemcpyDeviceToHost, c->stream));
    DFSPH_HIP(hipStreamSynchronize(c->stream));
    double sum = 0.0;
    for (uint32_t q = 0; q < wg; ++q) sum += lp[q];
    *loss = sum / double(n);
    return DFSPH_OK;
}

}  // extern "C"
