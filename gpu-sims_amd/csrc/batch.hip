// batch.hip -- many independent DFSPH scenes advanced together: one
// persistent launch per substep for every scene of the batch, captured once
// in a HIP graph and replayed.
//
// The reference runs one scene per process and records 26 dispatches per
// substep (main.cpp:2663-2867).  Here the scenes' particles, cells and
// boundary samples are packed into shared arrays (scene s owns slots
// [p_off, p_off + n), cells [cell_off, cell_off + cells)), and particle
// tiles of 256 never straddle two scenes.  One persistent kernel walks the
// stages of the substep -- hash, reorder, density, divergence solve, forces,
// density source, density solve, integration -- separated by grid barriers;
// the Jacobi loops run on the device until every scene meets its tolerance
// (or max_iter), with each scene's FP64 error reduced over its own tiles in
// the fixed order the single-scene path uses, so a scene advanced in a batch
// is bit-identical to the same scene advanced alone.
//
// Grid barriers count arrivals with device-scope atomics and wait with a
// bounded spin: a workgroup that is not co-resident (more workgroups than
// the device holds at once) makes the barrier time out, every workgroup then
// returns at its next barrier and the step reports an error instead of
// hanging.  The host sizes the grid from the kernel's occupancy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "hash_bodies.h"
#include "launch.h"
#include "sweep_bodies.h"

using namespace dfsph;

namespace {

constexpr uint32_t kTile = 256;          // particles per tile = threads per workgroup
constexpr uint32_t kSpinLimit = 1u << 22;   // default: ~0.3 s of polling before a barrier gives up

// Waves per SIMD the persistent kernel is compiled for: 6 (at most 80 VGPRs).
// Its stage bodies are inlined into one function, so its register count is
// the largest any stage needs beside what stays live across stages.
constexpr int kBatchWavesPerSimd = 6;

// Stage bodies are lambdas handed to the tile loops; inlined, the stages
// share one register allocation instead of each call taking the full budget.
#define DFSPH_INLINE __attribute__((always_inline))

// The batch and scene descriptors are written by the host and only read by
// the kernel.  Read through the constant address space (scalar loads that
// may be repeated rather than kept live), the buffer pointers loaded from
// them are known to be global, so their accesses take global instructions
// with a scalar base instead of flat ones with 64-bit per-lane addresses.
// The view is made from the address as an integer so that no pair of
// address-space casts folds back into a generic pointer.
template <class T>
using kptr = const __attribute__((address_space(4))) T*;
template <class T>
__device__ __forceinline__ kptr<T> kview(const T* p) {
    return reinterpret_cast<kptr<T>>(reinterpret_cast<uintptr_t>(p));
}

// Scene descriptor read by every tile of the scene (wave-uniform).
struct SceneDesc {
    SimConst k;        // k.n = end slot of the scene (p_off + n)
    SimConst k_psi;    // the same with bscale = rho0 (pressure scaled by the rest density)
    Buffers b;         // the packed buffers with this scene's status and cell tables in place
    Grid g;
    uint32_t p_off, n, tile_off, ntiles;
    uint32_t cell_off, cells, b_off, nb;
    float two_r;       // 2r, for the CFL step
    uint32_t pad[3];
};

// Barrier and loop control, beside the per-scene DeviceStatus array.
struct BatchCtl {
    uint32_t count;       // arrivals at the current barrier
    uint32_t gen;         // barrier generation
    uint32_t abort;       // 1: a barrier timed out
    uint32_t keys_ready;  // 1: the previous substep left the keys and cell counts of this one
    uint32_t done[2];     // scenes whose solve is done (per solver)
    uint32_t groups;      // XCDs the launch's workgroups run on
    uint32_t pad1;
    // Per XCD x (hardware id), one 128-byte line at [x * kXLine]: arrivals at
    // the current barrier, then the number of workgroups running there.
    uint32_t xline[16 * 32];
};
constexpr uint32_t kXLine = 32;

struct BatchDev {
    const SceneDesc* scenes;
    const uint32_t* tile_scene;   // scene of each particle tile
    const uint32_t* b_off;        // boundary offsets, S + 1 (init only)
    uint32_t S, T, N, cells;
    uint32_t cell_tiles;          // scan tiles over the packed cells
    uint32_t xcd_map;             // 1: contiguous tile ranges per XCD
    uint32_t spin_limit;          // polls of a barrier wait before the substep gives up
    Buffers b;                    // packed; b.status = DeviceStatus[S]
    BatchCtl* ctl;
};

// Per-step policy (kernel argument; the host re-captures the graph when it changes).
struct BatchPol {
    float dt_fixed, cfl, dt_max;
    int reference, warm, div_enabled, psi;
    int min_it[2], max_it[2];
    float tol[2];
    float res_bound[2];   // bound on the largest per-particle residual (0: none)
};

// a[sv] by select: a run-time index into the policy argument would keep a
// per-lane copy of it in scratch memory.
template <class T>
__device__ __forceinline__ T pick(const T (&a)[2], int sv) { return sv ? a[1] : a[0]; }

template <class T>
__device__ __forceinline__ T ld_rlx(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Hardware id of the XCD (XCC) the calling wave runs on, 0-15: HW_REG_XCC_ID
// (hardware register 20 of the gfx940 family), bits 3:0.
__device__ __forceinline__ uint32_t hw_xcc_id() {
    return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u;
}

// Per-workgroup barrier phase: 0 until the launch's first barrier is passed.
__device__ __forceinline__ int& barrier_phase() {
    __shared__ int phase;
    return phase;
}

// Registers the calling workgroup on its XCD (thread 0, before any barrier).
__device__ __forceinline__ void barrier_register(BatchCtl* ctl) {
    if (threadIdx.x == 0) {
        barrier_phase() = 0;
        uint32_t* pop = &ctl->xline[hw_xcc_id() * kXLine + 1u];
        if (__hip_atomic_fetch_add(pop, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u)
            __hip_atomic_fetch_add(&ctl->groups, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Grid-wide barrier with a bounded wait; false once any barrier timed out.
// Every wave first waits for its own stores to reach its XCD's L2.  After the
// launch's first barrier (flat: every workgroup releases and arrives on the
// global counter; it also makes every workgroup's XCD registration visible),
// barriers are two-level: one thread per workgroup arrives on its XCD's
// counter; the last of the XCD's workgroups writes that L2 back (agent-scope
// release -- it holds the stores of every local arrival) and arrives on the
// global counter, whose last arrival opens the next generation.  Either way
// each workgroup then invalidates its caches (agent-scope acquire).  Each
// fence is a cache-wide operation that the XCD serialises, so their number
// sets the barrier's cost; grouping by the hardware XCD id keeps the release
// correct however workgroups are placed.
__device__ __forceinline__ bool grid_sync(const BatchDev& d) {
    BatchCtl* ctl = d.ctl;
    __shared__ int ok;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        int res = 1;
        const uint32_t gen = ld_rlx(&ctl->gen);
        bool global = true;                    // this workgroup arrives on the global counter
        uint32_t arrivals = gridDim.x;
        if (barrier_phase() != 0) {
            uint32_t* xl = &ctl->xline[hw_xcc_id() * kXLine];
            const uint32_t members = ld_rlx(xl + 1);
            global = __hip_atomic_fetch_add(xl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u == members;
            if (global) __hip_atomic_store(xl, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            arrivals = ld_rlx(&ctl->groups);
        }
        bool wait = !global;
        if (global) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            const uint32_t arrived =
                __hip_atomic_fetch_add(&ctl->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
            if (arrived == arrivals) {
                __hip_atomic_store(&ctl->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(&ctl->gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                wait = true;
            }
        }
        if (wait) {
            uint32_t spins = 0;
            while (ld_rlx(&ctl->gen) == gen) {
                if (ld_rlx(&ctl->abort)) { res = 0; break; }
                if (++spins > d.spin_limit) {
                    __hip_atomic_store(&ctl->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    res = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        barrier_phase() = 1;
        ok = res && !ld_rlx(&ctl->abort);
    }
    __syncthreads();
    return ok != 0;
}

// Visit units [0, count) of a stage: with xcd_map, the workgroups of one XCD
// (workgroup w runs on XCD w % 8) take one contiguous eighth of the units, so
// each XCD's L2 holds one coherent slab of the packed arrays.
template <class F>
__device__ __forceinline__ void for_units(const BatchDev& d, uint32_t count, F&& fn) {
    const uint32_t w = blockIdx.x, G = gridDim.x;
    if (d.xcd_map) {
        const uint32_t xcd = w & 7u, local = w >> 3, per = G >> 3;
        const uint32_t lo = uint32_t((uint64_t(count) * xcd) >> 3);
        const uint32_t hi = uint32_t((uint64_t(count) * (xcd + 1u)) >> 3);
        for (uint32_t u = lo + local; u < hi; u += per) fn(u);
    } else {
        for (uint32_t u = w; u < count; u += G) fn(u);
    }
}

// A tile's scene: constants, grid and buffers are read in place from the
// scene's descriptor (wave-uniform addresses, so scalar loads); a per-lane
// copy of these structs would live in scratch memory.
struct TileCtx {
    uint32_t s, i;
    const SimConst& k;
    const Grid& g;
    const Buffers& b;
};

// Particle tiles: fn(tile, ctx) with the scene's constants, its cell tables
// and its status in place.
template <class F>
__device__ __forceinline__ void for_tiles(const BatchDev& d, const BatchPol& pol, F&& fn) {
    for_units(d, d.T, [&](uint32_t t) DFSPH_INLINE {
        const uint32_t s = kview(d.tile_scene)[t];
        const SceneDesc& sd = *(const SceneDesc*)(kview(d.scenes) + s);
        TileCtx c{s, sd.p_off + (t - sd.tile_off) * kTile + threadIdx.x, pol.psi ? sd.k_psi : sd.k, sd.g, sd.b};
        fn(t, c);
    });
}

// Fixed-order sum of p[0..count) (as the single-scene finish kernel does it),
// returned to every thread of the workgroup.
__device__ __forceinline__ double ordered_sum_all(const double* __restrict__ p, uint32_t count) {
    __shared__ double tot;
    const double t = sb::ordered_sum(p, count);
    if (threadIdx.x == 0) tot = t;
    __syncthreads();
    const double r = tot;
    __syncthreads();
    return r;
}

// Done flag of solve sv of a scene, read once per workgroup and broadcast, so
// every thread takes the same branch even while the scene's first tile is
// recording a new status in the same stage.
__device__ __forceinline__ bool done_uniform(const DeviceStatus* st, int sv) {
    __shared__ int flag;
    if (threadIdx.x == 0) flag = ld_rlx(&st->solve[sv].done);
    __syncthreads();
    const bool r = flag != 0;
    __syncthreads();
    return r;
}

// done_at of solve sv of a scene (0 while undecided, else the iteration count
// of the decision), read once per workgroup and broadcast.  One 32-bit word,
// so a read racing the stage that records the decision sees 0 or that stage's
// count, never a torn value.
__device__ __forceinline__ int done_at_uniform(const DeviceStatus* st, int sv) {
    __shared__ int v;
    if (threadIdx.x == 0) v = ld_rlx(&st->solve[sv].done_at);
    __syncthreads();
    const int r = v;
    __syncthreads();
    return r;
}

// Convergence of the tile's scene after `it` (>= 1) completed iterations of
// solve sv: the scene's per-tile partials summed in the same fixed order as
// the single-scene decision, recorded by the scene's first tile (the caller)
// during the pressure stage, so no separate finish stage (and barrier) is
// needed.  Called uniformly by the whole workgroup.
__device__ __forceinline__ void scene_decide(const BatchDev& d, const BatchPol& pol, int sv, const TileCtx c,
                                             uint32_t t, int it) {
    DeviceStatus* st = c.b.status;
    const SceneDesc& sd = d.scenes[c.s];
    const double total = ordered_sum_all(d.b.partials + sd.tile_off, sd.ntiles);
    __shared__ float rmax_s;
    const float rm = sb::max_bits(d.b.res_max_part + sd.tile_off, sd.ntiles);
    if (threadIdx.x == 0) rmax_s = rm;
    __syncthreads();
    const float rmax = rmax_s;
    __syncthreads();
    SolveStatus q{};
    sb::solve_update(q, total, rmax, sd.n, ld_rlx(&st->dt), sv, it - 1, pol.reference, pick(pol.min_it, sv),
                     pick(pol.max_it, sv), pick(pol.tol, sv), pick(pol.res_bound, sv));
    if (t == sd.tile_off && threadIdx.x == 0) {
        st->solve[sv] = q;
        if (q.done) atomicAdd(&d.ctl->done[sv], 1u);
    }
}

// Jacobi loop of solve `sv` for every scene; false if a barrier timed out.
// Pressure stage it takes the decision on iteration it - 1 and still runs for
// the scenes it decides, so on return every scene is decided and its acc4
// holds the acceleration of its final pressure (the velocity update is then
// pointwise); a scene decided at an earlier stage skips the stage.  Stage
// max_iter decides every scene left.
__device__ __forceinline__ bool solve_all(const BatchDev& d, const BatchPol& pol, int sv) {
    using namespace sb;
    const int mx = pick(pol.max_it, sv);
    for (int it = 0;; ++it) {
        for_tiles(d, pol, [&](uint32_t t, TileCtx c) DFSPH_INLINE {
            const int da = done_at_uniform(c.b.status, sv);
            if (da != 0 && da < it) return;
            // Only the scene's first tile takes (and records) the decision: the
            // stage runs for the scene either way, and the next stage reads it.
            if (it >= 1 && da == 0 && t == d.scenes[c.s].tile_off) scene_decide(d, pol, sv, c, t, it);
            pressure_accel_one<DFSPH_GRID_FINE, kFmtMask>(c.k, c.g, c.b, c.i);
        });
        if (!grid_sync(d)) return false;
        if (ld_rlx(&d.ctl->done[sv]) >= d.S || it >= mx) return true;
        for_tiles(d, pol, [&](uint32_t t, TileCtx c) DFSPH_INLINE {
            if (done_uniform(c.b.status, sv)) return;
            const float dt = ld_rlx(&c.b.status->dt);
            const double e = jacobi_one<DFSPH_GRID_FINE, kFmtMask>(c.k, c.g, c.b, c.i, sv, pol.reference, dt);
            const double bs = block_sum(e);
            const uint32_t mx = block_max_bits(__float_as_uint(float(e)));
            if (threadIdx.x == 0) {
                d.b.partials[t] = bs;
                d.b.res_max_part[t] = mx;
            }
        });
        if (!grid_sync(d)) return false;
    }
    return true;
}

// The batch descriptor and the policy are read from device memory: a
// by-value kernel argument whose address is taken (every helper takes them by
// reference) is copied to scratch memory per lane.
__global__ void __launch_bounds__(kTile, kBatchWavesPerSimd) k_batch_substep(const BatchDev* __restrict__ dp,
                                                         const BatchPol* __restrict__ pp) {
    using namespace sb;
    const BatchDev& d = *(const BatchDev*)kview(dp);
    const BatchPol& pol = *(const BatchPol*)kview(pp);
    barrier_register(d.ctl);
    const Buffers& B = d.b;
    __shared__ float4 stage_lds[kTile / 64][kStageCap];   // density pass: one candidate window per wave
    // Keys and cell counts come from the previous substep's integration stage,
    // except on the first substep (read by every workgroup before the first
    // barrier; cleared after it, set again at the end).
    const bool need_keys = ld_rlx(&d.ctl->keys_ready) == 0u;
    // Step start: each scene's dt from its CFL condition (needed after the hash).
    if (blockIdx.x == 0)
        for (uint32_t s = threadIdx.x; s < d.S; s += blockDim.x)
            begin_step_one(B.status + s, pol.dt_fixed, pol.cfl * d.scenes[s].two_r, pol.dt_max);

    // ---- hash: (keys + histogram), scan, scatter, per-cell order + reorder ----
    if (need_keys) {
        for_tiles(d, pol, [&](uint32_t, TileCtx c) DFSPH_INLINE {
            const SceneDesc& sd = d.scenes[c.s];
            hb::keys_count<DFSPH_GRID_FINE>(sd.g, sd.cell_off, static_cast<const float4*>(B.x_stage), c.i,
                                            c.i < sd.k.n, B.keys, B.cell_count);
        });
        if (!grid_sync(d)) return;
    }
    for_units(d, d.cell_tiles, [&](uint32_t ct) DFSPH_INLINE { hb::scan_tile_sum(d.cells, B.cell_count, B.scan_blocks, ct); });
    if (!grid_sync(d)) return;
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) __hip_atomic_store(&d.ctl->keys_ready, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        hb::scan_tile_sums(d.cell_tiles, B.scan_blocks);
    }
    if (!grid_sync(d)) return;
    for_units(d, d.cell_tiles, [&](uint32_t ct) DFSPH_INLINE {
        hb::scan_apply(d.cells, B.cell_count, B.scan_blocks, B.cell_start, ct);
    });
    if (!grid_sync(d)) return;
    for_tiles(d, pol, [&](uint32_t, TileCtx c) DFSPH_INLINE {
        hb::scatter_one(c.i, c.i < c.k.n, B.keys, B.cell_start, B.cell_count, B.perm, B.skey,
                        hb::CellKey{B.id_stage, static_cast<const float4*>(B.x_stage)});
    });
    if (!grid_sync(d)) return;
    // One thread per cell orders the cell's slots by (x, original index) (keys
    // written by the scatter); then
    // one thread per slot gathers its state from the stage arrays (coalesced
    // writes; scenes' slots are packed, so every slot below N is a particle).
    for_units(d, (d.cells + kTile - 1) / kTile, [&](uint32_t u) DFSPH_INLINE {
        const uint32_t cell = u * kTile + threadIdx.x;
        if (cell >= d.cells) return;
        hb::cell_order_one<8>(cell, B.cell_start, B.cell_count, B.perm, B.id_stage,
                           static_cast<const float4*>(B.x_stage), B.skey);
    });
    if (!grid_sync(d)) return;
    for_units(d, (d.N + kTile - 1) / kTile, [&](uint32_t u) DFSPH_INLINE {
        const uint32_t slot = u * kTile + threadIdx.x;
        if (slot < d.N) reorder_one(B, slot);
    });
    if (!grid_sync(d)) return;

    // ---- density, alpha, divergence source (+ start of the divergence solve) --
    for_tiles(d, pol, [&](uint32_t, TileCtx c) DFSPH_INLINE {
        uint32_t unused = 0;
        density_alpha_one<DFSPH_GRID_FINE, kFmtMask>(c.k, c.g, c.b, c.i, nullptr, unused, stage_lds[threadIdx.x >> 6]);
        if (c.i >= c.k.n) return;
        if (pol.div_enabled)
            begin_solve_one(c.k, c.b, c.i, 0, pol.warm, ld_rlx(&c.b.status->dt), ld_rlx(&c.b.status->dt_prev));
        else
            B.pdiv_stage[c.i] = 0.0f;
    });
    if (!grid_sync(d)) return;

    if (pol.div_enabled) {
        if (!solve_all(d, pol, 0)) return;
        for_tiles(d, pol, [&](uint32_t t, TileCtx c) DFSPH_INLINE {
            accel_apply_one<DFSPH_GRID_FINE, kFmtMask, true>(c.k, c.g, c.b, c.i, 0, ld_rlx(&c.b.status->dt),
                                                      B.pdiv_stage, nullptr);
        });
        if (!grid_sync(d)) return;
    }

    // ---- forces, density source (+ start of the density solve) -------------
    for_tiles(d, pol, [&](uint32_t, TileCtx c) DFSPH_INLINE {
        forces_one<DFSPH_GRID_FINE, kFmtMask, false>(c.k, c.g, c.b, c.i, ld_rlx(&c.b.status->dt));
    });
    if (!grid_sync(d)) return;
    for_tiles(d, pol, [&](uint32_t, TileCtx c) DFSPH_INLINE {
        const float dt = ld_rlx(&c.b.status->dt);
        density_adv_one<DFSPH_GRID_FINE, kFmtMask>(c.k, c.g, c.b, c.i, dt);
        begin_solve_one(c.k, c.b, c.i, 1, pol.warm, dt, ld_rlx(&c.b.status->dt_prev));
    });
    if (!grid_sync(d)) return;
    if (!solve_all(d, pol, 1)) return;

    // ---- final acceleration, integration, next substep's keys, max |v| -------
    for_tiles(d, pol, [&](uint32_t t, TileCtx c) DFSPH_INLINE {
        const StepPart p = accel_apply_one<DFSPH_GRID_FINE, kFmtMask, true>(
            c.k, c.g, c.b, c.i, 1, ld_rlx(&c.b.status->dt), B.pden_stage, B.id_stage);
        const SceneDesc& sd = d.scenes[c.s];
        hb::keys_count_at<DFSPH_GRID_FINE>(sd.g, sd.cell_off, p.x, c.i, c.i < sd.k.n, B.keys, B.cell_count);
        const StepPart q = block_step_part(p);
        if (threadIdx.x == 0) {
            B.step_part[2 * t] = q.v2bits;
            B.step_part[2 * t + 1] = q.bad;
        }
    });
    if (!grid_sync(d)) return;
    for (uint32_t s = blockIdx.x; s < d.S; s += gridDim.x) {
        const SceneDesc& sd = d.scenes[s];
        step_finish_one(B.status + s, B.step_part + 2 * sd.tile_off, sd.ntiles);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_store(&d.ctl->keys_ready, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Clears the barrier and loop counters before each persistent launch (a
// kernel node rather than a memset node in the captured graph).
__global__ void k_batch_reset(BatchCtl* ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->count = 0u;
        ctl->abort = 0u;
        ctl->done[0] = 0u;
        ctl->done[1] = 0u;
    }
    if (threadIdx.x == 0 && blockIdx.x == 0) ctl->groups = 0u;
    if (blockIdx.x == 0 && threadIdx.x < 16u) {
        ctl->xline[threadIdx.x * kXLine] = 0u;
        ctl->xline[threadIdx.x * kXLine + 1u] = 0u;
    }
}

__device__ __forceinline__ uint32_t boundary_scene(const BatchDev& d, uint32_t i) {
    uint32_t lo = 0, hi = d.S;   // b_off[lo] <= i < b_off[hi]
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (d.b_off[mid] <= i) lo = mid; else hi = mid;
    }
    return lo;
}

// Boundary keys for the packed samples (init only).
__global__ void __launch_bounds__(256) k_batch_bkeys(BatchDev d, const float4* __restrict__ bstage,
                                                     uint32_t nb_total, uint32_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < nb_total;
    uint32_t s = valid ? boundary_scene(d, i) : 0u;
    const SceneDesc& sd = d.scenes[s];
    hb::keys_count<DFSPH_GRID_FINE>(sd.g, sd.cell_off, bstage, i, valid, keys, d.b.cell_count);
}

// V_b per packed sample (compute_boundary_volume.comp.glsl:55-92), then the
// sample index made local to its scene.
__global__ void __launch_bounds__(256) k_batch_bvolume(BatchDev d, uint32_t nb_total) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb_total) return;
    const uint32_t s = boundary_scene(d, i);
    const SceneDesc& sd = d.scenes[s];
    const float4* __restrict__ bpos = static_cast<const float4*>(d.b.bpos4);
    const float4 pi = bpos[i];
    float delta = sb::kernel_w(0.0f, sd.k.knorm);
    for_each_candidate<DFSPH_GRID_FINE>(sd.g, d.b.bcell_start + sd.cell_off, pi.x, pi.y, pi.z, [&](uint32_t j) {
        if (j == i) return;
        const float4 pj = bpos[j];
        const float r2 = dist2_exact(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
        if (!(r2 < sd.k.r2_q1)) return;
        const float inv_r = sb::rsq(r2);
        const float q = r2 > 0.0f ? r2 * inv_r * sd.k.inv_h : 0.0f;
        delta += sb::kernel_w(q, sd.k.knorm);
    });
    d.b.bvol[i] = 1.0f / fmaxf(delta, 1.0e-9f);
}

__global__ void __launch_bounds__(256) k_batch_bfinish(BatchDev d, uint32_t nb_total) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb_total) return;
    static_cast<float4*>(d.b.bpos4)[i].w = d.b.bvol[i];
    d.b.bids[i] -= d.b_off[boundary_scene(d, i)];
}

#define BATCH_HIP(call)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return set_error(DFSPH_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline uint32_t blocks_for(uint64_t n, uint32_t t) { return uint32_t((n + t - 1) / t); }

}  // namespace

struct dfsph_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint32_t S = 0, N = 0, NB = 0, T = 0, cells = 0, cell_tiles = 0;
    uint32_t grid = 0, max_grid = 0, flags = 0;
    uint32_t spin_limit = kSpinLimit;
    // Set once a substep's grid barrier gave up: that substep stopped between
    // stages (cell counts, cursors and the keys flag half-updated), so no
    // further substep may run on this state.
    bool failed = false;
    std::vector<dfsph_params> params;
    std::vector<SceneDesc> desc;
    std::vector<uint32_t> b_off_h;
    SceneDesc* d_desc = nullptr;
    BatchDev* d_dev = nullptr;         // dev(), uploaded once the batch is built
    BatchPol* d_pol = nullptr;         // the policy of the enqueued substeps
    BatchPol up_pol{};                 // host copy of *d_pol (the upload's source)
    bool pol_up = false;
    uint32_t* d_tile_scene = nullptr;
    uint32_t* d_b_off = nullptr;
    Buffers b{};
    BatchCtl* ctl = nullptr;
    void* h_state = nullptr;           // pinned: DeviceStatus[S] + BatchCtl
    size_t state_bytes = 0;
    bool stepped = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    BatchPol graph_pol{};
    bool have_graph = false;
    void* export_tmp = nullptr;
    size_t export_tmp_bytes = 0;
    std::vector<void*> allocs;

    int alloc(void** p, size_t bytes) {
        if (hipMalloc(p, bytes ? bytes : 16) != hipSuccess)
            return set_error(DFSPH_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        allocs.push_back(*p);
        return DFSPH_OK;
    }
    template <class T>
    int alloc_n(T** p, size_t count) { return alloc(reinterpret_cast<void**>(p), count * sizeof(T)); }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        have_graph = false;
    }
    ~dfsph_batch() {
        drop_graph();
        for (void* p : allocs) (void)hipFree(p);
        if (h_state) (void)hipHostFree(h_state);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
    BatchDev dev() const {
        BatchDev d{};
        d.scenes = d_desc;
        d.tile_scene = d_tile_scene;
        d.b_off = d_b_off;
        d.S = S; d.T = T; d.N = N; d.cells = cells; d.cell_tiles = cell_tiles;
        d.xcd_map = (flags & DFSPH_BATCH_NO_XCD_MAP) ? 0u : 1u;
        d.spin_limit = spin_limit;
        d.b = b;
        d.ctl = ctl;
        return d;
    }
    const DeviceStatus* h_status() const { return static_cast<const DeviceStatus*>(h_state); }
    const BatchCtl* h_ctl() const {
        return reinterpret_cast<const BatchCtl*>(static_cast<const char*>(h_state) + sizeof(DeviceStatus) * S);
    }
    // The launches of one substep: counter reset, persistent kernel, state readback.
    int enqueue() {
        DFSPH_LAUNCH(k_batch_reset, dim3(1), dim3(64), 0, stream, ctl);
        // A cooperative launch: the runtime refuses a grid whose workgroups
        // cannot all be resident at once instead of letting the grid
        // barriers wait on workgroups that never start.
        const BatchDev* dv = static_cast<const BatchDev*>(d_dev);
        const BatchPol* pv = static_cast<const BatchPol*>(d_pol);
        void* kargs[] = {&dv, &pv};
        BATCH_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_batch_substep), dim3(grid), dim3(kTile),
                                             kargs, 0, stream));
        BATCH_HIP(hipGetLastError());
        BATCH_HIP(hipMemcpyAsync(h_state, b.status, state_bytes, hipMemcpyDeviceToHost, stream));
        return DFSPH_OK;
    }
};

namespace {

BatchPol make_pol(float dt, const dfsph_solver_policy& p) {
    BatchPol q{};
    const bool reference = p.mode == DFSPH_POLICY_REFERENCE;
    q.dt_fixed = dt;
    q.cfl = p.cfl;
    q.dt_max = p.dt_max;
    q.reference = reference ? 1 : 0;
    q.warm = reference ? DFSPH_WARM_ZERO : p.warm_start;
    q.div_enabled = p.div_enabled ? 1 : 0;
    q.psi = p.boundary_psi < 0 ? (reference ? 0 : 1) : p.boundary_psi;
    const int mn[2] = {p.min_iter_div, p.min_iter_density};
    const int mx[2] = {p.max_iter_div, p.max_iter_density};
    const float tl[2] = {p.tol_div, p.tol_density};
    for (int s = 0; s < 2; ++s) {
        if (reference) {   // fixed max(min_iter, 1) iterations (main.cpp:2774, :2830)
            q.min_it[s] = q.max_it[s] = std::max(mn[s], 1);
            q.tol[s] = 0.0f;
        } else {
            q.min_it[s] = mn[s];
            q.max_it[s] = mx[s];
            q.tol[s] = tl[s];
            // Same rule as the single-scene step; a batch has no redo from
            // zero, so a scene that misses it at max_iter is reported as not
            // converged.
            const float bound = p.residual_bound == 0.0f ? 25.0f * tl[s] : p.residual_bound;
            q.res_bound[s] = std::max(bound, 0.0f);
        }
    }
    return q;
}

}  // namespace

extern "C" {

int dfsph_batch_create(const dfsph_batch_config* cfg, const dfsph_params* params,
                       const dfsph_scene* scenes, uint32_t count, dfsph_batch** out) {
    if (!out) return set_error(DFSPH_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!cfg || !params || !scenes) return set_error(DFSPH_E_INVALID, "NULL argument");
    if (cfg->abi_version != DFSPH_ABI_VERSION)
        return set_error(DFSPH_E_INVALID, "abi_version mismatch");
    if (count == 0) return set_error(DFSPH_E_INVALID, "a batch needs at least one scene");
    if (cfg->workgroups % 8u) return set_error(DFSPH_E_INVALID, "workgroups must be a multiple of 8");
    uint64_t n_total = 0, cells_total = 0;
    std::vector<Derived> der(count);
    std::vector<Grid> grids(count);
    for (uint32_t s = 0; s < count; ++s) {
        if (int rc = check_params(&params[s])) return rc;
        if (params[s].grid != DFSPH_GRID_FINE)
            return set_error(DFSPH_E_INVALID, "batched scenes use the FINE grid");
        if (scenes[s].count == 0 || !scenes[s].positions)
            return set_error(DFSPH_E_INVALID, "scene " + std::to_string(s) + " is empty");
        if (scenes[s].count > params[s].capacity)
            return set_error(DFSPH_E_CAPACITY, "scene " + std::to_string(s) + " exceeds its capacity");
        der[s] = derive(params[s]);
        grids[s] = make_grid(params[s], der[s], DFSPH_GRID_FINE);
        n_total += scenes[s].count;
        cells_total += grids[s].cells;
    }
    if (n_total > 0x7FFFFFFFull || cells_total > 0x7FFFFFFFull)
        return set_error(DFSPH_E_CAPACITY, "batch too large");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_error(DFSPH_E_DEVICE, "no HIP device available");
    if (cfg->device < 0 || cfg->device >= ndev) return set_error(DFSPH_E_DEVICE, "device ordinal out of range");
    hipDeviceProp_t prop{};
    BATCH_HIP(hipGetDeviceProperties(&prop, cfg->device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(DFSPH_E_DEVICE, std::string("device is ") + prop.gcnArchName +
                                             "; this library is built for gfx950 only");
    BATCH_HIP(hipSetDevice(cfg->device));

    auto* c = new (std::nothrow) dfsph_batch();
    if (!c) return set_error(DFSPH_E_NOMEM, "batch allocation failed");
    auto bail = [&](int rc) { delete c; return rc; };
    c->device = cfg->device;
    c->flags = cfg->flags;
    c->spin_limit = cfg->spin_limit ? cfg->spin_limit : kSpinLimit;
    c->S = count;
    c->params.assign(params, params + count);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return bail(set_error(DFSPH_E_DEVICE, "stream/event creation failed"));

    // Layout: packed particles, tiles, cells and boundary samples.
    std::vector<std::vector<float>> bsamples(count);
    std::vector<uint32_t> tile_scene;
    c->desc.resize(count);
    c->b_off_h.assign(count + 1, 0u);
    uint32_t p_off = 0, cell_off = 0, b_off = 0;
    for (uint32_t s = 0; s < count; ++s) {
        if (scenes[s].boundary)
            bsamples[s] = boundary_samples(params[s].domain_min, params[s].domain_max,
                                           params[s].particle_radius, scenes[s].boundary_cap);
        const uint32_t n = scenes[s].count, nb = uint32_t(bsamples[s].size() / 3);
        SceneDesc& sd = c->desc[s];
        sd.k = make_sim_const(params[s], der[s], 1.0f, p_off + n, nb);
        sd.k.slot0 = p_off;
        sd.g = grids[s];
        sd.p_off = p_off;
        sd.n = n;
        sd.tile_off = uint32_t(tile_scene.size());
        sd.ntiles = blocks_for(n, kTile);
        sd.cell_off = cell_off;
        sd.cells = grids[s].cells;
        sd.b_off = b_off;
        sd.nb = nb;
        sd.two_r = 2.0f * params[s].particle_radius;
        for (uint32_t t = 0; t < sd.ntiles; ++t) tile_scene.push_back(s);
        c->b_off_h[s] = b_off;
        p_off += n;
        cell_off += sd.cells;
        b_off += nb;
    }
    c->b_off_h[count] = b_off;
    c->N = p_off;
    c->NB = b_off;
    c->T = uint32_t(tile_scene.size());
    c->cells = cell_off;
    c->cell_tiles = blocks_for(c->cells, hb::kScanTile);
    c->state_bytes = sizeof(DeviceStatus) * count + sizeof(BatchCtl);

    // Device buffers (packed).
    const size_t N = c->N;
    Buffers& b = c->b;
    int rc = DFSPH_OK;
    auto A4 = [&](void** p) { if (!rc) rc = c->alloc(p, N * 16); };
    auto AF = [&](float** p) { if (!rc) rc = c->alloc_n(p, N); };
    auto AU = [&](uint32_t** p, size_t cnt) { if (!rc) rc = c->alloc_n(p, cnt); };
    A4(&b.pos4); A4(&b.vel4); A4(&b.vel4b); A4(&b.acc4); A4(&b.gb4);
    A4(&b.x_stage); A4(&b.v_stage);
    AF(&b.rho); AF(&b.alpha); AF(&b.aor2); AF(&b.rho_dot); AF(&b.rho_adv); AF(&b.p);
    AF(&b.pdiv_prev); AF(&b.pden_prev); AF(&b.pdiv_stage); AF(&b.pden_stage);
    AU(&b.nbr, N); AU(&b.nbrb, N); AU(&b.ids, N); AU(&b.keys, N); AU(&b.id_stage, N); AU(&b.perm, N);
    if (!rc) rc = c->alloc_n(&b.skey, N);
    AU(&b.nrec, 12 * N);
    b.mask_stride = uint32_t(N);
    if (!rc) rc = c->alloc_n(&b.wave_cells, 2 * size_t(c->T) * (kTile / 64));
    AU(&b.cell_count, c->cells); AU(&b.cell_start, size_t(c->cells) + 1);
    AU(&b.bcell_start, size_t(c->cells) + 1);
    AU(&b.scan_blocks, size_t(c->cell_tiles) + 1);
    if (!rc) rc = c->alloc_n(&b.partials, c->T);
    if (!rc) rc = c->alloc_n(&b.res_max_part, c->T);
    if (!rc) rc = c->alloc_n(&b.step_part, 2 * size_t(c->T));
    if (!rc) rc = c->alloc(reinterpret_cast<void**>(&b.status), c->state_bytes);
    if (!rc) rc = c->alloc_n(&c->d_desc, count);
    if (!rc) rc = c->alloc_n(&c->d_tile_scene, c->T);
    if (!rc) rc = c->alloc_n(&c->d_b_off, size_t(count) + 1);
    if (!rc && c->NB) {
        if ((rc = c->alloc(&b.bpos4, size_t(c->NB) * 16))) {}
        else if ((rc = c->alloc_n(&b.bvol, c->NB))) {}
        else rc = c->alloc_n(&b.bids, c->NB);
    }
    if (rc) return bail(rc);
    c->ctl = reinterpret_cast<BatchCtl*>(reinterpret_cast<char*>(b.status) + sizeof(DeviceStatus) * count);
    if (hipHostMalloc(&c->h_state, c->state_bytes, hipHostMallocDefault) != hipSuccess)
        return bail(set_error(DFSPH_E_NOMEM, "pinned state allocation failed"));

    // Scene data: stage arrays in packed order; initial max |v| per scene.
    std::vector<float> x4(N * 4, 0.0f), v4(N * 4, 0.0f);
    std::vector<uint32_t> ids(N);
    std::vector<DeviceStatus> st(count);
    std::vector<uint8_t> seen;
    for (uint32_t s = 0; s < count; ++s) {
        const dfsph_scene& sc = scenes[s];
        const uint32_t o = c->desc[s].p_off, n = sc.count;
        float vmax2 = 0.0f;
        if (sc.ids) seen.assign(n, 0);
        for (uint32_t i = 0; i < n; ++i) {
            for (int a = 0; a < 3; ++a) {
                x4[4 * (o + i) + a] = sc.positions[3 * i + a];
                v4[4 * (o + i) + a] = sc.velocities ? sc.velocities[3 * i + a] : 0.0f;
            }
            const float vx = v4[4 * (o + i)], vy = v4[4 * (o + i) + 1], vz = v4[4 * (o + i) + 2];
            vmax2 = std::max(vmax2, (vx * vx + vy * vy) + vz * vz);
            const uint32_t id = sc.ids ? sc.ids[i] : i;
            if (sc.ids) {
                if (id >= n || seen[id]) return bail(set_error(DFSPH_E_INVALID, "scene " + std::to_string(s) +
                                                                                  ": ids must be a permutation"));
                seen[id] = 1;
            }
            ids[o + i] = id;
        }
        st[s] = DeviceStatus{};
        std::memcpy(&st[s].vmax2_bits, &vmax2, 4);
    }
    std::vector<char> state(c->state_bytes, 0);
    std::memcpy(state.data(), st.data(), sizeof(DeviceStatus) * count);
    BATCH_HIP(hipMemcpyAsync(b.x_stage, x4.data(), x4.size() * 4, hipMemcpyHostToDevice, c->stream));
    BATCH_HIP(hipMemcpyAsync(b.v_stage, v4.data(), v4.size() * 4, hipMemcpyHostToDevice, c->stream));
    BATCH_HIP(hipMemcpyAsync(b.id_stage, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, c->stream));
    BATCH_HIP(hipMemsetAsync(b.pdiv_stage, 0, N * 4, c->stream));
    BATCH_HIP(hipMemsetAsync(b.pden_stage, 0, N * 4, c->stream));
    BATCH_HIP(hipMemsetAsync(b.cell_count, 0, size_t(c->cells) * 4, c->stream));
    BATCH_HIP(hipMemcpyAsync(b.status, state.data(), c->state_bytes, hipMemcpyHostToDevice, c->stream));
    for (uint32_t s = 0; s < count; ++s) {
        SceneDesc& sd = c->desc[s];
        sd.k_psi = sd.k;
        sd.k_psi.bscale = sd.k.rho0;
        sd.b = b;
        sd.b.status = b.status + s;
        sd.b.cell_start = b.cell_start + sd.cell_off;
        sd.b.bcell_start = b.bcell_start + sd.cell_off;
        sd.b.wave_cells = b.wave_cells + 2 * size_t(sd.tile_off) * (kTile / 64);
    }
    BATCH_HIP(hipMemcpyAsync(c->d_desc, c->desc.data(), sizeof(SceneDesc) * count, hipMemcpyHostToDevice, c->stream));
    BATCH_HIP(hipMemcpyAsync(c->d_tile_scene, tile_scene.data(), tile_scene.size() * 4, hipMemcpyHostToDevice, c->stream));
    BATCH_HIP(hipMemcpyAsync(c->d_b_off, c->b_off_h.data(), c->b_off_h.size() * 4, hipMemcpyHostToDevice, c->stream));

    // Boundary: packed samples hashed into the packed cell tables, V_b per sample.
    if (c->NB) {
        std::vector<float> b4(size_t(c->NB) * 4, 0.0f);
        for (uint32_t s = 0; s < count; ++s)
            for (size_t q = 0; q < bsamples[s].size() / 3; ++q)
                for (int a = 0; a < 3; ++a) b4[4 * (c->b_off_h[s] + q) + a] = bsamples[s][3 * q + a];
        void* bstage = nullptr;
        uint32_t *bkeys = nullptr, *bperm = nullptr;
        if ((rc = c->alloc(&bstage, size_t(c->NB) * 16)) || (rc = c->alloc_n(&bkeys, c->NB)) ||
            (rc = c->alloc_n(&bperm, c->NB)))
            return bail(rc);
        BATCH_HIP(hipMemcpyAsync(bstage, b4.data(), b4.size() * 4, hipMemcpyHostToDevice, c->stream));
        const BatchDev d = c->dev();
        DFSPH_LAUNCH(k_batch_bkeys, dim3(blocks_for(c->NB, 256)), dim3(256), 0, c->stream, d,
                           static_cast<const float4*>(bstage), c->NB, bkeys);
        launch_scan(c->cells, b.cell_count, b.bcell_start, b.scan_blocks, c->stream);
        launch_scatter(c->NB, bkeys, b.bcell_start, b.cell_count, bperm, nullptr, nullptr, nullptr, c->cells, c->stream);
        SweepArgs a{};
        a.k.nb = c->NB;
        a.b = b;
        launch_reorder_boundary(a, bstage, bperm, c->stream);
        DFSPH_LAUNCH(k_batch_bvolume, dim3(blocks_for(c->NB, 256)), dim3(256), 0, c->stream, d, c->NB);
        DFSPH_LAUNCH(k_batch_bfinish, dim3(blocks_for(c->NB, 256)), dim3(256), 0, c->stream, d, c->NB);
        BATCH_HIP(hipGetLastError());
        BATCH_HIP(hipStreamSynchronize(c->stream));
        (void)hipFree(bstage); (void)hipFree(bkeys); (void)hipFree(bperm);
        c->allocs.erase(std::remove_if(c->allocs.begin(), c->allocs.end(), [&](void* p) {
            return p == bstage || p == bkeys || p == bperm; }), c->allocs.end());
    } else {
        BATCH_HIP(hipMemsetAsync(b.bcell_start, 0, (size_t(c->cells) + 1) * 4, c->stream));
    }

    // Persistent grid: every workgroup must be resident at once.
    int per_cu = 0;
    BATCH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_substep, kTile, 0));
    if (per_cu < 1) return bail(set_error(DFSPH_E_DEVICE, "persistent kernel does not fit on a CU"));
    c->max_grid = (uint32_t(per_cu) * uint32_t(prop.multiProcessorCount)) & ~7u;
    const uint32_t want = cfg->workgroups ? cfg->workgroups : c->max_grid;
    c->grid = std::max(8u, std::min(want, c->max_grid));
    if ((rc = c->alloc_n(&c->d_dev, 1)) || (rc = c->alloc_n(&c->d_pol, 1))) return bail(rc);
    const BatchDev dv = c->dev();
    BATCH_HIP(hipMemcpyAsync(c->d_dev, &dv, sizeof(BatchDev), hipMemcpyHostToDevice, c->stream));
    BATCH_HIP(hipStreamSynchronize(c->stream));
    *out = c;
    return DFSPH_OK;
}

int dfsph_batch_destroy(dfsph_batch* c) {
    if (!c) return DFSPH_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete c;
    return DFSPH_OK;
}

int dfsph_batch_trace_mark(dfsph_batch* c, uint32_t tag) {
    if (!c) return set_error(DFSPH_E_INVALID, "batch is NULL");
    if (hipSetDevice(c->device) != hipSuccess) return set_error(DFSPH_E_DEVICE, "hipSetDevice failed");
    launch_trace_mark(tag, c->stream);
    if (hipGetLastError() != hipSuccess) return set_error(DFSPH_E_DEVICE, "k_trace_mark launch failed");
    return DFSPH_OK;
}

int dfsph_batch_info(dfsph_batch* c, dfsph_batch_info_t* out) {
    if (!c || !out) return set_error(DFSPH_E_INVALID, "NULL argument");
    dfsph_batch_info_t r{};
    r.scenes = c->S;
    r.particles = c->N;
    r.boundary_particles = c->NB;
    r.tiles = c->T;
    r.cells = c->cells;
    r.workgroups = c->grid;
    r.max_workgroups = c->max_grid;
    *out = r;
    return DFSPH_OK;
}

int dfsph_batch_step(dfsph_batch* c, float dt, const dfsph_solver_policy* pol,
                     dfsph_step_stats* per_scene, dfsph_step_stats* total) {
    if (!c) return set_error(DFSPH_E_INVALID, "batch is NULL");
    if (c->failed)
        return set_error(DFSPH_E_DEVICE, "an earlier substep of this batch timed out at a grid barrier and left "
                                         "its state incomplete; destroy the batch");
    if (int rc = check_policy(pol)) return rc;
    if (!(dt >= 0.0f) || !std::isfinite(dt)) return set_error(DFSPH_E_INVALID, "dt must be finite and >= 0");
    if (pol->mode == DFSPH_POLICY_CONVERGED && pol->warm_start != DFSPH_WARM_ZERO &&
        pol->warm_start != DFSPH_WARM_PREVIOUS)
        return set_error(DFSPH_E_INVALID, "batched scenes take the ZERO or PREVIOUS warm start");
    if (pol->mode == DFSPH_POLICY_CONVERGED && pol->div_start != DFSPH_DIV_START_DEFAULT)
        return set_error(DFSPH_E_INVALID, "batched scenes start both solves as warm_start says (div_start 0)");
    BATCH_HIP(hipSetDevice(c->device));
    const BatchPol bp = make_pol(dt, *pol);
    if (!c->pol_up || std::memcmp(&bp, &c->up_pol, sizeof(bp)) != 0) {
        // The previous substep has finished (each step ends in a wait), so
        // neither the kernel nor the last upload still reads the source.
        c->up_pol = bp;
        BATCH_HIP(hipMemcpyAsync(c->d_pol, &c->up_pol, sizeof(BatchPol), hipMemcpyHostToDevice, c->stream));
        c->pol_up = true;
    }
    const bool use_graph = !(c->flags & DFSPH_BATCH_NO_GRAPH);
    if (use_graph && (!c->have_graph || std::memcmp(&bp, &c->graph_pol, sizeof(bp)) != 0)) {
        c->drop_graph();
        BATCH_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int erc = c->enqueue();
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(c->stream, &g);
        if (erc) { if (g) (void)hipGraphDestroy(g); return erc; }
        if (ec != hipSuccess) return set_error(DFSPH_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        c->graph = g;
        BATCH_HIP(hipGraphInstantiate(&c->exec, c->graph, nullptr, nullptr, 0));
        c->graph_pol = bp;
        c->have_graph = true;
    }
    BATCH_HIP(hipEventRecord(c->ev0, c->stream));
    if (use_graph) BATCH_HIP(hipGraphLaunch(c->exec, c->stream));
    else if (int rc = c->enqueue()) return rc;
    BATCH_HIP(hipEventRecord(c->ev1, c->stream));
    BATCH_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    BATCH_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stepped = true;
    if (c->h_ctl()->abort) {
        c->failed = true;
        return set_error(DFSPH_E_DEVICE, "a grid barrier of the persistent substep timed out "
                                         "(workgroups not co-resident); the step is incomplete and "
                                         "the batch refuses further substeps");
    }
    dfsph_step_stats tot{};
    tot.converged_div = tot.converged_density = 1;
    uint64_t bad = 0;
    for (uint32_t s = 0; s < c->S; ++s) {
        const DeviceStatus& h = c->h_status()[s];
        dfsph_step_stats r{};
        r.dt = h.dt;
        const bool div = pol->div_enabled != 0;
        r.iters_div = div ? h.solve[0].iters : 0;
        r.iters_density = h.solve[1].iters;
        r.converged_div = div ? h.solve[0].converged : 1;
        r.converged_density = h.solve[1].converged;
        r.err_div = div ? h.solve[0].err : 0.0;
        r.err_density = h.solve[1].err;
        float v2;
        std::memcpy(&v2, &h.vmax2_start, 4);
        r.max_speed = std::sqrt(double(v2));
        r.nonfinite = h.nonfinite;
        r.max_residual_div = div ? h.solve[0].res_max : 0.0f;
        r.max_residual_density = h.solve[1].res_max;
        r.gpu_ms = ms;
        r.launches = 2;
        r.host_syncs = 1;
        if (per_scene) per_scene[s] = r;
        tot.dt = s ? std::min(tot.dt, r.dt) : r.dt;
        tot.iters_div = std::max(tot.iters_div, r.iters_div);
        tot.iters_density = std::max(tot.iters_density, r.iters_density);
        tot.converged_div &= r.converged_div;
        tot.converged_density &= r.converged_density;
        tot.err_div = std::max(tot.err_div, r.err_div);
        tot.err_density = std::max(tot.err_density, r.err_density);
        tot.max_residual_div = std::max(tot.max_residual_div, r.max_residual_div);
        tot.max_residual_density = std::max(tot.max_residual_density, r.max_residual_density);
        tot.max_speed = std::max(tot.max_speed, r.max_speed);
        bad += h.nonfinite;
    }
    tot.nonfinite = bad;
    tot.gpu_ms = ms;
    tot.launches = 2;
    tot.host_syncs = 1;
    if (total) *total = tot;
    if (bad) return set_error(DFSPH_E_NONFINITE, std::to_string(bad) + " non-finite position/velocity components");
    if (pol->strict && (!tot.converged_div || !tot.converged_density))
        return set_error(DFSPH_E_NOT_CONVERGED, "a pressure solve stopped at max_iter above tolerance");
    return DFSPH_OK;
}

int dfsph_batch_field_size(dfsph_batch* c, uint32_t scene, int32_t field, size_t* bytes) {
    if (!c || !bytes) return set_error(DFSPH_E_INVALID, "NULL argument");
    if (scene >= c->S) return set_error(DFSPH_E_INVALID, "scene index out of range");
    const size_t n = c->desc[scene].n, nb = c->desc[scene].nb;
    switch (field) {
        case DFSPH_FIELD_POSITION: case DFSPH_FIELD_VELOCITY: case DFSPH_FIELD_PRESSURE_ACCEL:
            *bytes = n * 12; return DFSPH_OK;
        case DFSPH_FIELD_DENSITY: case DFSPH_FIELD_ALPHA_OVER_RHO2: case DFSPH_FIELD_DENSITY_ADV:
        case DFSPH_FIELD_DENSITY_CHANGE: case DFSPH_FIELD_PRESSURE_DIV: case DFSPH_FIELD_PRESSURE:
        case DFSPH_FIELD_ALPHA: case DFSPH_FIELD_NEIGHBOR_COUNT:
        case DFSPH_FIELD_BOUNDARY_NEIGHBOR_COUNT: case DFSPH_FIELD_ID:
            *bytes = n * 4; return DFSPH_OK;
        case DFSPH_FIELD_BOUNDARY_POSITION: *bytes = nb * 12; return DFSPH_OK;
        case DFSPH_FIELD_BOUNDARY_VOLUME: *bytes = nb * 4; return DFSPH_OK;
        default:
            return set_error(DFSPH_E_INVALID, "field " + std::to_string(field) +
                                                  " is not exported by batches (reference hash fields are single-scene only)");
    }
}

int dfsph_batch_export(dfsph_batch* c, uint32_t scene, int32_t field, void* dst, size_t bytes) {
    size_t need = 0;
    if (int rc = dfsph_batch_field_size(c, scene, field, &need)) return rc;
    if (!dst && need) return set_error(DFSPH_E_INVALID, "dst is NULL");
    if (bytes != need) return set_error(DFSPH_E_INVALID, "bytes must be " + std::to_string(need));
    const bool step_field = field != DFSPH_FIELD_POSITION && field != DFSPH_FIELD_VELOCITY &&
                            field != DFSPH_FIELD_BOUNDARY_POSITION &&
                            field != DFSPH_FIELD_BOUNDARY_VOLUME && field != DFSPH_FIELD_ID;
    if (step_field && !c->stepped)
        return set_error(DFSPH_E_INVALID, "field " + std::to_string(field) + " exists only after a step");
    if (!need) return DFSPH_OK;
    BATCH_HIP(hipSetDevice(c->device));
    if (c->export_tmp_bytes < need) {
        if (c->export_tmp) {
            (void)hipFree(c->export_tmp);
            c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), c->export_tmp), c->allocs.end());
            c->export_tmp = nullptr;
            c->export_tmp_bytes = 0;
        }
        if (int rc = c->alloc(&c->export_tmp, need)) return rc;
        c->export_tmp_bytes = need;
    }
    const SceneDesc& sd = c->desc[scene];
    const Buffers& b = c->b;
    const uint32_t o = sd.p_off, n = sd.n;
    const uint32_t* ids = b.id_stage + o;
    float* tf = static_cast<float*>(c->export_tmp);
    uint32_t* tu = static_cast<uint32_t*>(c->export_tmp);
    auto v4 = [&](void* p) { return static_cast<const char*>(p) + size_t(o) * 16; };
    switch (field) {
        case DFSPH_FIELD_POSITION: launch_to_original_vec(v4(b.x_stage), ids, n, tf, c->stream); break;
        case DFSPH_FIELD_VELOCITY: launch_to_original_vec(v4(b.v_stage), ids, n, tf, c->stream); break;
        case DFSPH_FIELD_DENSITY: launch_to_original_f32(b.rho + o, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_ALPHA_OVER_RHO2: launch_to_original_f32(b.aor2 + o, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_DENSITY_ADV: launch_to_original_f32(b.rho_adv + o, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_DENSITY_CHANGE: launch_to_original_f32(b.rho_dot + o, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_PRESSURE_DIV: launch_to_original_f32(b.pdiv_stage + o, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_PRESSURE: launch_to_original_f32(b.pden_stage + o, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_PRESSURE_ACCEL: launch_to_original_vec(v4(b.acc4), ids, n, tf, c->stream); break;
        case DFSPH_FIELD_ALPHA: launch_to_original_f32(b.alpha + o, ids, n, tf, c->stream); break;
        case DFSPH_FIELD_NEIGHBOR_COUNT: launch_to_original_u32(b.nbr + o, ids, n, tu, c->stream); break;
        case DFSPH_FIELD_BOUNDARY_NEIGHBOR_COUNT: launch_to_original_u32(b.nbrb + o, ids, n, tu, c->stream); break;
        case DFSPH_FIELD_ID: BATCH_HIP(hipMemcpyAsync(tu, ids, need, hipMemcpyDeviceToDevice, c->stream)); break;
        case DFSPH_FIELD_BOUNDARY_POSITION:
            launch_to_original_vec(static_cast<const char*>(b.bpos4) + size_t(sd.b_off) * 16, b.bids + sd.b_off,
                                   sd.nb, tf, c->stream);
            break;
        case DFSPH_FIELD_BOUNDARY_VOLUME:
            launch_to_original_w(static_cast<const char*>(b.bpos4) + size_t(sd.b_off) * 16, b.bids + sd.b_off,
                                 sd.nb, tf, c->stream);
            break;
        default: return set_error(DFSPH_E_INVALID, "unknown field");
    }
    BATCH_HIP(hipGetLastError());
    BATCH_HIP(hipMemcpyAsync(dst, c->export_tmp, need, hipMemcpyDeviceToHost, c->stream));
    BATCH_HIP(hipStreamSynchronize(c->stream));
    return DFSPH_OK;
}

}  // extern "C"
