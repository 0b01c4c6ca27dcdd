// sweeps.hip -- per-stage kernels of one DFSPH substep for one scene: thin
// wrappers over the per-particle bodies of sweep_bodies.h, one thread per
// particle over cell-sorted SoA state, one launch per stage.  Per-workgroup
// FP64 error partials are stored by tile index and summed in a fixed order by
// one workgroup (no global atomics; deterministic).
#include "sweep_bodies.h"
#include "hash_bodies.h"
#include "launch.h"

namespace dfsph {

namespace {

using namespace sb;

__device__ __forceinline__ uint32_t tile_of(int remap) {
    return xcd_tile(blockIdx.x, gridDim.x, remap);
}

// Capped at 72 VGPRs (7 waves per SIMD): the staged mask build holds the
// records' byte offsets and otherwise takes 79 (6 waves); the cap costs 4
// spilled VGPRs and runs the 1M density pass about 1.5 % faster.
template <int KIND, int FMT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7)))
k_density_alpha(SimConst k, Grid g, Buffers b, int remap, int warm_div) {
    __shared__ uint32_t list_lds[FMT == kFmtList ? kListSlots * 256 : 1];
    constexpr bool kStaged = KIND == DFSPH_GRID_FINE && FMT == kFmtMask;
    __shared__ float4 stage_lds[kStaged ? 4 : 1][kStaged ? kStageCap : 1];   // one window per wave
    const uint32_t i = tile_of(remap) * blockDim.x + threadIdx.x;
    uint32_t list_cnt = 0;
    density_alpha_one<KIND, FMT>(k, g, b, i, list_lds, list_cnt, kStaged ? stage_lds[threadIdx.x >> 6] : nullptr);
    // The divergence solve's start (zero, previous or caller pressure) needs
    // only this particle's new density: taken here instead of a launch of its own.
    if (warm_div >= 0) begin_solve_one(k, b, i, 0, warm_div, b.status->dt, b.status->dt_prev);
    if (FMT == kFmtList) {
        // Copy the staged lists out, one slot at a time across the workgroup.
        __shared__ uint32_t wg_max;
        if (threadIdx.x == 0) wg_max = 0;
        __syncthreads();
        if (list_cnt) atomicMax(&wg_max, list_cnt);
        __syncthreads();
        const uint32_t m = wg_max;
        for (uint32_t t = 0; t < m; ++t)
            if (i < k.n && t < list_cnt)
                b.nlist[size_t(t) * b.mask_stride + i] = list_lds[t * blockDim.x + threadIdx.x];
    }
}

// The block's windows of the LDS-staged walk (NE extra floats per slot).
template <int NE, bool ON>
struct WaveStage {
    static constexpr uint32_t kSlots = ON ? 4u * 2u * kWinCap<NE> : 1u;
    float4 a[kSlots];
    float2 b[kSlots];
    float c[NE == 4 && ON ? kSlots : 1u];
    __device__ StageWin<NE> win() { return StageWin<NE>{a, b, c, (threadIdx.x >> 6) * 2u * kWinCap<NE>}; }
};

template <int KIND, int FMT>
__global__ void __launch_bounds__(256) k_pressure_accel(SimConst k, Grid g, Buffers b, int remap,
                                                        int solver, int iter, SolveDecide d) {
    DeviceStatus* st = b.status;
    if (!d.decide) {
        if (st->solve[solver].done) return;
    } else {
        // Workgroup 0 may set done_at = iter during this launch; every
        // workgroup reads 0 or iter, and both mean "run".
        const int done_at = st->solve[solver].done_at;
        if (done_at != 0 && done_at < iter) return;
        if (iter >= 1 && blockIdx.x == 0 && done_at == 0) {
            const double total = ordered_sum(b.partials, gridDim.x);
            const float rmax = max_bits(b.res_max_part, gridDim.x);
            if (threadIdx.x == 0) {
                SolveStatus q = st->solve[solver];
                solve_update(q, total, rmax, k.n, st->dt, solver, iter - 1, d.reference, d.min_it, d.max_it,
                             d.tol, d.res_bound);
                st->solve[solver] = q;
                if (SolveMirror* mr = b.mirror) {
                    // Four tagged words, no fence (the reader matches tags).
                    uint64_t* w = mr->solve[solver].w;
                    const uint64_t tag = uint64_t(kPubTagBit | uint32_t(q.iters)) << 32;
                    const uint64_t eb = uint64_t(__double_as_longlong(q.err));
                    const uint64_t v[4] = {tag | uint32_t(q.done | (q.converged << 1)),
                                           tag | __float_as_uint(q.res_max), tag | uint32_t(eb),
                                           tag | uint32_t(eb >> 32)};
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        __hip_atomic_store(&w[u], v[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                }
            }
        }
    }
    pressure_accel_one<KIND, FMT>(k, g, b, tile_of(remap) * blockDim.x + threadIdx.x);
}

template <int KIND, int FMT, bool STAGED>
__global__ void __launch_bounds__(256) k_jacobi(SimConst k, Grid g, Buffers b, int remap,
                                                int solver, int reference) {
    if (b.status->solve[solver].done) return;
    const uint32_t tile = tile_of(remap);
    const uint32_t i = tile * blockDim.x + threadIdx.x;
    double e;
    if constexpr (STAGED) {
        __shared__ WaveStage<3, STAGED> st;
        e = jacobi_staged(k, g, b, i, solver, reference, b.status->dt, st.win());
    } else {
        e = jacobi_one<KIND, FMT>(k, g, b, i, solver, reference, b.status->dt);
    }
    const double bs = block_sum(e);
    const uint32_t mx = block_max_bits(__float_as_uint(float(e)));
    if (threadIdx.x == 0) {
        b.partials[tile] = bs;
        b.res_max_part[tile] = mx;
    }
}

// The convergence decision runs as its own one-workgroup launch: deciding in
// the last Jacobi workgroup instead needs a device-scope release fence per
// workgroup, which writes back the XCD's L2 (full of the sweep's dirty
// pressure lines) and cost far more than the launch.
// One workgroup: fixed-order FP64 sum of the per-tile partials of k_jacobi,
// then the convergence decision for the next launch to read.
__global__ void __launch_bounds__(256) k_solve_finish(DeviceStatus* st, const double* __restrict__ partials,
                                                      const uint32_t* __restrict__ res_max_part, uint32_t ntiles, uint32_t n, int solver, int iter,
                                                      int reference, int min_it, int max_it, float tol,
                                                      float res_bound) {
    if (st->solve[solver].done) return;
    const double total = ordered_sum(partials, ntiles);
    const float rmax = max_bits(res_max_part, ntiles);
    if (threadIdx.x == 0)
        solve_update(st->solve[solver], total, rmax, n, st->dt, solver, iter, reference, min_it, max_it, tol,
                     res_bound);
}

__global__ void __launch_bounds__(256) k_step_finish(DeviceStatus* st, const uint32_t* __restrict__ part,
                                                     uint32_t ntiles, StepMirror* mr, uint32_t ticket) {
    step_finish_block(st, part, ntiles, mr, ticket);
}

// keys, count (after the density solve, or null): the next cell build's keys
// and histogram from the integrated positions (its keys pass, fused).
template <int KIND, int FMT, bool FROM_ACC = false>
__global__ void __launch_bounds__(256) k_accel_apply(SimConst k, Grid g, Buffers b, int remap,
                                                     int solver, uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ count) {
    const uint32_t tile = tile_of(remap);
    const uint32_t i = tile * blockDim.x + threadIdx.x;
    const StepPart p = accel_apply_one<KIND, FMT, FROM_ACC>(k, g, b, i, solver, b.status->dt,
                                                            solver == 0 ? b.p_div : b.p_den, nullptr);
    if (solver == 1 && keys) hb::keys_count_at<KIND>(g, 0u, p.x, i, i < k.n, keys, count);
    if (solver == 1) {
        const StepPart t = block_step_part(p);
        if (threadIdx.x == 0) {
            b.step_part[2 * tile] = t.v2bits;
            b.step_part[2 * tile + 1] = t.bad;
        }
    }
}

template <int KIND, int FMT, bool STAGED, bool DIV>
__device__ __forceinline__ void forces_kernel(const SimConst& k, const Grid& g, const Buffers& b, int remap) {
    const uint32_t i = tile_of(remap) * blockDim.x + threadIdx.x;
    if constexpr (STAGED) {
        __shared__ WaveStage<4, STAGED> st;
        forces_staged<DIV>(k, g, b, i, b.status->dt, st.win());
    } else {
        forces_one<KIND, FMT, DIV>(k, g, b, i, b.status->dt);
    }
}
template <int KIND, int FMT, bool STAGED>
__global__ void __launch_bounds__(256) k_forces(SimConst k, Grid g, Buffers b, int remap) {
    forces_kernel<KIND, FMT, STAGED, false>(k, g, b, remap);
}
// The same applying the divergence solve's velocity update (final acceleration
// in acc4) as velocities are read.
template <int KIND, int FMT, bool STAGED>
__global__ void __launch_bounds__(256) k_forces_div(SimConst k, Grid g, Buffers b, int remap) {
    forces_kernel<KIND, FMT, STAGED, true>(k, g, b, remap);
}

// warm_den >= 0: the density solve's start (zero, previous or caller
// pressure) needs only this particle's density: taken here instead of a
// launch of its own.
template <int KIND, int FMT, bool STAGED>
__global__ void __launch_bounds__(256) k_density_adv(SimConst k, Grid g, Buffers b, int remap, int warm_den) {
    const uint32_t i = tile_of(remap) * blockDim.x + threadIdx.x;
    if constexpr (STAGED) {
        __shared__ WaveStage<3, STAGED> st;
        density_adv_staged(k, g, b, i, b.status->dt, st.win());
    } else {
        density_adv_one<KIND, FMT>(k, g, b, i, b.status->dt);
    }
    if (warm_den >= 0) begin_solve_one(k, b, i, 1, warm_den, b.status->dt, b.status->dt_prev);
}

__global__ void __launch_bounds__(256) k_begin_solve(SimConst k, Buffers b, int solver, int warm) {
    begin_solve_one(k, b, blockIdx.x * blockDim.x + threadIdx.x, solver, warm, b.status->dt,
                    b.status->dt_prev);
}

// Restart solve `solver` from zero pressure with a cleared status.
__global__ void __launch_bounds__(256) k_reset_solve(SimConst k, Buffers b, int solver) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) b.status->solve[solver] = SolveStatus{0.0, 0, 0, 0, 0.0f};
    if (i >= k.n) return;
    b.p[i] = 0.0f;
    static_cast<float4*>(b.pos4)[i].w = 0.0f;   // p / rho^2
}

// The gather into cell order; thread 0 of workgroup 0 also starts the step's
// status (dt by CFL from the previous step's max |v|), which nothing before
// the density pass reads.
__global__ void __launch_bounds__(256) k_reorder(Buffers b, uint32_t n, float dt_fixed, float cfl_2r,
                                                 float dt_max) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) begin_step_one(b.status, dt_fixed, cfl_2r, dt_max);
    if (k < n) reorder_one(b, k);
}

__global__ void __launch_bounds__(256) k_reorder_boundary(float4* __restrict__ dst,
                                                          uint32_t* __restrict__ bids,
                                                          const float4* __restrict__ src,
                                                          const uint32_t* __restrict__ perm,
                                                          uint32_t nb) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const uint32_t s = perm[k];
    const float4 x = src[s];
    dst[k] = make_float4(x.x, x.y, x.z, 0.0f);
    bids[k] = s;
}

// compute_boundary_volume.comp.glsl:55-92: V_b = 1 / max(W(0) + sum_{b' != b, q<1} W, 1e-9).
template <int KIND>
__global__ void __launch_bounds__(256) k_boundary_volume(SimConst k, Grid g, Buffers b) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k.nb) return;
    const float4* __restrict__ bpos = static_cast<const float4*>(b.bpos4);
    const float4 pi = bpos[i];
    float delta = kernel_w(0.0f, k.knorm);
    for_each_candidate<KIND>(g, b.bcell_start, pi.x, pi.y, pi.z, [&](uint32_t j) {
        if (j == i) return;
        const float4 pj = bpos[j];
        const float r2 = dist2_exact(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
        if (!(r2 < k.r2_q1)) return;
        const float inv_r = rsq(r2);
        const float q = r2 > 0.0f ? r2 * inv_r * k.inv_h : 0.0f;
        delta += kernel_w(q, k.knorm);
    });
    b.bvol[i] = 1.0f / fmaxf(delta, 1.0e-9f);
}

__global__ void __launch_bounds__(256) k_pack_bvol(float4* __restrict__ bpos,
                                                   const float* __restrict__ bvol, uint32_t nb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) bpos[i].w = bvol[i];
}

// ---- export helpers -----------------------------------------------------------
__global__ void k_to_orig_f32(const float* __restrict__ src, const uint32_t* __restrict__ ids,
                              uint32_t n, float* __restrict__ dst) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[ids[k]] = src[k];
}
__global__ void k_to_orig_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ ids,
                              uint32_t n, uint32_t* __restrict__ dst) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[ids[k]] = src[k];
}
__global__ void k_to_orig_vec(const float4* __restrict__ src, const uint32_t* __restrict__ ids,
                              uint32_t n, float* __restrict__ dst) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float4 v = src[k];
    const uint32_t o = 3u * ids[k];
    dst[o] = v.x; dst[o + 1] = v.y; dst[o + 2] = v.z;
}
__global__ void k_to_orig_w(const float4* __restrict__ src, const uint32_t* __restrict__ ids,
                            uint32_t n, float* __restrict__ dst) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[ids[k]] = src[k].w;
}
__global__ void k_permute_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ perm,
                              uint32_t n, uint32_t* __restrict__ dst) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[perm[k]];
}

inline uint32_t blocks_for(uint64_t n, uint32_t t) { return uint32_t((n + t - 1) / t); }
inline hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

#define DFSPH_LAUNCH_GRID(kernel, nthreads, stream, ...)                                          \
    do {                                                                                          \
        if (!(nthreads)) break;                                                                   \
        const uint32_t blk_ = a.block;                                                            \
        const dim3 grd_(blocks_for(nthreads, blk_));                                              \
        if (a.grid.kind == DFSPH_GRID_REFERENCE)                                                  \
            DFSPH_LAUNCH((kernel<DFSPH_GRID_REFERENCE, kFmtTest>), grd_, dim3(blk_), 0,     \
                               S(stream), __VA_ARGS__);                                           \
        else if (a.b.nlist)                                                                       \
            DFSPH_LAUNCH((kernel<DFSPH_GRID_FINE, kFmtList>), grd_, dim3(blk_), 0,          \
                               S(stream), __VA_ARGS__);                                           \
        else if (a.b.nrec)                                                                       \
            DFSPH_LAUNCH((kernel<DFSPH_GRID_FINE, kFmtMask>), grd_, dim3(blk_), 0,          \
                               S(stream), __VA_ARGS__);                                           \
        else                                                                                      \
            DFSPH_LAUNCH((kernel<DFSPH_GRID_FINE, kFmtTest>), grd_, dim3(blk_), 0,          \
                               S(stream), __VA_ARGS__);                                           \
    } while (0)

// The solver-phase sweeps: with neighbour masks and a.staged, the LDS-staged walk.
#define DFSPH_LAUNCH_GRID_S(kernel, nthreads, stream, ...)                                        \
    do {                                                                                          \
        if (!(nthreads)) break;                                                                   \
        const uint32_t blk_ = a.block;                                                            \
        const dim3 grd_(blocks_for(nthreads, blk_));                                              \
        if (a.grid.kind == DFSPH_GRID_REFERENCE)                                                  \
            DFSPH_LAUNCH((kernel<DFSPH_GRID_REFERENCE, kFmtTest, false>), grd_, dim3(blk_), 0,    \
                         S(stream), __VA_ARGS__);                                                 \
        else if (a.b.nlist)                                                                       \
            DFSPH_LAUNCH((kernel<DFSPH_GRID_FINE, kFmtList, false>), grd_, dim3(blk_), 0,         \
                         S(stream), __VA_ARGS__);                                                 \
        else if (a.b.nrec && a.staged)                                                            \
            DFSPH_LAUNCH((kernel<DFSPH_GRID_FINE, kFmtMask, true>), grd_, dim3(blk_), 0,          \
                         S(stream), __VA_ARGS__);                                                 \
        else if (a.b.nrec)                                                                        \
            DFSPH_LAUNCH((kernel<DFSPH_GRID_FINE, kFmtMask, false>), grd_, dim3(blk_), 0,         \
                         S(stream), __VA_ARGS__);                                                 \
        else                                                                                      \
            DFSPH_LAUNCH((kernel<DFSPH_GRID_FINE, kFmtTest, false>), grd_, dim3(blk_), 0,         \
                         S(stream), __VA_ARGS__);                                                 \
    } while (0)

}  // namespace

void launch_density_alpha(const SweepArgs& a, int warm_div, void* stream) {
    DFSPH_LAUNCH_GRID(k_density_alpha, a.k.n, stream, a.k, a.grid, a.b, a.xcd_remap, warm_div);
}
bool fused_decision(const SweepArgs& a) { return a.block == 256u; }
void launch_pressure_accel(const SweepArgs& a, int solver, int iter, const SolveDecide& d, void* stream) {
    DFSPH_LAUNCH_GRID(k_pressure_accel, a.k.n, stream, a.k, a.grid, a.b, a.xcd_remap, solver, iter, d);
}
namespace {

struct EmitterSet {
    float4 e[2 * DFSPH_MAX_EMITTERS];   // (pos, radius), (velocity, rate)
};

// PCG hash and unit-ball sample of shaders/apply_emitter.comp.glsl:29-42.
__device__ __forceinline__ uint32_t pcg_hash(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

__global__ void __launch_bounds__(256) k_emit(EmitterSet es, uint32_t count, uint32_t n0, uint32_t k, Buffers b) {
#pragma clang fp contract(off)
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= k) return;
    const uint32_t e = g % count, slot = n0 + g;
    const uint32_t seed = slot * 7919u + e * 31u;
    float u[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) u[a] = float(pcg_hash(seed + uint32_t(a))) / 4294967295.0f;
    float s[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) s[a] = u[a] * 2.0f - 1.0f;
    const float len = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    if (len > 1.0f)
#pragma unroll
        for (int a = 0; a < 3; ++a) s[a] = s[a] / len;
    const float4 pr = es.e[2 * e], vr = es.e[2 * e + 1];
    const float4 x = make_float4(pr.x + s[0] * pr.w, pr.y + s[1] * pr.w, pr.z + s[2] * pr.w, 0.0f);
    static_cast<float4*>(b.x_stage)[slot] = x;
    static_cast<float4*>(b.v_stage)[slot] = make_float4(vr.x, vr.y, vr.z, 0.0f);
    b.id_stage[slot] = slot;
    b.pdiv_stage[slot] = 0.0f;
    b.pden_stage[slot] = 0.0f;
    const float v2 = (vr.x * vr.x + vr.y * vr.y) + vr.z * vr.z;
    if (g < count) atomicMax(&b.status->vmax2_bits, __float_as_uint(v2));
}

}  // namespace

namespace {

// Boundary sample s: the last segment starting at or before s, then the
// sampler's expression for its coordinate (no contraction, as on the host).
__global__ void __launch_bounds__(256) k_boundary_sample(const BoundarySeg* __restrict__ segs, uint32_t nseg,
                                                         uint32_t total, float spacing, float4* __restrict__ out) {
#pragma clang fp contract(off)
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= total) return;
    uint32_t lo = 0, hi = nseg - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (segs[mid].start <= s) lo = mid;
        else hi = mid - 1u;
    }
    const BoundarySeg g = segs[lo];
    const float va = g.kind == 2 ? g.base : g.base + float(s - g.start) * spacing;
    float q[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 3; ++a) q[a] = a == g.ax_a ? va : a == g.ax_b ? g.c1 : g.c2;
    out[s] = make_float4(q[0], q[1], q[2], 0.0f);
}

}  // namespace

void launch_boundary_sample(const BoundarySeg* segs, uint32_t nseg, uint32_t total, float spacing, void* out,
                            void* stream) {
    if (!total || !nseg) return;
    DFSPH_LAUNCH(k_boundary_sample, dim3(blocks_for(total, 256)), dim3(256), 0, S(stream), segs, nseg, total,
                       spacing, static_cast<float4*>(out));
}

void launch_emit(const Buffers& b, const float* e8, uint32_t count, uint32_t n0, uint32_t k, void* stream) {
    EmitterSet es{};
    for (uint32_t e = 0; e < count && e < uint32_t(DFSPH_MAX_EMITTERS); ++e) {
        es.e[2 * e] = make_float4(e8[8 * e], e8[8 * e + 1], e8[8 * e + 2], e8[8 * e + 3]);
        es.e[2 * e + 1] = make_float4(e8[8 * e + 4], e8[8 * e + 5], e8[8 * e + 6], e8[8 * e + 7]);
    }
    DFSPH_LAUNCH(k_emit, dim3(blocks_for(k, 256)), dim3(256), 0, S(stream), es, count, n0, k, b);
}

void launch_jacobi(const SweepArgs& a, int solver, int reference, void* stream) {
    DFSPH_LAUNCH_GRID_S(k_jacobi, a.k.n, stream, a.k, a.grid, a.b, a.xcd_remap, solver, reference);
}
__global__ void __launch_bounds__(256) k_boundary_near(Grid g, const uint32_t* __restrict__ bcell_start,
                                                       uint8_t* bnear) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.cells) return;
    const int ax = int(g.axes[0]), ay = int(g.axes[1]), az = int(g.axes[2]);
    const int cx = int(c % g.axes[0]), cy = int((c / g.axes[0]) % g.axes[1]), cz = int(c / (g.axes[0] * g.axes[1]));
    uint8_t any = 0;
    for (int z = max(cz - 1, 0); z <= min(cz + 1, az - 1); ++z)
        for (int y = max(cy - 1, 0); y <= min(cy + 1, ay - 1); ++y) {
            const uint32_t row = (uint32_t(z) * uint32_t(ay) + uint32_t(y)) * uint32_t(ax);
            const uint32_t lo = row + uint32_t(max(cx - 1, 0)), hi = row + uint32_t(min(cx + 1, ax - 1));
            any |= bcell_start[hi + 1] > bcell_start[lo] ? 1 : 0;
        }
    bnear[c] = any;
}

void launch_boundary_near(const Grid& g, const uint32_t* bcell_start, uint8_t* bnear, void* stream) {
    DFSPH_LAUNCH(k_boundary_near, dim3(blocks_for(g.cells, 256)), dim3(256), 0, S(stream), g, bcell_start, bnear);
}

void launch_solve_finish(const SweepArgs& a, int solver, int iter, int reference, int min_it,
                         int max_it, float tol, float res_bound, void* stream) {
    if (a.k.n)
        DFSPH_LAUNCH(k_solve_finish, dim3(1), dim3(256), 0, S(stream), a.b.status,
                           a.b.partials, a.b.res_max_part, blocks_for(a.k.n, a.block), a.k.n, solver, iter, reference,
                           min_it, max_it, tol, res_bound);
}
void launch_accel_apply(const SweepArgs& a, int solver, uint32_t* keys, uint32_t* count, void* stream) {
    DFSPH_LAUNCH_GRID(k_accel_apply, a.k.n, stream, a.k, a.grid, a.b, a.xcd_remap, solver, keys, count);
}
// The pointwise variant does not walk neighbours, so one instantiation serves
// every format of a grid kind.
void launch_accel_apply_acc(const SweepArgs& a, int solver, uint32_t* keys, uint32_t* count, void* stream) {
    if (!a.k.n) return;
    if (a.grid.kind == DFSPH_GRID_REFERENCE)
        DFSPH_LAUNCH((k_accel_apply<DFSPH_GRID_REFERENCE, kFmtTest, true>), dim3(blocks_for(a.k.n, a.block)),
                     dim3(a.block), 0, S(stream), a.k, a.grid, a.b, a.xcd_remap, solver, keys, count);
    else
        DFSPH_LAUNCH((k_accel_apply<DFSPH_GRID_FINE, kFmtMask, true>), dim3(blocks_for(a.k.n, a.block)),
                     dim3(a.block), 0, S(stream), a.k, a.grid, a.b, a.xcd_remap, solver, keys, count);
}
void launch_step_finish(const SweepArgs& a, uint32_t ticket, void* stream) {
    DFSPH_LAUNCH(k_step_finish, dim3(1), dim3(256), 0, S(stream), a.b.status,
                       a.b.step_part, blocks_for(a.k.n, a.block), a.b.step_mirror, ticket);
}
void launch_forces(const SweepArgs& a, bool div_update, void* stream) {
    if (div_update)
        DFSPH_LAUNCH_GRID_S(k_forces_div, a.k.n, stream, a.k, a.grid, a.b, a.xcd_remap);
    else
        DFSPH_LAUNCH_GRID_S(k_forces, a.k.n, stream, a.k, a.grid, a.b, a.xcd_remap);
}
void launch_density_adv(const SweepArgs& a, int warm_den, void* stream) {
    DFSPH_LAUNCH_GRID_S(k_density_adv, a.k.n, stream, a.k, a.grid, a.b, a.xcd_remap, warm_den);
}
void launch_begin_solve(const SweepArgs& a, int solver, int warm, void* stream) {
    if (!a.k.n) return;
    DFSPH_LAUNCH(k_begin_solve, dim3(blocks_for(a.k.n, 256)), dim3(256), 0, S(stream), a.k,
                       a.b, solver, warm);
}
void launch_reset_solve(const SweepArgs& a, int solver, void* stream) {
    DFSPH_LAUNCH(k_reset_solve, dim3(blocks_for(a.k.n ? a.k.n : 1, 256)), dim3(256), 0, S(stream), a.k,
                       a.b, solver);
}
void launch_reorder(const SweepArgs& a, float dt_fixed, float cfl_2r, float dt_max, void* stream) {
    DFSPH_LAUNCH(k_reorder, dim3(blocks_for(a.k.n ? a.k.n : 1, 256)), dim3(256), 0, S(stream), a.b, a.k.n,
                       dt_fixed, cfl_2r, dt_max);
}
void launch_reorder_boundary(const SweepArgs& a, const void* bstage4, const uint32_t* perm,
                             void* stream) {
    if (!a.k.nb) return;
    DFSPH_LAUNCH(k_reorder_boundary, dim3(blocks_for(a.k.nb, 256)), dim3(256), 0, S(stream),
                       static_cast<float4*>(a.b.bpos4), a.b.bids,
                       static_cast<const float4*>(bstage4), perm, a.k.nb);
}
void launch_boundary_volume(const SweepArgs& a, void* stream) {
    if (!a.k.nb) return;
    const dim3 grd(blocks_for(a.k.nb, 256));
    if (a.grid.kind == DFSPH_GRID_REFERENCE)
        DFSPH_LAUNCH(k_boundary_volume<DFSPH_GRID_REFERENCE>, grd, dim3(256), 0, S(stream),
                           a.k, a.grid, a.b);
    else
        DFSPH_LAUNCH(k_boundary_volume<DFSPH_GRID_FINE>, grd, dim3(256), 0, S(stream), a.k,
                           a.grid, a.b);
    DFSPH_LAUNCH(k_pack_bvol, grd, dim3(256), 0, S(stream), static_cast<float4*>(a.b.bpos4),
                       a.b.bvol, a.k.nb);
}
void launch_to_original_f32(const float* src, const uint32_t* ids, uint32_t n, float* dst, void* stream) {
    if (n) DFSPH_LAUNCH(k_to_orig_f32, dim3(blocks_for(n, 256)), dim3(256), 0, S(stream), src, ids, n, dst);
}
void launch_to_original_u32(const uint32_t* src, const uint32_t* ids, uint32_t n, uint32_t* dst,
                            void* stream) {
    if (n) DFSPH_LAUNCH(k_to_orig_u32, dim3(blocks_for(n, 256)), dim3(256), 0, S(stream), src, ids, n, dst);
}
void launch_to_original_vec(const void* src4, const uint32_t* ids, uint32_t n, float* dst3, void* stream) {
    if (n)
        DFSPH_LAUNCH(k_to_orig_vec, dim3(blocks_for(n, 256)), dim3(256), 0, S(stream),
                           static_cast<const float4*>(src4), ids, n, dst3);
}
void launch_to_original_w(const void* src4, const uint32_t* ids, uint32_t n, float* dst, void* stream) {
    if (n)
        DFSPH_LAUNCH(k_to_orig_w, dim3(blocks_for(n, 256)), dim3(256), 0, S(stream),
                           static_cast<const float4*>(src4), ids, n, dst);
}
void launch_permute_u32(const uint32_t* src, const uint32_t* perm, uint32_t n, uint32_t* dst,
                        void* stream) {
    if (n) DFSPH_LAUNCH(k_permute_u32, dim3(blocks_for(n, 256)), dim3(256), 0, S(stream), src, perm, n, dst);
}

}  // namespace dfsph
