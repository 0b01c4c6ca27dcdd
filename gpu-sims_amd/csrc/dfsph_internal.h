// dfsph_internal.h -- shared declarations of the DFSPH library (host side and
// the launch wrappers of each kernel file).  Not part of the public ABI.
#pragma once

#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>

#include "../../include/dfsph.h"

namespace dfsph {

using Derived = dfsph_derived;

Derived derive(const dfsph_params& p);
// Why a domain cannot be simulated (non-finite or inverted bounds, a Morton
// table past 1024^3 cells, a FINE grid past 2^31 cells), or nullptr.
const char* domain_problem(const float dmin[3], const float dmax[3], float radius, float fine_cell_scale);
// Bounds the boundary sampler accepts: finite, dmax >= dmin, at most 2^24
// sample spacings per axis.
bool sampler_extent_ok(const float dmin[3], const float dmax[3], float radius);
// Why preset_fill cannot fill a preset at this radius, or nullptr.
const char* preset_problem(const dfsph_preset& p, float radius);
std::vector<float> boundary_samples(const float dmin[3], const float dmax[3], float radius,
                                    uint32_t cap);
// Boundary samples as segments of the sampler's emission order (face rows,
// edges, corners): sample start + i has coordinate ax_a = base + i * spacing
// (kind 0, 1) or base (kind 2, a corner), ax_b = c1, ax_c = c2.
struct BoundarySeg {
    uint32_t start, count;
    int32_t kind, ax_a, ax_b, ax_c;
    float base, c1, c2;
};
// Segments of boundary_samples(dmin, dmax, radius, cap), same order and cap;
// *total receives the sample count.
std::vector<BoundarySeg> boundary_segments(const float dmin[3], const float dmax[3], float radius, uint32_t cap,
                                           uint32_t* total);
const dfsph_preset* preset_table();
void preset_fill(const dfsph_preset& p, float radius, uint32_t capacity, float* x, float* v, uint32_t* flags,
                 uint32_t* count, uint32_t* brick_count);
void scene_block(uint32_t nx, uint32_t ny, uint32_t nz, float radius, const float origin[3],
                 uint64_t seed, float jitter, uint32_t id0, float* out);

// Grid description handed to the kernels by value.
struct Grid {
    int kind;          // DFSPH_GRID_REFERENCE or DFSPH_GRID_FINE
    float dmin[3];
    float cell;        // cell edge
    uint32_t axes[3];  // cells per axis
    uint32_t cells;    // table size (without the sentinel)
};

// Per-step constants handed to the sweep kernels by value.
struct SimConst {
    float h, inv_h, mass, knorm, gnorm, gh, m6g, rho0, relax, visc, coh, vort;
    float g[3];
    float dmin[3], dmax[3];
    float r2_q1, r2_eps_gt, r2_eps_lt;
    float bscale;      // boundary mass factor in alpha and the pressure force (1 or rho0)
    float alpha_cap;   // 10 x alpha of a full lattice neighbourhood (bounds learned guesses)
    // Neighbour sums over pairs use gradW = gfac * gs(t) * r_ij and
    // W = wfac * ws(t), t = h - r: gfac = m6g / h^2, wfac = 2 knorm / h^3,
    // h3half = h^3 / 2.
    float gfac, wfac, h3half;
    uint32_t n, nb;
    // First slot of the scene (0 for a single simulation; a batch packs its
    // scenes' particles back to back).  Waves count from it: wave w of the
    // scene holds slots slot0 + 64 w ...
    uint32_t slot0;
};

constexpr uint32_t kListSlots = 48;
constexpr uint32_t kListOverflow = 0xFFFFFFFFu;

// Solver status written by the device, read by the host after a solve.
struct SolveStatus {
    double err;        // FP64 error of the last completed iteration
    int32_t iters;     // completed iterations
    int32_t done;      // 1 once converged or at max_iter
    int32_t converged; // 1 if done by tolerance
    float res_max;     // largest per-particle residual of that iteration (same units as err)
    int32_t done_at;   // 0 while undecided; iteration + 1 of the decision once done
};

// Decisions of the fused solve path published to pinned, coherent host
// memory as they are taken, so the host can enqueue the next stage while the
// deciding launch is still running.  One decision is four 8-byte words, each
// stored whole at system scope and each carrying the same tag in its high
// half (bit 31 | completed iterations); low halves: done | converged << 1,
// res_max bits, err low, err high.  A reader that finds one tag in all four
// words holds one decision, whatever order the stores landed in, so the
// writer needs no fence (a system-scope fence writes back and invalidates the
// whole L2 of its XCD while the sweep runs).  The host zeroes a solve's words
// before its first launch; tag 0 means nothing published.
struct SolvePub {
    uint64_t w[4];
};
struct SolveMirror {
    SolvePub solve[2];
};
constexpr uint32_t kPubTagBit = 0x80000000u;

struct DeviceStatus {
    SolveStatus solve[2];    // 0 = divergence, 1 = density
    float dt;                // this step's dt
    float dt_prev;           // previous step's dt (warm-start scaling)
    uint32_t vmax2_bits;     // max |v|^2 of the step's final velocities (float bits)
    uint32_t vmax2_start;    // max |v|^2 at the start of the step
    uint64_t nonfinite;
    uint32_t pad0;
    uint32_t pad1;
};

// The substep's final status, published to pinned, coherent host memory by
// the substep's last launch (seqlock: seq is odd while `status` is written,
// then 2 x the substep's ticket), so the host learns the end of a substep by
// polling instead of a stream synchronisation.
struct StepMirror {
    DeviceStatus status;
    uint32_t seq;
    uint32_t pad[3];
};

// Device buffers of one context (sorted order unless marked "stage").
struct Buffers {
    // fluid, sorted order of the current step
    void* pos4 = nullptr;     // float4 x, y, z, p / max(rho^2, eps) (0 outside a solve)
    float* rho = nullptr;     // density
    void* vel4 = nullptr;     // float4 v, m / max(rho, 1e-3) (from the density pass on)
    void* vel4b = nullptr;    // float4 v after forces (ping-pong)
    void* acc4 = nullptr;     // float4 pressure acceleration
    void* gb4 = nullptr;      // float4 sum_b V_b gradW (boundary); with neighbour records,
                              // written only where the record's boundary flag is set
    float* alpha = nullptr;
    float* aor2 = nullptr;    // alpha / max(rho^2, eps)
    float* rho_dot = nullptr; // divergence source rho_dot
    float* rho_adv = nullptr; // density source rho*
    float* p = nullptr;       // pressure being solved
    float* p_div = nullptr;   // final divergence pressure
    float* p_den = nullptr;   // final density pressure
    float* pdiv_prev = nullptr;  // previous step's final pressures, this step's order
    float* pden_prev = nullptr;
    float* warm_ext_div = nullptr;  // caller guesses, indexed by original index
    float* warm_ext_den = nullptr;
    uint32_t* nbr = nullptr;
    uint32_t* nbrb = nullptr;
    // FINE grid: each particle's neighbour record, written by the density pass
    // as three 16-byte words at nrec[k * mask_stride + i] (k-major, so a wave's
    // loads of word k are contiguous): the bitmasks of the 9 x-runs (bit q of
    // run r marks slot base_r + q as a neighbour), then
    //   {m8, flags | off8 << 18, off0..off3, off4..off7}
    // with base_r = (first slot of the wave's window for run r) + off_r, one
    // byte each; flags bits 0-8 mark runs the sweeps re-test from their
    // candidates (a neighbour past bit 31), bits 9-17 runs whose first slot
    // is over 255 past the window's start (read from cell_start instead),
    // bit 31 boundary samples in reach.
    uint32_t* nrec = nullptr;
    uint32_t mask_stride = 0;
    // With records: per wave of 64 slots (counted from SimConst::slot0), the
    // cells of its first and last live particle as a pair of words at
    // wave_cells[2 w], which fix the wave's run windows.
    uint32_t* wave_cells = nullptr;
    // FINE grid, list format: neighbour slots of particle i at
    // nlist[t * mask_stride + i], t < ncnt[i] (slot-major, coalesced per slot);
    // ncnt[i] == kListOverflow means "re-test all candidates".
    uint32_t* nlist = nullptr;
    uint32_t* ncnt = nullptr;
    uint32_t* ids = nullptr;
    uint32_t* keys = nullptr; // sort keys in stage order
    // stage: particle data in the previous step's order, input of the sort
    void* x_stage = nullptr;  // float4
    void* v_stage = nullptr;  // float4
    uint32_t* id_stage = nullptr;
    float* pdiv_stage = nullptr;
    float* pden_stage = nullptr;
    // hash
    uint32_t* perm = nullptr;       // sorted slot -> stage index
    uint64_t* skey = nullptr;       // FINE grid: each slot's (x, original index) sort key (scatter output)
    uint32_t* cell_count = nullptr; // zero between steps
    uint32_t* cell_start = nullptr; // cells + 1
    uint32_t* scan_blocks = nullptr;
    // boundary, sorted by the step grid
    void* bpos4 = nullptr;          // float4 x, y, z, V_b
    float* bvol = nullptr;          // V_b before packing into bpos4.w
    uint32_t* bcell_start = nullptr;
    uint32_t* bids = nullptr;       // sample index of each sorted slot
    // FINE grid: per cell, 1 when a boundary sample lies in the cell's 3x3x3
    // neighbourhood (the density pass walks the boundary only there);
    // nullptr: every particle walks its boundary runs.
    uint8_t* bnear = nullptr;
    // reductions
    double* partials = nullptr;     // one FP64 error partial per workgroup
    uint32_t* res_max_part = nullptr; // per workgroup: largest residual (float bits)
    uint32_t* step_part = nullptr;  // per workgroup: max |v|^2 bits, non-finite count
    DeviceStatus* status = nullptr;
    SolveMirror* mirror = nullptr;   // device address of the pinned host mirror, or null
    StepMirror* step_mirror = nullptr;   // device address of the pinned substep mirror, or null
};

// ---- host helpers shared by the single-scene and batched paths ----
Grid make_grid(const dfsph_params& p, const Derived& d, int kind);
// Sweep constants of one scene; n is the end slot of the scene's particles.
SimConst make_sim_const(const dfsph_params& p, const Derived& d, float bscale, uint32_t n, uint32_t nb);
// Error reporting and argument checks of the C ABI (api.hip).
int set_error(int code, const std::string& msg);
int check_params(const dfsph_params* p);
int check_policy(const dfsph_solver_policy* pol);
// Context state for the capture format (api.hip): the particle radius, a
// particle's flags (original index), and resuming after dfsph_init_scene from
// a capture's final pressures (original order), flags and step count.
float ctx_particle_radius(const dfsph_ctx* c);
uint32_t ctx_particle_flags(const dfsph_ctx* c, uint32_t i);
int ctx_resume(dfsph_ctx* c, const float* p_density, const uint32_t* flags, uint64_t step);

// ---- launch wrappers (hash.hip) ----
// Keys for `n` float4 positions, counted into cell_count.
void launch_keys(const Grid& g, const void* pos4, uint32_t n, uint32_t* keys, uint32_t* cell_count,
                 void* stream);
// Exclusive scan of cell_count into cell_start[0..cells], resets cell_count to 0.
// A substep's finish carried by another launch (launch_step_finish's work:
// max |v|, non-finite count, status to the mirror when one is given).
struct StepFinish {
    DeviceStatus* status;
    const uint32_t* part;
    uint32_t ntiles;
    StepMirror* mirror;
    uint32_t ticket;
};
// finish (optional): one more workgroup of the first scan launch does it.
void launch_scan(uint32_t cells, uint32_t* cell_count, uint32_t* cell_start, uint32_t* scan_blocks,
                 void* stream, const StepFinish* finish = nullptr);
// perm[cell_start[key]+slot] = i, then per-cell order by ids[perm] (stable (key, id)),
// or, with stage positions xs4 (float4 per stage slot), by (x, id) so that each
// row of cells is sorted by x.  With skey (n entries) the scatter also writes
// each slot's sort key, which the per-cell order then reads instead of
// gathering it.
void launch_scatter(uint32_t n, const uint32_t* keys, const uint32_t* cell_start,
                    uint32_t* cell_count, uint32_t* perm, const uint32_t* ids, const void* xs4,
                    uint64_t* skey, uint32_t cells, void* stream);
size_t scan_scratch_elems(uint32_t cells);

// ---- launch wrappers (sweeps.hip) ----
struct SweepArgs {
    SimConst k;
    Grid grid;
    Buffers b;
    uint32_t block;     // threads per workgroup
    int xcd_remap;      // 1: map consecutive workgroups of one XCD to consecutive tiles
    int staged;         // 1: solver-phase sweeps walk LDS-staged run windows (masks only)
};
// Gather of the step's state into cell order; also starts the step's status
// (dt from dt_fixed, or by CFL from the previous max |v|: cfl_2r / |v|max,
// capped at dt_max).
void launch_reorder(const SweepArgs& a, float dt_fixed, float cfl_2r, float dt_max, void* stream);
// bnear[c] for every FINE cell c from the boundary's cell ranges.
void launch_boundary_near(const Grid& g, const uint32_t* bcell_start, uint8_t* bnear, void* stream);
// One empty k_trace_mark launch (a region marker for kernel traces).
void launch_trace_mark(uint32_t tag, void* stream);

// Screen-space render passes (render.hip): constants by value, images in
// device memory (row 0 at the top, width * height pixels).
struct RenderConst {
    float view[16], proj[16], inv_vp[16];   // column-major
    float cam[3];
    uint32_t w, h, n;
    float radius, point_scale, tpp, sigma_s, sigma_d;
    float tint[3], rough, zenith[3], horizon[3];
    float exposure, f0, absorption, thick_scale;
};
struct RenderBuffers {
    const void* x4;             // float4 positions, any order
    const uint32_t* ids;        // original index of each slot (a permutation of 0..n-1)
    void* proj4;                // float4 per original index: window x, y, point sizes
    void* view4;                // float4 per original index: view position, clip flag
    int* winner;                // per pixel: highest original index covering it, or -1
    unsigned long long* thick_acc;
    float *depth, *smooth_a, *smooth_b, *thick;
    void* rgba;                 // float4 per pixel
};
// Runs every pass; returns the image the composite read (depth or a smoothed one).
const float* launch_render(const RenderConst& k, const RenderBuffers& b, int iterations, int alternate,
                           void* stream);
// Inverse of a column-major 4x4 float matrix by cofactors; false if singular.
bool invert4(const float m[16], float out[16]);
// Column-major a * b.
void mul4(const float a[16], const float b[16], float out[16]);
// Density pass; with warm_div >= 0 (a dfsph_warm other than MLP) it also
// starts the divergence solve (begin_solve_one) for each particle.
void launch_density_alpha(const SweepArgs& a, int warm_div, void* stream);
// Pressure acceleration of Jacobi iteration `iter`.  With `decide` (256-thread
// workgroups only) workgroup 0 first takes the convergence decision of
// iteration iter - 1 from the Jacobi partials, and the launch then runs unless
// the solve was decided before that iteration, so the last launch of a solve
// leaves the acceleration of its final pressure in acc4 (see
// launch_accel_apply_final_acc).  Without `decide` it runs until the solve is
// done (the decision is launch_solve_finish's).
struct SolveDecide {
    int decide, reference, min_it, max_it;
    float tol, res_bound;
};
void launch_pressure_accel(const SweepArgs& a, int solver, int iter, const SolveDecide& d, void* stream);
bool fused_decision(const SweepArgs& a);
void launch_jacobi(const SweepArgs& a, int solver, int reference, void* stream);
// Boundary samples from their segments into float4 out[0..total) (w = 0).
void launch_boundary_sample(const BoundarySeg* segs, uint32_t nseg, uint32_t total, float spacing, void* out,
                            void* stream);
// Emitted particles into stage slots [n0, n0 + k): e8 holds `count` emitters
// as (pos, radius, velocity, rate); raises the status' max |v|^2.
void launch_emit(const Buffers& b, const float* e8, uint32_t count, uint32_t n0, uint32_t k, void* stream);
void launch_solve_finish(const SweepArgs& a, int solver, int iter, int reference, int min_it,
                         int max_it, float tol, float res_bound, void* stream);
// Max |v|^2 and non-finite count of the substep; with Buffers.step_mirror
// set, the substep's status is published there under `ticket` (> 0).
void launch_step_finish(const SweepArgs& a, uint32_t ticket, void* stream);
void launch_reorder_boundary(const SweepArgs& a, const void* bstage4, const uint32_t* perm, void* stream);
// keys, count (null: none): after the density solve, the next cell build's
// keys and histogram (cell_count must be zero) from the integrated positions.
void launch_accel_apply(const SweepArgs& a, int solver, uint32_t* keys, uint32_t* count, void* stream);
// The same with the acceleration already in acc4 (pointwise).
void launch_accel_apply_acc(const SweepArgs& a, int solver, uint32_t* keys, uint32_t* count, void* stream);
// div_update: apply the divergence solve's velocity update (final acceleration
// in acc4, pressure to p_div) as velocities are read, in place of its launch.
void launch_forces(const SweepArgs& a, bool div_update, void* stream);
// warm_den >= 0: also the density solve's start (begin_solve, solver 1).
void launch_density_adv(const SweepArgs& a, int warm_den, void* stream);
void launch_begin_solve(const SweepArgs& a, int solver, int warm, void* stream);
void launch_reset_solve(const SweepArgs& a, int solver, void* stream);
void launch_boundary_volume(const SweepArgs& a, void* stream);
// dst[ids[k]] = src[k]: sorted-order fields to original-index order.
void launch_to_original_f32(const float* src, const uint32_t* ids, uint32_t n, float* dst, void* stream);
void launch_to_original_u32(const uint32_t* src, const uint32_t* ids, uint32_t n, uint32_t* dst, void* stream);
void launch_to_original_vec(const void* src4, const uint32_t* ids, uint32_t n, float* dst3, void* stream);
void launch_to_original_w(const void* src4, const uint32_t* ids, uint32_t n, float* dst, void* stream);
// dst[k] = src[perm[k]] for uint32 (sorted index of the reference hash).
void launch_permute_u32(const uint32_t* src, const uint32_t* perm, uint32_t n, uint32_t* dst, void* stream);

}  // namespace dfsph
