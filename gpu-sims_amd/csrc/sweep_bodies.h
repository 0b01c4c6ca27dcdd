// sweep_bodies.h -- per-particle device bodies of the DFSPH neighbour sweeps,
// shared by the per-stage kernels of sweeps.hip (one scene) and the
// persistent batched substep of batch.hip.  One thread per particle over
// cell-sorted SoA state; i is the sorted slot.
//
// Reference kernels restated (same per-pair arithmetic, fused where the data
// flow allows):
//   density_alpha_one   density_alpha.comp.glsl:99-179 + compute_density_change.comp.glsl:82-151
//   pressure_accel_one  compute_pressure_accel.comp.glsl:84-157
//   jacobi_one          compute_aij_pj.comp.glsl:89-163 + jacobi_update_{divergence,density}.comp.glsl:47-59
//   accel_apply_one     compute_pressure_accel + apply_velocity.comp.glsl:45-51
//                       (+ integrate_forces.comp.glsl:65-100, mode 1, after the density solve)
//   forces_one          integrate_forces.comp.glsl:102-143 (mode 0)
//   density_adv_one     compute_density_adv.comp.glsl:80-153
// Boundary samples are static during a substep, so their gradient sum
// G_b = sum_b V_b gradW_ib is formed once (density pass); every later
// boundary term of the reference is a dot product with it
// (compute_pressure_accel.comp.glsl:151, compute_aij_pj.comp.glsl:154,
// compute_density_adv.comp.glsl:146).  Neighbourhood is decided from r^2 with
// thresholds equal to the reference's sqrt/divide tests, so neighbour sets are
// bit-identical to the reference's.
#pragma once

#include "grid.h"

namespace dfsph {
namespace sb {

constexpr float kEps = 1.0e-5f;   // DFSPH_ALPHA_EPS (main.cpp:121)

// The flag word of a neighbour record: bits 0-8 mark runs the sweeps re-test
// from their candidates (a neighbour past bit 31 of the mask); bits 9-17 mark
// runs whose first slot lies too far into the wave's window for its byte
// (the sweeps read it from cell_start instead); bits 18-25 hold run 8's slot
// offset; bit 31 marks a particle with boundary samples in reach (G_b != 0),
// other particles skip the G_b load.
constexpr uint32_t kHasBoundary = 0x80000000u;
constexpr uint32_t kRetestRuns = 0x1ffu;
constexpr uint32_t kFarShift = 9;
constexpr uint32_t kOff8Shift = 18;

// Neighbour formats of the sweeps after the density pass.
constexpr int kFmtTest = 0;   // test every candidate of the 27-cell neighbourhood
constexpr int kFmtMask = 1;   // FINE grid: 9 x 32-bit run masks
constexpr int kFmtList = 2;   // FINE grid: slot-major index lists

// Cubic spline W(q) for 0 <= q < 1 (density_alpha.comp.glsl:87-91), both
// branches evaluated and selected (no divergent branch in the pair loop).
__device__ __forceinline__ float kernel_w(float q, float k) {
    const float o = fmaxf(1.0f - q, 0.0f);
    const float inner = k * (6.0f * q * q * (q - 1.0f) + 1.0f);   // 6q^3 - 6q^2 + 1
    const float outer = 2.0f * k * o * o * o;
    return q < 0.5f ? inner : outer;
}

// Scale s with gradW = s * r_ij (density_alpha.comp.glsl:92-97); the q < 0.5
// branch uses poly/r = (18q - 12)/h, which needs no 1/r.
__device__ __forceinline__ float grad_scale(float q, float inv_r, const SimConst& k) {
    const float o = fmaxf(1.0f - q, 0.0f);
    const float inner = k.gh * (18.0f * q - 12.0f);
    const float outer = k.m6g * o * o * inv_r;
    return q < 0.5f ? inner : outer;
}

// 1/sqrt(r^2) with the hardware instruction: r^2 >= 1e-14 for every pair
// that reaches it, far above the denormal range the library call guards.
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

// The same functions for a decided neighbour in terms of t = h - r (one fused
// multiply-add from r^2 and 1/r), with their constant factors taken out of
// the neighbour sums and applied once per particle:
//   gradW_ij = gfac * gs * r_ij,  gs = 3t - h          (q < 1/2),  t^2 / r  (otherwise)
//   W_ij     = wfac * ws,         ws = h^3/2 - 3 r^2 t  (q < 1/2),  t^3      (otherwise)
// gfac = m6g / h^2, wfac = 2 knorm / h^3: gh (18q - 12) = gfac (3t - h),
// m6g (1 - q)^2 / r = gfac t^2 / r, knorm (6q^3 - 6q^2 + 1) = wfac (h^3/2 -
// 3 r^2 t), 2 knorm (1 - q)^3 = wfac t^3.  Both branches meet at q = 1/2.
// Six to eight instructions per pair instead of ten to twenty.
struct PairT {
    float t;      // h - r
    bool inner;   // q < 1/2
};
__device__ __forceinline__ PairT pair_t(float r2, float inv_r, const SimConst& k) {
    const float t = fmaf(-r2, inv_r, k.h);
    return {t, t > 0.5f * k.h};
}
__device__ __forceinline__ float grad_t(PairT p, float inv_r, const SimConst& k) {
    const float inner = fmaf(3.0f, p.t, -k.h);
    const float outer = p.t * p.t * inv_r;
    return p.inner ? inner : outer;
}
__device__ __forceinline__ float kernel_t(PairT p, float r2, const SimConst& k) {
    const float inner = fmaf(-3.0f * r2, p.t, k.h3half);
    const float outer = p.t * p.t * p.t;
    return p.inner ? inner : outer;
}

// Walk the neighbours recorded in run masks m[r] (slot base(r) + bit), calling
// body(j, ld(j)).  Runs are walked as mirror pairs (r, 8 - r), run r =
// (dz+1)*3 + (dy+1) with its mirror (-dz, -dy): a particle near one side of its
// cell has more neighbours in that side's runs and fewer in the mirror's, so a
// pair's count varies less across a wave than one run's, and the loops take
// about as many trips per wave as one flat loop over all runs would (two
// thirds of a loop per run).  Within a pair the next neighbour's payload is
// loaded before the current one is processed, so every trip has a gather in
// flight instead of waiting for its own.
// Neighbour record j of a float4 array with a 32-bit byte offset (j < 2^28
// slots), so the load takes a scalar base and a 32-bit vector offset instead
// of 64-bit address arithmetic per neighbour.
__device__ __forceinline__ float4 ld_nbr(const void* __restrict__ base, uint32_t j) {
    return *reinterpret_cast<const float4*>(static_cast<const char*>(base) + (j << 4u));
}

// Software-pipelined walk over one group's slots: next() yields the next slot
// while more() holds; the payload of the following slot is loaded before the
// current one is processed, alternating two payload registers (no copies).
template <class N, class M, class L, class B>
__device__ __forceinline__ void walk_pipelined(N&& next, M&& more, L&& ld, B&& body) {
    if (!more()) return;
    uint32_t j0 = next(), j1 = j0;
    auto p0 = ld(j0);
    auto p1 = p0;
    while (true) {
        bool go = more();
        if (go) {
            j1 = next();
            p1 = ld(j1);
        }
        body(j0, p0);
        if (!go) break;
        go = more();
        if (go) {
            j0 = next();
            p0 = ld(j0);
        }
        body(j1, p1);
        if (!go) break;
    }
}

// Walk of one mirror group, run a's bits then run b's (the order every walk
// of the sweeps follows).  Yields base + SCALE * bit: SCALE = 1 gives slots
// (ba, bb the runs' first slots); with a slot's address stride and
// pre-scaled bases it gives addresses directly (one multiply-add a
// neighbour).  The run in progress is held in cur; when it empties, run b
// takes its place once.
template <uint32_t SCALE, class L, class B>
__device__ __forceinline__ void walk_group(uint32_t ma, uint32_t mb, uint32_t ba, uint32_t bb, L&& ld, B&& body) {
    uint32_t cur = ma ? ma : mb, pend = ma ? mb : 0u, base = ma ? ba : bb;
    walk_pipelined(
        [&]() {
            const uint32_t j = base + SCALE * uint32_t(__builtin_ctz(cur));
            cur &= cur - 1u;
            const bool e = cur == 0u;
            cur = e ? pend : cur;
            base = e ? bb : base;
            pend = e ? 0u : pend;
            return j;
        },
        [&]() { return cur != 0u; }, ld, body);
}

// base(r): run r's first slot, asked for when its group is walked.
template <class BF, class L, class B>
__device__ __forceinline__ void walk_masks(const uint32_t (&m)[9], BF&& base, L&& ld, B&& body) {
    walk_group<1>(m[4], 0u, base(4), 0u, ld, body);
#pragma unroll
    for (int a = 0; a < 4; ++a) walk_group<1>(m[a], m[8 - a], base(a), base(8 - a), ld, body);
}

// Neighbour payload of the mask walk: the position record and a sweep's
// extra per-neighbour data.
template <class E>
struct NbrLoad {
    float4 pj;
    E e;
};
struct NoExtra {};

// Mask walk calling fn(j, pj, e, dx, dy, dz, r2) with r^2 from fused
// multiply-adds (neighbourhood is already decided); pj = ldp(j), e = ldx(j).
// ldp reads global memory; visit_staged walks the same masks over LDS copies
// with the same arithmetic.
template <class P, class BF, class X, class F>
__device__ __forceinline__ void visit_masks_p(P&& ldp, float4 pi, const uint32_t (&m)[9], BF&& base, X&& ldx, F&& fn) {
    using E = decltype(ldx(0u));
    walk_masks(
        m, base, [&](uint32_t j) { return NbrLoad<E>{ldp(j), ldx(j)}; },
        [&](uint32_t j, const NbrLoad<E>& q) {
            const float dx = pi.x - q.pj.x, dy = pi.y - q.pj.y, dz = pi.z - q.pj.z;
            fn(j, q.pj, q.e, dx, dy, dz, fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
        });
}
template <class BF, class X, class F>
__device__ __forceinline__ void visit_masks_x(const float4* __restrict__ pos, float4 pi, const uint32_t (&m)[9],
                                              BF&& base, X&& ldx, F&& fn) {
    visit_masks_p([&](uint32_t j) { return ld_nbr(pos, j); }, pi, m, base, static_cast<X&&>(ldx),
                  static_cast<F&&>(fn));
}

// The neighbour record of particle i (FINE grid, mask format, see
// Buffers::nrec): the run masks, each run's first slot as a byte offset from
// the start of the wave's window for that run, and the flag word.
struct NbrRecord {
    uint32_t m[9], off[9], flags, far;
};
__device__ __forceinline__ NbrRecord load_record(const Buffers& b, uint32_t i) {
    const uint4* __restrict__ rec = static_cast<const uint4*>(static_cast<const void*>(b.nrec));
    const size_t st = b.mask_stride;
    const uint4 w0 = rec[i], w1 = rec[st + i], w2 = rec[2 * st + i];
    NbrRecord r{{w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w, w2.x}, {}, w2.y, (w2.y >> kFarShift) & 0x1ffu};
#pragma unroll
    for (int q = 0; q < 8; ++q) r.off[q] = ((q < 4 ? w2.z : w2.w) >> (8 * (q & 3))) & 0xffu;
    r.off[8] = (w2.y >> kOff8Shift) & 0xffu;
    return r;
}
// Written in two parts: masks 0-7 once the neighbour test is done, the last
// word at the end of the density pass (flags: the flag word with run 8's
// offset in place; off_lo, off_hi: the offsets of runs 0-3 and 4-7, a byte
// each).
__device__ __forceinline__ void store_record_masks(const Buffers& b, uint32_t i, const uint32_t (&m)[9]) {
    uint4* __restrict__ rec = static_cast<uint4*>(static_cast<void*>(b.nrec));
    rec[i] = make_uint4(m[0], m[1], m[2], m[3]);
    rec[size_t(b.mask_stride) + i] = make_uint4(m[4], m[5], m[6], m[7]);
}
__device__ __forceinline__ void store_record_last(const Buffers& b, uint32_t i, uint32_t m8, uint32_t flags,
                                                  uint32_t off_lo, uint32_t off_hi) {
    uint4* __restrict__ rec = static_cast<uint4*>(static_cast<void*>(b.nrec));
    rec[2 * size_t(b.mask_stride) + i] = make_uint4(m8, flags, off_lo, off_hi);
}
__device__ __forceinline__ uint32_t record_flags(const Buffers& b, uint32_t i) {
    return b.nrec[(2 * size_t(b.mask_stride) + i) * 4u + 1u];
}

// The run windows of wave wv (slots [64 wv, 64 wv + 64)): run r of every
// particle of the wave lies in cells [c0 + off_r - 1, cL + off_r + 1] of the
// run's row offset off_r, c0 and cL the cells of the wave's first and last
// particle (written by the density pass).  start(r) is the window's first
// slot; the record's byte offsets count from it.
struct WaveWindows {
    int c0, cL, ax, ay, cells;
    const uint32_t* cell_start;
    __device__ __forceinline__ int row_off(int r) const { return ((r / 3 - 1) * ay + (r % 3 - 1)) * ax; }
    __device__ __forceinline__ uint32_t start(int r) const {
        return cell_start[min(max(c0 + row_off(r) - 1, 0), cells)];
    }
    // First slot and slot count of run r's window (count 0: no cell of it exists).
    __device__ __forceinline__ void window(int r, uint32_t& w0, uint32_t& wn) const {
        const int wlo = max(c0 + row_off(r) - 1, 0), whi = min(cL + row_off(r) + 1, cells - 1);
        w0 = 0u;
        wn = 0u;
        if (wlo <= whi) {
            w0 = cell_start[wlo];
            wn = cell_start[whi + 1] - w0;
        }
    }
};
// The windows of the wave holding slot i (every lane of a wave passes the
// same wave; the cells come from one scalar load).
__device__ __forceinline__ WaveWindows wave_windows(const SimConst& k, const Grid& g, const Buffers& b, uint32_t i) {
    const uint32_t wv = __builtin_amdgcn_readfirstlane((i - k.slot0) >> 6);
    const uint2 wc = reinterpret_cast<const uint2*>(b.wave_cells)[wv];
    return WaveWindows{int(wc.x), int(wc.y), int(g.axes[0]), int(g.axes[1]), int(g.cells), b.cell_start};
}
// First slot of run r of the particle at pi: the first cell of the unclipped
// run, row offset (r / 3 - 1, r % 3 - 1) from the particle's cell, as the
// density pass finds it.  Only asked for runs with neighbours, whose row
// exists.
__device__ __forceinline__ uint32_t run_first_slot(const Grid& g, const uint32_t* cell_start, float4 pi, int r) {
    const int cx = cell_coord(pi.x, g.dmin[0], g.cell, g.axes[0]);
    const int cy = cell_coord(pi.y, g.dmin[1], g.cell, g.axes[1]);
    const int cz = cell_coord(pi.z, g.dmin[2], g.cell, g.axes[2]);
    const uint32_t row = (uint32_t(cz + r / 3 - 1) * g.axes[1] + uint32_t(cy + r % 3 - 1)) * g.axes[0];
    return cell_start[row + uint32_t(cx > 0 ? cx - 1 : 0)];
}
// Run r's first slot from its byte offset (runs 0-3 in off_lo, 4-7 in
// off_hi, run 8 in off8) and the far flags: window start + offset, or from
// cell_start for a far run.
__device__ __forceinline__ uint32_t run_base(const Grid& g, const WaveWindows& ww, float4 pi, int r, uint32_t off_lo,
                                             uint32_t off_hi, uint32_t off8, uint32_t far) {
    if ((far >> r) & 1u) return run_first_slot(g, ww.cell_start, pi, r);
    const uint32_t o = r == 8 ? off8 : (((r < 4 ? off_lo : off_hi) >> (8 * (r & 3))) & 0xffu);
    return ww.start(r) + o;
}

// Visit the sweep neighbours of particle i -- j != i with 1e-7 <= r < h, the
// pair set every reference sweep after density_alpha uses -- calling
// fn(j, pj, e, dx, dy, dz, r2) with e = ldx(j), the sweep's extra data of the
// neighbour.  On the FINE grid with masks recorded, only the marked slots are
// loaded (neighbourhood already decided, so r^2 may be formed with fused
// multiply-adds) and the next neighbour's loads are issued before the current
// one is processed; otherwise every candidate of the 27-cell neighbourhood is
// tested in the reference's walk order with the exact r^2.  The summation
// order differs between formats, so their results agree to rounding, not bit
// for bit.
template <int KIND, int FMT, class X, class F>
__device__ __forceinline__ void visit_neighbors_x(const SimConst& k, const Grid& g, const Buffers& b,
                                                  uint32_t i, float4 pi, X&& ldx, F&& fn) {
    const float4* __restrict__ pos = static_cast<const float4*>(b.pos4);
    auto test = [&](uint32_t j) {
        const float4 pj = pos[j];
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float r2 = dist2_exact(dx, dy, dz);
        if (!(r2 < k.r2_q1) || j == i || r2 < k.r2_eps_lt) return;
        fn(j, pj, ldx(j), dx, dy, dz, r2);
    };
    if (FMT == kFmtList) {
        const uint32_t cnt = b.ncnt[i];
        if (cnt != kListOverflow) {
            // Batches of four: the four index loads (coalesced across the wave),
            // then the four gathers, are all in flight before any pair is computed.
            const uint32_t* __restrict__ lst = b.nlist + i;
            const size_t st = b.mask_stride;
            for (uint32_t t = 0; t < cnt; t += 4) {
                uint32_t j[4];
                float4 p[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) j[u] = t + u < cnt ? lst[size_t(t + u) * st] : i;
#pragma unroll
                for (int u = 0; u < 4; ++u) p[u] = pos[j[u]];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (t + u < cnt) {
                        const float dx = pi.x - p[u].x, dy = pi.y - p[u].y, dz = pi.z - p[u].z;
                        fn(j[u], p[u], ldx(j[u]), dx, dy, dz, dist2_exact(dx, dy, dz));
                    }
                }
            }
            return;
        }
        for_each_candidate<KIND>(g, b.cell_start, pi.x, pi.y, pi.z, test);
    } else if (FMT == kFmtMask) {
        // The particle's neighbour record: run masks and first-slot offsets in
        // three coalesced 16-byte loads.  A run whose neighbours reached past bit 31
        // (rare) re-tests its candidates here and is left out of the mask walk.
        const NbrRecord rec = load_record(b, i);
        const WaveWindows ww = wave_windows(k, g, b, i);
        uint32_t m[9], base[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            m[q] = rec.m[q];
            base[q] = ww.start(q) + rec.off[q];
        }
        if (rec.far) {
#pragma unroll
            for (int q = 0; q < 9; ++q)
                if ((rec.far >> q) & 1u) base[q] = run_first_slot(g, b.cell_start, pi, q);
        }
        if (const uint32_t ovf = rec.flags & kRetestRuns) {
            for_each_run_fine(g, b.cell_start, pi.x, pi.y, pi.z, [&](int r, uint32_t lo, uint32_t hi) {
                if (!((ovf >> r) & 1u)) return;
                const uint32_t e = b.cell_start[hi + 1];
                for (uint32_t j = b.cell_start[lo]; j < e; ++j) test(j);
#pragma unroll
                for (int q = 0; q < 9; ++q)
                    if (q == r) m[q] = 0u;
            });
        }
        visit_masks_x(pos, pi, m, [&](int r) { return base[r]; }, ldx, fn);
    } else {
        for_each_candidate<KIND>(g, b.cell_start, pi.x, pi.y, pi.z, test);
    }
}

// The same without extra neighbour data: fn(j, pj, dx, dy, dz, r2).
template <int KIND, int FMT, class F>
__device__ __forceinline__ void visit_neighbors(const SimConst& k, const Grid& g, const Buffers& b,
                                                uint32_t i, float4 pi, F&& fn) {
    visit_neighbors_x<KIND, FMT>(k, g, b, i, pi, [](uint32_t) { return NoExtra{}; },
                                 [&](uint32_t j, float4 pj, NoExtra, float dx, float dy, float dz, float r2) {
                                     fn(j, pj, dx, dy, dz, r2);
                                 });
}

// ---- LDS-staged mask walk (FINE grid, masks; the solver-phase sweeps) --------
// A wave's particles sit in consecutive cells c0..cL, so run r of every lane
// lies in the cell window [c0 + off_r - 1, cL + off_r + 1] (off_r the run's
// row offset), one contiguous slot range.  Before a mirror group (run a, then
// run 8 - a) is walked, the wave copies the windows of both runs from memory
// with coalesced loads into its own LDS region; every lane then reads its
// neighbours' payloads from there.  A gather from memory costs the texture
// path a tag lookup for every distinct line its lanes touch (about 24 per
// instruction here); the staging loads touch each line once and the LDS reads
// do not use that path.  The walk order and the arithmetic are those of
// visit_masks_x, so results are bit-identical to the unstaged walk; a group
// whose two windows exceed 2 * kWinCap slots together (a wave spanning
// unusually full rows: about 3 % of groups in a developed 1M dam break)
// walks from memory.
// Slots per staged window: about the ten cells a wave's runs span (8 + 2),
// with room for jitter.  Two windows of 24-28 B slots per wave keep a
// 256-thread workgroup under 22 KB of LDS (7-8 workgroups per CU).
template <int NE>
constexpr uint32_t kWinCap = 96;

// The block's LDS windows: per staged slot, a = (x, y, z, e0) (16 B),
// b = (e1, e2) (8 B) and, for NE = 4, c = e3 (4 B), each array holding
// 2 * kWinCap slots per wave (run a's window, then run b's) from slot
// ws = w * 2 * kWinCap for wave w.  The walk addresses a slot by u = slot * kUnit<NE> bytes, the
// stride of the narrowest array, so each array's address is a shift of u.
template <int NE>
constexpr uint32_t kUnit = NE == 4 ? 4u : 8u;
template <int NE>
struct StageWin {
    float4* a;     // the block's arrays (link-time LDS addresses)
    float2* b;
    float* c;
    uint32_t ws;   // this wave's first slot
};

template <int NE, class E>
__device__ __forceinline__ void stage_put(const StageWin<NE>& w, uint32_t t, const float4& pj, const E& e) {
    const uint32_t q = w.ws + t;
    w.a[q] = make_float4(pj.x, pj.y, pj.z, e.x);
    w.b[q] = make_float2(e.y, e.z);
    if constexpr (NE == 4) w.c[q] = e.w;
}
// Payload of the slot at u = slot * kUnit<NE>.
template <int NE, class E>
__device__ __forceinline__ NbrLoad<E> stage_get(const StageWin<NE>& w, uint32_t u) {
    const float4 a = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(w.a) + u * (16u / kUnit<NE>));
    const float2 b = *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(w.b) + u * (8u / kUnit<NE>));
    float c = 0.0f;
    if constexpr (NE == 4) c = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(w.c) + u);
    return NbrLoad<E>{make_float4(a.x, a.y, a.z, 0.0f), make_float4(a.w, b.x, b.y, c)};
}

// The sweep neighbours of particle i (the pair set of visit_neighbors_x) with
// fn(j, pj, e, dx, dy, dz, r2), e = ldx(j).  Every lane of the wave makes the
// call (k.n > 0); lanes past the count (live false) only help staging.
template <int NE, class X, class F>
__device__ __forceinline__ void visit_staged(const SimConst& k, const Grid& g, const Buffers& b, uint32_t i,
                                             bool live, float4 pi, const StageWin<NE>& w, X&& ldx, F&& fn) {
    using E = decltype(ldx(0u));
    const float4* __restrict__ pos = static_cast<const float4*>(b.pos4);
    const uint32_t lane = threadIdx.x & 63u;
    const NbrRecord rec = load_record(b, live ? i : k.n - 1u);
    uint32_t m[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) m[q] = live ? rec.m[q] : 0u;
    // Runs whose neighbours reached past bit 31 re-test their candidates
    // first, as visit_neighbors_x does.
    if (const uint32_t ovf = live ? (rec.flags & kRetestRuns) : 0u) {
        for_each_run_fine(g, b.cell_start, pi.x, pi.y, pi.z, [&](int r, uint32_t lo, uint32_t hi) {
            if (!((ovf >> r) & 1u)) return;
            const uint32_t e = b.cell_start[hi + 1];
            for (uint32_t j = b.cell_start[lo]; j < e; ++j) {
                const float4 pj = pos[j];
                const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                const float r2 = dist2_exact(dx, dy, dz);
                if (!(r2 < k.r2_q1) || j == i || r2 < k.r2_eps_lt) continue;
                fn(j, pj, ldx(j), dx, dy, dz, r2);
            }
#pragma unroll
            for (int q = 0; q < 9; ++q)
                if (q == r) m[q] = 0u;
        });
    }
    // The wave's run windows: the same cells the density pass staged, so a
    // record's byte offsets count from their starts.
    const WaveWindows ww = wave_windows(k, g, b, i);
    auto body = [&](uint32_t j, const NbrLoad<E>& q) {
        const float dx = pi.x - q.pj.x, dy = pi.y - q.pj.y, dz = pi.z - q.pj.z;
        fn(j, q.pj, q.e, dx, dy, dz, fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
    };
    auto ld_mem = [&](uint32_t j) { return NbrLoad<E>{ld_nbr(pos, j), ldx(j)}; };
    auto ld_lds = [&](uint32_t u) { return stage_get<NE, E>(w, u); };
    constexpr uint32_t U = kUnit<NE>;
#pragma unroll
    for (int gi = 0; gi < 5; ++gi) {
        // Group 0 is run 4 (the particle's own row); then the pairs (0, 8),
        // (1, 7), (2, 6), (3, 5).
        const int ra = gi == 0 ? 4 : gi - 1, rb = gi == 0 ? 4 : 9 - gi;
        const uint32_t ma = m[ra], mb = gi == 0 ? 0u : m[rb];
        if (!__any(int((ma | mb) != 0u))) continue;
        uint32_t wa, na, wb = 0u, nb = 0u;
        ww.window(ra, wa, na);
        if (gi != 0) ww.window(rb, wb, nb);
        // The two windows share the wave's 2 * kWinCap slots: run b's starts
        // where run a's ends.
        if (na + nb <= 2u * kWinCap<NE>) {
            for (uint32_t t = lane; t < na; t += 64u) stage_put<NE>(w, t, ld_nbr(pos, wa + t), ldx(wa + t));
            for (uint32_t t = lane; t < nb; t += 64u) stage_put<NE>(w, na + t, ld_nbr(pos, wb + t), ldx(wb + t));
            // The wave's LDS stores complete before any lane reads them, and
            // its reads before the next group's stores.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // Addresses u of the runs' first slots: the record's offsets count
            // from the windows' starts.
            walk_group<U>(ma, mb, (w.ws + rec.off[ra]) * U, (w.ws + rec.off[rb] + na) * U, ld_lds, body);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            // The only place a far first slot (past the byte's reach, so in a
            // window over the staging capacity) can occur.
            uint32_t ba = wa + rec.off[ra], bb = wb + rec.off[rb];
            if (rec.far) {
                if ((rec.far >> ra) & 1u) ba = run_first_slot(g, b.cell_start, pi, ra);
                if ((rec.far >> rb) & 1u) bb = run_first_slot(g, b.cell_start, pi, rb);
            }
            walk_group<1>(ma, mb, ba, bb, ld_mem, body);
        }
    }
}

// ---- density pass: rho, alpha, G_f, G_b, rho_dot, neighbour counts ------------
// FMT == kFmtList stages the particle's list in list_lds (slot-major, stride
// blockDim.x) and returns the staged length through list_cnt; the caller
// copies it out.
// Candidate positions of one run window staged in LDS per wave (FINE grid,
// mask format): entries per wave.
constexpr uint32_t kStageCap = 192;

// First t in [lo, hi) of an x-sorted LDS window with win[t].x > key (UPPER:
// >= key), else hi: a fixed six-step binary search (ranges of up to 63
// slots), branch-free; every read stays within [lo, hi].
template <bool UPPER>
__device__ __forceinline__ uint32_t window_bound(const float4* win, uint32_t lo, uint32_t hi, float key) {
    uint32_t n = hi - lo;
#pragma unroll
    for (int it = 0; it < 6; ++it) {
        const uint32_t half = n >> 1;
        const uint32_t mid = lo + half;
        const float x = win[n ? mid : lo].x;
        const bool right = n != 0u && (UPPER ? x < key : !(x > key));
        lo = right ? mid + 1u : lo;
        n = right ? n - half - 1u : half;
    }
    return lo;
}

// stage: this wave's kStageCap-entry LDS window (FINE grid with masks), or
// nullptr.  With stage set, every lane of the wave must make the call (lanes
// past the particle count only help staging).
template <int KIND, int FMT>
__device__ __forceinline__ void density_alpha_one(const SimConst& k, const Grid& g, const Buffers& b,
                                                  uint32_t i, uint32_t* list_lds, uint32_t& list_cnt,
                                                  float4* stage = nullptr) {
    const bool live = i < k.n;
    // The mask format always stages: its records count run starts from the
    // wave's windows (every caller passes this wave's LDS window).
    const bool staged_walk = KIND == DFSPH_GRID_FINE && FMT == kFmtMask;
    if (!staged_walk && !live) return;
    if (staged_walk && k.n == 0) return;
    const float4* __restrict__ pos = static_cast<const float4*>(b.pos4);
    const float4* vel = static_cast<const float4*>(b.vel4);   // .w is written at the end
    const uint32_t il = live ? i : k.n - 1u;                 // a valid slot for helper lanes
    const float4 pi = pos[il];
    const float4 vi = vel[il];
    float rho = 0.0f, gfx = 0.0f, gfy = 0.0f, gfz = 0.0f, sq = 0.0f, dsum = 0.0f;
    uint32_t cnt = 0;
    // Mask format: the record's last word, written once the boundary flag is
    // known.
    uint32_t rec_m8 = 0, rec_flags = 0, rec_off_lo = 0, rec_off_hi = 0;
    // Per-candidate work; returns true for a sweep neighbour.
    auto cand = [&](uint32_t j) -> bool {
        const float4 pj = pos[j];
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float r2 = dist2_exact(dx, dy, dz);
        if (!(r2 < k.r2_q1)) return false;
        const float inv_r = rsq(r2);
        const float q = r2 > 0.0f ? r2 * inv_r * k.inv_h : 0.0f;
        rho += k.mass * kernel_w(q, k.knorm);          // includes self (:130)
        if (j == i) return false;
        ++cnt;
        if (r2 < k.r2_eps_lt) return false;
        const float s = grad_scale(q, inv_r, k);
        const float wx = s * dx, wy = s * dy, wz = s * dz;
        if (r2 >= k.r2_eps_gt) {                        // r > 1e-7 (:132)
            const float mx = k.mass * wx, my = k.mass * wy, mz = k.mass * wz;
            gfx += mx; gfy += my; gfz += mz;
            sq += mx * mx + my * my + mz * mz;
        }
        const float4 vj = vel[j];                        // r >= 1e-7 (compute_density_change:292)
        dsum += (vi.x - vj.x) * wx + (vi.y - vj.y) * wy + (vi.z - vj.z) * wz;
        return true;
    };
    if (FMT == kFmtList) {
        uint32_t t = 0;
        for_each_run_fine(g, b.cell_start, pi.x, pi.y, pi.z, [&](int, uint32_t lo, uint32_t hi) {
            const uint32_t s = b.cell_start[lo], e = b.cell_start[hi + 1];
            for (uint32_t j = s; j < e; ++j) {
                if (cand(j)) {
                    if (t < kListSlots) list_lds[t * blockDim.x + threadIdx.x] = j;
                    ++t;
                }
            }
        });
        b.ncnt[i] = t <= kListSlots ? t : kListOverflow;
        list_cnt = t < kListSlots ? t : kListSlots;
    } else if (FMT == kFmtMask) {
        // Phase 1: the neighbour test over the clipped candidate runs only --
        // a bit per neighbour (slot base + bit, base = the unclipped run's
        // first slot, as the later sweeps compute it).  The self term and
        // coincident pairs (r < 1e-7: density and count only) are taken here.
        const float r2max = k.h * k.h * 1.001f;
        // Masks, and each run's first slot (the unclipped run's, as the later
        // sweeps find it) as a byte offset from the start of the wave's window
        // for that run: runs 0-3 in off_lo, 4-7 in off_hi, 8 in off8; a run
        // starting farther in is flagged in far.
        uint32_t m[9], ovf = 0u, far = 0u, off_lo = 0u, off_hi = 0u, off8 = 0u;
        int wave_c0 = 0;
#pragma unroll
        for (int q = 0; q < 9; ++q) m[q] = 0u;
        rho += k.mass * k.knorm;                         // j == i: W(0) (:130)
        // r < 1e-7 and j != i: density and count only (the self pair, r = 0,
        // takes the same rare branch and is skipped there).
        auto coincident = [&](uint32_t j, float r2) {
            if (j == i) return;
            const float inv_r = rsq(r2);
            const float q = r2 > 0.0f ? r2 * inv_r * k.inv_h : 0.0f;
            rho += k.mass * kernel_w(q, k.knorm);
            ++cnt;
        };
        // The neighbour test of run r over candidate slots [s, e) (mask bit j - s0),
        // candidate positions from ldp(j).
        auto run_test = [&](int r, uint32_t s0, uint32_t s, uint32_t e, auto&& ldp) {
            uint32_t mr = 0u;
            if (e - s0 <= 32u) {
                // Every slot of the run has a bit (no range test).  Four
                // consecutive candidates per trip, all four loads issued before
                // the tests, so the loads overlap.
                for (uint32_t j = s; j < e; j += 4u) {
                    float4 pc[4];
#pragma unroll
                    for (uint32_t u = 0; u < 4u; ++u) pc[u] = ldp(min(j + u, e - 1u));
#pragma unroll
                    for (uint32_t u = 0; u < 4u; ++u) {
                        const float r2 = dist2_exact(pi.x - pc[u].x, pi.y - pc[u].y, pi.z - pc[u].z);
                        if (j + u < e && r2 < k.r2_q1) {
                            if (!(r2 < k.r2_eps_lt)) mr |= 1u << (j + u - s0);
                            else coincident(j + u, r2);
                        }
                    }
                }
            } else {
                for (uint32_t j = s; j < e; ++j) {
                    const float4 pj = ldp(j);
                    const float r2 = dist2_exact(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
                    if (r2 < k.r2_q1) {
                        if (r2 < k.r2_eps_lt) {
                            coincident(j, r2);
                        } else {
                            const uint32_t off = j - s0;
                            if (off < 32u) mr |= 1u << off;
                            else ovf |= 1u << r;
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < 9; ++q)
                if (q == r) m[q] = mr;
        };
        // The same test of run r read from the LDS window (win[j] = slot j's
        // position, at least three readable entries past e), branch-free: a
        // slot becomes a mask bit when in [s, e), r^2 < r2_q1 and not
        // coincident; a coincident one (r < 1e-7, i itself excluded) only
        // flags the lane, which then re-tests the run for density and count.
        auto run_test_window = [&](int r, uint32_t s0, uint32_t s, uint32_t e, const float4* win) {
            uint32_t mr = 0u;
            bool coin = false;
            for (uint32_t j = s; j < e; j += 4u) {
                float4 pc[4];
#pragma unroll
                for (uint32_t u = 0; u < 4u; ++u) pc[u] = win[j + u];
                const uint32_t off = j - s0;
#pragma unroll
                for (uint32_t u = 0; u < 4u; ++u) {
                    const float r2 = dist2_exact(pi.x - pc[u].x, pi.y - pc[u].y, pi.z - pc[u].z);
                    const bool in = (j + u < e) & (r2 < k.r2_q1);
                    const bool tiny = r2 < k.r2_eps_lt;
                    mr |= uint32_t(in & !tiny) << ((off + u) & 31u);   // bits past e are 0
                    coin |= in & tiny & (r != 4 || j + u != i);
                }
            }
            if (coin) {
                for (uint32_t j = s; j < e; ++j) {
                    const float4 pj = win[j];
                    const float r2 = dist2_exact(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
                    if (r2 < k.r2_eps_lt) coincident(j, r2);   // skips i itself
                }
            }
#pragma unroll
            for (int q = 0; q < 9; ++q)
                if (q == r) m[q] = mr;
        };
        auto ld_global = [&](uint32_t j) { return ld_nbr(pos, j); };
        {
            // Wave-cooperative: a wave's particles sit in consecutive cells, so
            // run r of every lane lies in cells [c_0 + off_r - 1, c_L + off_r + 1]
            // (c_0, c_L: the cells of the wave's first and last live lanes,
            // off_r the run's row offset).  That window's candidates are loaded
            // once into LDS and every lane tests its own run from there; a
            // window over kStageCap slots (a wave spanning a long row) is read
            // from memory as before.  Runs in the same order, clipped as
            // for_each_run_clipped clips them.
            const uint32_t lane = threadIdx.x & 63u;
            const uint64_t lm = __ballot(live);
            if (!lm) return;                                   // whole wave past the count
            const int last = __popcll(lm) - 1;                 // live lanes are a prefix
            const int ax = int(g.axes[0]), ay = int(g.axes[1]), az = int(g.axes[2]);
            const int cx = cell_coord(pi.x, g.dmin[0], g.cell, g.axes[0]);
            const int cy = cell_coord(pi.y, g.dmin[1], g.cell, g.axes[1]);
            const int cz = cell_coord(pi.z, g.dmin[2], g.cell, g.axes[2]);
            const int cell = (cz * ay + cy) * ax + cx;
            const int c0 = __builtin_amdgcn_readlane(cell, 0);
            const int cL = __builtin_amdgcn_readlane(cell, last);
            if (lane == 0u)
                reinterpret_cast<uint2*>(b.wave_cells)[(i - k.slot0) >> 6] = make_uint2(uint32_t(c0), uint32_t(cL));
            const float lox = g.dmin[0] + float(cx) * g.cell, loy = g.dmin[1] + float(cy) * g.cell,
                        loz = g.dmin[2] + float(cz) * g.cell;
            const float xl = pi.x - lox, xr = lox + g.cell - pi.x;
            const float dyl = fmaxf(pi.y - loy, 0.0f), dyh = fmaxf(loy + g.cell - pi.y, 0.0f);
            const float dzl = fmaxf(pi.z - loz, 0.0f), dzh = fmaxf(loz + g.cell - pi.z, 0.0f);
#pragma unroll
            for (int dz = -1; dz <= 1; ++dz) {
#pragma unroll
                for (int dy = -1; dy <= 1; ++dy) {
                    const int r = (dz + 1) * 3 + (dy + 1);
                    const int off = (dz * ay + dy) * ax;
                    const int wlo = max(c0 + off - 1, 0), whi = min(cL + off + 1, int(g.cells) - 1);
                    uint32_t plo = 0u;
                    bool staged = false;
                    if (wlo <= whi) {                          // wave-uniform
                        plo = b.cell_start[wlo];
                        const uint32_t cnt = b.cell_start[whi + 1] - plo;
                        if (cnt + 3u <= kStageCap) {   // three entries of read slack
                            for (uint32_t t = lane; t < cnt; t += 64u) stage[t] = ld_nbr(pos, plo + t);
                            __builtin_amdgcn_wave_barrier();
                            staged = true;
                        }
                    }
                    const int z_ = cz + dz, y_ = cy + dy;
                    if (live && z_ >= 0 && z_ < az && y_ >= 0 && y_ < ay) {
                        const float ez = dz < 0 ? dzl : (dz > 0 ? dzh : 0.0f);
                        const float ey = dy < 0 ? dyl : (dy > 0 ? dyh : 0.0f);
                        const float d2 = ey * ey + ez * ez;
                        if (d2 < r2max) {
                            const float w2 = r2max - d2;
                            const uint32_t row = (uint32_t(z_) * uint32_t(ay) + uint32_t(y_)) * uint32_t(ax);
                            const bool left = cx > 0 && (xl < 0.0f || xl * xl < w2);
                            const bool right = cx + 1 < ax && (xr < 0.0f || xr * xr < w2);
                            const uint32_t lo_u = row + uint32_t(cx > 0 ? cx - 1 : 0);
                            const uint32_t lo = row + uint32_t(left ? cx - 1 : cx);
                            const uint32_t hi = row + uint32_t(right ? cx + 1 : cx);
                            const uint32_t s0 = b.cell_start[lo_u], s = b.cell_start[lo], e = b.cell_start[hi + 1];
                            if (staged && e - s0 <= 32u) {
                                // Cells are ordered by x, so the run is one x-sorted range and
                                // the slots within w of pi.x (the only ones that can be within
                                // h) are one window of it, found by two binary searches.  w
                                // carries the rounding of the positions compared.
                                const float wr = sqrtf(w2);
                                const float w = wr + 4.8e-7f * (fabsf(pi.x) + wr);
                                const uint32_t ta = window_bound<false>(stage, s - plo, e - plo, pi.x - w);
                                const uint32_t tb = window_bound<true>(stage, ta, e - plo, pi.x + w);
                                run_test_window(r, s0, plo + ta, plo + tb, stage - plo);
                            }
                            else if (staged) run_test(r, s0, s, e, [&](uint32_t j) { return stage[j - plo]; });
                            else run_test(r, s0, s, e, ld_global);
                            // The lane's run lies in the window, so s0 >= plo.
                            if (m[r]) {
                                const uint32_t o = s0 - plo;
                                if (o > 255u) far |= 1u << r;
                                else if (r < 4) off_lo |= o << (8 * r);
                                else if (r < 8) off_hi |= o << (8 * (r - 4));
                                else off8 = o;
                            }
                        }
                    }
                    __builtin_amdgcn_wave_barrier();           // the window is reused by the next run
                }
            }
            if (!live) return;
            wave_c0 = c0;
        }
        store_record_masks(b, i, m);
        rec_m8 = m[8];
        rec_flags = ovf | (far << kFarShift) | (off8 << kOff8Shift);
        rec_off_lo = off_lo;
        rec_off_hi = off_hi;
        // Phase 2: the neighbour work over the masks (mirror-paired walk);
        // overflowing runs re-test their candidates instead.  Sums in the
        // t-form (W = wfac ws, gradW = gfac gs r_ij), scaled once below.
        float rho_t = 0.0f, gx_t = 0.0f, gy_t = 0.0f, gz_t = 0.0f, sq_t = 0.0f, dsum_t = 0.0f;
        auto nbr = [&](uint32_t, float4, float4 vj, float dx, float dy, float dz, float r2) {
            const float inv_r = rsq(r2);
            const PairT pt = pair_t(r2, inv_r, k);
            rho_t += kernel_t(pt, r2, k);
            ++cnt;
            const float s = grad_t(pt, inv_r, k);
            const float wx = s * dx, wy = s * dy, wz = s * dz;
            if (r2 >= k.r2_eps_gt) {                        // r > 1e-7 (:132)
                gx_t += wx; gy_t += wy; gz_t += wz;
                sq_t += s * s * r2;                         // |m gradW|^2 / (m gfac)^2
            }
            dsum_t += (vi.x - vj.x) * wx + (vi.y - vj.y) * wy + (vi.z - vj.z) * wz;
        };
        auto ld_vel = [&](uint32_t j) { return ld_nbr(vel, j); };
        if (ovf) {
            for_each_run_fine(g, b.cell_start, pi.x, pi.y, pi.z, [&](int r, uint32_t lo, uint32_t hi) {
                if (!((ovf >> r) & 1u)) return;
                const uint32_t e = b.cell_start[hi + 1];
                for (uint32_t j = b.cell_start[lo]; j < e; ++j) {
                    const float4 pj = pos[j];
                    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                    const float r2 = dist2_exact(dx, dy, dz);
                    if (r2 < k.r2_q1 && j != i && !(r2 < k.r2_eps_lt)) nbr(j, pj, vel[j], dx, dy, dz, r2);
                }
#pragma unroll
                for (int q = 0; q < 9; ++q)
                    if (q == r) m[q] = 0u;
            });
        }
        const WaveWindows ww{wave_c0, wave_c0, int(g.axes[0]), int(g.axes[1]), int(g.cells), b.cell_start};
        visit_masks_x(pos, pi, m, [&](int r) { return run_base(g, ww, pi, r, off_lo, off_hi, off8, far); }, ld_vel,
                      nbr);
        const float mg = k.mass * k.gfac;
        rho += k.mass * k.wfac * rho_t;
        gfx += mg * gx_t; gfy += mg * gy_t; gfz += mg * gz_t;
        sq += mg * mg * sq_t;
        dsum += k.gfac * dsum_t;
    } else {
        for_each_candidate<KIND>(g, b.cell_start, pi.x, pi.y, pi.z, [&](uint32_t j) { cand(j); });
    }
    float gbx = 0.0f, gby = 0.0f, gbz = 0.0f;
    uint32_t cntb = 0;
    // A particle whose cell neighbourhood holds no boundary sample would find
    // none: its walk (18 cell-range loads) is skipped.
    const bool walk_boundary = k.nb && !(KIND == DFSPH_GRID_FINE && b.bnear &&
                                         !b.bnear[cell_key<DFSPH_GRID_FINE>(g, pi.x, pi.y, pi.z)]);
    if (walk_boundary) {
        const float4* __restrict__ bpos = static_cast<const float4*>(b.bpos4);
        auto bcand = [&](uint32_t j) {
            const float4 pj = ld_nbr(bpos, j);           // .w = V_b
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float r2 = dist2_exact(dx, dy, dz);
            if (!(r2 < k.r2_q1)) return;
            ++cntb;
            const float inv_r = rsq(r2);
            const float q = r2 > 0.0f ? r2 * inv_r * k.inv_h : 0.0f;
            rho += k.rho0 * pj.w * kernel_w(q, k.knorm); // (:165)
            if (r2 >= k.r2_eps_gt) {
                const float s = pj.w * grad_scale(q, inv_r, k);
                gbx += s * dx; gby += s * dy; gbz += s * dz;
            }
        };
        if (KIND == DFSPH_GRID_FINE && FMT == kFmtMask) {
            // Same visiting order as the full walk, minus cells out of reach.
            for_each_run_clipped(g, pi.x, pi.y, pi.z, k.h * k.h * 1.001f, [&](int, uint32_t, uint32_t lo, uint32_t hi) {
                const uint32_t e = b.bcell_start[hi + 1];
                for (uint32_t j = b.bcell_start[lo]; j < e; ++j) bcand(j);
            });
        } else {
            for_each_candidate<KIND>(g, b.bcell_start, pi.x, pi.y, pi.z, bcand);
        }
    }
    if (FMT == kFmtMask)
        store_record_last(b, i, rec_m8, rec_flags | (cntb ? kHasBoundary : 0u), rec_off_lo, rec_off_hi);
    const float gx = gfx + k.bscale * gbx, gy = gfy + k.bscale * gby, gz = gfz + k.bscale * gbz;
    const float alpha = 1.0f / fmaxf(gx * gx + gy * gy + gz * gz + sq, kEps);   // (:174-176)
    b.rho[i] = rho;
    static_cast<float4*>(b.vel4)[i].w = k.mass / fmaxf(rho, 1e-3f);   // the forces sweep's m/rho_j
    b.alpha[i] = alpha;
    b.aor2[i] = alpha / fmaxf(rho * rho, kEps);
    const float vol = k.mass / fmaxf(rho, kEps);
    b.rho_dot[i] = vol * dsum + (vi.x * gbx + vi.y * gby + vi.z * gbz);
    // Readers of the boundary gradient go through near_boundary(), which with
    // neighbour records reads the flag set above: interior particles skip it.
    if (FMT != kFmtMask || cntb) static_cast<float4*>(b.gb4)[i] = make_float4(gbx, gby, gbz, 0.0f);
    b.nbr[i] = cnt;
    b.nbrb[i] = cntb;
}

// Fluid part of the pressure acceleration: -sum_j m (pr_i + pr_j) gradW_ij.
struct AccelPair {
    const SimConst& k;
    float pr_i;
    float &ax, &ay, &az;
    __device__ __forceinline__ void operator()(uint32_t, float4 pj, NoExtra, float dx, float dy, float dz,
                                               float r2) const {
        const float inv_r = rsq(r2);
        const float f = (pr_i + pj.w) * grad_t(pair_t(r2, inv_r, k), inv_r, k);   // pos4.w = p/rho^2
        ax -= f * dx; ay -= f * dy; az -= f * dz;
    }
};
template <int KIND, int FMT>
__device__ __forceinline__ float3 accel_fluid(const SimConst& k, const Grid& g, const Buffers& b,
                                              uint32_t i, float4 pi, float pr_i) {
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    visit_neighbors_x<KIND, FMT>(k, g, b, i, pi, [](uint32_t) { return NoExtra{}; }, AccelPair{k, pr_i, ax, ay, az});
    const float c = k.mass * k.gfac;
    return make_float3(c * ax, c * ay, c * az);
}

template <int FMT>
__device__ __forceinline__ bool near_boundary(const SimConst& k, const Buffers& b, uint32_t i) {
    if (!k.nb) return false;
    return FMT != kFmtMask || (record_flags(b, i) & kHasBoundary) != 0u;
}

// Boundary part (compute_pressure_accel.comp.glsl:130-154): gated on |p/rho^2| > eps.
template <int FMT>
__device__ __forceinline__ float3 accel_boundary(const SimConst& k, const Buffers& b, uint32_t i,
                                                 float pr_i, float3 a) {
    if (fabsf(pr_i) > kEps && near_boundary<FMT>(k, b, i)) {
        const float4 gb = static_cast<const float4*>(b.gb4)[i];
        const float f = pr_i * k.bscale;
        a.x -= f * gb.x; a.y -= f * gb.y; a.z -= f * gb.z;
    }
    return a;
}

template <int KIND, int FMT>
__device__ __forceinline__ void pressure_accel_one(const SimConst& k, const Grid& g, const Buffers& b,
                                                   uint32_t i) {
    if (i >= k.n) return;
    const float4 pi = static_cast<const float4*>(b.pos4)[i];
    const float pr_i = pi.w;
    float3 a = accel_fluid<KIND, FMT>(k, g, b, i, pi, pr_i);
    a = accel_boundary<FMT>(k, b, i, pr_i, a);
    static_cast<float4*>(b.acc4)[i] = make_float4(a.x, a.y, a.z, 0.0f);
}
// ---- FP64 sums in a fixed order (deterministic) -----------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Workgroup sum; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double ws[16];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) ws[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        const int nw = int(blockDim.x + 63) / 64;
        for (int q = 0; q < nw; ++q) t += ws[q];
    }
    __syncthreads();
    return t;
}

// Workgroup max of non-negative floats given as bits; valid in thread 0.
__device__ __forceinline__ uint32_t block_max_bits(uint32_t v) {
    __shared__ uint32_t wm[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_down(v, o, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wm[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = int(blockDim.x + 63) / 64;
        for (int q = 1; q < nw; ++q) v = max(v, wm[q]);
    }
    __syncthreads();
    return v;
}

// Max of non-negative float bits p[0..count) by one workgroup; valid in thread 0.
__device__ __forceinline__ float max_bits(const uint32_t* __restrict__ p, uint32_t count) {
    uint32_t m = 0u;
    for (uint32_t q = threadIdx.x; q < count; q += blockDim.x) m = max(m, p[q]);
    return __uint_as_float(block_max_bits(m));
}

// Sum of p[0..count) by one 256-thread workgroup in a fixed order: thread t
// adds p[t], p[t+256], ... sequentially, then the workgroup tree.  The single
// scene and the batched substep both reduce their per-tile partials this
// way, so a scene's error is the same bits either way.  Valid in thread 0.
__device__ __forceinline__ double ordered_sum(const double* __restrict__ p, uint32_t count) {
    double t = 0.0;
    for (uint32_t q = threadIdx.x; q < count; q += 256u) t += p[q];
    return block_sum(t);
}

// Convergence bookkeeping after iteration `iter` of solve `solver` from the
// summed per-particle errors.  With res_bound > 0 the largest per-particle
// residual must also be within res_bound: the average alone accepts a few
// over-pressured particles, which a warm start can leave.
__device__ __forceinline__ void solve_update(SolveStatus& s, double total, float res_max, uint32_t n,
                                             float dt, int solver, int iter, int reference, int min_it,
                                             int max_it, float tol, float res_bound) {
    double err = n ? total / double(n) : 0.0;
    if (solver == 0) err *= double(dt);
    const int it = iter + 1;
    s.err = err;
    s.res_max = solver == 0 ? res_max * dt : res_max;
    s.iters = it;
    const bool bounded = !(res_bound > 0.0f) || s.res_max <= res_bound;
    const bool conv = !reference && err <= double(tol) && it >= min_it && bounded;
    s.converged = reference ? 1 : (conv ? 1 : 0);
    s.done = (conv || it >= max_it) ? 1 : 0;
    s.done_at = s.done ? it : 0;
}

// ---- Jacobi: aij*p, residual, relaxed update; returns the particle's error ---
// The neighbour term of aij*p for particle i: sum_j gs_ij (a_i - a_j).r_ij.
struct JacobiPair {
    const SimConst& k;
    float4 ai;
    float& sum;
    __device__ __forceinline__ void operator()(uint32_t, float4, float4 aj, float dx, float dy, float dz,
                                               float r2) const {
        const float inv_r = rsq(r2);
        const float s = grad_t(pair_t(r2, inv_r, k), inv_r, k);
        sum += s * ((ai.x - aj.x) * dx + (ai.y - aj.y) * dy + (ai.z - aj.z) * dz);
    }
};

__device__ __forceinline__ double jacobi_finish(const SimConst& k, const Buffers& b, uint32_t i, float4 ai,
                                                float sum, int solver, int reference, float dt, bool near_b);

template <int KIND, int FMT>
__device__ __forceinline__ double jacobi_one(const SimConst& k, const Grid& g, const Buffers& b, uint32_t i,
                                             int solver, int reference, float dt) {
    if (i >= k.n) return 0.0;
    const float4* __restrict__ acc = static_cast<const float4*>(b.acc4);
    const float4 pi = static_cast<const float4*>(b.pos4)[i];
    const float4 ai = acc[i];
    float sum = 0.0f;
    visit_neighbors_x<KIND, FMT>(k, g, b, i, pi, [&](uint32_t j) { return ld_nbr(acc, j); }, JacobiPair{k, ai, sum});
    return jacobi_finish(k, b, i, ai, sum, solver, reference, dt, near_boundary<FMT>(k, b, i));
}
// The same through this wave's LDS windows; every lane of the wave makes the call.
__device__ __forceinline__ double jacobi_staged(const SimConst& k, const Grid& g, const Buffers& b, uint32_t i,
                                                int solver, int reference, float dt, const StageWin<3>& w) {
    const bool live = i < k.n;
    const uint32_t il = live ? i : k.n - 1u;
    const float4* __restrict__ acc = static_cast<const float4*>(b.acc4);
    const float4 pi = static_cast<const float4*>(b.pos4)[il];
    const float4 ai = acc[il];
    float sum = 0.0f;
    visit_staged<3>(k, g, b, i, live, pi, w, [&](uint32_t j) { return ld_nbr(acc, j); }, JacobiPair{k, ai, sum});
    if (!live) return 0.0;
    return jacobi_finish(k, b, i, ai, sum, solver, reference, dt, near_boundary<kFmtMask>(k, b, i));
}

// The pointwise rest of the Jacobi update after the neighbour sum.
__device__ __forceinline__ double jacobi_finish(const SimConst& k, const Buffers& b, uint32_t i, float4 ai,
                                                float sum, int solver, int reference, float dt, bool near_b) {
    const float scale = solver == 0 ? dt : dt * dt;   // compute_aij_pj.comp.glsl:160-161
    const float rho = b.rho[i];
    sum *= k.mass * k.gfac / fmaxf(rho, kEps);
    if (near_b) {
        const float4 gb = static_cast<const float4*>(b.gb4)[i];
        sum += ai.x * gb.x + ai.y * gb.y + ai.z * gb.z;
    }
    sum *= scale;
    // Source: -rho_dot (jacobi_update_divergence:55) or 1 - rho* (jacobi_update_density:55).
    const float src = solver == 0 ? -b.rho_dot[i] : 1.0f - b.rho_adv[i];
    const float res = src - sum;
    float coef;
    if (reference) coef = b.aor2[i];
    else coef = b.alpha[i] * rho * rho * rho / scale;
    const float p_old = b.p[i];
    const float pn = fmaxf(p_old - k.relax * res * coef, 0.0f);
    b.p[i] = pn;
    static_cast<float4*>(b.pos4)[i].w = pn / fmaxf(rho * rho, kEps);
    // Complementarity residual of the projected solve: where p > 0 the
    // equation must hold either way; where p = 0 only compression counts.
    return p_old > 0.0f ? fabs(double(res)) : fmax(-double(res), 0.0);
}

// ---- final acceleration, v += dt a (+ position update) -----------------------
struct StepPart {
    uint32_t v2bits;   // max |v|^2 (float bits; non-negative floats order as integers)
    uint32_t bad;      // non-finite components
    float4 x;          // after the density solve: the integrated position
};

// p_out receives the solve's final pressure; after the density solve the
// integrated x, v go to the stage arrays and, when id_out is given, the
// particle's original index too (the next step's stage order).
// FROM_ACC: the acceleration of the final pressure is already in acc4 (the
// solve's last pressure-acceleration launch computed it the same way).
template <int KIND, int FMT, bool FROM_ACC = false>
__device__ __forceinline__ StepPart accel_apply_one(const SimConst& k, const Grid& g, const Buffers& b,
                                                    uint32_t i, int solver, float dt, float* __restrict__ p_out,
                                                    uint32_t* __restrict__ id_out) {
    StepPart r{0u, 0u, {}};
    if (i >= k.n) return r;
    const float4 pi = static_cast<const float4*>(b.pos4)[i];
    float3 a;
    if (FROM_ACC) {
        const float4 a4 = static_cast<const float4*>(b.acc4)[i];
        a = make_float3(a4.x, a4.y, a4.z);
    } else {
        const float pr_i = pi.w;
        a = accel_fluid<KIND, FMT>(k, g, b, i, pi, pr_i);
        a = accel_boundary<FMT>(k, b, i, pr_i, a);
        static_cast<float4*>(b.acc4)[i] = make_float4(a.x, a.y, a.z, 0.0f);
    }
    if (solver == 0) {
        float4* vel = static_cast<float4*>(b.vel4);
        const float4 v = vel[i];
        vel[i] = make_float4(v.x + dt * a.x, v.y + dt * a.y, v.z + dt * a.z, v.w);   // .w: m/rho for forces
        p_out[i] = b.p[i];
        return r;
    }
    const float4 v0 = static_cast<const float4*>(b.vel4b)[i];
    float v[3] = {v0.x + dt * a.x, v0.y + dt * a.y, v0.z + dt * a.z};
    float x[3] = {pi.x, pi.y, pi.z};
    const float margin = 0.125f * k.h;            // integrate_forces.comp.glsl:85
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float pn = x[c] + dt * v[c];
        if (pn < k.dmin[c] - margin) { pn = k.dmin[c]; v[c] = fmaxf(v[c], 0.0f); }
        else if (pn > k.dmax[c] + margin) { pn = k.dmax[c]; v[c] = fminf(v[c], 0.0f); }
        x[c] = pn;
        if (!isfinite(x[c]) || !isfinite(v[c])) ++r.bad;
    }
    r.x = make_float4(x[0], x[1], x[2], 0.0f);
    static_cast<float4*>(b.x_stage)[i] = r.x;
    static_cast<float4*>(b.v_stage)[i] = make_float4(v[0], v[1], v[2], 0.0f);
    p_out[i] = b.p[i];
    if (id_out) id_out[i] = b.ids[i];
    float s2;
    {
#pragma clang fp contract(off)
        s2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    }
    r.v2bits = __float_as_uint(s2);
    return r;
}

// Workgroup max / sum of StepPart; valid in thread 0 (order-free: integers).
__device__ __forceinline__ StepPart block_step_part(StepPart p) {
    __shared__ uint32_t smax[16], ssum[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        p.v2bits = max(p.v2bits, __shfl_down(p.v2bits, o, 64));
        p.bad += __shfl_down(p.bad, o, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { smax[w] = p.v2bits; ssum[w] = p.bad; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = int(blockDim.x + 63) / 64;
        for (int q = 1; q < nw; ++q) { p.v2bits = max(p.v2bits, smax[q]); p.bad += ssum[q]; }
    }
    __syncthreads();
    return p;
}

// ---- forces: v* = v + dt (g + viscosity + cohesion), race-free ----------------
// Viscosity and cohesion sums of particle i; vel4.w holds m / max(rho, 1e-3),
// written by the density pass.  COH false: the scene has no cohesion, whose
// sum would only be multiplied by zero.
template <bool COH>
struct ForcesPair {
    const SimConst& k;
    float4 vi;
    float &avx, &avy, &avz, &acx, &acy, &acz;
    __device__ __forceinline__ void operator()(uint32_t, float4, float4 vj, float dx, float dy, float dz,
                                               float r2) const {
        const float inv_r = rsq(r2);
        const float w = kernel_t(pair_t(r2, inv_r, k), r2, k);        // W / wfac
        const float fv = vj.w * w;                                      // (:125-129)
        avx += fv * (vj.x - vi.x); avy += fv * (vj.y - vi.y); avz += fv * (vj.z - vi.z);
        if (COH) {
            const float fc = w * inv_r;                                 // (:130)
            acx -= fc * dx; acy -= fc * dy; acz -= fc * dz;
        }
    }
};
__device__ __forceinline__ void forces_finish(const SimConst& k, const Buffers& b, uint32_t i, float4 vi, float dt,
                                              float avx, float avy, float avz, float acx, float acy, float acz) {
    const float cv = k.visc * k.wfac, cc = k.coh * k.mass * k.wfac;
    float ax = k.g[0] + cv * avx, ay = k.g[1] + cv * avy, az = k.g[2] + cv * avz;
    if (k.coh != 0.0f) { ax += cc * acx; ay += cc * acy; az += cc * acz; }
    static_cast<float4*>(b.vel4b)[i] = make_float4(vi.x + dt * ax, vi.y + dt * ay, vi.z + dt * az, 0.0f);
}
// Velocity of slot j for the forces sweep: vel4, or with DIV the divergence
// solve's velocity update v + dt a applied as it is read (a: the solve's final
// acceleration in acc4), the same arithmetic as accel_apply_one's; .w (m/rho)
// kept.
template <bool DIV>
__device__ __forceinline__ float4 forces_vel(const Buffers& b, uint32_t j, float dt) {
    const float4 v = ld_nbr(b.vel4, j);
    if constexpr (!DIV) {
        return v;
    } else {
        const float4 a = ld_nbr(b.acc4, j);
        return make_float4(v.x + dt * a.x, v.y + dt * a.y, v.z + dt * a.z, v.w);
    }
}
// With DIV the sweep also ends the divergence solve (its pressure to p_div):
// the velocity update then needs no launch of its own.
template <int KIND, int FMT, bool DIV>
__device__ __forceinline__ void forces_one(const SimConst& k, const Grid& g, const Buffers& b, uint32_t i,
                                           float dt) {
    if (i >= k.n) return;
    const float4 pi = static_cast<const float4*>(b.pos4)[i];
    const float4 vi = forces_vel<DIV>(b, i, dt);
    if (DIV) b.p_div[i] = b.p[i];
    float avx = 0.0f, avy = 0.0f, avz = 0.0f, acx = 0.0f, acy = 0.0f, acz = 0.0f;
    auto ld = [&](uint32_t j) { return forces_vel<DIV>(b, j, dt); };
    if (k.coh != 0.0f)
        visit_neighbors_x<KIND, FMT>(k, g, b, i, pi, ld, ForcesPair<true>{k, vi, avx, avy, avz, acx, acy, acz});
    else
        visit_neighbors_x<KIND, FMT>(k, g, b, i, pi, ld, ForcesPair<false>{k, vi, avx, avy, avz, acx, acy, acz});
    forces_finish(k, b, i, vi, dt, avx, avy, avz, acx, acy, acz);
}
// The same through this wave's LDS windows; every lane of the wave makes the call.
template <bool DIV>
__device__ __forceinline__ void forces_staged(const SimConst& k, const Grid& g, const Buffers& b, uint32_t i,
                                              float dt, const StageWin<4>& w) {
    const bool live = i < k.n;
    const uint32_t il = live ? i : k.n - 1u;
    const float4 pi = static_cast<const float4*>(b.pos4)[il];
    const float4 vi = forces_vel<DIV>(b, il, dt);
    if (DIV && live) b.p_div[i] = b.p[i];
    float avx = 0.0f, avy = 0.0f, avz = 0.0f, acx = 0.0f, acy = 0.0f, acz = 0.0f;
    auto ld = [&](uint32_t j) { return forces_vel<DIV>(b, j, dt); };
    if (k.coh != 0.0f)
        visit_staged<4>(k, g, b, i, live, pi, w, ld, ForcesPair<true>{k, vi, avx, avy, avz, acx, acy, acz});
    else
        visit_staged<4>(k, g, b, i, live, pi, w, ld, ForcesPair<false>{k, vi, avx, avy, avz, acx, acy, acz});
    if (live) forces_finish(k, b, i, vi, dt, avx, avy, avz, acx, acy, acz);
}

// ---- density source: rho* = rho/rho0 + dt V_i sum (v_i - v_j).gradW + dt v_i.G_b
struct AdvPair {
    const SimConst& k;
    float4 vi;
    float& delta;
    __device__ __forceinline__ void operator()(uint32_t, float4, float4 vj, float dx, float dy, float dz,
                                               float r2) const {
        const float inv_r = rsq(r2);
        const float s = grad_t(pair_t(r2, inv_r, k), inv_r, k);
        delta += s * ((vi.x - vj.x) * dx + (vi.y - vj.y) * dy + (vi.z - vj.z) * dz);
    }
};
template <int FMT>
__device__ __forceinline__ void density_adv_finish(const SimConst& k, const Buffers& b, uint32_t i, float4 vi,
                                                   float dt, float delta) {
    delta *= k.gfac;
    const float rho = b.rho[i];
    float adv = rho / fmaxf(k.rho0, 1e-7f) + dt * (k.mass / fmaxf(rho, kEps)) * delta;
    if (near_boundary<FMT>(k, b, i)) {
        const float4 gb = static_cast<const float4*>(b.gb4)[i];
        adv += dt * (vi.x * gb.x + vi.y * gb.y + vi.z * gb.z);
    }
    b.rho_adv[i] = adv;
}
template <int KIND, int FMT>
__device__ __forceinline__ void density_adv_one(const SimConst& k, const Grid& g, const Buffers& b, uint32_t i,
                                                float dt) {
    if (i >= k.n) return;
    const float4* __restrict__ vel = static_cast<const float4*>(b.vel4b);
    const float4 pi = static_cast<const float4*>(b.pos4)[i];
    const float4 vi = vel[i];
    float delta = 0.0f;
    visit_neighbors_x<KIND, FMT>(k, g, b, i, pi, [&](uint32_t j) { return ld_nbr(vel, j); }, AdvPair{k, vi, delta});
    density_adv_finish<FMT>(k, b, i, vi, dt, delta);
}
// The same through this wave's LDS windows; every lane of the wave makes the call.
__device__ __forceinline__ void density_adv_staged(const SimConst& k, const Grid& g, const Buffers& b, uint32_t i,
                                                   float dt, const StageWin<3>& w) {
    const bool live = i < k.n;
    const uint32_t il = live ? i : k.n - 1u;
    const float4* __restrict__ vel = static_cast<const float4*>(b.vel4b);
    const float4 pi = static_cast<const float4*>(b.pos4)[il];
    const float4 vi = vel[il];
    float delta = 0.0f;
    visit_staged<3>(k, g, b, i, live, pi, w, [&](uint32_t j) { return ld_nbr(vel, j); }, AdvPair{k, vi, delta});
    if (live) density_adv_finish<kFmtMask>(k, b, i, vi, dt, delta);
}

// ---- solve start: initial pressure of solve `solver` -------------------------
__device__ __forceinline__ void begin_solve_one(const SimConst& k, const Buffers& b, uint32_t i,
                                                int solver, int warm, float dt, float dt_prev) {
    if (i >= k.n) return;
    float p0 = 0.0f;
    if (warm == DFSPH_WARM_PREVIOUS) {
        // kappa-style scaling (dt_prev/dt)^k, never above 1: after a sudden
        // drop of the CFL step (one fast particle) scaling up every pressure
        // by a large factor over-pressures the whole scene, while starting
        // low only costs iterations.
        const float ratio = dt_prev > 0.0f ? fminf(dt_prev / dt, 1.0f) : 0.0f;
        p0 = solver == 0 ? b.pdiv_prev[i] * ratio : b.pden_prev[i] * ratio * ratio;
    } else if (warm == DFSPH_WARM_EXTERNAL) {
        const float* w = solver == 0 ? b.warm_ext_div : b.warm_ext_den;
        p0 = w ? w[b.ids[i]] : 0.0f;
    }
    const float rho = b.rho[i];
    b.p[i] = p0;
    static_cast<float4*>(b.pos4)[i].w = p0 / fmaxf(rho * rho, kEps);
}

// ---- reorder sorted slot ks from its stage slot perm[ks] ----------------------
__device__ __forceinline__ void reorder_one(const Buffers& b, uint32_t ks) {
    const uint32_t s = b.perm[ks];
    const float4 x = static_cast<const float4*>(b.x_stage)[s];
    static_cast<float4*>(b.pos4)[ks] = make_float4(x.x, x.y, x.z, 0.0f);
    static_cast<float4*>(b.vel4)[ks] = static_cast<const float4*>(b.v_stage)[s];
    b.ids[ks] = b.id_stage[s];
    b.pdiv_prev[ks] = b.pdiv_stage[s];
    b.pden_prev[ks] = b.pden_stage[s];
}

// Step start of one scene: dt from the CFL condition on the previous step's
// max |v| (or the fixed dt), solve status cleared.
__device__ __forceinline__ void begin_step_one(DeviceStatus* st, float dt_fixed, float cfl_2r, float dt_max) {
    const float vmax = sqrtf(__uint_as_float(st->vmax2_bits));
    float dt = dt_fixed;
    if (!(dt > 0.0f)) dt = vmax > 0.0f ? fminf(cfl_2r / vmax, dt_max) : dt_max;
    st->dt_prev = st->dt;
    st->dt = dt;
    st->vmax2_start = st->vmax2_bits;
    st->vmax2_bits = 0u;
    st->nonfinite = 0ull;
    for (int s = 0; s < 2; ++s) st->solve[s] = SolveStatus{0.0, 0, 0, 0, 0.0f};
}

// Max |v|^2 and the non-finite count of a scene from its per-tile partials
// (256-thread workgroup; thread 0 writes).
__device__ __forceinline__ void step_finish_one(DeviceStatus* st, const uint32_t* __restrict__ part,
                                                uint32_t ntiles) {
    StepPart p{0u, 0u, {}};
    for (uint32_t q = threadIdx.x; q < ntiles; q += blockDim.x) {
        p.v2bits = max(p.v2bits, part[2 * q]);
        p.bad += part[2 * q + 1];
    }
    p = block_step_part(p);
    if (threadIdx.x == 0) {
        st->vmax2_bits = p.v2bits;
        st->nonfinite = p.bad;
    }
}
// The same, then (with mr) the status published to the pinned mirror under
// ticket: the sequence word odd while the status is written, 2 ticket after.
__device__ __forceinline__ void step_finish_block(DeviceStatus* st, const uint32_t* __restrict__ part,
                                                  uint32_t ntiles, StepMirror* mr, uint32_t ticket) {
    step_finish_one(st, part, ntiles);
    if (mr && threadIdx.x == 0) {   // the thread that wrote the step's fields
        const uint32_t seq = 2u * ticket;
        __hip_atomic_store(&mr->seq, seq - 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __threadfence_system();
        mr->status = *st;
        __threadfence_system();
        __hip_atomic_store(&mr->seq, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

}  // namespace sb
}  // namespace dfsph
