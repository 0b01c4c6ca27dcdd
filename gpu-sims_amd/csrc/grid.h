// grid.h -- device-side cell arithmetic shared by the hash and sweep kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "dfsph_internal.h"

namespace dfsph {

// 10-bit Morton interleave (shaders/morton_code.comp.glsl:24-36).
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
__device__ __forceinline__ uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return (spread10(x) << 2) | (spread10(y) << 1) | spread10(z);
}

// Cell coordinate along one axis.  The REFERENCE grid reproduces the
// reference's IEEE (x - dmin)/cell, clamp and truncation bit for bit
// (shaders/morton_code.comp.glsl:44-47); the FINE grid only has to agree with
// itself, so it may use the same division (kept identical for simplicity).
// HIP's fp32 division is correctly rounded unless fast division is requested,
// and a subtraction followed by a division cannot be contracted.
__device__ __forceinline__ int cell_coord(float x, float dmin, float cell, uint32_t axes) {
    const float rel = (x - dmin) / cell;
    const float c = fminf(fmaxf(rel, 0.0f), float(axes) - 1.0f);
    return int(c);
}

// Sort key of a position.
template <int KIND>
__device__ __forceinline__ uint32_t cell_key(const Grid& g, float x, float y, float z) {
    const int cx = cell_coord(x, g.dmin[0], g.cell, g.axes[0]);
    const int cy = cell_coord(y, g.dmin[1], g.cell, g.axes[1]);
    const int cz = cell_coord(z, g.dmin[2], g.cell, g.axes[2]);
    if (KIND == DFSPH_GRID_REFERENCE) return morton3(uint32_t(cx), uint32_t(cy), uint32_t(cz));
    return (uint32_t(cz) * g.axes[1] + uint32_t(cy)) * g.axes[0] + uint32_t(cx);
}

// Visit every candidate slot of the 3x3x3 cell neighbourhood of (x, y, z).
// REFERENCE: 27 single-cell ranges in dz, dy, dx order (density_alpha.comp.glsl:112-122).
// FINE: 9 contiguous x-runs of up to three cells (linear cell order is x fastest).
template <int KIND, class F>
__device__ __forceinline__ void for_each_candidate(const Grid& g, const uint32_t* __restrict__ start,
                                                   float x, float y, float z, F&& fn) {
    const int cx = cell_coord(x, g.dmin[0], g.cell, g.axes[0]);
    const int cy = cell_coord(y, g.dmin[1], g.cell, g.axes[1]);
    const int cz = cell_coord(z, g.dmin[2], g.cell, g.axes[2]);
    const int ax = int(g.axes[0]), ay = int(g.axes[1]), az = int(g.axes[2]);
    for (int dz = -1; dz <= 1; ++dz) {
        const int z_ = cz + dz;
        if (z_ < 0 || z_ >= az) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y_ = cy + dy;
            if (y_ < 0 || y_ >= ay) continue;
            if (KIND == DFSPH_GRID_FINE) {
                const uint32_t row = (uint32_t(z_) * uint32_t(ay) + uint32_t(y_)) * uint32_t(ax);
                const uint32_t lo = row + uint32_t(cx > 0 ? cx - 1 : 0);
                const uint32_t hi = row + uint32_t(cx + 1 < ax ? cx + 1 : ax - 1);
                const uint32_t s = start[lo], e = start[hi + 1];
                for (uint32_t j = s; j < e; ++j) fn(j);
            } else {
                for (int dx = -1; dx <= 1; ++dx) {
                    const int x_ = cx + dx;
                    if (x_ < 0 || x_ >= ax) continue;
                    const uint32_t m = morton3(uint32_t(x_), uint32_t(y_), uint32_t(z_));
                    const uint32_t s = start[m], e = start[m + 1];
                    for (uint32_t j = s; j < e; ++j) fn(j);
                }
            }
        }
    }
}

// FINE grid only: visit the in-range x-runs as (run index (dz+1)*3+(dy+1), first
// slot, end slot).
template <class F>
__device__ __forceinline__ void for_each_run_fine(const Grid& g, const uint32_t* __restrict__ start,
                                                  float x, float y, float z, F&& fn) {
    const int cx = cell_coord(x, g.dmin[0], g.cell, g.axes[0]);
    const int cy = cell_coord(y, g.dmin[1], g.cell, g.axes[1]);
    const int cz = cell_coord(z, g.dmin[2], g.cell, g.axes[2]);
    const int ax = int(g.axes[0]), ay = int(g.axes[1]), az = int(g.axes[2]);
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz) {
        const int z_ = cz + dz;
        if (z_ < 0 || z_ >= az) continue;
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
            const int y_ = cy + dy;
            if (y_ < 0 || y_ >= ay) continue;
            const uint32_t row = (uint32_t(z_) * uint32_t(ay) + uint32_t(y_)) * uint32_t(ax);
            const uint32_t lo = row + uint32_t(cx > 0 ? cx - 1 : 0);
            const uint32_t hi = row + uint32_t(cx + 1 < ax ? cx + 1 : ax - 1);
            fn((dz + 1) * 3 + (dy + 1), lo, hi);
        }
    }
}

// FINE grid only: the in-range x-runs of (x, y, z)'s neighbourhood clipped to
// the sphere r^2 < r2max around it.  A run whose (y, z) row lies at least that
// far away is skipped, and a run's end cell is dropped when the sphere cannot
// reach across the particle's own cell face into it.  Distances are taken to
// the faces of the particle's cell (clamped at zero, so a particle outside
// its clamped cell still gets a lower bound), so no candidate within r2max is
// ever dropped; r2max carries a margin over h^2 that covers their rounding.
// fn(run index (dz+1)*3+(dy+1), first cell of the unclipped run, first cell,
// last cell).
template <class F>
__device__ __forceinline__ void for_each_run_clipped(const Grid& g, float x, float y, float z, float r2max,
                                                     F&& fn) {
    const int cx = cell_coord(x, g.dmin[0], g.cell, g.axes[0]);
    const int cy = cell_coord(y, g.dmin[1], g.cell, g.axes[1]);
    const int cz = cell_coord(z, g.dmin[2], g.cell, g.axes[2]);
    const int ax = int(g.axes[0]), ay = int(g.axes[1]), az = int(g.axes[2]);
    const float lox = g.dmin[0] + float(cx) * g.cell, loy = g.dmin[1] + float(cy) * g.cell,
                loz = g.dmin[2] + float(cz) * g.cell;
    const float xl = x - lox, xr = lox + g.cell - x;
    const float dyl = fmaxf(y - loy, 0.0f), dyh = fmaxf(loy + g.cell - y, 0.0f);
    const float dzl = fmaxf(z - loz, 0.0f), dzh = fmaxf(loz + g.cell - z, 0.0f);
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz) {
        const int z_ = cz + dz;
        if (z_ < 0 || z_ >= az) continue;
        const float ez = dz < 0 ? dzl : (dz > 0 ? dzh : 0.0f);
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
            const int y_ = cy + dy;
            if (y_ < 0 || y_ >= ay) continue;
            const float ey = dy < 0 ? dyl : (dy > 0 ? dyh : 0.0f);
            const float d2 = ey * ey + ez * ez;
            if (d2 >= r2max) continue;
            const float w2 = r2max - d2;
            const uint32_t row = (uint32_t(z_) * uint32_t(ay) + uint32_t(y_)) * uint32_t(ax);
            const bool left = cx > 0 && (xl < 0.0f || xl * xl < w2);
            const bool right = cx + 1 < ax && (xr < 0.0f || xr * xr < w2);
            const uint32_t lo_u = row + uint32_t(cx > 0 ? cx - 1 : 0);
            fn((dz + 1) * 3 + (dy + 1), lo_u, row + uint32_t(left ? cx - 1 : cx), row + uint32_t(right ? cx + 1 : cx));
        }
    }
}

// r^2 of a pair with the reference's evaluation order and no contraction:
// length(r_ij)^2 = (x*x + y*y) + z*z (density_alpha.comp.glsl:126).
__device__ __forceinline__ float dist2_exact(float dx, float dy, float dz) {
#pragma clang fp contract(off)
    return (dx * dx + dy * dy) + dz * dz;
}

// Workgroup index with the workgroups that one XCD receives (round-robin
// dispatch: workgroup i runs on XCD i % 8) mapped to one contiguous range of
// tiles, so that each XCD's L2 sees one spatially coherent slab.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t i, uint32_t nblk, int remap) {
    if (!remap) return i;
    const uint32_t per = nblk / 8u, rem = nblk % 8u;
    const uint32_t xcd = i % 8u, local = i / 8u;
    return xcd < rem ? xcd * (per + 1u) + local : rem * (per + 1u) + (xcd - rem) * per + local;
}

}  // namespace dfsph
