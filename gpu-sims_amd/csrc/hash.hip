// hash.hip -- per-stage kernels of the neighbour-grid build for one scene
// (bodies in hash_bodies.h): keys + histogram, three-phase exclusive scan,
// scatter, per-cell order by original index.  The reference's GPU scan is
// mis-wired and its atomic scatter order is not reproducible; this build is
// deterministic and follows the reference's CPU twin (main.cpp:1452-1493).
#include "hash_bodies.h"
#include "sweep_bodies.h"
#include "launch.h"

namespace dfsph {

namespace {

using namespace hb;

template <int KIND>
__global__ void __launch_bounds__(256) k_keys(Grid g, const float4* __restrict__ pos, uint32_t n,
                                              uint32_t* __restrict__ keys,
                                              uint32_t* __restrict__ count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    keys_count<KIND>(g, 0u, pos, i, i < n, keys, count);
}

// With fin.status set, the last workgroup does the substep's finish instead.
__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(uint32_t cells,
                                                             const uint32_t* __restrict__ count,
                                                             uint32_t* __restrict__ tile_sum, StepFinish fin) {
    if (fin.status && blockIdx.x == gridDim.x - 1) {
        sb::step_finish_block(fin.status, fin.part, fin.ntiles, fin.mirror, fin.ticket);
        return;
    }
    scan_tile_sum(cells, count, tile_sum, blockIdx.x);
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tile_sums(uint32_t ntiles,
                                                                 uint32_t* __restrict__ tile_sum) {
    scan_tile_sums(ntiles, tile_sum);
}

__global__ void __launch_bounds__(kScanThreads) k_scan_apply(uint32_t cells,
                                                             uint32_t* __restrict__ count,
                                                             const uint32_t* __restrict__ tile_sum,
                                                             uint32_t* __restrict__ start) {
    scan_apply(cells, count, tile_sum, start, blockIdx.x);
}

__global__ void __launch_bounds__(kScanThreads) k_scan_apply_direct(uint32_t cells,
                                                                    uint32_t* __restrict__ count,
                                                                    const uint32_t* __restrict__ tile_sum,
                                                                    uint32_t* __restrict__ start) {
    scan_apply_direct(cells, count, tile_sum, start, blockIdx.x);
}

__global__ void __launch_bounds__(256) k_scatter(uint32_t n, const uint32_t* __restrict__ keys,
                                                 const uint32_t* __restrict__ start,
                                                 uint32_t* __restrict__ cursor,
                                                 uint32_t* __restrict__ perm, uint64_t* __restrict__ skey,
                                                 const uint32_t* __restrict__ ids,
                                                 const float4* __restrict__ xs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    scatter_one(i, i < n, keys, start, cursor, perm, skey, CellKey{ids, xs});
}

__global__ void __launch_bounds__(256) k_cell_order(uint32_t cells, const uint32_t* __restrict__ start,
                                                    uint32_t* __restrict__ cursor,
                                                    uint32_t* __restrict__ perm,
                                                    const uint32_t* __restrict__ ids,
                                                    const float4* __restrict__ xs,
                                                    const uint64_t* __restrict__ skey) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cells) cell_order_one(c, start, cursor, perm, ids, xs, skey);
}

inline uint32_t blocks_for(uint64_t n, uint32_t t) { return uint32_t((n + t - 1) / t); }

}  // namespace

size_t scan_scratch_elems(uint32_t cells) { return blocks_for(cells, kScanTile) + 1; }

void launch_keys(const Grid& g, const void* pos4, uint32_t n, uint32_t* keys, uint32_t* cell_count,
                 void* stream) {
    if (!n) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const float4* p = static_cast<const float4*>(pos4);
    if (g.kind == DFSPH_GRID_REFERENCE)
        DFSPH_LAUNCH(k_keys<DFSPH_GRID_REFERENCE>, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                           g, p, n, keys, cell_count);
    else
        DFSPH_LAUNCH(k_keys<DFSPH_GRID_FINE>, dim3(blocks_for(n, 256)), dim3(256), 0, s, g,
                           p, n, keys, cell_count);
}

void launch_scan(uint32_t cells, uint32_t* cell_count, uint32_t* cell_start, uint32_t* scan_blocks,
                 void* stream, const StepFinish* finish) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t tiles = blocks_for(cells, kScanTile);
    const StepFinish fin = finish ? *finish : StepFinish{};
    DFSPH_LAUNCH(k_scan_tiles, dim3(tiles + (fin.status ? 1u : 0u)), dim3(kScanThreads), 0, s, cells, cell_count,
                       scan_blocks, fin);
    // Up to 4096 tiles (8M cells) the apply pass sums the unscanned tile sums
    // before its tile directly (two launches; at most 16 loads per thread);
    // beyond that the one-workgroup phase scans them first (three).
    if (tiles <= 4096u) {
        DFSPH_LAUNCH(k_scan_apply_direct, dim3(tiles), dim3(kScanThreads), 0, s, cells, cell_count,
                           scan_blocks, cell_start);
        return;
    }
    DFSPH_LAUNCH(k_scan_tile_sums, dim3(1), dim3(kScanThreads), 0, s, tiles, scan_blocks);
    DFSPH_LAUNCH(k_scan_apply, dim3(tiles), dim3(kScanThreads), 0, s, cells, cell_count,
                       scan_blocks, cell_start);
}

void launch_scatter(uint32_t n, const uint32_t* keys, const uint32_t* cell_start,
                    uint32_t* cell_count, uint32_t* perm, const uint32_t* ids, const void* xs4,
                    uint64_t* skey, uint32_t cells, void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    const float4* xs = static_cast<const float4*>(xs4);
    if (n)
        DFSPH_LAUNCH(k_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, keys,
                           cell_start, cell_count, perm, skey, ids, xs);
    DFSPH_LAUNCH(k_cell_order, dim3(blocks_for(cells, 256)), dim3(256), 0, s, cells,
                       cell_start, cell_count, perm, ids, xs, static_cast<const uint64_t*>(skey));
}

}  // namespace dfsph
