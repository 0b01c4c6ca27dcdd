// hash.hip -- neighbour-grid build: cell keys, cell histogram, exclusive scan
// with sentinel, and a scatter whose order is stable by (key, original index).
//
// Replaces the reference's morton_code -> cell_count -> prefix_sum_{local,
// block,block_l2,addback} -> scatter chain (main.cpp:2693-2744).  Semantics
// follow the reference's CPU twin (main.cpp:1452-1493): cell_start[c] is the
// exclusive scan of the counts, cell_start[cells] = N, and slots inside a
// cell are ordered by original index.  The reference's GPU scan is mis-wired
// and its atomic scatter order is not reproducible; this one is deterministic.
#include "grid.h"

namespace dfsph {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;                         // cells per thread
constexpr int kScanTile = kScanThreads * kScanItems;  // cells per workgroup

// Lanes of a wave holding the same key in consecutive lanes form a run (input
// in the previous cell order makes runs of ~8).  One atomic per run instead of
// one per lane: the head lane adds the run length.
struct WaveRun {
    uint32_t head_lane;   // first lane of this lane's run
    uint32_t length;      // valid for the head lane
    bool head;
};
__device__ __forceinline__ WaveRun wave_run(uint32_t key) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t prev = __shfl_up(key, 1, 64);
    const bool head = lane == 0u || prev != key;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = heads & ((lane == 63u) ? ~0ull : ((2ull << lane) - 1ull));
    const uint32_t head_lane = 63u - uint32_t(__clzll(upto));
    const uint64_t after = heads & ~((lane == 63u) ? ~0ull : ((2ull << lane) - 1ull));
    const uint32_t next = after ? uint32_t(__ffsll((unsigned long long)after)) - 1u : 64u;
    return WaveRun{head_lane, next - lane, head};
}

// ---- keys + histogram -------------------------------------------------------
template <int KIND>
__global__ void __launch_bounds__(256) k_keys(Grid g, const float4* __restrict__ pos, uint32_t n,
                                              uint32_t* __restrict__ keys,
                                              uint32_t* __restrict__ count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t k = 0xFFFFFFFFu;   // tail lanes: a key no cell has
    if (i < n) {
        const float4 p = pos[i];
        k = cell_key<KIND>(g, p.x, p.y, p.z);
        keys[i] = k;
    }
    const WaveRun run = wave_run(k);
    if (run.head && k != 0xFFFFFFFFu) atomicAdd(&count[k], run.length);
}

// ---- scan: per-tile sums, tile-sum scan, tile scan + add --------------------
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread across the workgroup; returns the
// workgroup total through *total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        if (w < wid) base += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(uint32_t cells,
                                                             const uint32_t* __restrict__ count,
                                                             uint32_t* __restrict__ tile_sum) {
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < cells) s += count[base + k];
    uint32_t tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

// One workgroup scans the tile sums in place (exclusive), in chunks.
__global__ void __launch_bounds__(kScanThreads) k_scan_tile_sums(uint32_t ntiles,
                                                                 uint32_t* __restrict__ tile_sum) {
    uint32_t carry = 0;
    for (uint32_t b = 0; b < ntiles; b += kScanThreads) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t v = i < ntiles ? tile_sum[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, &tot);
        if (i < ntiles) tile_sum[i] = carry + ex;
        carry += tot;
    }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_apply(uint32_t cells,
                                                             uint32_t* __restrict__ count,
                                                             const uint32_t* __restrict__ tile_sum,
                                                             uint32_t* __restrict__ start) {
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < cells ? count[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = tile_sum[blockIdx.x] + block_excl_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < cells) {
            start[base + k] = run;
            count[base + k] = 0u;      // reused as the scatter cursor
        }
        run += v[k];
    }
    // The thread owning the last cell writes the sentinel start[cells] = N.
    if (base < cells && base + kScanItems >= cells) start[cells] = run;
}

// ---- scatter + per-cell order by original index ------------------------------
__global__ void __launch_bounds__(256) k_scatter(uint32_t n, const uint32_t* __restrict__ keys,
                                                 const uint32_t* __restrict__ start,
                                                 uint32_t* __restrict__ cursor,
                                                 uint32_t* __restrict__ perm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t k = i < n ? keys[i] : 0xFFFFFFFFu;
    const WaveRun run = wave_run(k);
    uint32_t base = 0u;
    if (run.head && k != 0xFFFFFFFFu) base = start[k] + atomicAdd(&cursor[k], run.length);
    base = __shfl(base, int(run.head_lane), 64);
    if (i < n) perm[base + ((threadIdx.x & 63u) - run.head_lane)] = i;
}

// One thread per cell: insertion sort of the cell's slots by original index,
// then clear the cursor for the next build.
__global__ void __launch_bounds__(256) k_cell_order(uint32_t cells, const uint32_t* __restrict__ start,
                                                    uint32_t* __restrict__ cursor,
                                                    uint32_t* __restrict__ perm,
                                                    const uint32_t* __restrict__ ids) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const uint32_t s = start[c], e = start[c + 1];
    cursor[c] = 0u;
    for (uint32_t a = s + 1; a < e; ++a) {
        const uint32_t pa = perm[a];
        const uint32_t ia = ids ? ids[pa] : pa;
        uint32_t b = a;
        while (b > s) {
            const uint32_t pb = perm[b - 1];
            const uint32_t ib = ids ? ids[pb] : pb;
            if (ib <= ia) break;
            perm[b] = pb;
            --b;
        }
        perm[b] = pa;
    }
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
        hipLaunchKernelGGL(k_keys<DFSPH_GRID_REFERENCE>, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                           g, p, n, keys, cell_count);
    else
        hipLaunchKernelGGL(k_keys<DFSPH_GRID_FINE>, dim3(blocks_for(n, 256)), dim3(256), 0, s, g,
                           p, n, keys, cell_count);
}

void launch_scan(uint32_t cells, uint32_t* cell_count, uint32_t* cell_start, uint32_t* scan_blocks,
                 void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t tiles = blocks_for(cells, kScanTile);
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanThreads), 0, s, cells, cell_count,
                       scan_blocks);
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(kScanThreads), 0, s, tiles, scan_blocks);
    hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(kScanThreads), 0, s, cells, cell_count,
                       scan_blocks, cell_start);
}

void launch_scatter(uint32_t n, const uint32_t* keys, const uint32_t* cell_start,
                    uint32_t* cell_count, uint32_t* perm, const uint32_t* ids, uint32_t cells,
                    void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n)
        hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, keys,
                           cell_start, cell_count, perm);
    hipLaunchKernelGGL(k_cell_order, dim3(blocks_for(cells, 256)), dim3(256), 0, s, cells,
                       cell_start, cell_count, perm, ids);
}

}  // namespace dfsph
