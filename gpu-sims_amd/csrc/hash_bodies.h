// hash_bodies.h -- device bodies of the neighbour-grid build (cell keys, cell
// histogram, exclusive scan with sentinel, scatter stable by (key, original
// index)), shared by the per-stage kernels of hash.hip and the persistent
// batched substep of batch.hip.
//
// Replaces the reference's morton_code -> cell_count -> prefix_sum_{local,
// block,block_l2,addback} -> scatter chain (main.cpp:2693-2744).  Semantics
// follow the reference's CPU twin (main.cpp:1452-1493): cell_start[c] is the
// exclusive scan of the counts, cell_start[cells] = N, and slots inside a
// cell are ordered by original index.
#pragma once

#include "grid.h"

namespace dfsph {
namespace hb {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;                         // cells per thread
constexpr int kScanTile = kScanThreads * kScanItems;  // cells per workgroup tile
constexpr uint32_t kNoKey = 0xFFFFFFFFu;

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

// Key of particle i at p (cell index + cell_base) into keys[i] and the cell
// histogram.  Called by every lane of a wave; lanes with valid == false take
// part in the run detection only.
template <int KIND>
__device__ __forceinline__ void keys_count_at(const Grid& g, uint32_t cell_base, float4 p, uint32_t i, bool valid,
                                              uint32_t* __restrict__ keys, uint32_t* __restrict__ count) {
    uint32_t k = kNoKey;
    if (valid) {
        k = cell_base + cell_key<KIND>(g, p.x, p.y, p.z);
        keys[i] = k;
    }
    const WaveRun run = wave_run(k);
    if (run.head && k != kNoKey) atomicAdd(&count[k], run.length);
}
template <int KIND>
__device__ __forceinline__ void keys_count(const Grid& g, uint32_t cell_base, const float4* __restrict__ pos,
                                           uint32_t i, bool valid, uint32_t* __restrict__ keys,
                                           uint32_t* __restrict__ count) {
    keys_count_at<KIND>(g, cell_base, valid ? pos[i] : float4{}, i, valid, keys, count);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread across a 256-thread workgroup;
// returns the workgroup total through *total.
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

// Scan phase 1: sum of the counts of cell tile `tile` into tile_sum[tile].
__device__ __forceinline__ void scan_tile_sum(uint32_t cells, const uint32_t* __restrict__ count,
                                              uint32_t* __restrict__ tile_sum, uint32_t tile) {
    const uint32_t base = tile * kScanTile + threadIdx.x * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < cells) s += count[base + k];
    uint32_t tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) tile_sum[tile] = tot;
}

// Scan phase 2 (one workgroup): tile sums scanned in place (exclusive).
__device__ __forceinline__ void scan_tile_sums(uint32_t ntiles, uint32_t* __restrict__ tile_sum) {
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

// Scan phase 3: cell_start of tile `tile`, counts cleared (they become the
// scatter cursors); the thread owning the last cell writes start[cells] = N.
__device__ __forceinline__ void scan_apply(uint32_t cells, uint32_t* __restrict__ count,
                                           const uint32_t* __restrict__ tile_sum,
                                           uint32_t* __restrict__ start, uint32_t tile) {
    const uint32_t base = tile * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < cells ? count[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = tile_sum[tile] + block_excl_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < cells) {
            start[base + k] = run;
            count[base + k] = 0u;
        }
        run += v[k];
    }
    if (base < cells && base + kScanItems >= cells) start[cells] = run;
}

// Scan phases 2 and 3 in one pass: the workgroup of tile `tile` sums the
// unscanned tile sums of the tiles before it itself (a few hundred loads
// spread over the workgroup), so no one-workgroup phase between is needed.
__device__ __forceinline__ void scan_apply_direct(uint32_t cells, uint32_t* __restrict__ count,
                                                  const uint32_t* __restrict__ tile_sum,
                                                  uint32_t* __restrict__ start, uint32_t tile) {
    __shared__ uint32_t wpre[kScanThreads / 64];
    uint32_t pre = 0;
    for (uint32_t q = threadIdx.x; q < tile; q += kScanThreads) pre += tile_sum[q];
    for (int o = 32; o > 0; o >>= 1) pre += __shfl_xor(pre, o, 64);
    if ((threadIdx.x & 63) == 0) wpre[threadIdx.x >> 6] = pre;
    __syncthreads();
    uint32_t prefix = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) prefix += wpre[w];
    const uint32_t base = tile * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < cells ? count[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = prefix + block_excl_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < cells) {
            start[base + k] = run;
            count[base + k] = 0u;
        }
        run += v[k];
    }
    if (base < cells && base + kScanItems >= cells) start[cells] = run;
}

// Sort keys of a cell's slots (perm entries are stage indices): the
// original index (ids[p], or p itself when ids is null), or, with stage
// positions xs, (x, original index) -- x as its order-preserving bit pattern
// in the high word -- so that a row of cells is sorted by x.
struct CellKey {
    const uint32_t* ids;
    const float4* xs;
    __device__ __forceinline__ uint64_t operator()(uint32_t p) const {
        const uint64_t id = ids ? ids[p] : p;
        if (!xs) return id;
        const uint32_t u = __float_as_uint(xs[p].x);
        const uint32_t ord = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
        return (uint64_t(ord) << 32) | id;
    }
};

// perm[start[key] + slot] = i (and, with skey, skey[that slot] = kf(i), read
// coalesced here instead of gathered by the per-cell sort).  Called by every
// lane of a wave.
__device__ __forceinline__ void scatter_one(uint32_t i, bool valid, const uint32_t* __restrict__ keys,
                                            const uint32_t* __restrict__ start,
                                            uint32_t* __restrict__ cursor, uint32_t* __restrict__ perm,
                                            uint64_t* __restrict__ skey = nullptr, CellKey kf = CellKey{}) {
    const uint32_t k = valid ? keys[i] : kNoKey;
    const WaveRun run = wave_run(k);
    uint32_t base = 0u;
    if (run.head && k != kNoKey) base = start[k] + atomicAdd(&cursor[k], run.length);
    base = __shfl(base, int(run.head_lane), 64);
    if (valid) {
        const uint32_t slot = base + ((threadIdx.x & 63u) - run.head_lane);
        perm[slot] = i;
        if (skey) skey[slot] = kf(i);
    }
}

// Batcher's odd-even merge sort network over N (power of two) (key, value)
// pairs in registers: compile-time indices only, so nothing spills.
template <int N, class Key>
__device__ __forceinline__ void sort_network(Key (&key)[N], uint32_t (&val)[N]) {
#pragma unroll
    for (int p = 1; p < N; p <<= 1)
#pragma unroll
        for (int k = p; k >= 1; k >>= 1)
#pragma unroll
            for (int j = k % p; j <= N - 1 - k; j += 2 * k)
#pragma unroll
                for (int i = 0; i <= (k - 1 < N - j - k - 1 ? k - 1 : N - j - k - 1); ++i)
                    if ((i + j) / (2 * p) == (i + j + k) / (2 * p)) {
                        const int a = i + j, b = i + j + k;
                        const bool sw = key[a] > key[b];
                        const Key ka = key[a], kb = key[b];
                        const uint32_t va = val[a], vb = val[b];
                        key[a] = sw ? kb : ka;
                        key[b] = sw ? ka : kb;
                        val[a] = sw ? vb : va;
                        val[b] = sw ? va : vb;
                    }
}


// Cell c's slots sorted by key (no two slots share one: the original index
// is part of it), then the cursor is cleared for the next build.  Cells of up
// to kRegSort (8 or 16) slots load every slot at once and sort in registers;
// larger ones use an insertion sort in place.  Either way the order is the
// same; the 16-slot network takes about 100 VGPRs, so a kernel that must
// stay under that (the persistent batch kernel) sorts in registers up to 8.
template <int N>
__device__ __forceinline__ void cell_sort_regs(uint32_t s, uint32_t n, uint32_t* __restrict__ perm, CellKey kf,
                                               const uint64_t* __restrict__ skey) {
    uint64_t key[N];
    uint32_t val[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
        val[q] = uint32_t(q) < n ? perm[s + q] : 0u;
        key[q] = uint32_t(q) < n ? (skey ? skey[s + q] : kf(val[q])) : ~0ull;   // padding sorts last
    }
    sort_network<N>(key, val);
#pragma unroll
    for (int q = 0; q < N; ++q)
        if (uint32_t(q) < n) perm[s + q] = val[q];
}

template <int kRegSort = 16>
__device__ __forceinline__ void cell_order_one(uint32_t c, const uint32_t* __restrict__ start,
                                               uint32_t* __restrict__ cursor, uint32_t* __restrict__ perm,
                                               const uint32_t* __restrict__ ids,
                                               const float4* __restrict__ xs = nullptr,
                                               const uint64_t* __restrict__ skey = nullptr) {
    const CellKey kf{ids, xs};
    const uint32_t s = start[c], e = start[c + 1];
    cursor[c] = 0u;
    const uint32_t n = e - s;
    if (n <= 1u) return;
    if (n <= 8u) {
        cell_sort_regs<8>(s, n, perm, kf, skey);
        return;
    }
    if (kRegSort >= 16 && n <= 16u) {
        cell_sort_regs<16>(s, n, perm, kf, skey);
        return;
    }
    for (uint32_t a = s + 1; a < e; ++a) {
        const uint32_t pa = perm[a];
        const uint64_t ka = kf(pa);
        uint32_t b = a;
        while (b > s) {
            const uint32_t pb = perm[b - 1];
            if (kf(pb) <= ka) break;
            perm[b] = pb;
            --b;
        }
        perm[b] = pa;
    }
}

}  // namespace hb
}  // namespace dfsph
