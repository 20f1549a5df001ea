// The per-step pipeline kernels for gfx950.
//
//   hash      calc_grid.comp: agent -> cell key and its offset in its bucket: an atomic count per cell; one
//             inactive bucket per block of 256 agents, counted and ranked by the block's ballots
//   scan      two passes over the bucket counts (tile sums, then tile-local exclusive scans): bucket starts
//             (the cell ranges) and the counts cleared for the next step
//   scatter   agent -> its bucket at start + offset
//   fixgather each bucket's run sorted by agent id (runs are a few agents), so the permutation is the
//             stable counting sort's: ascending (cell, agent), inactive agents at the tail in agent order
//             (the reference's SortMode::Counting, cpu_crowd_sim.cpp:997-1006); the state is gathered into
//             slot order (reorder.comp)
//   choice    PHYSICS_PASS 0, the queue-aware exit choice (exit_choice.cpp:683-791): prefix, evaluate,
//             thresholds, apply (+ every agent's e_goal)
//   predict   PHYSICS_PASS 1 (cpu_crowd_sim.cpp:655-777)
//   project   PHYSICS_PASS 2 (cpu_crowd_sim.cpp:781-831): the first iteration also collects each slot's
//             contact candidates, the later ones sum over them
//   final     PHYSICS_PASS 3 (cpu_crowd_sim.cpp:835-866) + evacuation events
//
// Every kernel has a plain instance (one scene, its tables as kernel arguments) and a batched one (many
// scenes: per-agent / per-slot scenario ids select the scene's tables, its cell range and its agent range;
// cell keys are global cell indices, so the counting sort is segmented by scenario). A neighbour ring row
// [x0, x1] of cells is one contiguous slot range [bstart[row + x0], bstart[row + x1 + 1]): the slots are
// sorted by cell, so the row's slots in slot order are the ring scan's order (dy, dx, slot).
// Data is SoA FP32. The kernels are latency-bound gathers and integer atomics: no MFMA.
#include "crowd_kernels.h"
#include "seated_step.h"

#include <hip/hip_ext.h>

namespace swarm {

// base[key] += the number of lanes of this wave that pass 'valid' with this key: one atomic per distinct key of
// the wave instead of one per lane (integer sums: the result does not depend on the order).
// Every lane of the wave must reach the call.
__device__ inline void waveCountAdd(uint32_t* base, uint32_t key, bool valid) {
    uint64_t pending = __ballot(valid);
    const int lane = __lane_id();
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t k = __shfl(key, leader, 64);
        const uint64_t same = __ballot(valid && key == k) & pending;
        if (lane == leader) atomicAdd(&base[k], uint32_t(__popcll(same)));
        pending &= ~same;
    }
}

// As waveCountAdd for keys that are mostly distinct within a wave (the exit choice's distance bins): the first
// 'rounds' keys go through the wave's leaders, every lane still pending adds its own one. Each valid lane is counted
// exactly once either way.
__device__ inline void waveCountAddFew(uint32_t* base, uint32_t key, bool valid, int rounds) {
    uint64_t pending = __ballot(valid);
    const int lane = __lane_id();
    for (int r = 0; r < rounds && pending; ++r) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t k = __shfl(key, leader, 64);
        const uint64_t same = __ballot(valid && key == k) & pending;
        if (lane == leader) atomicAdd(&base[k], uint32_t(__popcll(same)));
        pending &= ~same;
    }
    if ((pending >> lane) & 1ull) atomicAdd(&base[key], 1u);
}

__device__ inline uint32_t bucketOf(uint32_t key, uint32_t agent, uint32_t C) {
    return key == INACTIVE_CELL ? C + agent / uint32_t(HASH_BLOCK) : key;
}

constexpr int CHOICE_AHEAD_ROUNDS = 2;   // wave leaders for the distance bins before lanes add their own

// ---------------- grid passes ----------------
// Agent i's offset in its bucket. The block's inactive agents are its own bucket (C + block): their offsets are
// their rank by agent id, from the block's ballots, and the bucket's count is one store. Cell buckets take an
// atomic offset. Every thread of the block must call it, with i = (the block's number in agent order) HASH_BLOCK +
// threadIdx.x.
__device__ inline void bucketOffsets(int i, int n, uint32_t k, uint32_t* __restrict__ loc, uint32_t* __restrict__ bcnt, uint32_t C) {
    __shared__ uint32_t waveInactive[HASH_BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool inactive = i < n && k == INACTIVE_CELL;
    const uint64_t m = __ballot(inactive);
    if (lane == 0) waveInactive[wave] = uint32_t(__popcll(m));
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int w = 0; w < HASH_BLOCK / 64; ++w) {
        before += w < wave ? waveInactive[w] : 0u;
        total += waveInactive[w];
    }
    if (inactive) loc[i] = before + uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
    else if (i < n) loc[i] = atomicAdd(&bcnt[k], 1u);
    if (threadIdx.x == 0) bcnt[C + uint32_t(i) / uint32_t(HASH_BLOCK)] = total;   // the block of agents i belongs to
}

template <bool B>
__global__ void __launch_bounds__(HASH_BLOCK) k_hash(int n, SceneView s0, BatchView bt, float* __restrict__ ax, float* __restrict__ ay,
                                                     float* __restrict__ aw, float* __restrict__ az, uint32_t* __restrict__ key,
                                                     uint32_t* __restrict__ loc, uint32_t* __restrict__ bcnt, uint32_t C) {
    const int i = xcdBlock() * HASH_BLOCK + threadIdx.x;
    uint32_t k = INACTIVE_CELL;
    if (i < n) {
        float x = ax[i], y = ay[i], w = aw[i];
        if (isnan(x)) {   // calc_grid.comp resets a NaN position in the input state (to (0, 0, 0), active)
            x = 0.0f; y = 0.0f; w = 1.0f;
            ax[i] = x; ay[i] = y; aw[i] = w;
            if (az) az[i] = 0.0f;
        }
        if (w != 0.0f) {
            const ScnRef r = scnOfAgent<B>(bt, uint32_t(i));
            k = r.cellBase + hashCell(sceneOf<B>(s0, bt, r.sc), x, y);
        }
        key[i] = k;
    }
    bucketOffsets(i, n, k, loc, bcnt, C);
}

// the sort stage alone on given keys (CrowdSim::sortGrid): the bucket offsets of key[i]
__global__ void __launch_bounds__(HASH_BLOCK) k_key_count(int n, const uint32_t* __restrict__ key, uint32_t* __restrict__ loc,
                                                          uint32_t* __restrict__ bcnt, uint32_t C) {
    const int i = blockIdx.x * HASH_BLOCK + threadIdx.x;
    const uint32_t k = i < n ? key[i] : INACTIVE_CELL;
    bucketOffsets(i, n, k < C ? k : INACTIVE_CELL, loc, bcnt, C);   // a key that is no cell counts in none
}

// tile sums of the bucket counts
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_reduce(const uint32_t* __restrict__ bcnt, uint32_t nb, uint32_t* __restrict__ tileSum) {
    __shared__ uint32_t red[SCAN_BLOCK / 64];
    const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) sum += base + k < nb ? bcnt[base + k] : 0u;
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < SCAN_BLOCK / 64; ++w) t += red[w];
        tileSum[blockIdx.x] = t;
    }
}

// exclusive scan of the bucket counts: the bucket starts, the active count, and the counts cleared for the next step
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_apply(uint32_t* __restrict__ bcnt, uint32_t nb, uint32_t C,
                                                          const uint32_t* __restrict__ tileSum, uint32_t* __restrict__ bstart,
                                                          uint32_t* __restrict__ ctrl) {
    __shared__ uint32_t wsum[SCAN_BLOCK / 64];
    __shared__ uint32_t tilePrefix;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t pre = 0;   // the preceding tiles
    for (uint32_t t = threadIdx.x; t < blockIdx.x; t += SCAN_BLOCK) pre += tileSum[t];
    for (int off = 32; off > 0; off >>= 1) pre += __shfl_down(pre, off, 64);
    if (lane == 0) wsum[wave] = pre;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < SCAN_BLOCK / 64; ++w) t += wsum[w];
        tilePrefix = t;
    }
    __syncthreads();
    const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) { v[k] = base + k < nb ? bcnt[base + k] : 0u; local += v[k]; }
    uint32_t incl = local;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    __syncthreads();
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t waveOff = 0;
    for (int w = 0; w < wave; ++w) waveOff += wsum[w];
    uint32_t run = tilePrefix + waveOff + incl - local;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const uint32_t j = base + k;
        if (j < nb) {
            bstart[j] = run;
            if (j == C) ctrl[CTRL_NACTIVE] = run;
            bcnt[j] = 0u;
        }
        run += v[k];
    }
}

__global__ void __launch_bounds__(HASH_BLOCK) k_scatter(int n, const uint32_t* __restrict__ key, const uint32_t* __restrict__ loc,
                                                        const uint32_t* __restrict__ bstart, uint32_t* __restrict__ tmp, uint32_t C) {
    const int i = xcdBlock() * HASH_BLOCK + threadIdx.x;
    if (i >= n) return;
    tmp[bstart[bucketOf(key[i], uint32_t(i), C)] + loc[i]] = uint32_t(i);
}

template <bool B>
__global__ void __launch_bounds__(HASH_BLOCK) k_fix_gather(int n, BatchView bt, const uint32_t* __restrict__ key, const uint32_t* __restrict__ tmp,
                                                           const uint32_t* __restrict__ bstart, uint32_t C, uint32_t* __restrict__ sid,
                                                           uint32_t* __restrict__ skey, const float* __restrict__ ax,
                                                           const float* __restrict__ ay, const float* __restrict__ rad,
                                                           float* __restrict__ sx, float* __restrict__ sy, float* __restrict__ sr) {
    const int s = xcdBlock() * HASH_BLOCK + threadIdx.x;
    if (s >= n) return;
    const uint32_t id = tmp[s];
    const uint32_t k = key[id];
    uint32_t f = uint32_t(s);   // an inactive bucket is in agent order already
    if (k != INACTIVE_CELL) {
        const uint32_t st = bstart[k], en = bstart[k + 1];
        uint32_t rank = 0;
        for (uint32_t t = st; t < en; ++t) rank += tmp[t] < id ? 1u : 0u;
        f = st + rank;
    }
    sid[f] = id;
    skey[f] = k;
    if constexpr (B) bt.slotScen[f] = bt.agentScen[id];
    if (k != INACTIVE_CELL) {
        sx[f] = ax[id];
        sy[f] = ay[id];
        sr[f] = rad[id];
    }
}

// ---------------- the GPU reference's bitonic network (the optional sort mode) ----------------
// bitonic_sort.comp / bitonic_sort_shared.comp over the hash output in agent order: stages k = 2 .. n, passes
// j = k/2 .. 1; each pair (lo, lo + j), lo with bit j clear, is ascending where (lo & k) == 0; the keys are compared
// and the pair swapped when out of its direction. The network is the reference's pass for pass, so the
// permutation (unstable within a cell) is its own. A key rides in the high word of a 64-bit pair with its agent in
// the low word; only high words are compared. Passes with j < BITONIC_TILE run on a block's tile in LDS.
// FULL: the whole 64-bit pair is compared (a total order on (key, agent): the stable order of the counting sort)
template <bool FULL> __device__ inline bool outOfOrder(uint64_t a, uint64_t b, bool asc) {
    if constexpr (FULL) return asc ? a > b : a < b;
    const uint32_t ka = uint32_t(a >> 32), kb = uint32_t(b >> 32);
    return asc ? ka > kb : ka < kb;
}
template <bool FULL> __device__ inline void bitonicCx(uint64_t* __restrict__ t, uint32_t base, uint32_t p, uint32_t k, uint32_t j) {
    const uint32_t lo = ((p & ~(j - 1u)) << 1u) | (p & (j - 1u));   // pair p: a 0 inserted at bit log2(j)
    const uint32_t hi = lo | j;
    const bool asc = ((base + lo) & k) == 0u;
    const uint64_t a = t[lo], b = t[hi];
    if (outOfOrder<FULL>(a, b, asc)) { t[lo] = b; t[hi] = a; }
}

// Inside a tile a thread owns its elements 2t and 2t + 1 in registers. A pass with j = 1 pairs them; a pass with
// 2 <= j <= 64 pairs them with those of thread t ^ (j / 2), in the same wave (read by shuffles, each thread keeping
// its side of the compare-exchange); a pass with j >= 128 goes through the tile in LDS.
__device__ inline uint64_t shflXor64(uint64_t v, int d) {
    const uint32_t lo = uint32_t(__shfl_xor(int(uint32_t(v)), d)), hi = uint32_t(__shfl_xor(int(uint32_t(v >> 32)), d));
    return (uint64_t(hi) << 32) | uint64_t(lo);
}
template <bool FULL> __device__ inline void regPass(uint64_t& r0, uint64_t& r1, uint32_t base, uint32_t t, uint32_t k, uint32_t j) {
    const uint32_t e0 = 2u * t;
    if (j == 1u) {
        const bool asc = ((base + e0) & k) == 0u;
        if (outOfOrder<FULL>(r0, r1, asc)) { const uint64_t x = r0; r0 = r1; r1 = x; }
        return;
    }
    const int d = int(j >> 1);
    const uint64_t o0 = shflXor64(r0, d), o1 = shflXor64(r1, d);
    const bool meLo = (e0 & j) == 0u;                    // the same for e0 + 1: j >= 2
    const bool asc = ((base + (meLo ? e0 : (e0 ^ j))) & k) == 0u;   // k >= 4: bit 0 of the pair's lower index is moot
    auto pick = [&](uint64_t mine, uint64_t other) {
        const uint64_t lo = meLo ? mine : other, hi = meLo ? other : mine;
        return outOfOrder<FULL>(lo, hi, asc) ? other : mine;
    };
    r0 = pick(r0, o0);
    r1 = pick(r1, o1);
}

// stage k's passes j = jTop .. 1 on the tile at 'base' held in registers (r0, r1); the passes j >= 128 in LDS
template <bool FULL>
__device__ inline void tileStage(uint64_t* __restrict__ lds, uint64_t& r0, uint64_t& r1, uint32_t base, uint32_t t, uint32_t k, uint32_t jTop) {
    uint32_t j = jTop;
    if (j >= 128u) {
        lds[2u * t] = r0;
        lds[2u * t + 1u] = r1;
        __syncthreads();
        for (; j >= 128u; j >>= 1) {
            bitonicCx<FULL>(lds, base, t, k, j);   // pair t of the tile (blockDim = tile / 2 pairs)
            __syncthreads();
        }
        r0 = lds[2u * t];
        r1 = lds[2u * t + 1u];
        __syncthreads();   // the tile is read before the next stage writes it
    }
    for (; j > 0; j >>= 1) regPass<FULL>(r0, r1, base, t, k, j);
}

// the stages 2 .. tile of each tile of 'tile' pairs (tile = min(n, BITONIC_TILE), tile / 2 threads a block), from
// the keys in agent order (key != nullptr) or from the pairs as they are
template <bool FULL>
__global__ void __launch_bounds__(BITONIC_BLOCK) k_bitonic_tile(int tile, const uint32_t* __restrict__ key, uint64_t* __restrict__ pairs) {
    __shared__ uint64_t lds[BITONIC_TILE];
    const uint32_t t = threadIdx.x;
    const uint32_t base = blockIdx.x * uint32_t(tile);
    const uint32_t e = base + 2u * t;
    uint64_t r0 = key ? (uint64_t(key[e]) << 32) | uint64_t(e) : pairs[e];
    uint64_t r1 = key ? (uint64_t(key[e + 1u]) << 32) | uint64_t(e + 1u) : pairs[e + 1u];
    for (uint32_t k = 2; k <= uint32_t(tile); k <<= 1) tileStage<FULL>(lds, r0, r1, base, t, k, k >> 1);
    pairs[e] = r0;
    pairs[e + 1u] = r1;
}

// stage k's passes j = BITONIC_TILE/2 .. 1 on each tile (n > BITONIC_TILE; BITONIC_TILE / 2 threads a block)
template <bool FULL>
__global__ void __launch_bounds__(BITONIC_BLOCK) k_bitonic_merge(uint32_t k, uint64_t* __restrict__ pairs) {
    __shared__ uint64_t lds[BITONIC_TILE];
    const uint32_t t = threadIdx.x;
    const uint32_t base = blockIdx.x * uint32_t(BITONIC_TILE);
    const uint32_t e = base + 2u * t;
    uint64_t r0 = pairs[e], r1 = pairs[e + 1u];
    tileStage<FULL>(lds, r0, r1, base, t, k, BITONIC_TILE / 2);
    pairs[e] = r0;
    pairs[e + 1u] = r1;
}

// one pass (k, j >= BITONIC_TILE) over all n/2 pairs in HBM
template <bool FULL>
__global__ void __launch_bounds__(256) k_bitonic_pass(uint32_t half, uint32_t k, uint32_t j, uint64_t* __restrict__ pairs) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p < half) bitonicCx<FULL>(pairs, 0u, p, k, j);
}

// the sorted pairs to the slot order's ids and keys, and the state gathered into slot order (as k_fix_gather)
template <bool B>
__global__ void __launch_bounds__(HASH_BLOCK) k_bitonic_gather(int n, BatchView bt, const uint64_t* __restrict__ pairs, uint32_t* __restrict__ sid,
                                                               uint32_t* __restrict__ skey, const float* __restrict__ ax,
                                                               const float* __restrict__ ay, const float* __restrict__ rad,
                                                               float* __restrict__ sx, float* __restrict__ sy, float* __restrict__ sr) {
    const int f = blockIdx.x * HASH_BLOCK + threadIdx.x;
    if (f >= n) return;
    const uint64_t v = pairs[f];
    const uint32_t id = uint32_t(v), k = uint32_t(v >> 32);
    sid[f] = id;
    skey[f] = k;
    if constexpr (B) bt.slotScen[f] = bt.agentScen[id];
    if (k != INACTIVE_CELL) {
        sx[f] = ax[id];
        sy[f] = ay[id];
        sr[f] = rad[id];
    }
}

// the network over n (a power of two >= 2) pairs: built from the keys (key != nullptr) or sorted in place
template <bool FULL>
static void launchBitonicNet(hipStream_t st, int n, const uint32_t* key, uint64_t* pairs) {
    const int tile = n < BITONIC_TILE ? n : BITONIC_TILE;
    k_bitonic_tile<FULL><<<n / tile, tile / 2, 0, st>>>(tile, key, pairs);
    for (uint32_t k = 2u * uint32_t(tile); k <= uint32_t(n); k <<= 1) {
        for (uint32_t j = k >> 1; j >= uint32_t(tile); j >>= 1)
            k_bitonic_pass<FULL><<<(n / 2 + 255) / 256, 256, 0, st>>>(uint32_t(n) / 2u, k, j, pairs);
        k_bitonic_merge<FULL><<<n / BITONIC_TILE, BITONIC_TILE / 2, 0, st>>>(k, pairs);
    }
}
static void launchBitonic(hipStream_t st, int n, const uint32_t* key, uint64_t* pairs) { launchBitonicNet<false>(st, n, key, pairs); }

// ---------------- the stable (key, agent) sort of caller keys on the network ----------------
__global__ void __launch_bounds__(256) k_fill_u32(uint32_t n, uint32_t* __restrict__ p, uint32_t v) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = v;
}
// (key << 32 | index) for the n keys, then all-ones padding to the network's power of two (it sorts last)
__global__ void __launch_bounds__(256) k_pairs_init(uint32_t n, uint32_t npow, const uint32_t* __restrict__ key, uint64_t* __restrict__ pairs) {
    const uint32_t e = blockIdx.x * 256u + threadIdx.x;
    if (e < npow) pairs[e] = e < n ? (uint64_t(key[e]) << 32) | uint64_t(e) : ~0ull;
}
__global__ void __launch_bounds__(256) k_pairs_out(uint32_t n, const uint64_t* __restrict__ pairs, uint32_t* __restrict__ sortedKeys,
                                                   uint32_t* __restrict__ ids) {
    const uint32_t f = blockIdx.x * 256u + threadIdx.x;
    if (f >= n) return;
    const uint64_t v = pairs[f];
    if (sortedKeys) sortedKeys[f] = uint32_t(v >> 32);
    if (ids) ids[f] = uint32_t(v);
}
// the reference's cell table (reorder.comp:20-28: start 0 for an empty cell) from the bucket starts
__global__ void __launch_bounds__(256) k_cells_out(uint32_t C, const uint32_t* __restrict__ bstart, uint32_t* __restrict__ start,
                                                   uint32_t* __restrict__ count) {
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= C) return;
    const uint32_t a = bstart[c], cnt = bstart[c + 1u] - a;
    if (start) start[c] = cnt ? a : 0u;
    if (count) count[c] = cnt;
}

uint32_t keySortPow2(uint32_t n) {
    uint32_t p = 2u;
    while (p < n) p <<= 1;
    return p;
}
uint32_t keySortBuckets(uint32_t n, uint32_t C) { return C + uint32_t((int(n) + HASH_BLOCK - 1) / HASH_BLOCK); }
uint32_t keySortTiles(uint32_t n, uint32_t C) { return uint32_t((int(keySortBuckets(n, C)) + SCAN_TILE - 1) / SCAN_TILE); }

void launchStableKeySort(hipStream_t st, uint32_t n, uint32_t C, const uint32_t* key, const KeySortScratch& s, uint32_t* sortedKeys,
                         uint32_t* ids, uint32_t* cellStart, uint32_t* cellCount) {
    const uint32_t nb = keySortBuckets(n, C), npow = keySortPow2(n);
    const int tiles = int(keySortTiles(n, C));
    k_fill_u32<<<(nb + 255) / 256, 256, 0, st>>>(nb, s.bcnt, 0u);
    k_key_count<<<(int(n) + HASH_BLOCK - 1) / HASH_BLOCK, HASH_BLOCK, 0, st>>>(int(n), key, s.loc, s.bcnt, C);
    k_scan_reduce<<<tiles, SCAN_BLOCK, 0, st>>>(s.bcnt, nb, s.tileSum);
    k_scan_apply<<<tiles, SCAN_BLOCK, 0, st>>>(s.bcnt, nb, C, s.tileSum, s.bstart, s.ctrl);
    k_pairs_init<<<(npow + 255) / 256, 256, 0, st>>>(n, npow, key, s.pairs);
    launchBitonicNet<true>(st, int(npow), nullptr, s.pairs);
    k_pairs_out<<<(n + 255) / 256, 256, 0, st>>>(n, s.pairs, sortedKeys, ids);
    if (cellStart || cellCount) k_cells_out<<<(C + 255) / 256, 256, 0, st>>>(C, s.bstart, cellStart, cellCount);
}

void launchSortedGather(hipStream_t st, const StepLaunch& L, AgentHalf rd, const uint64_t* pairs, StepBuffers& b, const float* rad) {
    if (L.batched)
        k_bitonic_gather<true><<<(L.n + HASH_BLOCK - 1) / HASH_BLOCK, HASH_BLOCK, 0, st>>>(L.n, L.bt, pairs, b.sid, b.skey, rd.x, rd.y, rad, b.sx, b.sy, b.sr);
    else
        k_bitonic_gather<false><<<(L.n + HASH_BLOCK - 1) / HASH_BLOCK, HASH_BLOCK, 0, st>>>(L.n, L.bt, pairs, b.sid, b.skey, rd.x, rd.y, rad, b.sx, b.sy, b.sr);
}

// ---------------- exit choice (PHYSICS_PASS 0) ----------------
__device__ inline bool evalStep(uint32_t step) { return step > 0u && step % uint32_t(CHOICE_STEPS) == 0u; }

// per-scenario choice state
struct ChoiceRef {
    size_t aheadBase;   // this scenario's ahead block (2 x stride x 1024)
    uint32_t queueBase, counterBase, histBase, thresholdBase;
};
__device__ inline ChoiceRef choiceRef(uint32_t sc, int stride) {
    return ChoiceRef{ size_t(sc) * 2u * size_t(stride) * CHOICE_AHEAD_BINS, sc * uint32_t(QUEUE_STRIDE), sc * uint32_t(COUNTER_STRIDE),
                      sc * uint32_t(stride) * uint32_t(CHOICE_GAIN_BINS), sc * uint32_t(THRESHOLD_STRIDE) };
}

// prefix stage, blockIdx.x the scenario, one wave per exit (CHOICE_EXIT_WAVES exits per block, blockIdx.y their
// group): the read parity's queue-ahead histograms to exclusive prefix sums, each exit's histogram total added to the read
// parity's queue count (the apply stage counts there only the agents it does not bin), the write parity cleared, the
// write parity's queue counts cleared (by the scenario's first group)
constexpr int CHOICE_EXIT_WAVES = 4;
template <bool B>
__global__ void __launch_bounds__(64 * CHOICE_EXIT_WAVES) k_choice_prefix(SceneView s0, BatchView bt, ChoiceBuffers c,
                                                                          const uint32_t* __restrict__ ctrl) {
    const uint32_t sc = blockIdx.x;
    const int E = sceneOf<B>(s0, bt, sc).E;
    const uint32_t step = ctrl[CTRL_STEP];
    const uint32_t par = step & 1u;
    const ChoiceRef r = choiceRef(sc, c.exitStride);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int e = int(blockIdx.y) * CHOICE_EXIT_WAVES + wave;
    if (e < E) {
        constexpr int PER = CHOICE_AHEAD_BINS / 64;   // a lane's run of bins, PER / 4 16-byte words
        uint32_t* aRead = c.ahead + r.aheadBase + size_t(par) * size_t(c.exitStride) * CHOICE_AHEAD_BINS;
        uint32_t* aWrite = c.ahead + r.aheadBase + size_t(1u - par) * size_t(c.exitStride) * CHOICE_AHEAD_BINS;
        uint4* a = reinterpret_cast<uint4*>(aRead + size_t(e) * CHOICE_AHEAD_BINS + lane * PER);
        uint4 v[PER / 4];
        uint32_t local = 0;
#pragma unroll
        for (int k = 0; k < PER / 4; ++k) { v[k] = a[k]; local += v[k].x + v[k].y + v[k].z + v[k].w; }
        uint32_t incl = local;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off, 64);
            if (lane >= off) incl += o;
        }
        uint32_t run = incl - local;
#pragma unroll
        for (int k = 0; k < PER / 4; ++k) {
            uint4 w;
            w.x = run; run += v[k].x;
            w.y = run; run += v[k].y;
            w.z = run; run += v[k].z;
            w.w = run; run += v[k].w;
            a[k] = w;
        }
        const uint32_t total = __shfl(incl, 63, 64);
        if (lane == 0) c.queue[r.queueBase + par * uint32_t(CHOICE_QUEUE_STRIDE) + uint32_t(e)] += total;
        uint4* wrp = reinterpret_cast<uint4*>(aWrite + size_t(e) * CHOICE_AHEAD_BINS + lane * PER);
#pragma unroll
        for (int k = 0; k < PER / 4; ++k) wrp[k] = make_uint4(0u, 0u, 0u, 0u);
    }
    const uint32_t qWrite = uint32_t(CHOICE_QUEUE_STRIDE) - par * uint32_t(CHOICE_QUEUE_STRIDE);
    if (blockIdx.y == 0 && threadIdx.x < CHOICE_QUEUE_STRIDE) c.queue[r.queueBase + qWrite + threadIdx.x] = 0u;
}

// evaluate stage (evaluation steps only), per agent. LDS: a seated venue's own step with its cell table in LDS
template <bool B, bool LDS = false>
__global__ void __launch_bounds__(HASH_BLOCK) k_choice_eval(int n, SceneView s0, BatchView bt, const float* __restrict__ ax,
                                                            const float* __restrict__ ay, const float* __restrict__ aw, ChoiceBuffers c,
                                                            const uint32_t* __restrict__ ctrl) {
    const uint32_t step = ctrl[CTRL_STEP];
    if (!evalStep(step)) return;
    static_assert(!(B && LDS), "the venue's own step runs unbatched");
    if constexpr (LDS) s0.venueCells = venueCellsInLds(s0);
    const int i = blockIdx.x * HASH_BLOCK + threadIdx.x;
    bool evaluated = false, eligible = false;
    uint32_t hkey = 0, ckey = 0;
    if (i < n) {
        c.pending[i] = 0u;
        const ScnRef sr = scnOfAgent<B>(bt, uint32_t(i));
        const ChoiceRef r = choiceRef(sr.sc, c.exitStride);
        ckey = r.counterBase;
        if (aw[i] != 0.0f) {
            const SceneView& s = sceneOf<B>(s0, bt, sr.sc);
            const uint32_t* aRead = c.ahead + r.aheadBase + size_t(step & 1u) * size_t(c.exitStride) * CHOICE_AHEAD_BINS;
            const V2 p{ ax[i], ay[i] };
            const int cell = choiceCell(s, p);
            const uint32_t* sl = navSlots(s, cell);
            const uint16_t* dq = navDists(s, cell);
            const uint32_t current = c.assign[i];
            const int curK = slotExitOf(sl[0]) == CHOICE_EMPTY ? -1 : slotOfExit(s, cell, current);
            if (curK >= 0) {
                evaluated = true;
                float eBest;
                const int best = bestCandidate(s, cell, aRead, eBest);
                const uint32_t qc = dq[curK];
                const float eCur = estimatedTime(s, slotDistOf(s, qc), current, aRead[size_t(current) * CHOICE_AHEAD_BINS + aheadBin(qc)]);
                const float gain = eCur - eBest;
                const float m = CHOICE_REL_MARGIN * eCur;
                if (best != curK && gain > (m < CHOICE_ABS_MARGIN_S ? CHOICE_ABS_MARGIN_S : m)) {
                    const uint32_t g = uint32_t(gain);
                    const uint32_t bin = g < uint32_t(CHOICE_GAIN_BINS - 1) ? g : uint32_t(CHOICE_GAIN_BINS - 1);
                    c.pending[i] = 0x80000000u | (slotExitOf(sl[best]) << 16) | bin;
                    eligible = true;
                    hkey = r.histBase + current * uint32_t(CHOICE_GAIN_BINS) + bin;
                }
            }
        }
    }
    waveCountAdd(c.hist, hkey, eligible);
    waveCountAdd(c.counters, ckey + uint32_t(CNT_EVALUATIONS), evaluated);
    waveCountAdd(c.counters, ckey + uint32_t(CNT_ELIGIBLE), eligible);
}

// thresholds stage (evaluation steps only), blockIdx.x the scenario, one wave per exit (groups as the prefix stage).
// Walking an exit's gain histogram down from its top bin, the bins are taken while their running total stays within
// the cap; the threshold is the lowest bin taken, or the top non-empty bin when that alone exceeds the cap (none: the
// bin count). With S(b) the count in bins b and above, that is the lowest non-empty bin with S(b) <= cap, else the
// highest non-empty bin: a suffix sum across the wave and two reductions. The histogram is then cleared.
template <bool B>
__global__ void __launch_bounds__(64 * CHOICE_EXIT_WAVES) k_choice_thresh(SceneView s0, BatchView bt, ChoiceBuffers c,
                                                                          const uint32_t* __restrict__ ctrl) {
    const uint32_t step = ctrl[CTRL_STEP];
    if (!evalStep(step)) return;
    const uint32_t sc = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int e = int(blockIdx.y) * CHOICE_EXIT_WAVES + wave;
    if (e >= sceneOf<B>(s0, bt, sc).E) return;
    static_assert(CHOICE_GAIN_BINS == 4 * 64, "a lane holds four bins");
    const ChoiceRef r = choiceRef(sc, c.exitStride);
    const uint32_t qRead = (step & 1u) * uint32_t(CHOICE_QUEUE_STRIDE);
    uint4* h = reinterpret_cast<uint4*>(c.hist + r.histBase + uint32_t(e) * CHOICE_GAIN_BINS) + lane;   // bins 4 lane .. 4 lane + 3
    const uint32_t queued = c.queue[r.queueBase + qRead + uint32_t(e)];
    const uint32_t share = uint32_t(float(queued) * CHOICE_SWITCH_SHARE);
    const uint32_t cap = share > CHOICE_MIN_SWITCHES ? share : CHOICE_MIN_SWITCHES;
    const uint4 v = *h;
    const uint32_t cnt[4] = { v.x, v.y, v.z, v.w };
    // the lanes above this one, summed (a suffix scan down the wave)
    uint32_t above = cnt[0] + cnt[1] + cnt[2] + cnt[3];
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_down(above, off, 64);
        if (lane + off < 64) above += o;
    }
    above -= cnt[0] + cnt[1] + cnt[2] + cnt[3];
    int lowOk = CHOICE_GAIN_BINS, highNonEmpty = -1;
    uint32_t S = above;
    for (int j = 3; j >= 0; --j) {
        S += cnt[j];
        const int b = 4 * lane + j;
        if (cnt[j] != 0u) {
            if (S <= cap) lowOk = b;
            if (highNonEmpty < 0) highNonEmpty = b;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const int lo = __shfl_xor(lowOk, off, 64), hi = __shfl_xor(highNonEmpty, off, 64);
        lowOk = lo < lowOk ? lo : lowOk;
        highNonEmpty = hi > highNonEmpty ? hi : highNonEmpty;
    }
    const uint32_t thr = lowOk < CHOICE_GAIN_BINS ? uint32_t(lowOk) : (highNonEmpty >= 0 ? uint32_t(highNonEmpty) : uint32_t(CHOICE_GAIN_BINS));
    if (lane == 0) c.threshold[r.thresholdBase + uint32_t(e)] = thr;
    *h = make_uint4(0u, 0u, 0u, 0u);
}

// apply stage (every step), per slot: the agent's exit for this step, its counts for the next, and its e_goal. LDS: a
// seated venue's own step with its cell table in LDS
template <bool B, bool LDS = false>
__global__ void __launch_bounds__(HASH_BLOCK) k_choice_apply(int n, SceneView s0, BatchView bt, const uint32_t* __restrict__ skey,
                                                             const uint32_t* __restrict__ sid, const float* __restrict__ sx,
                                                             const float* __restrict__ sy, const float* __restrict__ srad, ChoiceBuffers c,
                                                             float* __restrict__ egx, float* __restrict__ egy, const uint32_t* __restrict__ ctrl,
                                                             const float* __restrict__ az) {
    static_assert(!(B && LDS), "the venue's own step runs unbatched");
    if constexpr (LDS) s0.venueCells = venueCellsInLds(s0);
    const int slot = xcdBlock() * HASH_BLOCK + threadIdx.x;
    const bool live = slot < n && skey[slot] != INACTIVE_CELL;
    const uint32_t step = ctrl[CTRL_STEP];
    const bool ev = evalStep(step);
    const uint32_t par = step & 1u;
    const uint32_t qWrite = uint32_t(CHOICE_QUEUE_STRIDE) - par * uint32_t(CHOICE_QUEUE_STRIDE);
    bool forced = false, switched = false, counted = false, binned = false;
    uint32_t qkey = 0, akey = 0, ckey = 0;
    if (live) {
        const ScnRef sr = scnOfSlot<B>(bt, uint32_t(slot));
        const SceneView& s = sceneOf<B>(s0, bt, sr.sc);
        const ChoiceRef r = choiceRef(sr.sc, c.exitStride);
        ckey = r.counterBase;
        const uint32_t* aRead = c.ahead + r.aheadBase + size_t(par) * size_t(c.exitStride) * CHOICE_AHEAD_BINS;
        const uint32_t id = sid[slot];
        const V2 p{ sx[slot], sy[slot] };
        const int cell = choiceCell(s, p);
        const uint32_t* sl = navSlots(s, cell);
        const uint32_t current = c.assign[id];
        uint32_t ex = current;
        const bool hasPath = slotExitOf(sl[0]) != CHOICE_EMPTY;
        const int curK = hasPath ? slotOfExit(s, cell, current) : -1;
        if (hasPath && curK < 0) {
            float eBest;
            ex = slotExitOf(sl[bestCandidate(s, cell, aRead, eBest)]);
            forced = current != CHOICE_NONE;
        } else if (ev && curK >= 0) {
            const uint32_t pd = c.pending[id];
            if ((pd & 0x80000000u) && (pd & 0xFFFFu) >= c.threshold[r.thresholdBase + current]) {
                ex = (pd >> 16) & 63u;
                switched = true;
            }
        }
        c.assign[id] = ex;
        counted = ex != CHOICE_NONE;
        qkey = r.queueBase + qWrite + (counted ? ex : 0u);
        const int k = slotOfExit(s, cell, ex);
        if (k >= 0) {
            binned = true;
            akey = uint32_t(r.aheadBase + size_t(1u - par) * size_t(c.exitStride) * CHOICE_AHEAD_BINS) + ex * uint32_t(CHOICE_AHEAD_BINS) +
                   aheadBin(navDists(s, cell)[k]);
        }
        const V2 dir = choiceDirection(s, cell, p, ex);
        // the wall term: the plan's wall boxes, or on a venue's own step its exact wall at the centre's height
        const V2 e = s.venue3d ? seatedstep::venueGoalWallRule(seatedstep::Venue3{ *s.venue, s.venueCells }, seatedstep::V3{ p.x, p.y, az[id] }, dir)
                               : goalWallRule(s, p, FEELER_REACH_RADII * srad[slot], dir);
        egx[slot] = e.x;
        egy[slot] = e.y;
    }
    // an exit's queue count is its binned agents (the next prefix stage sums its histogram) plus these
    waveCountAdd(c.queue, qkey, counted && !binned);
    waveCountAddFew(c.ahead, akey, binned, CHOICE_AHEAD_ROUNDS);
    waveCountAdd(c.counters, ckey + uint32_t(CNT_FORCED), forced);
    waveCountAdd(c.counters, ckey + uint32_t(CNT_SWITCHES), switched);
}

// e_goal from a forced direction table (TEST_DRIVE_TABLE), or zero without a table and without exits
__global__ void __launch_bounds__(HASH_BLOCK) k_goal_drive(int n, const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sid,
                                                           const float* __restrict__ dx, const float* __restrict__ dy,
                                                           float* __restrict__ egx, float* __restrict__ egy) {
    const int slot = blockIdx.x * HASH_BLOCK + threadIdx.x;
    if (slot >= n || skey[slot] == INACTIVE_CELL) return;
    const uint32_t id = sid[slot];
    egx[slot] = dx ? dx[id] : 0.0f;
    egy[slot] = dy ? dy[id] : 0.0f;
}

// ---------------- PHYSICS_PASS 1: predict ----------------
// Q > 1 (not REC): the slot's Q adjacent lanes share the perception ring. Lane l takes the row's slots
// k0 + u Q + l, so (u, l) order is the row's slot order: every lane adds the separation terms in that order through
// shuffles within the slot's lanes (the single lane's sum bit for bit); the headway gap is a minimum over the
// agents ahead, the same in any order, combined across the lanes. Every lane then carries the same state through
// the kinematics; the wall contacts split their feelers and contacts over the lanes (resolveStaticQ); lane 0 writes.
template <bool B, bool REC = false, int Q = 1>
__device__ inline void predictSlot(int slot, const SceneView& s0, const BatchView& bt, const ModelParams& mp, const StepBuffers& b,
                                   const ProjBuffers& out, const TapeStep& tp = TapeStep{}, int lane = 0) {
    static_assert(Q == 1 || !REC, "the taped predict runs one lane per slot");
    const uint32_t key = b.skey[slot];
    swarm_predict_rec* rec = nullptr;
    if constexpr (REC) {
        rec = tp.pred + slot;
        rec->flags = 0u;
        rec->n_events = 0u;
        rec->plane_kind = 0u;
        tp.capBits[slot] = 0ull;
        tp.planeBits[slot] = 0ull;
    }
    if (key == INACTIVE_CELL) return;
    const ScnRef sr = scnOfSlot<B>(bt, uint32_t(slot));
    const SceneView& s = sceneOf<B>(s0, bt, sr.sc);
    const uint32_t self = b.sid[slot] - sr.agentBase;   // the scene's own agent ids (pair hash, tie-break)
    const V2 p{ b.sx[slot], b.sy[slot] };
    const float ri = b.sr[slot];
    if constexpr (REC) rec->flags = SWARM_REC_ACTIVE;
    if (parkTest(s, p)) {
        out.p[slot] = make_float2(PARK_COORD, PARK_COORD);
        out.acc[slot] = 0.0f;
        b.pw[slot] = 0.0f; b.pvx[slot] = 0.0f; b.pvy[slot] = 0.0f;
        if constexpr (REC) rec->flags |= SWARM_REC_PARKED;
        return;
    }
    const uint32_t agent = b.sid[slot];
    const V2 v{ b.avx[agent], b.avy[agent] };
    const V2 eGoal{ b.egx[slot], b.egy[slot] };
    const uint32_t cid = key - sr.cellBase;
    const int cx = int(cid % uint32_t(s.hashNX)), cy = int(cid / uint32_t(s.hashNX));
    V2 sep{ 0.0f, 0.0f };
    float sAhead = 1e9f;   // the smallest gap (centre distance - contact distance) to an agent ahead
    uint32_t leader = NBR_OVERFLOW;
    const int R = s.percRing;
    const int x0 = cx - R < 0 ? 0 : cx - R;
    const int x1 = cx + R > s.hashNX - 1 ? s.hashNX - 1 : cx + R;
    for (int yy = -R; yy <= R; yy++) {
        const int ny = cy + yy;
        if (ny < 0 || ny >= s.hashNY) continue;
        const uint32_t row = sr.cellBase + uint32_t(ny * s.hashNX);
        const uint32_t st = b.bstart[row + uint32_t(x0)], en = b.bstart[row + uint32_t(x1) + 1u];
        // the row's slots in chunks: a chunk's loads are in flight together; the sums keep the ring order
        for (uint32_t k0 = st; k0 < en; k0 += uint32_t(Q * PRED_CHUNK)) {
            float qx[PRED_CHUNK], qy[PRED_CHUNK], qr[PRED_CHUNK];
#pragma unroll
            for (int u = 0; u < PRED_CHUNK; ++u) {
                const uint32_t kk = k0 + uint32_t(u * Q + lane);
                const uint32_t k = kk < en ? kk : en - 1u;
                qx[u] = b.sx[k]; qy[u] = b.sy[k]; qr[u] = b.sr[k];
            }
#pragma unroll
            for (int u = 0; u < PRED_CHUNK; ++u) {
                if constexpr (Q > 1) {
                    if (k0 + uint32_t(u * Q) >= en) break;   // uniform over the slot's lanes
                    const uint32_t k = k0 + uint32_t(u * Q + lane);
                    const V2 offset = p - V2{ qx[u], qy[u] };
                    const float dxy2 = dot2(offset, offset);
                    const bool hit = k < en && k != uint32_t(slot) && dxy2 < R_PERC2;
                    V2 term{ 0.0f, 0.0f };
                    if (hit) {
                        const uint32_t other = b.sid[k] - sr.agentBase;
                        const float D = ri + qr[u];
                        term = separationForce(offset, self, other, D + SEP_MARGIN);
                        const float ahead = -dot2(offset, eGoal);
                        const float lateral = fabsf(offset.x * eGoal.y - offset.y * eGoal.x);
                        const float gap = sqrtf(dxy2) - D;
                        if (ahead > 0.0f && lateral < D && gap < sAhead) {
                            const V2 eOther{ b.egx[k], b.egy[k] };
                            const float aheadOther = dot2(offset, eOther);
                            const bool mutual = aheadOther > 0.0f && fabsf(offset.x * eOther.y - offset.y * eOther.x) < D;
                            if (!mutual || ahead > aheadOther || (ahead == aheadOther && self > other)) sAhead = gap;
                        }
                    }
#pragma unroll
                    for (int l = 0; l < Q; ++l) {
                        const int h = __shfl(int(hit), l, Q);
                        const float tx = __shfl(term.x, l, Q);
                        const float ty = __shfl(term.y, l, Q);
                        if (h) sep = sep + V2{ tx, ty };
                    }
                    continue;
                }
                const uint32_t k = k0 + uint32_t(u);
                if (k >= en) break;
                if (k == uint32_t(slot)) continue;
                const V2 offset = p - V2{ qx[u], qy[u] };
                const float dxy2 = dot2(offset, offset);
                if (dxy2 < R_PERC2) {
                    const uint32_t other = b.sid[k] - sr.agentBase;
                    const float D = ri + qr[u];
                    sep = sep + separationForce(offset, self, other, D + SEP_MARGIN);
                    const float ahead = -dot2(offset, eGoal);
                    const float lateral = fabsf(offset.x * eGoal.y - offset.y * eGoal.x);
                    const float dxy = sqrtf(dxy2);
                    const float gap = dxy - D;
                    if (ahead > 0.0f && lateral < D && gap < sAhead) {
                        const V2 eOther{ b.egx[k], b.egy[k] };
                        const float aheadOther = dot2(offset, eOther);
                        const bool mutual = aheadOther > 0.0f && fabsf(offset.x * eOther.y - offset.y * eOther.x) < D;
                        if (!mutual || ahead > aheadOther || (ahead == aheadOther && self > other)) {
                            sAhead = gap;
                            if constexpr (REC) leader = k;
                        }
                    }
                }
            }
        }
    }
    if constexpr (Q > 1) {
#pragma unroll
        for (int d = 1; d < Q; d <<= 1) {
            const float o = __shfl_xor(sAhead, d, Q);
            sAhead = o < sAhead ? o : sAhead;
        }
    }
    const float stair = stairFactor(s, p);
    // a seated venue's plan: the free speed and the headway time gap by the floor's slope along e_goal on the gangway
    // inclines (physics.comp VENUE_INCLINE_SPEED / VENUE_STAIR_HEADWAY; cpu_crowd_sim.cpp:727-733)
    const float slope = s.venue ? seated::floorSlope(*s.venue, seated::F2{ p.x, p.y }, seated::F2{ eGoal.x, eGoal.y }) : 0.0f;
    const float vFree = s.venue ? inclineFreeSpeed(b.v0[agent], slope) : stair * b.v0[agent];
    float vc = s.venue && slope != 0.0f ? sAhead / inclineHeadwayTime(slope) : sAhead * HEADWAY_RATE;
    const bool vcZero = vc < 0.0f, vcFree = !vcZero && vFree < vc;
    vc = vcZero ? 0.0f : (vcFree ? vFree : vc);
    const float vDes = vc * len2(eGoal);
    const V2 eDir = safeNormalize(eGoal + sep);

    V2 aObs{ 0.0f, 0.0f };
    bool obst = false;
    const float dTorso = mapSdf(s, p);
    if (dTorso < OBSTACLE_RANGE) {
        const V2 norm = safeNormalize(calcGradient(s, p));
        if (len2(norm) > OBSTACLE_MIN_NORMAL) {
            aObs = safeNormalize(norm) * (OBSTACLE_RANGE - dTorso) * W_OBS * mp.obstacleGain;
            obst = true;
        }
    }
    const V2 vEq = eDir * vDes + aObs * mp.tau;
    V2 nv = vEq + (v - vEq) * mp.relaxStep;
    const float along = dot2(nv, eGoal);
    const bool alongCap = along > vDes;
    if (alongCap) nv = nv - eGoal * (along - vDes);
    const float vmax = b.v0[agent];
    const bool clamp = len2(nv) > vmax;
    if (clamp) nv = safeNormalize(nv) * vmax;
    V2 np = p + nv * DT;

    float wall[3];
    bool grounded;
    if constexpr (REC) {
        uint32_t f = SWARM_REC_ACTIVE;
        if (leader != NBR_OVERFLOW) { f |= SWARM_REC_LEADER; rec->leader = leader; }
        f |= vcZero ? SWARM_REC_VC_ZERO : (vcFree ? SWARM_REC_VC_FREE : SWARM_REC_VC_RATE);
        if (!nonzero(eDir)) f |= SWARM_REC_EDIR_ZERO;
        if (obst) f |= SWARM_REC_OBST;
        if (alongCap) f |= SWARM_REC_ALONG;
        if (clamp) f |= SWARM_REC_CLAMP;
        // the goal anchor behind e_goal (choiceDirection), or a constant e_goal
        bool efixed = tp.efixed != 0;
        if (!efixed) {
            const int cell = choiceCell(s, p);
            const uint32_t ex = tp.assign[agent];
            const int kk = slotOfExit(s, cell, ex);
            if (kk < 0) efixed = true;
            else {
                const V2 a = anchorCentre(s, navSlots(s, cell)[kk] & CHOICE_ANCHOR_MASK);
                rec->anchor_x = a.x; rec->anchor_y = a.y;
                if (!(len2(a - p) > GOAL_DEADZONE)) efixed = true;
                else {   // the wall rule: e = unit(e0 - n (e0 . n)) with n frozen; nothing left: a constant 0
                    bool wrule = false;
                    V2 wn{ 0.0f, 0.0f };
                    const V2 e1 = goalWallRule(s, p, FEELER_REACH_RADII * ri, choiceDirection(s, cell, p, ex), &wrule, &wn);
                    if (wrule && !nonzero(e1)) efixed = true;
                    else if (wrule) { f |= SWARM_REC_GOAL_WALL; rec->goal_wall_x = wn.x; rec->goal_wall_y = wn.y; }
                }
            }
        }
        if (efixed) f |= SWARM_REC_EFIXED;
        rec->flags = f;
        rec->egoal_x = eGoal.x; rec->egoal_y = eGoal.y;
        rec->stair = stair;
        rec->sep_x = sep.x; rec->sep_y = sep.y;
        WallTape wt{ &rec->flags, &rec->n_events, rec->ev, &rec->sweep_lo, tp.err };
        PlaneRec pl{};
        grounded = resolveStatic(s, p, np, nv, ri, wall, wt, &pl);
        if (wall[0] == 0.0f && wall[1] == 0.0f && !wallFree(s, np)) wallPlaneProbe(s, np, ri, wall, wt, &pl);
        rec->plane_kind = uint32_t(pl.kind == 3 ? 2 : pl.kind);
        rec->plane_nx = pl.n.x; rec->plane_ny = pl.n.y;
        rec->pc_ux = pl.c.u.x; rec->pc_uy = pl.c.u.y; rec->pc_gx = pl.c.g.x; rec->pc_gy = pl.c.g.y;
        rec->pc_x0 = pl.c.x0.x; rec->pc_y0 = pl.c.x0.y; rec->pc_t0 = pl.c.t0;
    } else if constexpr (Q > 1) {   // every lane holds the same state; the wall contacts share the lanes' work
        grounded = resolveStaticQ<Q>(s, p, np, nv, ri, wall, lane);
        if (wall[0] == 0.0f && wall[1] == 0.0f && !wallFree(s, np)) wallPlaneProbeQ<Q>(s, np, ri, wall, lane);
        if (lane != 0) return;
    } else {
        grounded = resolveStatic(s, p, np, nv, ri, wall);
        if (wall[0] == 0.0f && wall[1] == 0.0f && !wallFree(s, np)) wallPlaneProbe(s, np, ri, wall);
    }
    out.p[slot] = make_float2(np.x, np.y);
    out.acc[slot] = 0.0f;
    b.pw[slot] = 1.0f; b.pvx[slot] = nv.x; b.pvy[slot] = nv.y;
    const V2 wn{ wall[0], wall[1] };
    b.planes[slot] = make_float4(wn.x, wn.y, dot2(np, wn) - wall[2], (nonzero(wn) ? 1.0f : 0.0f) + (grounded ? 2.0f : 0.0f));
}

// the step's candidate-budget words (per scenario): whole at the start, or broken from the start without lists
__device__ inline void resetListBudget(const StepBuffers& b) {
    if (b.listMoves < 0.0f) return;
    for (int i = int(blockIdx.x * blockDim.x + threadIdx.x); i < b.nScen; i += int(gridDim.x * blockDim.x))
        b.listBroken[i] = b.listMoves == 0.0f ? 1u : 0u;
}

// the physics kernels loop over slots with the grid's stride: one slot per thread, or a persistent grid
template <bool B, bool REC = false>
__global__ void __launch_bounds__(PHYS_MAX_BLOCK) k_predict(int n, SceneView s0, BatchView bt, ModelParams mp, StepBuffers b, ProjBuffers out,
                                                            TapeStep tp = TapeStep{}) {
    resetListBudget(b);
    // the taped forward clears every slot's record; otherwise the inactive slots (the last ones) are left alone
    const int end = REC ? n : int(b.ctrl[CTRL_NACTIVE]);
    for (int slot = xcdBlock() * blockDim.x + threadIdx.x; slot < end; slot += gridDim.x * blockDim.x)
        predictSlot<B, REC>(slot, s0, bt, mp, b, out, tp);
}

// predict with Q lanes per slot (the perception ring shared, then lane 0 alone)
template <bool B, int Q>
__global__ void __launch_bounds__(PHYS_MAX_BLOCK) k_predict_lanes(int n, SceneView s0, BatchView bt, ModelParams mp, StepBuffers b,
                                                                  ProjBuffers out) {
    static_assert(Q == 2 || Q == 4 || Q == 8, "lanes per slot");
    resetListBudget(b);
    const int total = int(b.ctrl[CTRL_NACTIVE]) * Q;   // the inactive slots (the last ones) are left alone
    for (int t = xcdBlock() * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x)
        predictSlot<B, false, Q>(t / Q, s0, bt, mp, b, out, TapeStep{}, t % Q);
}

// ---------------- PHYSICS_PASS 2: the Jacobi projection ----------------
// The end of projectSlot (cpu_crowd_sim.cpp:805-821): the mean correction, its cap, the wall plane, the accumulator
// capped / planed: whether the cap and the wall plane applied (the adjoint's tape)
// pl: the slot's wall plane (planes[slot])
__device__ inline void finishProjection(const float4 pl, float bx, float by, float acc, V2 corr, int nContacts,
                                        float ri, float& ox, float& oy, float& oacc, bool* capped = nullptr, bool* planed = nullptr) {
    corr = corr * (PROJ_RELAX / float(nContacts > 1 ? nContacts : 1));
    const float maxShift = PROJ_MAX_SHIFT_RADII * ri;
    const float len = len2(corr);
    const bool cap = len > maxShift;
    if (cap) corr = corr * (maxShift / len);
    V2 q = V2{ bx, by } + corr;
    bool pld = false;
    if (int(pl.w) & 1) {
        const float sd = pl.z - dot2(q, V2{ pl.x, pl.y });
        pld = sd > 0.0f;
        if (pld) q = q + V2{ pl.x, pl.y } * sd;
    }
    const V2 d = q - V2{ bx, by };
    ox = q.x; oy = q.y;
    oacc = acc + len2(d);
    if (capped) { *capped = cap; *planed = pld; }
}

// contactCorrection (cpu_crowd_sim.cpp:141-148) for a pair already known to overlap (dxy2 < D^2); the other
// agent's id only for a stacked pair
__device__ inline V2 overlapCorrection(V2 offset, float dxy2, float D, uint32_t self, const uint32_t* __restrict__ sid, uint32_t k,
                                       uint32_t agentBase) {
    const float dxy = sqrtf(dxy2);
    const float shift = 0.5f * (D - dxy);
    return dxy2 <= SEP_STACKED_XY2 ? shift * pairDirection(self, sid[k] - agentBase) : shift * (offset / dxy);
}
// The same with the slot's own id read only for a stacked pair (the list iterations do not load it otherwise)
__device__ inline V2 overlapCorrectionAt(V2 offset, float dxy2, float D, const uint32_t* __restrict__ sid, uint32_t slot, uint32_t k,
                                         uint32_t agentBase) {
    const float dxy = sqrtf(dxy2);
    const float shift = 0.5f * (D - dxy);
    if (dxy2 <= SEP_STACKED_XY2) return shift * pairDirection(sid[slot] - agentBase, sid[k] - agentBase);
    return shift * (offset / dxy);
}

// One iteration for one slot by the ring scan (cpu_crowd_sim.cpp:781-822); with 'collect', also the slot's contact
// candidates: every projection iteration moves an agent by at most PROJ_MAX_SHIFT_RADII r (the capped correction; the
// wall plane is a projection onto a half-space holding the predicted position, which moves it no farther), so over
// the P iterations of a step a pair can touch only if its predicted distance is below D + P (shift_i + shift_j).
// The candidates are the ring's active slots within that bound (with 0.1% of float slack), in ring order, so the
// later iterations' sums over them are the ring scan's sums bit for bit. With a displacement budget (b.listMoves > 0)
// the bound is D + listMoves (shift_i + shift_j): it holds while no path is longer than listMoves shifts, which the
// iterations check (a scenario past it scans the ring for the rest of its step).
__device__ inline void projectRing(const SceneView& s, const StepBuffers& b, const ProjBuffers& in, int n, int slot, uint32_t key,
                                   const ScnRef& sr, int P, bool collect, float& ox, float& oy, float& oacc, uint64_t* mask = nullptr,
                                   bool* capped = nullptr, bool* planed = nullptr) {
    const float2 bp = in.p[slot];
    const float acc = in.acc[slot];
    const uint32_t self = b.sid[slot] - sr.agentBase;
    const float ri = b.sr[slot];
    const float4 pl = b.planes[slot];
    const uint32_t cid = key - sr.cellBase;
    const int cx = int(cid % uint32_t(s.hashNX)), cy = int(cid / uint32_t(s.hashNX));
    V2 corr{ 0.0f, 0.0f };
    int nContacts = 0;
    uint32_t cnt = 0;
    const size_t stride = size_t(n);
    const float Pf = b.listMoves > 0.0f ? b.listMoves : float(P);   // the moves the candidates' reach allows for
    const int R = s.projRing;
    const int x0 = cx - R < 0 ? 0 : cx - R;
    const int x1 = cx + R > s.hashNX - 1 ? s.hashNX - 1 : cx + R;
    for (int yy = -R; yy <= R; yy++) {
        const int ny = cy + yy;
        if (ny < 0 || ny >= s.hashNY) continue;
        const uint32_t row = sr.cellBase + uint32_t(ny * s.hashNX);
        const uint32_t st = b.bstart[row + uint32_t(x0)], en = b.bstart[row + uint32_t(x1) + 1u];
        for (uint32_t k0 = st; k0 < en; k0 += PRED_CHUNK) {   // chunks of loads in flight together, ring order kept
            float2 qp[PRED_CHUNK];
            float qw[PRED_CHUNK], qr[PRED_CHUNK];
#pragma unroll
            for (int u = 0; u < PRED_CHUNK; ++u) {
                const uint32_t k = k0 + uint32_t(u) < en ? k0 + uint32_t(u) : en - 1u;
                qp[u] = in.p[k]; qw[u] = b.pw[k]; qr[u] = b.sr[k];
            }
#pragma unroll
            for (int u = 0; u < PRED_CHUNK; ++u) {
                const uint32_t k = k0 + uint32_t(u);
                if (k >= en) break;
                if (k == uint32_t(slot) || qw[u] == 0.0f) continue;
                const V2 offset = V2{ bp.x, bp.y } - V2{ qp[u].x, qp[u].y };
                const float D = ri + qr[u];
                const float dxy2 = dot2(offset, offset);
                if (dxy2 < contact2(D)) {
                    corr = corr + overlapCorrection(offset, dxy2, D, self, b.sid, k, sr.agentBase);
                    ++nContacts;
                    if (mask && cnt < uint32_t(NBR_CAP)) *mask |= 1ull << cnt;   // the contact's index among the candidates
                }
                if (collect) {
                    const float reach = (D + Pf * (PROJ_MAX_SHIFT_RADII * D)) * 1.001f + 1e-6f;
                    if (dxy2 < reach * reach) {
                        if (cnt < uint32_t(NBR_CAP)) b.nbr[size_t(cnt) * stride + size_t(slot)] = k;
                        ++cnt;
                    }
                }
            }
        }
    }
    if (collect) b.nbrCount[slot] = cnt <= uint32_t(NBR_CAP) ? cnt : NBR_OVERFLOW;
    finishProjection(pl, bp.x, bp.y, acc, corr, nContacts, ri, ox, oy, oacc, capped, planed);
}

// A slot's own words for one projection iteration over its candidate list. The iteration is a short chain of
// dependent load rounds per slot (the slot's words, the candidates' indices, the candidates' positions); loading
// every word of the slot at once, ahead of the kernel's branches on them, makes the first link one round. The
// slot's key (only for the ring scan) and id (only for a stacked pair) are read where they are needed.
struct SlotHead {
    float2 p;
    float acc, ri;
    float4 plane;
};
__device__ inline SlotHead loadHead(const StepBuffers& b, const ProjBuffers& in, int slot) {
    return SlotHead{ in.p[slot], in.acc[slot], b.sr[slot], b.planes[slot] };
}
// The compiler otherwise moves each load into the branch that first uses it, one round per branch: this empty
// statement uses every value at the point of the call, so all the loads issue before it and the wave waits once.
__device__ inline void holdHead(float pw, uint32_t listed, const SlotHead& h) {
    asm volatile("" ::"v"(pw), "v"(listed), "v"(h.p.x), "v"(h.p.y), "v"(h.acc), "v"(h.ri), "v"(h.plane.x), "v"(h.plane.y),
                 "v"(h.plane.z), "v"(h.plane.w));
}

// One iteration for one slot over its contact candidates (cnt >= 1), CH of them loaded together
template <int CH = PROJ_CHUNK>
__device__ inline void projectList(const StepBuffers& b, const ProjBuffers& in, int n, int slot, uint32_t cnt, const ScnRef& sr,
                                   const SlotHead& h, float& ox, float& oy, float& oacc, uint64_t* mask = nullptr,
                                   bool* capped = nullptr, bool* planed = nullptr) {
    const float2 bp = h.p;
    const float acc = h.acc;
    const float ri = h.ri;
    V2 corr{ 0.0f, 0.0f };
    int nContacts = 0;
    const size_t stride = size_t(n);
    // candidates in chunks: the chunk's index loads, then its positions and radii, are all in flight before the first
    // use; the sum still runs in candidate order. A chunk's entries past the count load the last candidate again (no
    // branch around the loads) and are never used. A pair's contact distance is the radii's sum, as the collecting
    // scan formed it (the lists hold indices only).
    for (uint32_t j0 = 0; j0 < cnt; j0 += CH) {
        uint32_t kk[CH];
        float dd[CH];
        float2 pp[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            const uint32_t j = j0 + uint32_t(u) < cnt ? j0 + uint32_t(u) : cnt - 1u;
            kk[u] = b.nbr[size_t(j) * stride + size_t(slot)];
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            pp[u] = in.p[kk[u]];
            dd[u] = ri + b.sr[kk[u]];
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            if (j0 + uint32_t(u) >= cnt) break;
            const V2 offset = V2{ bp.x, bp.y } - V2{ pp[u].x, pp[u].y };
            const float dxy2 = dot2(offset, offset);
            if (dxy2 < contact2(dd[u])) {
                corr = corr + overlapCorrectionAt(offset, dxy2, dd[u], b.sid, uint32_t(slot), kk[u], sr.agentBase);
                ++nContacts;
                if (mask) *mask |= 1ull << (j0 + uint32_t(u));
            }
        }
    }
    finishProjection(h.plane, bp.x, bp.y, acc, corr, nContacts, ri, ox, oy, oacc, capped, planed);
}

// One iteration over the candidates that also drops those out of reach for the rest of the step: with 'rem'
// iterations left (this one included) each agent moves by at most rem PROJ_MAX_SHIFT_RADII r, so a pair farther apart
// than D + rem PROJ_MAX_SHIFT_RADII D (with the collection's slack) cannot touch again. The kept candidates keep
// their order, so every later sum is the unpruned sum bit for bit (the dropped terms are never contacts).
template <int CH = PROJ_CHUNK>
__device__ inline void projectListPrune(const StepBuffers& b, const ProjBuffers& in, int n, int slot, uint32_t cnt, const ScnRef& sr,
                                        const SlotHead& h, int rem, float& ox, float& oy, float& oacc) {
    const float2 bp = h.p;
    const float acc = h.acc;
    const float ri = h.ri;
    V2 corr{ 0.0f, 0.0f };
    int nContacts = 0;
    const size_t stride = size_t(n);
    // a budget caps what the rest of the step can move each agent (a broken budget never reaches here)
    const float remf = b.listMoves > 0.0f ? fminf(float(rem), b.listMoves) : float(rem);
    uint32_t kept = 0;
    for (uint32_t j0 = 0; j0 < cnt; j0 += CH) {
        uint32_t kk[CH];
        float dd[CH];
        float2 pp[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            const bool in_ = j0 + uint32_t(u) < cnt;
            kk[u] = in_ ? b.nbr[size_t(j0 + uint32_t(u)) * stride + size_t(slot)] : uint32_t(slot);
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            pp[u] = in.p[kk[u]];
            dd[u] = ri + b.sr[kk[u]];
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            if (j0 + uint32_t(u) >= cnt) break;
            const V2 offset = V2{ bp.x, bp.y } - V2{ pp[u].x, pp[u].y };
            const float dxy2 = dot2(offset, offset);
            if (dxy2 < contact2(dd[u])) {
                corr = corr + overlapCorrectionAt(offset, dxy2, dd[u], b.sid, uint32_t(slot), kk[u], sr.agentBase);
                ++nContacts;
            }
            const float reach = (dd[u] + remf * (PROJ_MAX_SHIFT_RADII * dd[u])) * 1.001f + 1e-6f;
            if (dxy2 < reach * reach) {
                // reads of this chunk are done; later chunks sit above 'kept'
                if (kept != j0 + uint32_t(u)) b.nbr[size_t(kept) * stride + size_t(slot)] = kk[u];
                ++kept;
            }
        }
    }
    if (kept != cnt) b.nbrCount[slot] = kept;
    finishProjection(h.plane, bp.x, bp.y, acc, corr, nContacts, ri, ox, oy, oacc);
}

// candidates loaded together in a list iteration: the plain step's launches hold one wave per SIMD and gain from
// more loads in flight; the batched step's hold many and are bound by bandwidth (same-box: 8 for the plain bowl,
// 313 us per step against 323 us at 4; 4 for the 256-scenario ensemble, 1.279 ms against 1.352 ms at 8)
template <bool B> __host__ __device__ constexpr int listChunk() { return B ? PROJ_CHUNK_ENSEMBLE : PROJ_CHUNK; }

// A later iteration with Q lanes per slot (the production step's iterations that do not prune). Lane l of a slot
// loads and tests candidates j = l, l + Q, ..., PROJ_CHUNK of them per round, so Q x PROJ_CHUNK candidates are in
// flight together; every lane then adds the contact terms in candidate order (j ascending), reading the other lanes'
// terms by shuffles within the slot's Q lanes, so the sum is projectList's bit for bit. Lane 0 finishes and writes.
// A slot's Q lanes are adjacent lanes of one wave and follow the same path.
// REC: iteration 'it' of the adjoint's taped forward (contact masks, cap and plane bits, as k_project<B, true>)
template <bool B, int Q, bool REC = false>
__global__ void __launch_bounds__(PHYS_MAX_BLOCK) k_project_lanes(int n, SceneView s0, BatchView bt, StepBuffers b, ProjBuffers in,
                                                                  ProjBuffers out, int P, TapeStep tp = TapeStep{}, int it = 0) {
    static_assert(Q == 2 || Q == 4 || Q == 8, "lanes per slot");
    const int total = int(b.ctrl[CTRL_NACTIVE]) * Q;   // the inactive slots (the last ones) are left alone
    const size_t stride = size_t(n);
    // blockDim and the grid stride are multiples of Q, so a slot's lanes take every branch together
    for (int t = xcdBlock() * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int slot = t / Q;
        const int lane = t % Q;
        const uint32_t key = b.skey[slot];
        if (key == INACTIVE_CELL) continue;
        const float2 bp = in.p[slot];
        const float acc = in.acc[slot];
        if (b.pw[slot] == 0.0f) {   // parked this step: carried
            if (lane == 0) {
                out.p[slot] = bp; out.acc[slot] = acc;
                if constexpr (REC) tp.mask[size_t(it - 1) * stride + size_t(slot)] = 0ull;
            }
            continue;
        }
        const ScnRef sr = scnOfSlot<B>(bt, uint32_t(slot));
        const bool broken = b.listMoves >= 0.0f && b.listBroken[sr.sc] != 0u;
        const uint32_t cnt = broken ? NBR_OVERFLOW : b.nbrCount[slot];
        if (cnt == NBR_OVERFLOW) {   // the ring scan, on lane 0
            if (lane == 0) {
                float x, y, oacc;
                if constexpr (REC) {
                    uint64_t mask = 0ull;
                    bool capped = false, planed = false;
                    projectRing(sceneOf<B>(s0, bt, sr.sc), b, in, n, slot, key, sr, P, false, x, y, oacc, &mask, &capped, &planed);
                    atomicOr(tp.err, broken ? 4u : 1u);
                    tp.mask[size_t(it - 1) * stride + size_t(slot)] = mask;
                } else {
                    projectRing(sceneOf<B>(s0, bt, sr.sc), b, in, n, slot, key, sr, P, false, x, y, oacc);
                    if (!broken && b.listMoves > 0.0f && oacc > b.listMoves * (PROJ_MAX_SHIFT_RADII * b.sr[slot]))
                        atomicOr(&b.listBroken[sr.sc], 1u);
                }
                out.p[slot] = make_float2(x, y);
                out.acc[slot] = oacc;
            }
            continue;
        }
        const uint32_t self = b.sid[slot] - sr.agentBase;
        const float ri = b.sr[slot];
        V2 corr{ 0.0f, 0.0f };
        int nContacts = 0;
        uint64_t mask = 0ull;   // REC: the contacts' indices among the candidates
        for (uint32_t j0 = 0; j0 < cnt; j0 += uint32_t(Q * PROJ_CHUNK)) {
            uint32_t kk[PROJ_CHUNK];
            float dd[PROJ_CHUNK];
            float2 pp[PROJ_CHUNK];
#pragma unroll
            for (int c = 0; c < PROJ_CHUNK; ++c) {
                const uint32_t j = j0 + uint32_t(c * Q + lane);
                kk[c] = j < cnt ? b.nbr[size_t(j) * stride + size_t(slot)] : uint32_t(slot);
            }
#pragma unroll
            for (int c = 0; c < PROJ_CHUNK; ++c) {
                pp[c] = in.p[kk[c]];
                dd[c] = ri + b.sr[kk[c]];
            }
            float tx[PROJ_CHUNK], ty[PROJ_CHUNK];
            int hit[PROJ_CHUNK];
#pragma unroll
            for (int c = 0; c < PROJ_CHUNK; ++c) {
                const uint32_t j = j0 + uint32_t(c * Q + lane);
                const V2 offset = V2{ bp.x, bp.y } - V2{ pp[c].x, pp[c].y };
                const float dxy2 = dot2(offset, offset);
                hit[c] = j < cnt && dxy2 < contact2(dd[c]);
                V2 term{ 0.0f, 0.0f };
                if (hit[c]) term = overlapCorrection(offset, dxy2, dd[c], self, b.sid, kk[c], sr.agentBase);
                tx[c] = term.x;
                ty[c] = term.y;
            }
            // candidate order: c outer, lane inner
#pragma unroll
            for (int c = 0; c < PROJ_CHUNK; ++c) {
#pragma unroll
                for (int l = 0; l < Q; ++l) {
                    const int h = __shfl(hit[c], l, Q);
                    const float x = __shfl(tx[c], l, Q);
                    const float y = __shfl(ty[c], l, Q);
                    if (h) {
                        corr = corr + V2{ x, y };
                        ++nContacts;
                        if constexpr (REC) mask |= 1ull << (j0 + uint32_t(c * Q + l));
                    }
                }
            }
        }
        if (lane == 0) {
            float x, y, oacc;
            bool capped = false, planed = false;
            finishProjection(b.planes[slot], bp.x, bp.y, acc, corr, nContacts, ri, x, y, oacc, &capped, &planed);
            out.p[slot] = make_float2(x, y);
            out.acc[slot] = oacc;
            if (b.listMoves > 0.0f && oacc > b.listMoves * (PROJ_MAX_SHIFT_RADII * ri)) atomicOr(&b.listBroken[sr.sc], 1u);
            if constexpr (REC) {
                tp.mask[size_t(it - 1) * stride + size_t(slot)] = mask;
                if (capped) tp.capBits[slot] |= 1ull << (it - 1);
                if (planed) tp.planeBits[slot] |= 1ull << (it - 1);
            }
        }
    }
}

// The first iteration (ring scan + candidate collection, projectRing with 'collect') with Q lanes per slot. In each
// ring row lane l of a slot takes the row's slots k0 + c Q + l (c < PROJ_CHUNK), so the (c, l) order of a round is
// the row's slot order: the contact terms are added in that order through shuffles within the slot's lanes, and the
// candidates within reach are written at their rank in that order (the slot's ballot bits below the lane), so the
// sum, the list and its count are projectRing's bit for bit. Lane 0 finishes and writes. A slot's Q lanes are
// adjacent lanes of one wave and follow the same path.
template <bool B, int Q>
__global__ void __launch_bounds__(PHYS_MAX_BLOCK) k_project_collect(int n, SceneView s0, BatchView bt, StepBuffers b, ProjBuffers in,
                                                                    ProjBuffers out, int P) {
    static_assert(Q == 2 || Q == 4 || Q == 8, "lanes per slot");
    const int total = int(b.ctrl[CTRL_NACTIVE]) * Q;   // the inactive slots (the last ones) are left alone
    const size_t stride = size_t(n);
    const int groupLo = __lane_id() & ~(Q - 1);   // the slot's first lane in the wave (blockDim is a multiple of 64)
    const float Pf = b.listMoves > 0.0f ? b.listMoves : float(P);   // the moves the candidates' reach allows for
    for (int t = xcdBlock() * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int slot = t / Q;
        const int lane = t % Q;
        const uint32_t key = b.skey[slot];
        if (key == INACTIVE_CELL) continue;
        const float2 bp = in.p[slot];
        const float acc = in.acc[slot];
        if (b.pw[slot] == 0.0f) {   // parked this step: carried
            if (lane == 0) { out.p[slot] = bp; out.acc[slot] = acc; }
            continue;
        }
        const ScnRef sr = scnOfSlot<B>(bt, uint32_t(slot));
        const SceneView& s = sceneOf<B>(s0, bt, sr.sc);
        const uint32_t self = b.sid[slot] - sr.agentBase;
        const float ri = b.sr[slot];
        const uint32_t cid = key - sr.cellBase;
        const int cx = int(cid % uint32_t(s.hashNX)), cy = int(cid / uint32_t(s.hashNX));
        const int R = s.projRing;
        const int x0 = cx - R < 0 ? 0 : cx - R;
        const int x1 = cx + R > s.hashNX - 1 ? s.hashNX - 1 : cx + R;
        V2 corr{ 0.0f, 0.0f };
        int nContacts = 0;
        uint32_t cnt = 0;
        for (int yy = -R; yy <= R; yy++) {
            const int ny = cy + yy;
            if (ny < 0 || ny >= s.hashNY) continue;
            const uint32_t row = sr.cellBase + uint32_t(ny * s.hashNX);
            const uint32_t st = b.bstart[row + uint32_t(x0)], en = b.bstart[row + uint32_t(x1) + 1u];
            for (uint32_t k0 = st; k0 < en; k0 += uint32_t(Q * PROJ_CHUNK)) {
                float2 qp[PROJ_CHUNK];
                float qw[PROJ_CHUNK], qr[PROJ_CHUNK];
#pragma unroll
                for (int c = 0; c < PROJ_CHUNK; ++c) {
                    const uint32_t k = k0 + uint32_t(c * Q + lane);
                    const uint32_t kc = k < en ? k : en - 1u;
                    qp[c] = in.p[kc]; qw[c] = b.pw[kc]; qr[c] = b.sr[kc];
                }
#pragma unroll
                for (int c = 0; c < PROJ_CHUNK; ++c) {
                    if (k0 + uint32_t(c * Q) >= en) break;   // uniform over the slot's lanes
                    const uint32_t k = k0 + uint32_t(c * Q + lane);
                    const bool live = k < en && k != uint32_t(slot) && qw[c] != 0.0f;
                    const V2 offset = V2{ bp.x, bp.y } - V2{ qp[c].x, qp[c].y };
                    const float D = ri + qr[c];
                    const float dxy2 = dot2(offset, offset);
                    const bool hit = live && dxy2 < contact2(D);
                    V2 term{ 0.0f, 0.0f };
                    if (hit) term = overlapCorrection(offset, dxy2, D, self, b.sid, k, sr.agentBase);
                    const float reach = (D + Pf * (PROJ_MAX_SHIFT_RADII * D)) * 1.001f + 1e-6f;
                    const bool take = live && dxy2 < reach * reach;
                    const uint32_t bits = uint32_t(__ballot(take) >> groupLo) & ((1u << Q) - 1u);
                    if (take) {
                        const uint32_t at = cnt + uint32_t(__popc(bits & ((1u << lane) - 1u)));
                        if (at < uint32_t(NBR_CAP)) b.nbr[size_t(at) * stride + size_t(slot)] = k;
                    }
                    cnt += uint32_t(__popc(bits));
#pragma unroll
                    for (int l = 0; l < Q; ++l) {
                        const int h = __shfl(int(hit), l, Q);
                        const float tx = __shfl(term.x, l, Q);
                        const float ty = __shfl(term.y, l, Q);
                        if (h) {
                            corr = corr + V2{ tx, ty };
                            ++nContacts;
                        }
                    }
                }
            }
        }
        if (lane == 0) {
            b.nbrCount[slot] = cnt <= uint32_t(NBR_CAP) ? cnt : NBR_OVERFLOW;
            float x, y, oacc;
            finishProjection(b.planes[slot], bp.x, bp.y, acc, corr, nContacts, ri, x, y, oacc);
            out.p[slot] = make_float2(x, y);
            out.acc[slot] = oacc;
            if (b.listMoves > 0.0f && oacc > b.listMoves * (PROJ_MAX_SHIFT_RADII * ri)) atomicOr(&b.listBroken[sr.sc], 1u);
        }
    }
}

// iteration 'first' (1: ring scan + candidates) or a later one (the candidates), slot order in -> out. REC: iteration
// 'it' of the adjoint's taped forward (contact masks, cap and plane bits)
// 'prune' > 0 (not REC): the candidates are pruned for the 'prune' iterations left, this one included
template <bool B, bool REC = false>
__global__ void __launch_bounds__(PHYS_MAX_BLOCK) k_project(int n, SceneView s0, BatchView bt, StepBuffers b, ProjBuffers in,
                                                            ProjBuffers out, int P, int first, TapeStep tp = TapeStep{}, int it = 0,
                                                            int prune = 0) {
    const int active = int(b.ctrl[CTRL_NACTIVE]);   // the inactive slots follow the active ones and are left alone
    for (int slot = xcdBlock() * blockDim.x + threadIdx.x; slot < active; slot += gridDim.x * blockDim.x) {
        // every word of the slot in one load round, before the branches on them
        const float pw = b.pw[slot];
        const uint32_t listed = b.nbrCount[slot];
        const SlotHead h = loadHead(b, in, slot);
        holdHead(pw, listed, h);
        float x, y, acc;
        uint64_t mask = 0ull;
        bool capped = false, planed = false;
        if (pw == 0.0f) {   // parked this step: carried
            x = h.p.x; y = h.p.y; acc = h.acc;
        } else {
            const ScnRef sr = scnOfSlot<B>(bt, uint32_t(slot));
            const bool broken = !first && b.listMoves >= 0.0f && b.listBroken[sr.sc] != 0u;
            const uint32_t cnt = first || broken ? NBR_OVERFLOW : listed;
            if constexpr (REC) {
                if (cnt == NBR_OVERFLOW) {
                    projectRing(sceneOf<B>(s0, bt, sr.sc), b, in, n, slot, b.skey[slot], sr, P, first != 0, x, y, acc, &mask, &capped,
                                &planed);
                    // the tape's masks index the candidate lists: a broken budget (bit 2; taped again without one) or an
                    // overflowing list (bit 0) leaves this step's untaped
                    if (broken) atomicOr(tp.err, 4u);
                    else if (!first || b.nbrCount[slot] == NBR_OVERFLOW) atomicOr(tp.err, 1u);
                } else {
                    projectList<listChunk<B>()>(b, in, n, slot, cnt, sr, h, x, y, acc, &mask, &capped, &planed);
                }
            } else {
                if (cnt == NBR_OVERFLOW) projectRing(sceneOf<B>(s0, bt, sr.sc), b, in, n, slot, b.skey[slot], sr, P, first != 0, x, y, acc);
                else if (prune > 0) projectListPrune<listChunk<B>()>(b, in, n, slot, cnt, sr, h, prune, x, y, acc);
                else projectList<listChunk<B>()>(b, in, n, slot, cnt, sr, h, x, y, acc);
            }
            // the budget: a path past listMoves shifts makes the scenario's later iterations scan the ring
            if (!broken && b.listMoves > 0.0f && acc > b.listMoves * (PROJ_MAX_SHIFT_RADII * h.ri)) atomicOr(&b.listBroken[sr.sc], 1u);
        }
        out.p[slot] = make_float2(x, y);
        out.acc[slot] = acc;
        if constexpr (REC) {
            tp.mask[size_t(it - 1) * size_t(n) + size_t(slot)] = mask;
            if (capped) tp.capBits[slot] |= 1ull << (it - 1);
            if (planed) tp.planeBits[slot] |= 1ull << (it - 1);
        }
    }
}

// ---------------- PHYSICS_PASS 3: final ----------------
// REC: the adjoint's taped forward (the last iteration's contacts, the final pass's wall pushes)
template <bool B, bool REC = false>
__global__ void __launch_bounds__(PHYS_MAX_BLOCK) k_final(int n, SceneView s0, BatchView bt, ModelParams mp, StepBuffers b, ProjBuffers in,
                                                          int useList, AgentHalf rd, AgentHalf wr, EvacBuffers ev, uint32_t* __restrict__ ctrl,
                                                          TapeStep tp = TapeStep{}, ProjBuffers last = ProjBuffers{}) {
    const uint32_t step = ctrl[CTRL_STEP];
    const int active = int(ctrl[CTRL_NACTIVE]);
    for (int slot = xcdBlock() * blockDim.x + threadIdx.x; slot < n; slot += gridDim.x * blockDim.x) {
        if constexpr (REC) { tp.fin[slot].flags = 0u; tp.fin[slot].n_events = 0u; }
        if (slot >= active) {   // an inactive slot (the inactive key sorts last): its agent is carried over
            const uint32_t agent = b.sid[slot];
            wr.x[agent] = rd.x[agent]; wr.y[agent] = rd.y[agent]; wr.w[agent] = rd.w[agent];
            wr.vx[agent] = rd.vx[agent]; wr.vy[agent] = rd.vy[agent]; wr.spd[agent] = rd.spd[agent];
            continue;
        }
        // every word of the slot in one load round, before the branches on them
        const uint32_t key = b.skey[slot];
        const uint32_t agent = b.sid[slot];
        const float pw = b.pw[slot];
        const uint32_t listed = b.nbrCount[slot];
        const SlotHead h = loadHead(b, in, slot);
        const float sx = b.sx[slot], sy = b.sy[slot], pvx = b.pvx[slot], pvy = b.pvy[slot];
        holdHead(pw, listed, h);
        asm volatile("" ::"v"(key), "v"(agent), "v"(sx), "v"(sy), "v"(pvx), "v"(pvy));
        if (key == INACTIVE_CELL) {
            wr.x[agent] = rd.x[agent]; wr.y[agent] = rd.y[agent]; wr.w[agent] = rd.w[agent];
            wr.vx[agent] = rd.vx[agent]; wr.vy[agent] = rd.vy[agent]; wr.spd[agent] = rd.spd[agent];
        } else {
            const ScnRef sr = scnOfSlot<B>(bt, uint32_t(slot));
            const SceneView& s = sceneOf<B>(s0, bt, sr.sc);
            float x, y, acc;
            if (pw == 0.0f || mp.projIters == 0) {
                x = h.p.x; y = h.p.y; acc = h.acc;
                if constexpr (REC) tp.mask[size_t(mp.projIters - 1) * size_t(n) + size_t(slot)] = 0ull;
            } else {
                const bool broken = b.listMoves >= 0.0f && b.listBroken[sr.sc] != 0u;
                const uint32_t cnt = useList && !broken ? listed : NBR_OVERFLOW;
                if constexpr (REC) {
                    uint64_t mask = 0ull;
                    bool capped = false, planed = false;
                    if (cnt == NBR_OVERFLOW) {
                        projectRing(s, b, in, n, slot, key, sr, mp.projIters, false, x, y, acc, &mask, &capped, &planed);
                        atomicOr(tp.err, broken ? 4u : 1u);
                    } else {
                        projectList<listChunk<B>()>(b, in, n, slot, cnt, sr, h, x, y, acc, &mask, &capped, &planed);
                    }
                    const int it = mp.projIters;
                    tp.mask[size_t(it - 1) * size_t(n) + size_t(slot)] = mask;
                    if (capped) tp.capBits[slot] |= 1ull << (it - 1);
                    if (planed) tp.planeBits[slot] |= 1ull << (it - 1);
                } else {
                    if (cnt == NBR_OVERFLOW) projectRing(s, b, in, n, slot, key, sr, mp.projIters, false, x, y, acc);
                    else projectList<listChunk<B>()>(b, in, n, slot, cnt, sr, h, x, y, acc);
                }
            }
            if constexpr (REC) { last.p[slot] = make_float2(x, y); last.acc[slot] = acc; }
            if (pw == 0.0f) {   // parked this step: an evacuation event
                // the crossing: the input position (inside) and the previous step's input, which the write half
                // still holds
                ev.inside[agent] = make_float2(sx, sy);
                ev.before[agent] = make_float2(wr.x[agent], wr.y[agent]);
                wr.x[agent] = x; wr.y[agent] = y; wr.w[agent] = 0.0f;
                wr.vx[agent] = 0.0f; wr.vy[agent] = 0.0f; wr.spd[agent] = acc;
                const int e = exitOf(s, V2{ sx, sy });
                const int ex = e >= 0 ? e : s.E;
                ev.step[agent] = int(step);
                ev.exit[agent] = ex;
                atomicAdd(&ev.perExit[sr.sc * uint32_t(PER_EXIT_STRIDE) + uint32_t(ex)], 1u);
            } else {
                V2 np{ x, y }, nv{ pvx, pvy };
                if (acc > 0.0f) {
                    float wall[3];
                    if constexpr (REC) {
                        swarm_final_rec* fr = tp.fin + slot;
                        fr->flags = SWARM_REC_MOVED;
                        WallTape wt{ &fr->flags, &fr->n_events, fr->ev, &fr->sweep_lo, tp.err };
                        PlaneRec pl{};
                        resolveStatic(s, V2{ sx, sy }, np, nv, h.ri, wall, wt, &pl);
                    } else {
                        resolveStatic(s, V2{ sx, sy }, np, nv, h.ri, wall);
                    }
                }
                wr.x[agent] = np.x; wr.y[agent] = np.y; wr.w[agent] = 1.0f;
                wr.vx[agent] = nv.x; wr.vy[agent] = nv.y; wr.spd[agent] = len2(nv);
            }
        }
    }
    // the last block to finish advances the step counter: every block has read it by then
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&ctrl[CTRL_DONE_BLOCKS], 1u) == gridDim.x - 1u) {
            ctrl[CTRL_DONE_BLOCKS] = 0u;
            ctrl[CTRL_STEP] = step + 1u;
            *reinterpret_cast<unsigned long long*>(ctrl + CTRL_ACTIVE_SUM_LO) += ctrl[CTRL_NACTIVE];
        }
    }
}

// ---------------- the adjoint's loss terms ----------------
// wave sum of v, one double atomic per wave
__device__ inline void waveAddDouble(double* acc, float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v != 0.0f) atomicAdd(acc, double(v));
}

// The overlap term at a step's end: each slot's contact candidates that overlap at the final positions (both
// active), each pair added once (from its lower slot) as (D - d)^2 * weight; MASK: the overlapping candidates' bits
template <bool MASK>
__global__ void __launch_bounds__(256) k_overlap(int n, StepBuffers b, AgentHalf wr, float weight, double* loss, TapeStep tp) {
    const int slot = xcdBlock() * int(blockDim.x) + int(threadIdx.x);
    float part = 0.0f;
    if (slot < n) {
        uint64_t bits = 0ull;
        if (b.skey[slot] != INACTIVE_CELL && b.pw[slot] != 0.0f) {
            const uint32_t cnt = b.nbrCount[slot];
            const uint32_t ai = b.sid[slot];
            const V2 pi{ wr.x[ai], wr.y[ai] };
            const float ri = b.sr[slot];
            for (uint32_t k = 0; cnt != NBR_OVERFLOW && k < cnt; ++k) {
                const uint32_t j = b.nbr[size_t(k) * size_t(n) + size_t(slot)];
                const uint32_t aj = b.sid[j];
                const V2 d = pi - V2{ wr.x[aj], wr.y[aj] };
                const float D = ri + b.sr[j];
                const float d2 = dot2(d, d);
                if (d2 < contact2(D)) {
                    bits |= 1ull << k;
                    if (uint32_t(slot) < j) {
                        const float ov = D - sqrtf(d2);
                        part += weight * ov * ov;
                    }
                }
            }
        }
        if constexpr (MASK) tp.overlap[slot] = bits;
    }
    waveAddDouble(loss, part);
}

// The exit term after the last step: an active agent's D = g(anchor) + |x - anchor| for its exit (the goal anchor
// of its cell; g(anchor) = the cell's distance less the centre's distance to the anchor), D^2 into the loss
template <bool B>
__global__ void __launch_bounds__(256) k_exit_loss(int n, SceneView s0, BatchView bt, AgentHalf st, const uint32_t* __restrict__ assign,
                                                   swarm_exit_rec* __restrict__ out, double* loss) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    float part = 0.0f;
    if (i < n) {
        swarm_exit_rec r{ 0.0f, 0.0f, 0.0f, 0u };
        const uint32_t ex = assign[i];
        if (st.w[i] != 0.0f && ex != CHOICE_NONE) {
            const ScnRef sr = scnOfAgent<B>(bt, uint32_t(i));
            const SceneView& s = sceneOf<B>(s0, bt, sr.sc);
            const V2 p{ st.x[i], st.y[i] };
            const int cell = choiceCell(s, p);
            const int k = slotOfExit(s, cell, ex);
            if (k >= 0) {
                const V2 a = anchorCentre(s, navSlots(s, cell)[k] & CHOICE_ANCHOR_MASK);
                const V2 c = anchorCentre(s, uint32_t(cell));
                const float base = slotDistOf(s, navDists(s, cell)[k]) - len2(c - a);
                const float D = base + len2(p - a);
                r = swarm_exit_rec{ a.x, a.y, base, 1u };
                part = D * D;
            }
        }
        out[i] = r;
    }
    waveAddDouble(loss, part);
}

// a plain streaming copy of known size: the HBM traffic counters' calibration and the measured copy bandwidth
__global__ void __launch_bounds__(256) k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst, size_t n4) {
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n4; i += size_t(gridDim.x) * 256) dst[i] = src[i];
}

// the step counter of an exit-choice pass run on its own (CrowdSim::stepExitChoice)
__global__ void k_step_increment(uint32_t* __restrict__ ctrl) {
    if (threadIdx.x == 0) ctrl[CTRL_STEP] += 1u;
}

// ---------------- host-side launchers ----------------
static inline int blocks(int n, int b) { return (n + b - 1) / b; }

void launchStepIncrement(hipStream_t st, uint32_t* ctrl) { k_step_increment<<<1, 64, 0, st>>>(ctrl); }

__global__ void __launch_bounds__(HASH_BLOCK) k_nav_probe(int n, SceneView s, const float* __restrict__ x, const float* __restrict__ y,
                                                          const float* __restrict__ r, const uint32_t* __restrict__ ex,
                                                          int32_t* __restrict__ cell, float* __restrict__ dir, float* __restrict__ goal) {
    const int i = blockIdx.x * HASH_BLOCK + threadIdx.x;
    if (i >= n) return;
    const V2 p{ x[i], y[i] };
    const int c = choiceCell(s, p);
    const V2 d = choiceDirection(s, c, p, ex[i]);
    const V2 g = goalWallRule(s, p, FEELER_REACH_RADII * r[i], d);
    cell[i] = c;
    dir[2 * i] = d.x; dir[2 * i + 1] = d.y;
    goal[2 * i] = g.x; goal[2 * i + 1] = g.y;
}

void launchNavProbe(hipStream_t st, const SceneView& s, int n, const float* x, const float* y, const float* r, const uint32_t* ex,
                    int32_t* cell, float* dir, float* goal) {
    k_nav_probe<<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, s, x, y, r, ex, cell, dir, goal);
}

// up to TAPE_COPY_MAX arrays of n words copied in one launch (the taped step's state copies)
__global__ void __launch_bounds__(256) k_copy_words(int n, TapeCopy c) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    const int a = int(blockIdx.y);
    if (i < n && a < c.count) c.dst[a][i] = c.src[a][i];
}
void launchTapeCopy(hipStream_t st, int n, const TapeCopy& c) {
    if (n <= 0 || c.count <= 0) return;
    k_copy_words<<<dim3(unsigned(blocks(n, 256)), unsigned(c.count)), 256, 0, st>>>(n, c);
}

void launchCopyProbe(hipStream_t st, const void* src, void* dst, size_t bytes) {
    k_copy_probe<<<4096, 256, 0, st>>>(static_cast<const float4*>(src), static_cast<float4*>(dst), bytes / 16);
}

template <bool B>
static void gridPasses(hipStream_t st, const StepLaunch& L, AgentHalf rd, const GridBuffers& g, StepBuffers& b, const float* rad,
                       uint32_t* ctrl) {
    const int n = L.n;
    const uint32_t C = L.cells;
    const uint32_t nb = C + uint32_t(blocks(n, HASH_BLOCK));
    const int tiles = blocks(int(nb), SCAN_TILE);
    k_hash<B><<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, L.s, L.bt, rd.x, rd.y, rd.w, rd.z, g.key, g.loc, g.bcnt, C);
    k_scan_reduce<<<tiles, SCAN_BLOCK, 0, st>>>(g.bcnt, nb, g.tileSum);
    k_scan_apply<<<tiles, SCAN_BLOCK, 0, st>>>(g.bcnt, nb, C, g.tileSum, g.bstart, ctrl);
    if (L.bitonic) {   // the cell ranges from the counts, the permutation from the reference's network
        launchBitonic(st, n, g.key, g.pairs);
        k_bitonic_gather<B><<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, L.bt, g.pairs, b.sid, b.skey, rd.x, rd.y, rad, b.sx, b.sy, b.sr);
        return;
    }
    k_scatter<<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, g.key, g.loc, g.bstart, g.tmp, C);
    k_fix_gather<B><<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, L.bt, g.key, g.tmp, g.bstart, C, b.sid, b.skey, rd.x, rd.y, rad,
                                                                   b.sx, b.sy, b.sr);
}

void launchSortKeys(hipStream_t st, const StepLaunch& L, AgentHalf rd, const GridBuffers& g, StepBuffers& b, const float* rad,
                    uint32_t* ctrl) {
    const int n = L.n;
    const uint32_t C = L.cells;
    const uint32_t nb = C + uint32_t(blocks(n, HASH_BLOCK));
    const int tiles = blocks(int(nb), SCAN_TILE);
    k_key_count<<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, g.key, g.loc, g.bcnt, C);
    k_scan_reduce<<<tiles, SCAN_BLOCK, 0, st>>>(g.bcnt, nb, g.tileSum);
    k_scan_apply<<<tiles, SCAN_BLOCK, 0, st>>>(g.bcnt, nb, C, g.tileSum, g.bstart, ctrl);
    if (L.bitonic) {
        launchBitonic(st, n, g.key, g.pairs);
        if (L.batched)
            k_bitonic_gather<true><<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, L.bt, g.pairs, b.sid, b.skey, rd.x, rd.y, rad, b.sx, b.sy, b.sr);
        else
            k_bitonic_gather<false><<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, L.bt, g.pairs, b.sid, b.skey, rd.x, rd.y, rad, b.sx, b.sy, b.sr);
        return;
    }
    k_scatter<<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, g.key, g.loc, g.bstart, g.tmp, C);
    if (L.batched)
        k_fix_gather<true><<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, L.bt, g.key, g.tmp, g.bstart, C, b.sid, b.skey, rd.x, rd.y, rad,
                                                                          b.sx, b.sy, b.sr);
    else
        k_fix_gather<false><<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, L.bt, g.key, g.tmp, g.bstart, C, b.sid, b.skey, rd.x, rd.y,
                                                                           rad, b.sx, b.sy, b.sr);
}

void launchGridPasses(hipStream_t st, const StepLaunch& L, AgentHalf rd, const GridBuffers& g, StepBuffers& b, const float* rad,
                      uint32_t* ctrl) {
    if (L.batched) gridPasses<true>(st, L, rd, g, b, rad, ctrl);
    else gridPasses<false>(st, L, rd, g, b, rad, ctrl);
}

template <bool B>
static void choicePasses(hipStream_t st, const StepLaunch& L, AgentHalf rd, const ChoiceBuffers& c, StepBuffers& b, uint32_t* ctrl) {
    const int n = L.n;
    const dim3 exitGrid(unsigned(L.scenarios), unsigned((std::max(L.maxExits, 1) + CHOICE_EXIT_WAVES - 1) / CHOICE_EXIT_WAVES));
    k_choice_prefix<B><<<exitGrid, 64 * CHOICE_EXIT_WAVES, 0, st>>>(L.s, L.bt, c, ctrl);
    // a seated venue's own step reads its cell table (the surface rule, the venue's wall term) from LDS where it fits
    int lds = 0;
    if constexpr (!B) lds = L.s.venue3d ? venueLdsBytes(L.s) : 0;
    if constexpr (!B) {
        if (lds > 0) k_choice_eval<false, true><<<blocks(n, HASH_BLOCK), HASH_BLOCK, lds, st>>>(n, L.s, L.bt, rd.x, rd.y, rd.w, c, ctrl);
    }
    if (lds == 0) k_choice_eval<B><<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, L.s, L.bt, rd.x, rd.y, rd.w, c, ctrl);
    k_choice_thresh<B><<<exitGrid, 64 * CHOICE_EXIT_WAVES, 0, st>>>(L.s, L.bt, c, ctrl);
    if constexpr (!B) {
        if (lds > 0)
            k_choice_apply<false, true><<<blocks(n, HASH_BLOCK), HASH_BLOCK, lds, st>>>(n, L.s, L.bt, b.skey, b.sid, b.sx, b.sy, b.sr, c,
                                                                                         b.egx, b.egy, ctrl, rd.z);
    }
    if (lds == 0)
        k_choice_apply<B><<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, L.s, L.bt, b.skey, b.sid, b.sx, b.sy, b.sr, c, b.egx, b.egy,
                                                                         ctrl, rd.z);
}

void launchChoice(hipStream_t st, const StepLaunch& L, AgentHalf rd, const ChoiceBuffers& c, StepBuffers& b, const float* driveX,
                  const float* driveY, uint32_t* ctrl) {
    if (!L.batched && (driveX || L.s.E == 0)) {
        k_goal_drive<<<blocks(L.n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(L.n, b.skey, b.sid, driveX, driveY, b.egx, b.egy);
        return;
    }
    if (L.batched) choicePasses<true>(st, L, rd, c, b, ctrl);
    else choicePasses<false>(st, L, rd, c, b, ctrl);
}

template <bool B>
static void physicsPasses(hipStream_t st, const StepLaunch& L, StepBuffers& b, ProjBuffers pa, ProjBuffers pb, AgentHalf rd, AgentHalf wr,
                          EvacBuffers ev, uint32_t* ctrl, hipEvent_t* stamps, hipEvent_t* iterStamps) {
    const int n = L.n;
    const int bs = L.physBlock > 0 ? L.physBlock : PHYS_BLOCK;
    const int full = blocks(n, bs);
    const int g = L.physGrid > 0 && L.physGrid < full ? L.physGrid : full;   // a persistent grid, or one slot per thread
    const int P = L.mp.projIters;
    if (L.predictLanes > 1) {
        const int q = L.predictLanes;
        const int full = blocks(n * q, bs);
        const int gq = L.physGrid > 0 && L.physGrid < full ? L.physGrid : full;
        if (q == 2) k_predict_lanes<B, 2><<<gq, bs, 0, st>>>(n, L.s, L.bt, L.mp, b, pa);
        else if (q == 4) k_predict_lanes<B, 4><<<gq, bs, 0, st>>>(n, L.s, L.bt, L.mp, b, pa);
        else k_predict_lanes<B, 8><<<gq, bs, 0, st>>>(n, L.s, L.bt, L.mp, b, pa);
    } else {
        k_predict<B><<<g, bs, 0, st>>>(n, L.s, L.bt, L.mp, b, pa);
    }
    if (stamps) (void)hipEventRecord(stamps[0], st);
    ProjBuffers in = pa, out = pb;
    for (int it = 1; it < P; ++it) {
        // pruning at iterations 1 + k pruneEvery (k >= 1); the final pass is iteration P
        const int prune = L.pruneEvery > 0 && it > 1 && (it - 1) % L.pruneEvery == 0 ? P - it + 1 : 0;
        if (iterStamps && it > 1 && (prune > 0 || L.projLanes <= 1)) {
            // timed: the dispatch itself records the pair of events, so they bracket the kernel's execution
            hipExtLaunchKernelGGL(k_project<B, false>, dim3(g), dim3(bs), 0, st, iterStamps[2 * (it - 2)], iterStamps[2 * (it - 2) + 1], 0u,
                                  n, L.s, L.bt, b, in, out, P, 0, TapeStep{}, 0, prune);
            const ProjBuffers t = in; in = out; out = t;
            continue;
        }
        if (iterStamps && it > 1) (void)hipEventRecord(iterStamps[2 * (it - 2)], st);
        if (it == 1 && L.collectLanes > 1) {
            const int q = L.collectLanes;
            const int full = blocks(n * q, bs);
            const int gq = L.physGrid > 0 && L.physGrid < full ? L.physGrid : full;
            if (q == 2) k_project_collect<B, 2><<<gq, bs, 0, st>>>(n, L.s, L.bt, b, in, out, P);
            else if (q == 4) k_project_collect<B, 4><<<gq, bs, 0, st>>>(n, L.s, L.bt, b, in, out, P);
            else k_project_collect<B, 8><<<gq, bs, 0, st>>>(n, L.s, L.bt, b, in, out, P);
        } else if (it == 1 || prune > 0 || L.projLanes <= 1) {
            k_project<B><<<g, bs, 0, st>>>(n, L.s, L.bt, b, in, out, P, it == 1 ? 1 : 0, TapeStep{}, 0, prune);
        } else {
            const int q = L.projLanes;
            const int full = blocks(n * q, bs);
            const int gq = L.physGrid > 0 && L.physGrid < full ? L.physGrid : full;
            if (q == 2) k_project_lanes<B, 2><<<gq, bs, 0, st>>>(n, L.s, L.bt, b, in, out, P);
            else if (q == 4) k_project_lanes<B, 4><<<gq, bs, 0, st>>>(n, L.s, L.bt, b, in, out, P);
            else k_project_lanes<B, 8><<<gq, bs, 0, st>>>(n, L.s, L.bt, b, in, out, P);
        }
        if (iterStamps && it > 1) (void)hipEventRecord(iterStamps[2 * (it - 2) + 1], st);
        const ProjBuffers t = in; in = out; out = t;
    }
    if (stamps) (void)hipEventRecord(stamps[1], st);
    k_final<B><<<g, bs, 0, st>>>(n, L.s, L.bt, L.mp, b, in, P >= 2 ? 1 : 0, rd, wr, ev, ctrl);
}

void launchPhysics(hipStream_t st, const StepLaunch& L, StepBuffers& b, ProjBuffers pa, ProjBuffers pb, AgentHalf rd, AgentHalf wr,
                   EvacBuffers ev, uint32_t* ctrl, hipEvent_t* stamps, hipEvent_t* iterStamps) {
    if (L.batched) physicsPasses<true>(st, L, b, pa, pb, rd, wr, ev, ctrl, stamps, iterStamps);
    else physicsPasses<false>(st, L, b, pa, pb, rd, wr, ev, ctrl, stamps, iterStamps);
}

template <bool B>
static void physicsRecord(hipStream_t st, const StepLaunch& L, StepBuffers& bIn, float2* q, float* acc0, float* acc1, AgentHalf rd,
                          AgentHalf wr, EvacBuffers ev, uint32_t* ctrl, const TapeStep& tp) {
    StepBuffers b = bIn;   // its listMoves: the caller's budget for the taped forward
    const int n = L.n;
    const int bs = L.physBlock > 0 ? L.physBlock : PHYS_BLOCK;
    const int g = blocks(n, bs);
    const int P = L.mp.projIters;
    float* acc[2] = { acc0, acc1 };
    auto at = [&](int it) { return ProjBuffers{ q + size_t(it) * size_t(n), acc[it & 1] }; };
    k_predict<B, true><<<g, bs, 0, st>>>(n, L.s, L.bt, L.mp, b, at(0), tp);
    const int lanes = L.projLanes;
    const int gq = blocks(n * lanes, bs);
    for (int it = 1; it < P; ++it) {
        if (it == 1 || lanes <= 1) k_project<B, true><<<g, bs, 0, st>>>(n, L.s, L.bt, b, at(it - 1), at(it), P, it == 1 ? 1 : 0, tp, it);
        else if (lanes == 2) k_project_lanes<B, 2, true><<<gq, bs, 0, st>>>(n, L.s, L.bt, b, at(it - 1), at(it), P, tp, it);
        else if (lanes == 4) k_project_lanes<B, 4, true><<<gq, bs, 0, st>>>(n, L.s, L.bt, b, at(it - 1), at(it), P, tp, it);
        else k_project_lanes<B, 8, true><<<gq, bs, 0, st>>>(n, L.s, L.bt, b, at(it - 1), at(it), P, tp, it);
    }
    k_final<B, true><<<g, bs, 0, st>>>(n, L.s, L.bt, L.mp, b, at(P - 1), 1, rd, wr, ev, ctrl, tp, at(P));
}

void launchPhysicsRecord(hipStream_t st, const StepLaunch& L, StepBuffers& b, float2* q, float* acc0, float* acc1, AgentHalf rd,
                         AgentHalf wr, EvacBuffers ev, uint32_t* ctrl, const TapeStep& tp) {
    if (L.batched) physicsRecord<true>(st, L, b, q, acc0, acc1, rd, wr, ev, ctrl, tp);
    else physicsRecord<false>(st, L, b, q, acc0, acc1, rd, wr, ev, ctrl, tp);
}

void launchOverlap(hipStream_t st, const StepLaunch& L, StepBuffers& b, AgentHalf wr, float weight, double* loss, const TapeStep* tp) {
    const int g = blocks(L.n, 256);
    if (tp) k_overlap<true><<<g, 256, 0, st>>>(L.n, b, wr, weight, loss, *tp);
    else k_overlap<false><<<g, 256, 0, st>>>(L.n, b, wr, weight, loss, TapeStep{});
}

void launchExitLoss(hipStream_t st, const StepLaunch& L, AgentHalf s, const uint32_t* assign, swarm_exit_rec* out, double* loss) {
    const int g = blocks(L.n, 256);
    if (L.batched) k_exit_loss<true><<<g, 256, 0, st>>>(L.n, L.s, L.bt, s, assign, out, loss);
    else k_exit_loss<false><<<g, 256, 0, st>>>(L.n, L.s, L.bt, s, assign, out, loss);
}

} // namespace swarm
