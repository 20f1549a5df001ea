// The per-step pipeline kernels for gfx950 (one scene, N agents).
//
//   hash      calc_grid.comp: agent -> cell key, a per-bucket atomic count (cells, then one inactive bucket
//             per 256 agents)
//   scan      two passes over the bucket counts (tile sums, then tile-local exclusive scans): bucket starts,
//             the cell table, and the counts cleared for the next step
//   scatter   agent -> its bucket at start + atomic offset
//   fixgather each bucket's run sorted by agent id (runs are a few agents), so the permutation is the
//             stable counting sort's: ascending (cell, agent), inactive agents at the tail in agent order
//             (the reference's SortMode::Counting, cpu_crowd_sim.cpp:997-1006); the state is gathered into
//             slot order (reorder.comp)
//   choice    PHYSICS_PASS 0, the queue-aware exit choice (exit_choice.cpp:683-791): prefix, evaluate,
//             thresholds, apply (+ every agent's e_goal)
//   predict   PHYSICS_PASS 1 (cpu_crowd_sim.cpp:655-777)
//   project   PHYSICS_PASS 2, one Jacobi iteration (cpu_crowd_sim.cpp:781-831)
//   final     PHYSICS_PASS 3 (cpu_crowd_sim.cpp:835-866) + evacuation events
//
// Data is SoA FP32. Kernels are memory-latency bound gathers and integer atomics: no MFMA.
#include "crowd_kernels.h"

namespace swarm {

// base[key] += the number of lanes of this wave that pass 'valid' with this key: one atomic per distinct key of
// the wave instead of one per lane (the counts are integer sums, so the result does not depend on the order).
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

__device__ inline uint32_t bucketOf(uint32_t key, uint32_t agent, uint32_t C) {
    return key == INACTIVE_CELL ? C + agent / uint32_t(HASH_BLOCK) : key;
}

__global__ void __launch_bounds__(HASH_BLOCK) k_hash(int n, SceneView s, float* __restrict__ ax, float* __restrict__ ay,
                                                     float* __restrict__ aw, uint32_t* __restrict__ key,
                                                     uint32_t* __restrict__ loc, uint32_t* __restrict__ bcnt, uint32_t C) {
    const int i = blockIdx.x * HASH_BLOCK + threadIdx.x;
    if (i >= n) return;
    float x = ax[i], y = ay[i], w = aw[i];
    if (isnan(x)) {   // calc_grid.comp resets a NaN position in the input state
        x = 0.0f; y = 0.0f; w = 1.0f;
        ax[i] = x; ay[i] = y; aw[i] = w;
    }
    const uint32_t k = w == 0.0f ? INACTIVE_CELL : hashCell(s, x, y);
    key[i] = k;
    loc[i] = atomicAdd(&bcnt[bucketOf(k, uint32_t(i), C)], 1u);
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

// exclusive scan of the bucket counts: bucket starts, the cell table (start 0 for an empty cell, as the
// reference's cleared table), the active count, and the counts cleared for the next step
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_apply(uint32_t* __restrict__ bcnt, uint32_t nb, uint32_t C,
                                                          const uint32_t* __restrict__ tileSum,
                                                          uint32_t* __restrict__ bstart, uint32_t* __restrict__ cellStart,
                                                          uint32_t* __restrict__ cellCount, uint32_t* __restrict__ ctrl) {
    __shared__ uint32_t wsum[SCAN_BLOCK / 64];
    __shared__ uint32_t tilePrefix;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // prefix of the preceding tiles
    uint32_t pre = 0;
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
    // wave-inclusive scan of the per-thread sums
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
            if (j < C) {
                cellStart[j] = v[k] ? run : 0u;
                cellCount[j] = v[k];
            }
            if (j == C) ctrl[CTRL_NACTIVE] = run;
            bcnt[j] = 0u;
        }
        run += v[k];
    }
}

__global__ void __launch_bounds__(HASH_BLOCK) k_scatter(int n, const uint32_t* __restrict__ key, const uint32_t* __restrict__ loc,
                                                        const uint32_t* __restrict__ bstart, uint32_t* __restrict__ tmp, uint32_t C) {
    const int i = blockIdx.x * HASH_BLOCK + threadIdx.x;
    if (i >= n) return;
    tmp[bstart[bucketOf(key[i], uint32_t(i), C)] + loc[i]] = uint32_t(i);
}

__global__ void __launch_bounds__(HASH_BLOCK) k_fix_gather(int n, const uint32_t* __restrict__ key, const uint32_t* __restrict__ tmp,
                                                           const uint32_t* __restrict__ bstart, uint32_t C,
                                                           uint32_t* __restrict__ sid, uint32_t* __restrict__ skey,
                                                           const float* __restrict__ ax, const float* __restrict__ ay,
                                                           const float* __restrict__ rad, float* __restrict__ sx,
                                                           float* __restrict__ sy, float* __restrict__ sr) {
    const int s = blockIdx.x * HASH_BLOCK + threadIdx.x;
    if (s >= n) return;
    const uint32_t id = tmp[s];
    const uint32_t k = key[id];
    const uint32_t b = bucketOf(k, id, C);
    const uint32_t st = bstart[b], en = bstart[b + 1];
    uint32_t rank = 0;
    for (uint32_t t = st; t < en; ++t) rank += tmp[t] < id ? 1u : 0u;
    const uint32_t f = st + rank;
    sid[f] = id;
    skey[f] = k;
    if (k != INACTIVE_CELL) {
        sx[f] = ax[id];
        sy[f] = ay[id];
        sr[f] = rad[id];
    }
}

// ---------------- exit choice (PHYSICS_PASS 0) ----------------
__device__ inline bool evalStep(uint32_t step) { return step > 0u && step % uint32_t(CHOICE_STEPS) == 0u; }

// prefix stage: the read parity's queue-ahead histograms to exclusive prefix sums, the write parity cleared
__global__ void __launch_bounds__(1024) k_choice_prefix(int E, uint32_t* __restrict__ ahead, uint32_t* __restrict__ queue,
                                                        const uint32_t* __restrict__ ctrl) {
    const uint32_t step = ctrl[CTRL_STEP];
    const uint32_t par = step & 1u;
    uint32_t* aRead = ahead + size_t(par) * CHOICE_QUEUE_STRIDE * CHOICE_AHEAD_BINS;
    uint32_t* aWrite = ahead + size_t(1u - par) * CHOICE_QUEUE_STRIDE * CHOICE_AHEAD_BINS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int PER = CHOICE_AHEAD_BINS / 64;
    for (int e = wave; e < E; e += 16) {
        uint32_t* a = aRead + size_t(e) * CHOICE_AHEAD_BINS + lane * PER;
        uint32_t v[PER];
        uint32_t local = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) { v[k] = a[k]; local += v[k]; }
        uint32_t incl = local;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off, 64);
            if (lane >= off) incl += o;
        }
        uint32_t run = incl - local;
#pragma unroll
        for (int k = 0; k < PER; ++k) { a[k] = run; run += v[k]; }
        uint32_t* wr = aWrite + size_t(e) * CHOICE_AHEAD_BINS + lane * PER;
#pragma unroll
        for (int k = 0; k < PER; ++k) wr[k] = 0u;
    }
    const uint32_t qWrite = uint32_t(CHOICE_QUEUE_STRIDE) - par * uint32_t(CHOICE_QUEUE_STRIDE);
    if (threadIdx.x < CHOICE_QUEUE_STRIDE) queue[qWrite + threadIdx.x] = 0u;
}

// evaluate stage (evaluation steps only), per agent
__global__ void __launch_bounds__(HASH_BLOCK) k_choice_eval(int n, SceneView s, const float* __restrict__ ax,
                                                            const float* __restrict__ ay, const float* __restrict__ aw,
                                                            const uint32_t* __restrict__ assign, const uint32_t* __restrict__ ahead,
                                                            uint32_t* __restrict__ pending, uint32_t* __restrict__ hist,
                                                            uint32_t* __restrict__ counters, const uint32_t* __restrict__ ctrl) {
    const uint32_t step = ctrl[CTRL_STEP];
    if (!evalStep(step)) return;
    const int i = blockIdx.x * HASH_BLOCK + threadIdx.x;
    bool evaluated = false, eligible = false;
    uint32_t hkey = 0;
    if (i < n) {
        pending[i] = 0u;
        if (aw[i] != 0.0f) {
            const uint32_t* aRead = ahead + size_t(step & 1u) * CHOICE_QUEUE_STRIDE * CHOICE_AHEAD_BINS;
            const V2 p{ ax[i], ay[i] };
            const int c = choiceCell(s, p);
            const uint32_t* sl = s.navSlot + size_t(c) * size_t(s.K);
            const uint16_t* dq = s.navDist + size_t(c) * size_t(s.K);
            const uint32_t current = assign[i];
            const int curK = slotExitOf(sl[0]) == CHOICE_EMPTY ? -1 : slotOfExit(s, c, current);
            if (curK >= 0) {
                evaluated = true;
                float eBest;
                const int best = bestCandidate(s, c, aRead, eBest);
                const uint32_t qc = dq[curK];
                const float eCur = estimatedTime(s, slotDistOf(s, qc), current, aRead[size_t(current) * CHOICE_AHEAD_BINS + aheadBin(qc)]);
                const float gain = eCur - eBest;
                const float m = CHOICE_REL_MARGIN * eCur;
                if (best != curK && gain > (m < CHOICE_ABS_MARGIN_S ? CHOICE_ABS_MARGIN_S : m)) {
                    const uint32_t g = uint32_t(gain);
                    const uint32_t bin = g < uint32_t(CHOICE_GAIN_BINS - 1) ? g : uint32_t(CHOICE_GAIN_BINS - 1);
                    pending[i] = 0x80000000u | (slotExitOf(sl[best]) << 16) | bin;
                    eligible = true;
                    hkey = current * uint32_t(CHOICE_GAIN_BINS) + bin;
                }
            }
        }
    }
    waveCountAdd(hist, hkey, eligible);
    waveCountAdd(counters, uint32_t(CNT_EVALUATIONS), evaluated);
    waveCountAdd(counters, uint32_t(CNT_ELIGIBLE), eligible);
}

// thresholds stage (evaluation steps only), one thread per exit
__global__ void k_choice_thresh(int E, uint32_t* __restrict__ hist, const uint32_t* __restrict__ queue,
                                uint32_t* __restrict__ threshold, const uint32_t* __restrict__ ctrl) {
    const uint32_t step = ctrl[CTRL_STEP];
    if (!evalStep(step)) return;
    const int e = threadIdx.x;
    if (e >= E) return;
    const uint32_t qRead = (step & 1u) * uint32_t(CHOICE_QUEUE_STRIDE);
    uint32_t* h = hist + size_t(e) * CHOICE_GAIN_BINS;
    const uint32_t queued = queue[qRead + uint32_t(e)];
    const uint32_t share = uint32_t(float(queued) * CHOICE_SWITCH_SHARE);
    const uint32_t cap = share > CHOICE_MIN_SWITCHES ? share : CHOICE_MIN_SWITCHES;
    uint32_t cum = 0, thr = uint32_t(CHOICE_GAIN_BINS);
    for (int b = CHOICE_GAIN_BINS - 1; b >= 0; --b) {
        const uint32_t cnt = h[b];
        if (cnt == 0) continue;
        if (cum + cnt > cap) {
            if (cum == 0) thr = uint32_t(b);
            break;
        }
        cum += cnt;
        thr = uint32_t(b);
    }
    threshold[e] = thr;
    for (int b = 0; b < CHOICE_GAIN_BINS; ++b) h[b] = 0u;
}

// apply stage (every step), per slot: the agent's exit for this step, its counts for the next, and its e_goal
__global__ void __launch_bounds__(HASH_BLOCK) k_choice_apply(int n, SceneView s, const uint32_t* __restrict__ skey,
                                                             const uint32_t* __restrict__ sid, const float* __restrict__ sx,
                                                             const float* __restrict__ sy, uint32_t* __restrict__ assign,
                                                             uint32_t* __restrict__ ahead, const uint32_t* __restrict__ pending,
                                                             const uint32_t* __restrict__ threshold, uint32_t* __restrict__ queue,
                                                             uint32_t* __restrict__ counters, float* __restrict__ egx,
                                                             float* __restrict__ egy, const uint32_t* __restrict__ ctrl) {
    const int slot = blockIdx.x * HASH_BLOCK + threadIdx.x;
    const bool live = slot < n && skey[slot] != INACTIVE_CELL;
    const uint32_t step = ctrl[CTRL_STEP];
    const bool ev = evalStep(step);
    const uint32_t par = step & 1u;
    const uint32_t* aRead = ahead + size_t(par) * CHOICE_QUEUE_STRIDE * CHOICE_AHEAD_BINS;
    uint32_t* aWrite = ahead + size_t(1u - par) * CHOICE_QUEUE_STRIDE * CHOICE_AHEAD_BINS;
    const uint32_t qWrite = uint32_t(CHOICE_QUEUE_STRIDE) - par * uint32_t(CHOICE_QUEUE_STRIDE);
    bool forced = false, switched = false, counted = false, binned = false;
    uint32_t ex = CHOICE_NONE, bin = 0;
    if (live) {
        const uint32_t id = sid[slot];
        const V2 p{ sx[slot], sy[slot] };
        const int c = choiceCell(s, p);
        const uint32_t* sl = s.navSlot + size_t(c) * size_t(s.K);
        const uint32_t current = assign[id];
        ex = current;
        const bool hasPath = slotExitOf(sl[0]) != CHOICE_EMPTY;
        const int curK = hasPath ? slotOfExit(s, c, current) : -1;
        if (hasPath && curK < 0) {
            float eBest;
            ex = slotExitOf(sl[bestCandidate(s, c, aRead, eBest)]);
            forced = current != CHOICE_NONE;
        } else if (ev && curK >= 0) {
            const uint32_t pd = pending[id];
            if ((pd & 0x80000000u) && (pd & 0xFFFFu) >= threshold[current]) {
                ex = (pd >> 16) & 63u;
                switched = true;
            }
        }
        assign[id] = ex;
        counted = ex != CHOICE_NONE;
        const int k = slotOfExit(s, c, ex);
        if (k >= 0) {
            binned = true;
            bin = ex * uint32_t(CHOICE_AHEAD_BINS) + aheadBin(s.navDist[size_t(c) * size_t(s.K) + size_t(k)]);
        }
        const V2 e = choiceDirection(s, c, p, ex);
        egx[slot] = e.x;
        egy[slot] = e.y;
    }
    // the counts of the next step: one atomic per distinct key in the wave (integer sums: order-free)
    waveCountAdd(queue + qWrite, ex, counted);
    waveCountAdd(aWrite, bin, binned);
    waveCountAdd(counters, uint32_t(CNT_FORCED), forced);
    waveCountAdd(counters, uint32_t(CNT_SWITCHES), switched);
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
__global__ void __launch_bounds__(PHYS_BLOCK) k_predict(int n, SceneView s, ModelParams mp, StepBuffers b, ProjBuffers out) {
    const int slot = blockIdx.x * PHYS_BLOCK + threadIdx.x;
    if (slot >= n) return;
    const uint32_t cid = b.skey[slot];
    if (cid == INACTIVE_CELL) return;
    const uint32_t self = b.sid[slot];
    const V2 p{ b.sx[slot], b.sy[slot] };
    const float ri = b.sr[slot];
    if (parkTest(s, p)) {
        out.x[slot] = PARK_COORD; out.y[slot] = PARK_COORD; out.acc[slot] = 0.0f;
        b.pw[slot] = 0.0f; b.pvx[slot] = 0.0f; b.pvy[slot] = 0.0f;
        return;
    }
    const V2 v{ b.avx[self], b.avy[self] };
    const V2 eGoal{ b.egx[slot], b.egy[slot] };
    const int cx = int(cid % uint32_t(s.hashNX)), cy = int(cid / uint32_t(s.hashNX));
    V2 sep{ 0.0f, 0.0f };
    float sAhead = 1e9f;
    const int R = s.percRing;
    for (int yy = -R; yy <= R; yy++) {
        const int ny = cy + yy;
        if (ny < 0 || ny >= s.hashNY) continue;
        for (int xx = -R; xx <= R; xx++) {
            const int nx = cx + xx;
            if (nx < 0 || nx >= s.hashNX) continue;
            const int c = nx + ny * s.hashNX;
            const uint32_t st = b.cellStart[c], en = st + b.cellCount[c];
            for (uint32_t k = st; k < en; k++) {
                if (k == uint32_t(slot)) continue;
                const V2 offset = p - V2{ b.sx[k], b.sy[k] };
                const float dxy2 = dot2(offset, offset);
                if (dxy2 < R_PERC2) {
                    const uint32_t other = b.sid[k];
                    const float D = ri + b.sr[k];
                    sep = sep + separationForce(offset, self, other, D + SEP_MARGIN);
                    const float ahead = -dot2(offset, eGoal);
                    const float lateral = fabsf(offset.x * eGoal.y - offset.y * eGoal.x);
                    const float dxy = sqrtf(dxy2);
                    const float gap = dxy - D;
                    if (ahead > 0.0f && lateral < D && gap < sAhead) {
                        const V2 eOther{ b.egx[k], b.egy[k] };
                        const float aheadOther = dot2(offset, eOther);
                        const bool mutual = aheadOther > 0.0f && fabsf(offset.x * eOther.y - offset.y * eOther.x) < D;
                        if (!mutual || ahead > aheadOther || (ahead == aheadOther && self > other)) sAhead = gap;
                    }
                }
            }
        }
    }
    const float vFree = stairFactor(s, p) * b.v0[self];
    float vc = sAhead * HEADWAY_RATE;
    vc = vc < 0.0f ? 0.0f : (vFree < vc ? vFree : vc);
    const float vDes = vc * len2(eGoal);
    const V2 eDir = safeNormalize(eGoal + sep);

    V2 aObs{ 0.0f, 0.0f };
    const float dTorso = mapSdf(s, p);
    if (dTorso < OBSTACLE_RANGE) {
        const V2 norm = safeNormalize(calcGradient(s, p));
        if (len2(norm) > OBSTACLE_MIN_NORMAL) aObs = safeNormalize(norm) * (OBSTACLE_RANGE - dTorso) * W_OBS * mp.obstacleGain;
    }
    const V2 vEq = eDir * vDes + aObs * mp.tau;
    V2 nv = vEq + (v - vEq) * mp.relaxStep;
    const float along = dot2(nv, eGoal);
    if (along > vDes) nv = nv - eGoal * (along - vDes);
    const float vmax = b.v0[self];
    if (len2(nv) > vmax) nv = safeNormalize(nv) * vmax;
    V2 np = p + nv * DT;

    float wall[3];
    const bool grounded = resolveStatic(s, p, np, nv, ri, wall);
    if (wall[0] == 0.0f && wall[1] == 0.0f && !wallFree(s, np)) wallPlaneProbe(s, np, ri, wall);
    out.x[slot] = np.x; out.y[slot] = np.y; out.acc[slot] = 0.0f;
    b.pw[slot] = 1.0f; b.pvx[slot] = nv.x; b.pvy[slot] = nv.y;
    const V2 wn{ wall[0], wall[1] };
    b.planes[slot] = make_float4(wn.x, wn.y, dot2(np, wn) - wall[2], (nonzero(wn) ? 1.0f : 0.0f) + (grounded ? 2.0f : 0.0f));
}

// ---------------- the projection's contact candidates ----------------
// Every projection iteration moves an agent by at most PROJ_MAX_SHIFT_RADII r (the capped correction; the wall
// plane is a projection onto a half-space holding the predicted position, which moves it no farther), so over the
// step's P iterations a pair can touch only if its predicted distance is below D + P (shift_i + shift_j). The
// candidates are the ring's slots that pass this bound (with 0.1% of float slack), in ring order: summing the
// contacts over them gives the ring scan's sums bit for bit, at a fraction of its loads. Parked slots never
// count (the ring scan skips them too).
__global__ void __launch_bounds__(PHYS_BLOCK) k_contact_list(int n, SceneView s, ModelParams mp, StepBuffers b, ProjBuffers pred) {
    const int slot = blockIdx.x * PHYS_BLOCK + threadIdx.x;
    if (slot >= n) return;
    const uint32_t cid = b.skey[slot];
    if (cid == INACTIVE_CELL || b.pw[slot] == 0.0f) return;
    const V2 p{ pred.x[slot], pred.y[slot] };
    const float ri = b.sr[slot];
    const float P = float(mp.projIters);
    const int cx = int(cid % uint32_t(s.hashNX)), cy = int(cid / uint32_t(s.hashNX));
    const int R = s.projRing;
    const size_t stride = size_t(n);
    uint32_t cnt = 0;
    for (int yy = -R; yy <= R; yy++) {
        const int ny = cy + yy;
        if (ny < 0 || ny >= s.hashNY) continue;
        for (int xx = -R; xx <= R; xx++) {
            const int nx = cx + xx;
            if (nx < 0 || nx >= s.hashNX) continue;
            const int c = nx + ny * s.hashNX;
            const uint32_t st = b.cellStart[c], en = st + b.cellCount[c];
            for (uint32_t k = st; k < en; k++) {
                if (k == uint32_t(slot) || b.pw[k] == 0.0f) continue;
                const float rk = b.sr[k];
                const float reach = ((ri + rk) + P * (PROJ_MAX_SHIFT_RADII * (ri + rk))) * 1.001f + 1e-6f;
                const V2 d = p - V2{ pred.x[k], pred.y[k] };
                if (dot2(d, d) < reach * reach) {
                    if (cnt < uint32_t(NBR_CAP)) b.nbr[size_t(cnt) * stride + size_t(slot)] = k;
                    ++cnt;
                }
            }
        }
    }
    b.nbrCount[slot] = cnt <= uint32_t(NBR_CAP) ? cnt : NBR_OVERFLOW;
}

// ---------------- PHYSICS_PASS 2: one Jacobi iteration ----------------
// projectSlot (cpu_crowd_sim.cpp:781-822); returns the slot's new position and displacement accumulator
__device__ inline void projectOne(const SceneView& s, const StepBuffers& b, const ProjBuffers& in, int n, int slot, uint32_t cid,
                                  float& ox, float& oy, float& oacc) {
    const float bx = in.x[slot], by = in.y[slot];
    ox = bx; oy = by; oacc = in.acc[slot];
    if (b.pw[slot] == 0.0f) return;
    const uint32_t self = b.sid[slot];
    const float ri = b.sr[slot];
    V2 corr{ 0.0f, 0.0f };
    int nContacts = 0;
    const uint32_t cnt = b.nbrCount[slot];
    if (cnt != NBR_OVERFLOW) {
        // the candidates in ring order: the same contacts summed in the same order as the ring scan below
        const size_t stride = size_t(n);
        for (uint32_t j = 0; j < cnt; ++j) {
            const uint32_t k = b.nbr[size_t(j) * stride + size_t(slot)];
            const V2 offset = V2{ bx, by } - V2{ in.x[k], in.y[k] };
            const float D = ri + b.sr[k];
            const float dxy2 = dot2(offset, offset);
            if (dxy2 < contact2(D)) {   // contactCorrection, inlined: the pair's id only for a stacked pair
                const float dxy = sqrtf(dxy2);
                const float shift = 0.5f * (D - dxy);
                corr = corr + (dxy2 <= SEP_STACKED_XY2 ? shift * pairDirection(self, b.sid[k]) : shift * (offset / dxy));
                ++nContacts;
            }
        }
    } else {
        const int cx = int(cid % uint32_t(s.hashNX)), cy = int(cid / uint32_t(s.hashNX));
        const int R = s.projRing;
        for (int yy = -R; yy <= R; yy++) {
            const int ny = cy + yy;
            if (ny < 0 || ny >= s.hashNY) continue;
            for (int xx = -R; xx <= R; xx++) {
                const int nx = cx + xx;
                if (nx < 0 || nx >= s.hashNX) continue;
                const int c = nx + ny * s.hashNX;
                const uint32_t st = b.cellStart[c], en = st + b.cellCount[c];
                for (uint32_t k = st; k < en; k++) {
                    if (k == uint32_t(slot)) continue;
                    const V2 offset = V2{ bx, by } - V2{ in.x[k], in.y[k] };
                    const float D = ri + b.sr[k];
                    if (b.pw[k] != 0.0f && dot2(offset, offset) < contact2(D)) {
                        corr = corr + contactCorrection(offset, self, b.sid[k], D);
                        ++nContacts;
                    }
                }
            }
        }
    }
    corr = corr * (PROJ_RELAX / float(nContacts > 1 ? nContacts : 1));
    const float maxShift = PROJ_MAX_SHIFT_RADII * ri;
    const float len = len2(corr);
    if (len > maxShift) corr = corr * (maxShift / len);
    V2 q = V2{ bx, by } + corr;
    const float4 pl = b.planes[slot];
    if (int(pl.w) & 1) {
        const float sd = pl.z - dot2(q, V2{ pl.x, pl.y });
        if (sd > 0.0f) q = q + V2{ pl.x, pl.y } * sd;
    }
    const V2 d = q - V2{ bx, by };
    ox = q.x; oy = q.y;
    oacc = in.acc[slot] + len2(d);
}

__global__ void __launch_bounds__(PHYS_BLOCK) k_project(int n, SceneView s, StepBuffers b, ProjBuffers in, ProjBuffers out) {
    const int slot = blockIdx.x * PHYS_BLOCK + threadIdx.x;
    if (slot >= n) return;
    const uint32_t cid = b.skey[slot];
    if (cid == INACTIVE_CELL) return;
    float x, y, acc;
    projectOne(s, b, in, n, slot, cid, x, y, acc);
    out.x[slot] = x; out.y[slot] = y; out.acc[slot] = acc;
}

// ---------------- PHYSICS_PASS 3: final ----------------
__global__ void __launch_bounds__(PHYS_BLOCK) k_final(int n, SceneView s, ModelParams mp, StepBuffers b, ProjBuffers in,
                                                      AgentHalf rd, AgentHalf wr, EvacBuffers ev, uint32_t* __restrict__ ctrl) {
    const int slot = blockIdx.x * PHYS_BLOCK + threadIdx.x;
    const uint32_t step = ctrl[CTRL_STEP];
    if (slot < n) {
        const uint32_t cid = b.skey[slot];
        const uint32_t self = b.sid[slot];
        if (cid == INACTIVE_CELL) {
            wr.x[self] = rd.x[self]; wr.y[self] = rd.y[self]; wr.w[self] = rd.w[self];
            wr.vx[self] = rd.vx[self]; wr.vy[self] = rd.vy[self]; wr.spd[self] = rd.spd[self];
        } else {
            float x, y, acc;
            if (mp.projIters > 0) projectOne(s, b, in, n, slot, cid, x, y, acc);
            else { x = in.x[slot]; y = in.y[slot]; acc = in.acc[slot]; }
            if (b.pw[slot] == 0.0f) {   // parked this step: an evacuation event
                wr.x[self] = x; wr.y[self] = y; wr.w[self] = 0.0f;
                wr.vx[self] = 0.0f; wr.vy[self] = 0.0f; wr.spd[self] = acc;
                const int e = exitOf(s, V2{ b.sx[slot], b.sy[slot] });
                const int ex = e >= 0 ? e : s.E;
                ev.step[self] = int(step);
                ev.exit[self] = ex;
                atomicAdd(&ev.perExit[ex], 1u);
            } else {
                V2 np{ x, y }, nv{ b.pvx[slot], b.pvy[slot] };
                if (acc > 0.0f) {
                    float wall[3];
                    resolveStatic(s, V2{ b.sx[slot], b.sy[slot] }, np, nv, b.sr[slot], wall);
                }
                wr.x[self] = np.x; wr.y[self] = np.y; wr.w[self] = 1.0f;
                wr.vx[self] = nv.x; wr.vy[self] = nv.y; wr.spd[self] = len2(nv);
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {   // every kernel of this step has read the step counter
        ctrl[CTRL_STEP] = step + 1u;
        *reinterpret_cast<unsigned long long*>(ctrl + CTRL_ACTIVE_SUM_LO) += ctrl[CTRL_NACTIVE];
    }
}

// a plain streaming copy of known size: the HBM traffic counters' calibration and the measured copy bandwidth
__global__ void __launch_bounds__(256) k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst, size_t n4) {
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n4; i += size_t(gridDim.x) * 256) dst[i] = src[i];
}

void launchCopyProbe(hipStream_t st, const void* src, void* dst, size_t bytes) {
    const size_t n4 = bytes / 16;
    k_copy_probe<<<4096, 256, 0, st>>>(static_cast<const float4*>(src), static_cast<float4*>(dst), n4);
}

// the step counter of an exit-choice pass run on its own (CrowdSim::stepExitChoice)
__global__ void k_step_increment(uint32_t* __restrict__ ctrl) {
    if (threadIdx.x == 0) ctrl[CTRL_STEP] += 1u;
}

// ---------------- host-side launchers ----------------
static inline int blocks(int n, int b) { return (n + b - 1) / b; }

void launchStepIncrement(hipStream_t st, uint32_t* ctrl) { k_step_increment<<<1, 64, 0, st>>>(ctrl); }

void launchGridPasses(hipStream_t st, int n, const SceneView& s, AgentHalf rd, const GridBuffers& g, StepBuffers& b,
                      const float* rad, uint32_t* ctrl) {
    const uint32_t C = uint32_t(s.hashNX) * uint32_t(s.hashNY);
    const uint32_t nb = C + uint32_t(blocks(n, HASH_BLOCK));
    const int tiles = blocks(int(nb), SCAN_TILE);
    k_hash<<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, s, rd.x, rd.y, rd.w, g.key, g.loc, g.bcnt, C);
    k_scan_reduce<<<tiles, SCAN_BLOCK, 0, st>>>(g.bcnt, nb, g.tileSum);
    k_scan_apply<<<tiles, SCAN_BLOCK, 0, st>>>(g.bcnt, nb, C, g.tileSum, g.bstart, b.cellStart, b.cellCount, ctrl);
    k_scatter<<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, g.key, g.loc, g.bstart, g.tmp, C);
    k_fix_gather<<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, g.key, g.tmp, g.bstart, C, b.sid, b.skey, rd.x, rd.y, rad,
                                                                b.sx, b.sy, b.sr);
}

void launchChoice(hipStream_t st, int n, const SceneView& s, AgentHalf rd, const ChoiceBuffers& c, StepBuffers& b,
                  const float* driveX, const float* driveY, uint32_t* ctrl) {
    if (driveX || s.E == 0) {
        k_goal_drive<<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, b.skey, b.sid, driveX, driveY, b.egx, b.egy);
        return;
    }
    k_choice_prefix<<<1, 1024, 0, st>>>(s.E, c.ahead, c.queue, ctrl);
    k_choice_eval<<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, s, rd.x, rd.y, rd.w, c.assign, c.ahead, c.pending, c.hist,
                                                                 c.counters, ctrl);
    k_choice_thresh<<<1, 64, 0, st>>>(s.E, c.hist, c.queue, c.threshold, ctrl);
    k_choice_apply<<<blocks(n, HASH_BLOCK), HASH_BLOCK, 0, st>>>(n, s, b.skey, b.sid, b.sx, b.sy, c.assign, c.ahead, c.pending,
                                                                  c.threshold, c.queue, c.counters, b.egx, b.egy, ctrl);
}

void launchPhysics(hipStream_t st, int n, const SceneView& s, const ModelParams& mp, StepBuffers& b, ProjBuffers pa,
                   ProjBuffers pb, AgentHalf rd, AgentHalf wr, EvacBuffers ev, uint32_t* ctrl, hipEvent_t* stamps) {
    const int g = blocks(n, PHYS_BLOCK);
    k_predict<<<g, PHYS_BLOCK, 0, st>>>(n, s, mp, b, pa);
    if (mp.projIters > 0) k_contact_list<<<g, PHYS_BLOCK, 0, st>>>(n, s, mp, b, pa);
    if (stamps) (void)hipEventRecord(stamps[0], st);
    ProjBuffers in = pa, out = pb;
    for (int it = 1; it < mp.projIters; ++it) {
        k_project<<<g, PHYS_BLOCK, 0, st>>>(n, s, b, in, out);
        const ProjBuffers t = in; in = out; out = t;
    }
    if (stamps) (void)hipEventRecord(stamps[1], st);
    k_final<<<g, PHYS_BLOCK, 0, st>>>(n, s, mp, b, in, rd, wr, ev, ctrl);
}

} // namespace swarm
