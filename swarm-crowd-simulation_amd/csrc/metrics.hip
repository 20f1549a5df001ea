// The safety-metrics passes (gfx950). metrics.h describes them; the arithmetic follows metrics_deposit.comp:76-98
// (cloud-in-cell in fixed point) and metrics_finalize.comp:38-57 (per-cell accumulators).
#include "metrics.h"

namespace swarm {

namespace {

__device__ inline int clampCell(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

__device__ inline void depositOne(uint32_t* mass, int cell, uint32_t w, float speedMps) {
    if (w == 0u) return;
    const float sm = float(w) * speedMps + 0.5f;   // rounded after the multiply and after the add (no contraction)
    atomicAdd(&mass[2 * cell], w);
    atomicAdd(&mass[2 * cell + 1], uint32_t(floorf(sm)));
}

// One agent's part of the deposit: its evacuation (if any) and its cloud-in-cell corners. Returns false for an agent
// that is not active in the output.
struct Corners {
    int x0, x1, y0, y1;        // clamped cell columns and rows (rows already times resX)
    uint32_t qx, qy;           // the position within the cell in 1/256
    float s;                   // speed in m/s
};

__device__ inline bool agentCorners(const MetricGrid& g, uint32_t i, const float* __restrict__ px, const float* __restrict__ py,
                                    const float* __restrict__ pw, const float* __restrict__ cx, const float* __restrict__ cy,
                                    const float* __restrict__ cw, const float* __restrict__ cvx, const float* __restrict__ cvy, int E,
                                    const float* __restrict__ exits, const MetricBuffers& mb, uint32_t step, Corners& k) {
    const float wNow = cw[i];
    if (pw[i] != 0.0f && wNow == 0.0f) {   // left in this step, through the exit at its last active position
        const float x = px[i], y = py[i];
        int slot = E;
        for (int e = 0; e < E; ++e) {
            const float* r = exits + 4 * e;
            if (x >= r[0] && x <= r[2] && y >= r[1] && y <= r[3]) { slot = e; break; }
        }
        atomicAdd(&mb.counters[4 + slot], 1u);
        atomicAdd(&mb.counters[2], 1u);
        mb.evacStep[i] = step;
    }
    if (wNow == 0.0f) return false;
    const float gx = cx[i] * g.cellsPerUnit + g.offX, gy = cy[i] * g.cellsPerUnit + g.offY;
    const float fx = floorf(gx), fy = floorf(gy);
    const float tx = gx - fx, ty = gy - fy;   // exact (Sterbenz)
    k.qx = uint32_t(floorf(tx * float(METRIC_SUBCELL) + 0.5f));
    k.qy = uint32_t(floorf(ty * float(METRIC_SUBCELL) + 0.5f));
    const int ix = int(fx), iy = int(fy);
    k.x0 = clampCell(ix, g.resX - 1); k.x1 = clampCell(ix + 1, g.resX - 1);
    k.y0 = clampCell(iy, g.resY - 1) * g.resX; k.y1 = clampCell(iy + 1, g.resY - 1) * g.resX;
    const float vx = cvx[i], vy = cvy[i];
    k.s = sqrtf(vx * vx + vy * vy) * g.speedPerUnit;
    return true;
}

// Each block takes METRIC_DEPOSIT_PER_THREAD x 256 consecutive slots. Slots follow the step's hash-cell order, so a
// block's agents stand in a few rows of the floor: when the span of cells they touch fits in LDS, the block sums its
// masses there (LDS atomics) and adds each touched cell to HBM once; otherwise each corner goes to HBM directly.
template <int A>
__global__ void __launch_bounds__(256) k_metrics_deposit(int n, MetricGrid g, const uint32_t* __restrict__ ids,
                                                         const float* __restrict__ px, const float* __restrict__ py,
                                                         const float* __restrict__ pw, const float* __restrict__ cx,
                                                         const float* __restrict__ cy, const float* __restrict__ cw,
                                                         const float* __restrict__ cvx, const float* __restrict__ cvy, int E,
                                                         const float* __restrict__ exits, MetricBuffers mb) {
    __shared__ uint32_t win[2 * METRIC_DEPOSIT_WINDOW];   // mass, speed mass of the block's cells lo .. lo + span - 1
    __shared__ int redLo[8], redHi[8];
    __shared__ uint32_t redAct[8];
    const uint32_t step = mb.counters[0] + 1u;
    const int base = int(blockIdx.x) * (256 * A) + int(threadIdx.x);
    Corners k[A];
    bool on[A];
    int lo = 0x7FFFFFFF, hi = -1;
    uint32_t act = 0u;
#pragma unroll
    for (int a = 0; a < A; ++a) {
        const int t = base + a * 256;
        on[a] = false;
        if (t < n) {
            const uint32_t i = ids ? ids[t] : uint32_t(t);
            on[a] = agentCorners(g, i, px, py, pw, cx, cy, cw, cvx, cvy, E, exits, mb, step, k[a]);
        }
        if (on[a]) {
            ++act;
            lo = min(lo, k[a].x0 + k[a].y0);
            hi = max(hi, k[a].x1 + k[a].y1);
        }
    }
    // the block's cell span and active count (wave reductions, then one slot per wave)
    for (int d = 32; d > 0; d >>= 1) {
        lo = min(lo, __shfl_xor(lo, d));
        hi = max(hi, __shfl_xor(hi, d));
        act += __shfl_xor(act, d);
    }
    const int wave = int(threadIdx.x) >> 6;
    if ((threadIdx.x & 63) == 0) { redLo[wave] = lo; redHi[wave] = hi; redAct[wave] = act; }
    __syncthreads();
    lo = min(min(redLo[0], redLo[1]), min(redLo[2], redLo[3]));
    hi = max(max(redHi[0], redHi[1]), max(redHi[2], redHi[3]));
    if (threadIdx.x == 0) {
        const uint32_t blockAct = redAct[0] + redAct[1] + redAct[2] + redAct[3];
        if (blockAct != 0u) atomicAdd(&mb.counters[3], blockAct);
    }
    if (hi < lo) return;   // no active agent in the block (uniform)
    const int span = hi - lo + 1;
    if (span > METRIC_DEPOSIT_WINDOW) {   // spread-out block: straight to HBM
#pragma unroll
        for (int a = 0; a < A; ++a) {
            if (!on[a]) continue;
            const Corners& c = k[a];
            const uint32_t rx = METRIC_SUBCELL - c.qx, ry = METRIC_SUBCELL - c.qy;
            depositOne(mb.mass, c.x0 + c.y0, rx * ry, c.s);
            depositOne(mb.mass, c.x1 + c.y0, c.qx * ry, c.s);
            depositOne(mb.mass, c.x0 + c.y1, rx * c.qy, c.s);
            depositOne(mb.mass, c.x1 + c.y1, c.qx * c.qy, c.s);
        }
        return;
    }
    for (int j = int(threadIdx.x); j < 2 * span; j += 256) win[j] = 0u;
    __syncthreads();
#pragma unroll
    for (int a = 0; a < A; ++a) {
        if (!on[a]) continue;
        const Corners& c = k[a];
        const uint32_t rx = METRIC_SUBCELL - c.qx, ry = METRIC_SUBCELL - c.qy;
        depositOne(win, c.x0 + c.y0 - lo, rx * ry, c.s);
        depositOne(win, c.x1 + c.y0 - lo, c.qx * ry, c.s);
        depositOne(win, c.x0 + c.y1 - lo, rx * c.qy, c.s);
        depositOne(win, c.x1 + c.y1 - lo, c.qx * c.qy, c.s);
    }
    __syncthreads();
    for (int j = int(threadIdx.x); j < span; j += 256) {
        const uint32_t m = win[2 * j];
        if (m == 0u) continue;   // a touched cell has mass: every deposited weight is non-zero
        atomicAdd(&mb.mass[2 * (lo + j)], m);
        atomicAdd(&mb.mass[2 * (lo + j) + 1], win[2 * j + 1]);
    }
}

__device__ inline void add64(uint32_t* v, uint32_t b) {   // (lo, hi) + b
    const uint32_t lo = v[0] + b;
    v[1] += lo < b ? 1u : 0u;
    v[0] = lo;
}

__global__ void __launch_bounds__(256) k_metrics_finalize(int cells, MetricBuffers mb) {
    const int c = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (c == 0) {
        mb.counters[1] = mb.counters[3];
        mb.counters[3] = 0u;
        mb.counters[0] += 1u;
    }
    if (c >= cells) return;
    const uint32_t m = mb.mass[2 * c], sm = mb.mass[2 * c + 1];
    mb.mass[2 * c] = 0u;
    mb.mass[2 * c + 1] = 0u;
    const uint4 thr = reinterpret_cast<const uint4*>(mb.thr)[c];
    uint32_t* s = mb.cells + size_t(METRIC_CELL_WORDS) * size_t(c);
    s[0] = m;
    s[1] = sm;
    s[2] = s[2] > m ? s[2] : m;
    s[3] += m != 0u ? 1u : 0u;
    s[4] += m >= thr.x ? 1u : 0u;
    s[5] += m >= thr.y ? 1u : 0u;
    s[6] += m >= thr.z ? 1u : 0u;
    s[7] += m >= thr.w ? 1u : 0u;
    add64(s + 8, m);
    add64(s + 10, sm);
    if (m >= thr.z) add64(s + 12, m);   // the crush threshold (index 2)
}

} // namespace

void launchMetricsDeposit(hipStream_t st, int n, const MetricGrid& g, const uint32_t* ids, const float* px, const float* py,
                          const float* pw, const float* cx, const float* cy, const float* cw, const float* cvx, const float* cvy, int E,
                          const float* exits, MetricBuffers mb) {
    if (n <= 0) return;
    constexpr int A = METRIC_DEPOSIT_PER_THREAD;
    k_metrics_deposit<A><<<(n + 256 * A - 1) / (256 * A), 256, 0, st>>>(n, g, ids, px, py, pw, cx, cy, cw, cvx, cvy, E, exits, mb);
}

void launchMetricsFinalize(hipStream_t st, const MetricGrid& g, MetricBuffers mb) {
    const int cells = g.resX * g.resY;
    k_metrics_finalize<<<(cells + 255) / 256, 256, 0, st>>>(cells, mb);
}

} // namespace swarm
