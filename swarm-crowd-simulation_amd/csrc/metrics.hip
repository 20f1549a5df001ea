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

__global__ void __launch_bounds__(256) k_metrics_deposit(int n, MetricGrid g, const uint32_t* __restrict__ ids,
                                                         const float* __restrict__ px, const float* __restrict__ py,
                                                         const float* __restrict__ pw, const float* __restrict__ cx,
                                                         const float* __restrict__ cy, const float* __restrict__ cw,
                                                         const float* __restrict__ cvx, const float* __restrict__ cvy, int E,
                                                         const float* __restrict__ exits, MetricBuffers mb) {
    __shared__ uint32_t groupActive;
    if (threadIdx.x == 0) groupActive = 0u;
    __syncthreads();
    const int t = int(blockIdx.x * blockDim.x + threadIdx.x);
    const uint32_t step = mb.counters[0] + 1u;
    if (t < n) {
        const uint32_t i = ids ? ids[t] : uint32_t(t);
        if (pw[i] != 0.0f && cw[i] == 0.0f) {   // left in this step, through the exit at its last active position
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
        if (cw[i] != 0.0f) {
            atomicAdd(&groupActive, 1u);
            const float gx = cx[i] * g.cellsPerUnit + g.offX, gy = cy[i] * g.cellsPerUnit + g.offY;
            const float fx = floorf(gx), fy = floorf(gy);
            const float tx = gx - fx, ty = gy - fy;   // exact (Sterbenz)
            const uint32_t qx = uint32_t(floorf(tx * float(METRIC_SUBCELL) + 0.5f)), qy = uint32_t(floorf(ty * float(METRIC_SUBCELL) + 0.5f));
            const uint32_t rx = METRIC_SUBCELL - qx, ry = METRIC_SUBCELL - qy;
            const int ix = int(fx), iy = int(fy);
            const int x0 = clampCell(ix, g.resX - 1), x1 = clampCell(ix + 1, g.resX - 1);
            const int y0 = clampCell(iy, g.resY - 1) * g.resX, y1 = clampCell(iy + 1, g.resY - 1) * g.resX;
            const float vx = cvx[i], vy = cvy[i];
            const float s = sqrtf(vx * vx + vy * vy) * g.speedPerUnit;
            depositOne(mb.mass, x0 + y0, rx * ry, s);
            depositOne(mb.mass, x1 + y0, qx * ry, s);
            depositOne(mb.mass, x0 + y1, rx * qy, s);
            depositOne(mb.mass, x1 + y1, qx * qy, s);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && groupActive != 0u) atomicAdd(&mb.counters[3], groupActive);
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
    k_metrics_deposit<<<(n + 255) / 256, 256, 0, st>>>(n, g, ids, px, py, pw, cx, cy, cw, cvx, cvy, E, exits, mb);
}

void launchMetricsFinalize(hipStream_t st, const MetricGrid& g, MetricBuffers mb) {
    const int cells = g.resX * g.resY;
    k_metrics_finalize<<<(cells + 255) / 256, 256, 0, st>>>(cells, mb);
}

} // namespace swarm
