// The safety metrics of a step (metrics_deposit.comp and metrics_finalize.comp, swarm/metrics.hpp:1-33): write-only
// side outputs that read the step's input and output agent state and never write it.
//   deposit   per agent (taken in the step's slot order, so a block's agents stand in a few rows of the floor):
//             an agent active in the input and inactive in the output is an evacuation (its exit's counter, the
//             total, its evacuation step); every active output agent deposits one person onto the 4 floor cells
//             around it (cloud-in-cell, positions quantised to 1/256 of a cell, integer weights summing to 65536)
//             and the same weights times its speed in m/s, summed per block in LDS and added to each touched cell
//             once; the active count, one atomic per block
//   finalize  per floor cell: last, peak, occupied steps, steps at or above each threshold mass, the 64-bit sums of
//             mass and speed mass, exposure (mass while at or above the crush threshold); then the step's mass is
//             cleared and the counters published
// Every operation is an integer one (or the reference's float order before the integer conversion), so the results
// do not depend on the order of the atomics.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace swarm {

constexpr int METRIC_EXIT_SLOTS = 64;                        // the floor's exits, then 'none of them'
constexpr int METRIC_COUNTER_WORDS = 4 + METRIC_EXIT_SLOTS;  // step, active, evacuated, activeAccum, exits[]
constexpr int METRIC_CELL_WORDS = 16;                        // GpuMetricCell (gpu_types.hpp:127-137)
constexpr uint32_t METRIC_SUBCELL = 256u;                    // metrics.hpp:77
constexpr uint32_t METRIC_MASS_ONE = METRIC_SUBCELL * METRIC_SUBCELL;
constexpr uint32_t METRIC_NEVER = 0xFFFFFFFFu;
constexpr int METRIC_CRUSH_INDEX = 2;
constexpr int METRIC_DEPOSIT_PER_THREAD = 1;                 // slots per deposit thread (256 threads a block)
constexpr int METRIC_DEPOSIT_WINDOW = 4096;                  // cells a deposit block sums in LDS (2 words each)

struct MetricGrid {
    int resX, resY;
    float cellsPerUnit, offX, offY;   // g = xy * cellsPerUnit + off: the position in cells from cell 0's centre
    float speedPerUnit;               // m/s per model speed unit
};

struct MetricBuffers {
    uint32_t* counters;        // METRIC_COUNTER_WORDS
    uint32_t* evacStep;        // agent order: the metric step it left at (0: not yet)
    uint32_t* mass;            // per cell: mass, speed mass (zero between steps)
    uint32_t* cells;           // per cell METRIC_CELL_WORDS
    const uint32_t* thr;       // per cell 4 threshold masses
};

// the step's input (x, y, w) and output (x, y, w, vx, vy) in agent order; ids: the slot order's agents (nullptr:
// agent order); exits: E rectangles (lox, loy, hix, hiy)
void launchMetricsDeposit(hipStream_t st, int n, const MetricGrid& g, const uint32_t* ids, const float* px, const float* py,
                          const float* pw, const float* cx, const float* cy, const float* cw, const float* cvx, const float* cvy, int E,
                          const float* exits, MetricBuffers mb);
void launchMetricsFinalize(hipStream_t st, const MetricGrid& g, MetricBuffers mb);

} // namespace swarm
