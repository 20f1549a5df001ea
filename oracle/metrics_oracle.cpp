// TEST INFRASTRUCTURE ONLY: the oracle of the safety-metrics passes, a CPU restatement of the reference's CPU
// reference of metrics_deposit.comp and metrics_finalize.comp (src/metrics.cpp):
//   cicDeposit           src/metrics.cpp:250-261   cloud-in-cell weights in fixed point, the shader's float order
//   speedMass            src/metrics.cpp:263-266   floor(float(weight) * speed + 0.5)
//   horizontalSpeedMps   src/metrics.cpp:268       |v_xy| * SPEED_M_S_PER_UNIT
//   cellInfo thresholds  src/metrics.cpp:232-246   ceil(threshold * walkable * MASS_ONE), clamped to [1, NEVER]
//   MetricsReference     src/metrics.cpp:429-480   deposit (evacuations, active count, mass, speed mass), then the
//                                                  per-cell finalize (last, peak, occupied, steps above, 64-bit sums,
//                                                  exposure)
// Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline load it; the product never does.
//
// Generalisations of this build's floors (documented at the call sites): the grid has resX x resY cells with
// g = xy * cellsPerUnit + off per axis (the reference's centred square grid: res x res, off = res / 2 - 0.5), and
// an evacuated agent's exit slot is the floor's exit rectangle holding its last active position (E: none of them),
// where the reference maps the terrace's avenue gaps (metrics.hpp:98-106).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

constexpr uint32_t SUBCELL = 256u;                 // metrics.hpp:77
constexpr uint32_t MASS_ONE = SUBCELL * SUBCELL;   // metrics.hpp:78
constexpr uint32_t NEVER = 0xFFFFFFFFu;            // metrics.hpp:96
constexpr double THRESHOLD_P_M2[3] = { 1.08, 2.17, 5.0 };   // metrics.hpp:82-94: LoS E, LoS F, crush
constexpr int CRUSH_INDEX = 2;
constexpr int EXIT_SLOTS = 64;                     // counters: step, active, evacuated, activeAccum, exits[64]
constexpr int COUNTER_WORDS = 4 + EXIT_SLOTS;
constexpr int CELL_WORDS = 16;                     // gpu_types.hpp:127-137 (GpuMetricCell)

struct Grid {
    int resX, resY;
    float cellsPerUnit, offX, offY;
};

struct Cic {
    int cell[4];
    uint32_t weight[4];
};

// src/metrics.cpp:250-261: multiply, round, add, round; the position quantized to 1/SUBCELL of a cell
Cic cicDeposit(const Grid& g, float x, float y) {
    const float gx = x * g.cellsPerUnit + g.offX, gy = y * g.cellsPerUnit + g.offY;
    const float fx = std::floor(gx), fy = std::floor(gy);
    const float tx = gx - fx, ty = gy - fy;
    const uint32_t qx = uint32_t(std::floor(tx * float(SUBCELL) + 0.5f)), qy = uint32_t(std::floor(ty * float(SUBCELL) + 0.5f));
    const uint32_t rx = SUBCELL - qx, ry = SUBCELL - qy;
    const int ix = int(fx), iy = int(fy);
    auto idx = [&](int cx, int cy) { return std::clamp(cx, 0, g.resX - 1) + std::clamp(cy, 0, g.resY - 1) * g.resX; };
    return Cic{ { idx(ix, iy), idx(ix + 1, iy), idx(ix, iy + 1), idx(ix + 1, iy + 1) }, { rx * ry, qx * ry, rx * qy, qx * qy } };
}

// src/metrics.cpp:263-266
uint32_t speedMass(uint32_t weight, float speedMps) {
    const float sm = float(weight) * speedMps + 0.5f;
    return uint32_t(std::floor(sm));
}

// metrics.hpp:80 (units::metresPerSecond(1.0) in double, rounded once)
float speedPerUnit() { return float(1.0 * double(31.25f) / double(17.490671f)); }

void add64(uint32_t* v, uint32_t b) {   // src/metrics.cpp:422-426
    const uint64_t s = (uint64_t(v[0]) | (uint64_t(v[1]) << 32)) + b;
    v[0] = uint32_t(s);
    v[1] = uint32_t(s >> 32);
}

} // namespace

extern "C" {

void oracle_metrics_cic(int n, const float* xy, int resX, int resY, float cellsPerUnit, float offX, float offY, uint32_t* cells,
                        uint32_t* weights) {
    const Grid g{ resX, resY, cellsPerUnit, offX, offY };
    for (int i = 0; i < n; ++i) {
        const Cic d = cicDeposit(g, xy[2 * i], xy[2 * i + 1]);
        for (int k = 0; k < 4; ++k) {
            cells[4 * i + k] = uint32_t(d.cell[k]);
            weights[4 * i + k] = d.weight[k];
        }
    }
}

void oracle_metrics_speed_mass(int n, const uint32_t* weight, const float* mps, uint32_t* out) {
    for (int i = 0; i < n; ++i) out[i] = speedMass(weight[i], mps[i]);
}

float oracle_metrics_speed_per_unit() { return speedPerUnit(); }

// src/metrics.cpp:232-246: included cells get their threshold masses, the others (and slot 3) NEVER
void oracle_metrics_thresholds(int cells, const float* walkableM2, const uint8_t* included, uint32_t* thr4) {
    for (int c = 0; c < cells; ++c) {
        for (int k = 0; k < 4; ++k) thr4[4 * c + k] = NEVER;
        if (!included[c]) continue;
        for (int k = 0; k < 3; ++k) {
            const double m = std::ceil(THRESHOLD_P_M2[k] * double(walkableM2[c]) * double(MASS_ONE));
            thr4[4 * c + k] = uint32_t(std::clamp(m, 1.0, double(NEVER)));
        }
    }
}

// One recorded step (src/metrics.cpp:429-480). prev: x, y, w of the step's input; cur: x, y, w, vx, vy of its output
// (model units). rects: E exit rectangles (lox, loy, hix, hiy). counters: COUNTER_WORDS; evacStep: n; cells:
// CELL_WORDS per cell; thr4: 4 per cell.
void oracle_metrics_record(int n, const float* px, const float* py, const float* pw, const float* cx, const float* cy, const float* cw,
                           const float* cvx, const float* cvy, int resX, int resY, float cellsPerUnit, float offX, float offY,
                           const uint32_t* thr4, int E, const float* rects, uint32_t* counters, uint32_t* evacStep, uint32_t* cells) {
    const Grid g{ resX, resY, cellsPerUnit, offX, offY };
    const int C = resX * resY;
    const uint32_t step = counters[0] + 1u;
    const float speedUnit = speedPerUnit();
    std::vector<uint32_t> mass(size_t(C), 0u), speedM(size_t(C), 0u);
    uint32_t active = 0;
    for (int i = 0; i < n; ++i) {
        if (pw[i] != 0.0f && cw[i] == 0.0f) {
            int slot = E;
            for (int e = 0; e < E; ++e) {
                const float* r = rects + 4 * e;
                if (px[i] >= r[0] && px[i] <= r[2] && py[i] >= r[1] && py[i] <= r[3]) { slot = e; break; }
            }
            ++counters[4 + slot];
            ++counters[2];
            evacStep[i] = step;
        }
        if (cw[i] != 0.0f) {
            ++active;
            const Cic d = cicDeposit(g, cx[i], cy[i]);
            const float s = std::sqrt(cvx[i] * cvx[i] + cvy[i] * cvy[i]) * speedUnit;   // src/metrics.cpp:268
            for (int k = 0; k < 4; ++k) {
                if (d.weight[k] == 0u) continue;
                mass[size_t(d.cell[k])] += d.weight[k];
                speedM[size_t(d.cell[k])] += speedMass(d.weight[k], s);
            }
        }
    }
    counters[0] = step;
    counters[1] = active;
    counters[3] = 0u;
    for (int c = 0; c < C; ++c) {
        const uint32_t m = mass[size_t(c)], sm = speedM[size_t(c)];
        uint32_t* s = cells + size_t(CELL_WORDS) * size_t(c);
        const uint32_t* t = thr4 + 4 * size_t(c);
        s[0] = m;
        s[1] = sm;
        s[2] = std::max(s[2], m);
        s[3] += m != 0u ? 1u : 0u;
        for (int k = 0; k < 4; ++k) s[4 + k] += m >= t[k] ? 1u : 0u;
        add64(s + 8, m);
        add64(s + 10, sm);
        if (m >= t[CRUSH_INDEX]) add64(s + 12, m);
    }
}

int oracle_metrics_counter_words() { return COUNTER_WORDS; }
int oracle_metrics_cell_words() { return CELL_WORDS; }

} // extern "C"
