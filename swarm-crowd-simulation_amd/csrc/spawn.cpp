// Spawners: the reference's seat spawn (seatedSpawnSeats, src/crowd_sim.cpp:652-674) and, for the open floors, a uniform
// placement with rejection followed by the reference's spawn relaxation (relaxSpawn, src/crowd_sim.cpp:544-597) adapted to
// per-agent radii on a flat floor.
//
// Seats: the same subset of seats as the reference for the same (seat count, agents, load, seed): a partial Fisher-Yates
// over the seat indices on mt19937's raw output (the generator the C++ standard specifies bit for bit), the chosen seats
// in seat order, padded with -1.
// Open floors: every draw comes from mt19937's raw output (portable, unlike the standard distributions): a position
// uniform over the floor's spawn band (the room's interior less the clearance r_i from its walls, outside its holes), at
// most SPAWN_ATTEMPTS draws until it overlaps no agent placed before; then Gauss-Seidel sweeps in agent order over every
// overlapping pair (from a cell grid rebuilt each sweep, agent order within a cell): the pair moves apart along its
// offset to the contact distance r_i + r_j, half each (a stacked pair along the lowbias32 hash direction), each agent
// clamped to its band; until no pair overlaps by more than SPAWN_RELAX_TOL or SPAWN_RELAX_SWEEPS sweeps.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <random>
#include <stdexcept>
#include <vector>

#include "scene.hpp"
#include "swarm_params.h"

namespace swarm {

std::vector<int> spawnSeats(int seatCount, int numAgents, double load, uint32_t seed) {
    if (seatCount < 0 || numAgents < 0) throw std::invalid_argument("spawnSeats: negative count");
    const int m = int(std::lround(std::clamp(load, 0.0, 1.0) * double(seatCount)));
    if (m > numAgents) throw std::invalid_argument("spawnSeats: more seats chosen than agents");
    std::vector<int> idx(static_cast<size_t>(seatCount));
    for (int i = 0; i < seatCount; ++i) idx[size_t(i)] = i;
    std::mt19937 gen(seed);
    for (int i = 0; i < m && m < seatCount; ++i) {
        const int j = i + int(uint32_t(gen()) % uint32_t(seatCount - i));
        std::swap(idx[size_t(i)], idx[size_t(j)]);
    }
    std::vector<int> chosen(idx.begin(), idx.begin() + m);
    std::sort(chosen.begin(), chosen.end());
    chosen.resize(size_t(numAgents), -1);
    return chosen;
}

namespace {

constexpr int SPAWN_ATTEMPTS = 64;
constexpr int SPAWN_RELAX_SWEEPS = 2000;
constexpr float SPAWN_RELAX_TOL = 1e-6f;   // model units (0.03 mm)
constexpr double CLEAR_M = 0.01;           // the band keeps this much beyond r_i from every wall

struct Floor {   // metres: the interior box and the holes (obstacle boxes) inside it
    double x0, y0, x1, y1;
    std::vector<Rect> holes;
};

Floor floorOf(int layout) {
    switch (layout) {
    case LAYOUT_ROOM: return { 0.5, 0.5, 40.5, 30.5, {} };
    case LAYOUT_CORRIDOR: return { 0.5, 0.5, 60.5, 20.5, {} };
    case LAYOUT_PILLAR: return { 0.5, 0.5, 30.5, 30.5, { Rect{ 13.5f, 13.5f, 17.5f, 17.5f } } };
    default: throw std::invalid_argument("spawnUniform: the uniform spawn is for the open floors (room, corridor, pillar room)");
    }
}

double u01(std::mt19937& g) { return (double(g()) + 0.5) * (1.0 / 4294967296.0); }   // (0, 1)

// the band of an agent of radius r (model units): inside the interior shrunk by r + CLEAR, outside every hole grown by it
struct Band {
    float x0, y0, x1, y1;
    std::vector<Rect> holes;
    bool ok(float x, float y) const {
        if (x < x0 || x > x1 || y < y0 || y > y1) return false;
        for (const Rect& h : holes)
            if (x > h.lox && x < h.hix && y > h.loy && y < h.hiy) return false;
        return true;
    }
    // the nearest point of the band (clamp to the box, then out of a hole through its nearest side)
    void clamp(float& x, float& y) const {
        x = std::min(std::max(x, x0), x1);
        y = std::min(std::max(y, y0), y1);
        for (const Rect& h : holes) {
            if (!(x > h.lox && x < h.hix && y > h.loy && y < h.hiy)) continue;
            const float dl = x - h.lox, dr = h.hix - x, db = y - h.loy, dt = h.hiy - y;
            const float m = std::min(std::min(dl, dr), std::min(db, dt));
            if (m == dl) x = h.lox; else if (m == dr) x = h.hix; else if (m == db) y = h.loy; else y = h.hiy;
        }
    }
};

Band bandOf(const Floor& f, float r) {
    const double c = double(r) * M_PER_UNIT + CLEAR_M;   // metres
    Band b{ float((f.x0 + c) / M_PER_UNIT), float((f.y0 + c) / M_PER_UNIT), float((f.x1 - c) / M_PER_UNIT), float((f.y1 - c) / M_PER_UNIT), {} };
    for (const Rect& h : f.holes)
        b.holes.push_back(Rect{ float((h.lox - c) / M_PER_UNIT), float((h.loy - c) / M_PER_UNIT), float((h.hix + c) / M_PER_UNIT),
                                float((h.hiy + c) / M_PER_UNIT) });
    return b;
}

}  // namespace

int relaxSpawnDiscs(std::vector<float>& x, std::vector<float>& y, const std::vector<float>& r,
                    const std::function<void(size_t, float&, float&)>& clampTo, float& maxOverlap) {
    const int n = int(x.size());
    float rmax = 0.0f, xmin = 1e30f, ymin = 1e30f, xmax = -1e30f, ymax = -1e30f;
    for (int i = 0; i < n; ++i) {
        rmax = std::max(rmax, r[size_t(i)]);
        xmin = std::min(xmin, x[size_t(i)]); xmax = std::max(xmax, x[size_t(i)]);
        ymin = std::min(ymin, y[size_t(i)]); ymax = std::max(ymax, y[size_t(i)]);
    }
    maxOverlap = 0.0f;
    if (n < 2) return 0;
    const float cell = 2.0f * rmax;
    const int rx = std::max(1, int((xmax - xmin) / cell) + 1), ry = std::max(1, int((ymax - ymin) / cell) + 1);
    auto cellOf = [&](float px, float py) {
        const int cx = std::min(std::max(int((px - xmin) / cell), 0), rx - 1);
        const int cy = std::min(std::max(int((py - ymin) / cell), 0), ry - 1);
        return cx + cy * rx;
    };
    std::vector<int> start(size_t(rx) * size_t(ry) + 1), order(static_cast<size_t>(n)), cellIdx(static_cast<size_t>(n));
    int sweep = 0;
    for (; sweep < SPAWN_RELAX_SWEEPS; ++sweep) {
        std::fill(start.begin(), start.end(), 0);
        for (int i = 0; i < n; ++i) {
            cellIdx[size_t(i)] = cellOf(x[size_t(i)], y[size_t(i)]);
            ++start[size_t(cellIdx[size_t(i)]) + 1];
        }
        for (size_t c = 1; c < start.size(); ++c) start[c] += start[c - 1];
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int i = 0; i < n; ++i) order[size_t(fill[size_t(cellIdx[size_t(i)])]++)] = i;
        maxOverlap = 0.0f;
        for (int i = 0; i < n; ++i) {
            const int ci = cellIdx[size_t(i)], cx = ci % rx, cy = ci / rx;
            for (int yy = std::max(cy - 1, 0); yy <= std::min(cy + 1, ry - 1); ++yy)
                for (int xx = std::max(cx - 1, 0); xx <= std::min(cx + 1, rx - 1); ++xx)
                    for (int s = start[size_t(xx + yy * rx)]; s < start[size_t(xx + yy * rx) + 1]; ++s) {
                        const int j = order[size_t(s)];
                        if (j <= i) continue;   // each pair once per sweep
                        const float ox = x[size_t(i)] - x[size_t(j)], oy = y[size_t(i)] - y[size_t(j)];
                        const float d = std::sqrt(ox * ox + oy * oy);
                        const float D = r[size_t(i)] + r[size_t(j)];
                        if (d >= D) continue;
                        maxOverlap = std::max(maxOverlap, D - d);
                        const float shift = 0.5f * (D - d);
                        float dx, dy;
                        if (d > 1e-6f * D) {
                            dx = ox / d;
                            dy = oy / d;
                        } else {   // a stacked pair: the lowbias32 hash direction (the reference's spawnHash)
                            const float angle = float(hash32(hash32(uint32_t(i)) ^ uint32_t(j))) * HASH_TO_RADIANS;
                            sincos_cephes(angle, &dy, &dx);
                        }
                        float xi = x[size_t(i)] + shift * dx, yi = y[size_t(i)] + shift * dy;
                        float xj = x[size_t(j)] - shift * dx, yj = y[size_t(j)] - shift * dy;
                        clampTo(size_t(i), xi, yi);
                        clampTo(size_t(j), xj, yj);
                        x[size_t(i)] = xi; y[size_t(i)] = yi;
                        x[size_t(j)] = xj; y[size_t(j)] = yj;
                    }
        }
        if (maxOverlap <= SPAWN_RELAX_TOL) return sweep + 1;
    }
    return sweep;
}

Agents spawnUniform(int layout, size_t active, size_t total, uint32_t seed, double speedMean, double speedSd, int* sweeps,
                    float* maxOverlap) {
    if (total < active) throw std::invalid_argument("spawnUniform: total < active");
    const Floor f = floorOf(layout);
    std::mt19937 gen(seed);
    Agents a;
    a.x.assign(total, PARK_COORD); a.y.assign(total, PARK_COORD); a.r.resize(total); a.v0.resize(total); a.w.assign(total, 0.0f);
    for (size_t i = 0; i < total; ++i) {
        a.r[i] = float((0.20 + 0.05 * u01(gen)) / M_PER_UNIT);
        const double g = std::sqrt(-2.0 * std::log(u01(gen))) * std::cos(2.0 * 3.14159265358979323846 * u01(gen));   // Box-Muller
        a.v0[i] = float(std::min(std::max(speedMean + speedSd * g, 0.5), 2.0) * S_PER_UNIT / M_PER_UNIT);
    }
    // uniform with rejection: a cell grid of the agents placed so far
    const float cell = float(0.5 / M_PER_UNIT);
    const int rx = int((f.x1 - f.x0) / 0.5) + 2, ry = int((f.y1 - f.y0) / 0.5) + 2;
    std::vector<std::vector<uint32_t>> grid(size_t(rx) * size_t(ry));
    auto cellOf = [&](float px, float py) {
        const int cx = std::min(std::max(int((px - float(f.x0 / M_PER_UNIT)) / cell), 0), rx - 1);
        const int cy = std::min(std::max(int((py - float(f.y0 / M_PER_UNIT)) / cell), 0), ry - 1);
        return std::pair<int, int>(cx, cy);
    };
    std::vector<Band> bands;
    bands.reserve(active);
    for (size_t i = 0; i < active; ++i) {
        bands.push_back(bandOf(f, a.r[i]));
        const Band& b = bands.back();
        float px = 0.0f, py = 0.0f;
        for (int t = 0; t < SPAWN_ATTEMPTS; ++t) {
            px = b.x0 + float(u01(gen)) * (b.x1 - b.x0);
            py = b.y0 + float(u01(gen)) * (b.y1 - b.y0);
            if (!b.ok(px, py)) continue;
            const auto c = cellOf(px, py);
            bool free = true;
            for (int yy = std::max(c.second - 1, 0); yy <= std::min(c.second + 1, ry - 1) && free; ++yy)
                for (int xx = std::max(c.first - 1, 0); xx <= std::min(c.first + 1, rx - 1) && free; ++xx)
                    for (uint32_t j : grid[size_t(xx) + size_t(yy) * size_t(rx)]) {
                        const float dx = px - a.x[j], dy = py - a.y[j];
                        if (dx * dx + dy * dy < (a.r[i] + a.r[j]) * (a.r[i] + a.r[j])) { free = false; break; }
                    }
            if (free) break;
        }
        b.clamp(px, py);   // the last draw stands when every draw overlapped (the relaxation moves it)
        a.x[i] = px;
        a.y[i] = py;
        a.w[i] = 1.0f;
        const auto c = cellOf(px, py);
        grid[size_t(c.first) + size_t(c.second) * size_t(rx)].push_back(uint32_t(i));
    }
    std::vector<float> xs(a.x.begin(), a.x.begin() + long(active)), ys(a.y.begin(), a.y.begin() + long(active));
    std::vector<float> rs(a.r.begin(), a.r.begin() + long(active));
    float mo = 0.0f;
    const int sw = relaxSpawnDiscs(xs, ys, rs, [&](size_t k, float& px, float& py) { bands[k].clamp(px, py); }, mo);
    std::copy(xs.begin(), xs.end(), a.x.begin());
    std::copy(ys.begin(), ys.end(), a.y.begin());
    if (sweeps) *sweeps = sw;
    if (maxOverlap) *maxOverlap = mo;
    return a;
}

}  // namespace swarm
