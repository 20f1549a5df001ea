// Scene preparation: signed-distance grid, navigation grid, per-exit any-angle fields, candidate table,
// stair field and agent placements. Host code; the results are the step's read-only inputs.
#include "scene.hpp"
#include "swarm_params.h"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <limits>
#include <queue>
#include <thread>
#include <stdexcept>
#include <utility>

namespace swarm {

void NavGrid::centre(int ix, int iy, float& x, float& y) const {   // nav::Grid::centre (nav_field.cpp:32-35)
    const double c = double(cell);
    x = float(double(originX) + (double(ix) + 0.5) * c);
    y = float(double(originY) + (double(iy) + 0.5) * c);
}

namespace {

constexpr int DIR_X[8] = { 1, 1, 0, -1, -1, -1, 0, 1 };   // nav_field.hpp:68-69: counter-clockwise from +x
constexpr int DIR_Y[8] = { 0, 1, 1, 1, 0, -1, -1, -1 };
inline int opposite(int k) { return (k + 4) & 7; }
inline int dirIndex(int dx, int dy) {
    for (int k = 0; k < 8; ++k)
        if (DIR_X[k] == dx && DIR_Y[k] == dy) return k;
    return -1;
}

// A uniform bucket index over rectangles: every rectangle is listed in the buckets its box, grown by
// 'pad', overlaps, so a query at a point lists every rectangle within 'pad' of it.
struct RectIndex {
    float x0 = 0, y0 = 0, b = 1;
    int nx = 1, ny = 1;
    std::vector<std::vector<int>> cells;
    const std::vector<Rect>* rects = nullptr;
    void build(const std::vector<Rect>& rs, float xmin, float ymin, float xmax, float ymax, float bucket, float pad) {
        rects = &rs;
        x0 = xmin - pad; y0 = ymin - pad; b = bucket;
        nx = std::max(1, int(std::ceil((xmax - xmin + 2 * pad) / bucket)) + 1);
        ny = std::max(1, int(std::ceil((ymax - ymin + 2 * pad) / bucket)) + 1);
        cells.assign(size_t(nx) * ny, {});
        for (int i = 0; i < int(rs.size()); ++i) {
            const Rect& r = rs[size_t(i)];
            const int ax = clampX(int(std::floor((r.lox - pad - x0) / b))), bx = clampX(int(std::floor((r.hix + pad - x0) / b)));
            const int ay = clampY(int(std::floor((r.loy - pad - y0) / b))), by = clampY(int(std::floor((r.hiy + pad - y0) / b)));
            for (int y = ay; y <= by; ++y)
                for (int x = ax; x <= bx; ++x) cells[size_t(y) * nx + x].push_back(i);
        }
    }
    int clampX(int v) const { return std::min(std::max(v, 0), nx - 1); }
    int clampY(int v) const { return std::min(std::max(v, 0), ny - 1); }
    const std::vector<int>& at(float x, float y) const {
        return cells[size_t(clampY(int(std::floor((y - y0) / b)))) * nx + clampX(int(std::floor((x - x0) / b)))];
    }
};

// RectGeometry's float arithmetic (nav_field.cpp:106-125)
bool segmentHitsRect(float ax, float ay, float bx, float by, const Rect& r) {
    float t0 = 0.0f, t1 = 1.0f;
    const float a[2] = { ax, ay }, d[2] = { bx - ax, by - ay };
    const float lo[2] = { r.lox, r.loy }, hi[2] = { r.hix, r.hiy };
    for (int k = 0; k < 2; ++k) {
        if (std::abs(d[k]) < 1e-12f) {
            if (a[k] < lo[k] || a[k] > hi[k]) return false;
            continue;
        }
        float u0 = (lo[k] - a[k]) / d[k], u1 = (hi[k] - a[k]) / d[k];
        if (u0 > u1) std::swap(u0, u1);
        t0 = std::max(t0, u0);
        t1 = std::min(t1, u1);
        if (t0 > t1) return false;
    }
    return true;
}
float rectDistance(float px, float py, const Rect& r) {
    const float qx = std::min(std::max(px, r.lox), r.hix), qy = std::min(std::max(py, r.loy), r.hiy);
    const float dx = px - qx, dy = py - qy;
    return std::sqrt(dx * dx + dy * dy);
}
// the exact signed distance of a box (negative inside), in double
double sdBoxD(double px, double py, const Rect& r) {
    const double cx = 0.5 * (double(r.lox) + double(r.hix)), cy = 0.5 * (double(r.loy) + double(r.hiy));
    const double hx = 0.5 * (double(r.hix) - double(r.lox)), hy = 0.5 * (double(r.hiy) - double(r.loy));
    const double qx = std::abs(px - cx) - hx, qy = std::abs(py - cy) - hy;
    const double ox = std::max(qx, 0.0), oy = std::max(qy, 0.0);
    return std::sqrt(ox * ox + oy * oy) + std::min(std::max(qx, qy), 0.0);
}

// Worker count: the argument, else OMP_NUM_THREADS, else the hardware's, at most 64
int resolveThreads(int threads) {
    if (threads > 0) return threads;
    if (const char* e = std::getenv("OMP_NUM_THREADS")) {
        const int v = std::atoi(e);
        if (v > 0) return std::min(v, 64);
    }
    const unsigned h = std::thread::hardware_concurrency();
    return std::max(1, std::min(int(h ? h : 1), 64));
}

// body(i) for i in [0, n) on 'threads' workers, chunked dynamically; the results may not depend on the order
template <class Body> void parallelFor(int threads, long long n, long long chunk, const Body& body) {
    if (threads <= 1 || n <= chunk) {
        for (long long i = 0; i < n; ++i) body(i);
        return;
    }
    std::atomic<long long> next{ 0 };
    auto work = [&] {
        for (;;) {
            const long long a = next.fetch_add(chunk);
            if (a >= n) return;
            const long long b = std::min(n, a + chunk);
            for (long long i = a; i < b; ++i) body(i);
        }
    };
    std::vector<std::thread> pool;
    const int t = int(std::min<long long>(threads, (n + chunk - 1) / chunk));
    for (int k = 1; k < t; ++k) pool.emplace_back(work);
    work();
    for (std::thread& th : pool) th.join();
}

} // namespace

NavField buildNavGrid(const SceneDesc& d, int threads) {   // nav::buildGrid (nav_field.cpp:184-235) on a RectGeometry
    threads = resolveThreads(threads);
    NavField f;
    f.grid = d.nav;
    const NavGrid& g = d.nav;
    const size_t n = g.cells();
    f.cls.assign(n, NAV_CLASS_NOFIT);
    f.out.assign(n, 0);
    f.exitRegion.assign(n, -1);
    f.exitCount = int(d.exits.size());
    RectIndex idx;
    const float pad = std::max(d.navBody, 1.5f * g.cell) + 2.0f * g.cell;
    idx.build(d.walls, g.originX, g.originY, g.originX + g.cell * float(g.resX), g.originY + g.cell * float(g.resY),
              std::max(8.0f * g.cell, pad), pad);
    const float body = d.navBody;
    parallelFor(threads, g.resY, 8, [&](long long iyL) {
        const int iy = int(iyL);
        for (int ix = 0; ix < g.resX; ++ix) {
            const size_t i = size_t(g.index(ix, iy));
            float cx, cy;
            g.centre(ix, iy, cx, cy);
            bool fit = true;   // RectGeometry::fits (nav_field.cpp:134-138)
            for (int r : idx.at(cx, cy)) {
                const Rect& w = d.walls[size_t(r)];
                if (w.contains(cx, cy) || rectDistance(cx, cy, w) < body) { fit = false; break; }
            }
            f.cls[i] = fit ? NAV_CLASS_FIT : NAV_CLASS_NOFIT;
            int e = -1;   // RectGeometry::exitAt (nav_field.cpp:155-159)
            for (size_t k = 0; k < d.exits.size(); ++k)
                if (d.exits[k].contains(cx, cy)) { e = int(k); break; }
            f.exitRegion[i] = int8_t(e);
        }
    });
    parallelFor(threads, g.resY, 8, [&](long long iyL) {
        const int iy = int(iyL);
        for (int ix = 0; ix < g.resX; ++ix) {
            const size_t i = size_t(g.index(ix, iy));
            float cx, cy;
            g.centre(ix, iy, cx, cy);
            uint8_t mask = 0;
            for (int k = 0; k < 8; ++k) {
                const int nx = ix + DIR_X[k], ny = iy + DIR_Y[k];
                if (nx < 0 || ny < 0 || nx >= g.resX || ny >= g.resY) continue;
                const size_t j = size_t(g.index(nx, ny));
                if (DIR_X[k] != 0 && DIR_Y[k] != 0 && f.cls[i] == NAV_CLASS_FIT && f.cls[j] == NAV_CLASS_FIT &&
                    (f.cls[size_t(g.index(nx, iy))] != NAV_CLASS_FIT || f.cls[size_t(g.index(ix, ny))] != NAV_CLASS_FIT))
                    continue;
                float tx, ty;
                g.centre(nx, ny, tx, ty);
                bool legal = true;   // RectGeometry::moveLegal (nav_field.cpp:140-146), no ledges
                for (int r : idx.at(cx, cy))
                    if (segmentHitsRect(cx, cy, tx, ty, d.walls[size_t(r)])) { legal = false; break; }
                if (legal) mask = uint8_t(mask | (1u << k));
            }
            f.out[i] = mask;
        }
    });
    return f;
}

namespace {
// nav::lineLegal (nav_field.cpp:239-263): every cell transition of the supercover of the segment between
// the centres of cells a and b is a legal edge and none leads from a FIT cell into a NOFIT one
bool lineLegal(const NavField& f, int a, int b) {
    const int res = f.grid.resX;
    int x = a % res, y = a / res;
    const int x1 = b % res, y1 = b / res;
    const int dx = std::abs(x1 - x), dy = std::abs(y1 - y), sx = x1 > x ? 1 : -1, sy = y1 > y ? 1 : -1;
    long long ix = 0, iy = 0;
    while (ix < dx || iy < dy) {
        const long long cx = (2 * ix + 1) * dy, cy = (2 * iy + 1) * dx;
        int nx = x, ny = y;
        if (cx < cy) { nx += sx; ++ix; }
        else if (cx > cy) { ny += sy; ++iy; }
        else { nx += sx; ny += sy; ++ix; ++iy; }
        const int k = dirIndex(nx - x, ny - y);
        const size_t c = size_t(f.grid.index(x, y)), n = size_t(f.grid.index(nx, ny));
        if (!((f.out[c] >> k) & 1u) || (f.cls[c] == NAV_CLASS_FIT && f.cls[n] != NAV_CLASS_FIT)) return false;
        x = nx;
        y = ny;
    }
    return true;
}
} // namespace

// nav::lazyThetaStar (nav_field.cpp:268-342): Lazy Theta* over the reversed legal edges from the FIT cells of
// the exit region(s); a cell reached from a closed cell takes that cell's anchor, checked lazily on expansion
void lazyThetaStar(const NavField& f, int exitFilter, std::vector<double>& g, std::vector<int32_t>& par,
                   std::vector<uint8_t>& closed) {
    const NavGrid& grid = f.grid;
    const size_t n = grid.cells();
    const int res = grid.resX;
    constexpr double INF = std::numeric_limits<double>::infinity();
    g.assign(n, INF);
    par.assign(n, -1);
    closed.assign(n, 0);
    const double h = double(grid.cell);
    auto cellDist = [&](int a, int b) {
        const double dx = double(a % res - b % res), dy = double(a / res - b / res);
        return h * std::sqrt(dx * dx + dy * dy);
    };
    auto classOk = [&](size_t from, size_t to) { return f.cls[from] != NAV_CLASS_FIT || f.cls[to] == NAV_CLASS_FIT; };
    using Entry = std::pair<double, int32_t>;
    std::priority_queue<Entry, std::vector<Entry>, std::greater<Entry>> open;
    for (size_t i = 0; i < n; ++i)
        if (f.exitRegion[i] >= 0 && (exitFilter < 0 || f.exitRegion[i] == exitFilter) && f.cls[i] == NAV_CLASS_FIT) {
            g[i] = 0.0;
            par[i] = int32_t(i);
            open.push({ 0.0, int32_t(i) });
        }
    while (!open.empty()) {
        const auto [dcur, v] = open.top();
        open.pop();
        if (closed[size_t(v)] || dcur > g[size_t(v)]) continue;
        const int vx = v % res, vy = v / res;
        if (par[size_t(v)] != v && !lineLegal(f, v, par[size_t(v)])) {
            double best = INF;
            int32_t bp = -1;
            for (int k = 0; k < 8; ++k) {
                const int wx = vx + DIR_X[k], wy = vy + DIR_Y[k];
                if (wx < 0 || wy < 0 || wx >= res || wy >= grid.resY || !((f.out[size_t(v)] >> k) & 1u)) continue;
                const size_t w = size_t(grid.index(wx, wy));
                if (!closed[w] || !classOk(size_t(v), w)) continue;
                const double c = g[w] + cellDist(v, int(w));
                if (c < best) { best = c; bp = int32_t(w); }
            }
            g[size_t(v)] = best;
            par[size_t(v)] = bp;
        }
        closed[size_t(v)] = 1;
        if (par[size_t(v)] < 0) continue;
        const int32_t p = par[size_t(v)];
        for (int k = 0; k < 8; ++k) {
            const int ux = vx + DIR_X[k], uy = vy + DIR_Y[k];
            if (ux < 0 || uy < 0 || ux >= res || uy >= grid.resY) continue;
            const size_t u = size_t(grid.index(ux, uy));
            if (closed[u] || !((f.out[u] >> opposite(k)) & 1u) || !classOk(u, size_t(v))) continue;
            const double c = g[size_t(p)] + cellDist(int(u), p);
            if (c < g[u]) {
                g[u] = c;
                par[u] = p;
                open.push({ c, int32_t(u) });
            }
        }
    }
}

namespace {
// The metric grid and each cell's walkable area: 1 m cells from the whole metre at or below the domain's lower
// corner; samples at the centres of a 40 x 40 sub-grid per cell, walkable when inside the domain and outside every
// wall rectangle. One strip of cells (40 sample rows) is rasterised at a time.
void buildMetricGrid(const SceneDesc& d, Scene& s) {
    constexpr int SUB = 40;
    const double mpu = d.metresPerUnit;
    const double x0m = std::floor(double(d.xmin) * mpu), y0m = std::floor(double(d.ymin) * mpu);
    const int rx = int(std::ceil(double(d.xmax) * mpu - x0m)), ry = int(std::ceil(double(d.ymax) * mpu - y0m));
    s.metricResX = std::max(rx, 1);
    s.metricResY = std::max(ry, 1);
    s.metricCellsPerUnit = float(mpu);                 // 1 m cells
    s.metricOffX = float(-x0m - 0.5);                  // whole metres - 0.5: exact in float
    s.metricOffY = float(-y0m - 0.5);
    s.metricWalkableM2.assign(size_t(s.metricResX) * size_t(s.metricResY), 0.0f);
    const int W = s.metricResX * SUB;
    std::vector<uint8_t> free(size_t(W) * SUB);
    const double step = 1.0 / SUB;                     // metres between samples
    for (int cy = 0; cy < s.metricResY; ++cy) {
        // samples of this strip: x = x0m + (i + 0.5) step, y = y0m + cy + (j + 0.5) step (metres)
        for (int j = 0; j < SUB; ++j) {
            const double ym = y0m + cy + (j + 0.5) * step;
            const float yu = float(ym / mpu);
            for (int i = 0; i < W; ++i) {
                const float xu = float((x0m + (i + 0.5) * step) / mpu);
                free[size_t(j) * W + i] = xu >= d.xmin && xu <= d.xmax && yu >= d.ymin && yu <= d.ymax;
            }
        }
        const double stripLo = y0m + cy, stripHi = stripLo + 1.0;
        for (const Rect& r : d.walls) {
            const double lo = double(r.loy) * mpu, hi = double(r.hiy) * mpu;
            if (hi < stripLo || lo > stripHi) continue;
            const int j0 = std::max(0, int(std::ceil((lo - stripLo) / step - 0.5)));
            const int j1 = std::min(SUB - 1, int(std::floor((hi - stripLo) / step - 0.5)));
            const int i0 = std::max(0, int(std::ceil((double(r.lox) * mpu - x0m) / step - 0.5)));
            const int i1 = std::min(W - 1, int(std::floor((double(r.hix) * mpu - x0m) / step - 0.5)));
            for (int j = j0; j <= j1; ++j)
                for (int i = i0; i <= i1; ++i) free[size_t(j) * W + i] = 0;
        }
        for (int cx = 0; cx < s.metricResX; ++cx) {
            int count = 0;
            for (int j = 0; j < SUB; ++j)
                for (int i = cx * SUB; i < (cx + 1) * SUB; ++i) count += free[size_t(j) * W + i];
            s.metricWalkableM2[size_t(cy) * s.metricResX + cx] = float(count * step * step);
        }
    }
}
// The wall index of the exact wall query: a hash cell lists (ascending) every wall whose box lies within
// FEELER_REACH_RADII r_max + half a cell of the cell's rectangle, so a query from any point the hash puts in the cell
// (rounding at the cell's edges included) finds every wall within its reach there
void buildWallBins(const SceneDesc& d, Scene& s) {
    s.walls = d.walls;
    const size_t cells = size_t(s.hashNX) * size_t(s.hashNY);
    const double cw = double(s.hashSpanX) / double(s.hashNX), ch = double(s.hashSpanY) / double(s.hashNY);
    const double reach = double(FEELER_REACH_RADII) * double(s.maxRadius) + 0.5 * std::min(cw, ch);
    std::vector<uint32_t> count(cells + 1, 0u);
    auto span = [&](const Rect& r, int& x0, int& x1, int& y0, int& y1) {
        x0 = std::max(0, int(std::floor((double(r.lox) - reach - double(s.xmin)) / cw)));
        x1 = std::min(s.hashNX - 1, int(std::floor((double(r.hix) + reach - double(s.xmin)) / cw)));
        y0 = std::max(0, int(std::floor((double(r.loy) - reach - double(s.ymin)) / ch)));
        y1 = std::min(s.hashNY - 1, int(std::floor((double(r.hiy) + reach - double(s.ymin)) / ch)));
    };
    auto near = [&](const Rect& r, int cx, int cy) {   // the gap between the wall box and the cell's rectangle < reach
        const double lx = double(s.xmin) + cx * cw, hx = lx + cw, ly = double(s.ymin) + cy * ch, hy = ly + ch;
        const double gx = std::max({ 0.0, double(r.lox) - hx, lx - double(r.hix) });
        const double gy = std::max({ 0.0, double(r.loy) - hy, ly - double(r.hiy) });
        return gx * gx + gy * gy < reach * reach;
    };
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            for (size_t c = 0; c < cells; ++c) count[c + 1] += count[c];
            s.wallBinStart = count;
            s.wallBinList.assign(count[cells], 0u);
        }
        std::vector<uint32_t> fill(pass == 1 ? s.wallBinStart : std::vector<uint32_t>());
        for (size_t k = 0; k < d.walls.size(); ++k) {
            int x0, x1, y0, y1;
            span(d.walls[k], x0, x1, y0, y1);
            for (int cy = y0; cy <= y1; ++cy)
                for (int cx = x0; cx <= x1; ++cx) {
                    if (!near(d.walls[k], cx, cy)) continue;
                    const size_t c = size_t(cy) * size_t(s.hashNX) + size_t(cx);
                    if (pass == 0) ++count[c + 1];
                    else s.wallBinList[fill[c]++] = uint32_t(k);
                }
        }
    }
}
} // namespace

NavField buildSeatedNavGrid(const seated::Layout& L, const NavGrid& g, int threads, std::vector<float>* floorZ) {
    threads = resolveThreads(threads);
    NavField f;
    f.grid = g;
    const size_t n = g.cells();
    std::vector<float> fz(n);
    f.cls.assign(n, NAV_CLASS_NOFIT);
    f.out.assign(n, 0);
    f.exitRegion.assign(n, -1);
    f.exitCount = seated::exitCount(L);
    const float mid = seated::FENCE_HALF_MID;
    parallelFor(threads, g.resY, 8, [&](long long iyL) {
        const int iy = int(iyL);
        for (int ix = 0; ix < g.resX; ++ix) {
            const size_t i = size_t(g.index(ix, iy));
            float cx, cy;
            g.centre(ix, iy, cx, cy);
            const seated::F2 c{ cx, cy };
            const float z = seated::floorHeight(L, c);
            fz[i] = z;
            // fits: outside the fence's collision band, and the exact walls of the plane at the body's top (the centre at
            // floor + BODY_RADIUS, the plane wallAbove higher) none closer than BODY_RADIUS
            bool fit = seated::sdFence(c) >= 0.0f;
            if (fit) fit = seated::wall(L, seated::F3{ cx, cy, z + BODY_RADIUS + L.wallAbove }, BODY_RADIUS).z >= BODY_RADIUS;
            f.cls[i] = fit ? NAV_CLASS_FIT : NAV_CLASS_NOFIT;
            // exitAt: past the outer wall's outer face, the avenue to the domain's edge, a door up to its landing's end
            const seated::F2 fo = seated::fold(c);
            int e = -1;
            if (fo.y > L.outerFace) {
                e = seated::exitByLat(L, c);
                if (e >= 4 && !(fo.y <= L.landingEnd)) e = -1;
            }
            f.exitRegion[i] = int8_t(e);
        }
    });
    auto moveLegal = [&](seated::F2 from, float fromZ, seated::F2 to, float toZ) {
        if (seated::sdBox2(from, mid) * seated::sdBox2(to, mid) < 0.0f) return false;   // crosses the fence's midline
        if (toZ < fromZ - MAX_STEP_UP) return false;                                     // a drop deeper than a step
        if (toZ > fromZ + MAX_STEP_UP) return false;                                     // a rise above a step
        seated::F2 nrm;
        if (seated::stepUpScan(L, from, to, fromZ, nrm) < 1.0f) return false;            // a climb over a vertical face
        int rowTo = -1, rowFrom = -1, band = 0;
        if (seated::kindAt(L, to, &rowTo, &band) == seated::TREAD) {                      // rows are exits only
            if (seated::kindAt(L, from, &rowFrom, &band) != seated::TREAD || rowFrom != rowTo) return false;
        }
        return true;
    };
    parallelFor(threads, g.resY, 8, [&](long long iyL) {
        const int iy = int(iyL);
        for (int ix = 0; ix < g.resX; ++ix) {
            const size_t i = size_t(g.index(ix, iy));
            float cx, cy;
            g.centre(ix, iy, cx, cy);
            uint8_t mask = 0;
            for (int k = 0; k < 8; ++k) {
                const int nx = ix + DIR_X[k], ny = iy + DIR_Y[k];
                if (nx < 0 || ny < 0 || nx >= g.resX || ny >= g.resY) continue;
                const size_t j = size_t(g.index(nx, ny));
                if (DIR_X[k] != 0 && DIR_Y[k] != 0 && f.cls[i] == NAV_CLASS_FIT && f.cls[j] == NAV_CLASS_FIT &&
                    (f.cls[size_t(g.index(nx, iy))] != NAV_CLASS_FIT || f.cls[size_t(g.index(ix, ny))] != NAV_CLASS_FIT))
                    continue;
                float tx, ty;
                g.centre(nx, ny, tx, ty);
                if (moveLegal(seated::F2{ cx, cy }, fz[i], seated::F2{ tx, ty }, fz[j])) mask = uint8_t(mask | (1u << k));
            }
            f.out[i] = mask;
        }
    });
    if (floorZ) *floorZ = std::move(fz);
    return f;
}

void buildCandidateTable(const NavField& f, const std::vector<double>& widthM, const std::vector<int>& exitType, double metresPerUnit,
                         int threads, Scene& s) {
    threads = resolveThreads(threads);
    const size_t n = f.grid.cells();
    const int E = f.exitCount;
    if (E > CHOICE_MAX_EXITS) throw std::invalid_argument("buildCandidateTable: more than 63 exits");
    if (n > size_t(CHOICE_ANCHOR_MASK) + 1) throw std::invalid_argument("buildCandidateTable: more nav cells than a 25-bit anchor holds");
    s.E = E;
    s.K = std::max(1, std::min(CHOICE_K, E));
    const int K = s.K;
    int types = 1;
    for (int e = 0; e < E; ++e) types = std::max(types, exitType[size_t(e)] + 1);
    if (types > 2) throw std::invalid_argument("buildCandidateTable: at most two exit types");
    constexpr int TM = CHOICE_TYPE_MIN_MAX;
    std::vector<float> dist(n * size_t(K), std::numeric_limits<float>::infinity());
    s.navSlot.assign(n * size_t(K), CHOICE_EMPTY << CHOICE_EXIT_SHIFT);
    std::vector<float> typeDist;
    std::vector<uint32_t> typeSlot;
    if (types == 2) {
        typeDist.assign(n * 2 * TM, std::numeric_limits<float>::infinity());
        typeSlot.assign(n * 2 * TM, 0u);
    }
    // the per-exit searches in batches (each holds about 13 bytes per cell), merged in exit order: a later exit goes
    // before the first slot it beats, so ties keep the lower exit and the table does not depend on the batching
    const int batch = std::max(1, std::min(std::min(threads, E), int(std::max<size_t>(1, (size_t(4) << 30) / std::max<size_t>(n * 13, 1)))));
    std::vector<std::vector<double>> gE(static_cast<size_t>(batch));
    std::vector<std::vector<int32_t>> parE(static_cast<size_t>(batch));
    std::vector<std::vector<uint8_t>> closedE(static_cast<size_t>(batch));
    for (int e0 = 0; e0 < E; e0 += batch) {
        const int e1 = std::min(E, e0 + batch);
        parallelFor(std::max(1, std::min(threads, e1 - e0)), e1 - e0, 1,
                    [&](long long k) { lazyThetaStar(f, e0 + int(k), gE[size_t(k)], parE[size_t(k)], closedE[size_t(k)]); });
        for (int e = e0; e < e1; ++e) {
            const std::vector<double>& g = gE[size_t(e - e0)];
            const std::vector<int32_t>& par = parE[size_t(e - e0)];
            const std::vector<uint8_t>& closed = closedE[size_t(e - e0)];
            const int ty = exitType[size_t(e)], mt = CHOICE_TYPE_MIN[ty];
            parallelFor(threads, (long long)n, 4096, [&](long long ii) {
                const size_t i = size_t(ii);
                if (f.cls[i] != NAV_CLASS_FIT || !closed[i] || !std::isfinite(g[i])) return;
                const float dd = float(g[i]);
                const uint32_t slot = (uint32_t(e) << CHOICE_EXIT_SHIFT) | uint32_t(par[i]);
                float* cd = &dist[i * size_t(K)];
                uint32_t* cs = &s.navSlot[i * size_t(K)];
                int m = K;
                while (m > 0 && dd < cd[m - 1]) --m;
                if (m < K) {
                    for (int r = K - 1; r > m; --r) { cs[r] = cs[r - 1]; cd[r] = cd[r - 1]; }
                    cs[m] = slot;
                    cd[m] = dd;
                }
                if (types == 2) {   // the nearest of this exit's type
                    float* td = &typeDist[(i * 2 + size_t(ty)) * TM];
                    uint32_t* ts = &typeSlot[(i * 2 + size_t(ty)) * TM];
                    int q = mt;
                    while (q > 0 && dd < td[q - 1]) --q;
                    if (q < mt) {
                        for (int r = mt - 1; r > q; --r) { td[r] = td[r - 1]; ts[r] = ts[r - 1]; }
                        td[q] = dd;
                        ts[q] = slot;
                    }
                }
            });
        }
    }
    // the type rule on every cell whose CHOICE_K slots are all taken: the guaranteed nearest of each type first, then the
    // nearest others, sorted by (distance, exit)
    if (types == 2 && K == CHOICE_K) {
        parallelFor(threads, (long long)n, 4096, [&](long long ii) {
            const size_t i = size_t(ii);
            uint32_t* cs = &s.navSlot[i * size_t(K)];
            float* cd = &dist[i * size_t(K)];
            if ((cs[K - 1] >> CHOICE_EXIT_SHIFT) == CHOICE_EMPTY) return;
            struct Cand { float d; uint32_t s; };
            Cand res[CHOICE_K];
            int nres = 0;
            auto taken = [&](uint32_t e) {
                for (int m = 0; m < nres; ++m)
                    if ((res[m].s >> CHOICE_EXIT_SHIFT) == e) return true;
                return false;
            };
            for (int ty = 0; ty < 2; ++ty)
                for (int m = 0; m < CHOICE_TYPE_MIN[ty] && nres < K; ++m) {
                    const size_t j = (i * 2 + size_t(ty)) * TM + size_t(m);
                    if (std::isfinite(typeDist[j])) res[nres++] = { typeDist[j], typeSlot[j] };
                }
            for (int m = 0; m < K && nres < K; ++m)
                if (!taken(cs[m] >> CHOICE_EXIT_SHIFT)) res[nres++] = { cd[m], cs[m] };
            std::sort(res, res + nres, [](const Cand& a, const Cand& b) {
                return a.d < b.d || (a.d == b.d && (a.s >> CHOICE_EXIT_SHIFT) < (b.s >> CHOICE_EXIT_SHIFT));
            });
            for (int m = 0; m < nres; ++m) { cs[m] = res[m].s; cd[m] = res[m].d; }
        });
    }
    // capacities (setChoiceCapacities, exit_choice.cpp:128-141)
    s.invCapS.assign(size_t(E), 0.0f);
    for (int e = 0; e < E; ++e) {
        const double w = widthM[size_t(e)];
        s.invCapS[size_t(e)] = w > 0.0 ? float(60.0 / (w * 83.6)) : 0.0f;
    }
    const bool modelUnits = metresPerUnit == M_PER_UNIT;
    s.secondsPerUnit = modelUnits ? CHOICE_S_PER_UNIT : float(metresPerUnit / 1.34);
    s.distQuantum = modelUnits ? CHOICE_DIST_QUANTUM
                               : std::ldexp(1.0f, -int(std::ceil(std::log2(metresPerUnit / (double(CHOICE_DIST_QUANTUM) * M_PER_UNIT)))));
    s.navInfo.assign(n, 0u);
    s.navDist.assign(n * size_t(K), uint16_t(CHOICE_DIST_EMPTY));
    for (size_t i = 0; i < n; ++i) {
        const bool fit = f.cls[i] == NAV_CLASS_FIT;
        const bool has = fit && (s.navSlot[i * size_t(K)] >> CHOICE_EXIT_SHIFT) != CHOICE_EMPTY;
        s.navInfo[i] = (fit ? NAV_FIT : 0u) | (has ? NAV_HAS : 0u);
        for (int k = 0; k < K; ++k) {
            const float dd = dist[i * size_t(K) + size_t(k)];
            if (!has || !std::isfinite(dd)) continue;
            const double q = std::nearbyint(double(dd) / double(s.distQuantum));
            if (q >= double(CHOICE_DIST_EMPTY)) throw std::invalid_argument("buildCandidateTable: a candidate distance beyond the 16-bit range");
            s.navDist[i * size_t(K) + size_t(k)] = uint16_t(q);
        }
        if (!has)
            for (int k = 0; k < K; ++k) s.navSlot[i * size_t(K) + size_t(k)] = CHOICE_EMPTY << CHOICE_EXIT_SHIFT;
    }
}

Scene buildScene(const SceneDesc& d, float maxRadius, float maxV0, int threads) {
    const auto t0 = std::chrono::steady_clock::now();
    threads = resolveThreads(threads);
    if (d.exits.size() > size_t(CHOICE_MAX_EXITS)) throw std::invalid_argument("buildScene: more than 63 exits");
    if (!(d.xmax > d.xmin && d.ymax > d.ymin && d.hashCell > 0.0f)) throw std::invalid_argument("buildScene: empty domain");
    Scene s;
    s.xmin = d.xmin; s.ymin = d.ymin; s.xmax = d.xmax; s.ymax = d.ymax;
    s.hasVenue = d.hasVenue ? 1 : 0;
    s.venue = d.venue;
    s.hashNX = int(std::lround(double(d.xmax - d.xmin) / double(d.hashCell)));
    s.hashNY = int(std::lround(double(d.ymax - d.ymin) / double(d.hashCell)));
    s.hashSpanX = d.xmax - d.xmin;
    s.hashSpanY = d.ymax - d.ymin;
    s.maxRadius = maxRadius;
    s.maxV0 = maxV0;
    // The predict ring covers R_PERC (sim_params.hpp:516); the projection ring covers every pair that can come
    // into contact during the step: R cell >= D_max + 2 (v0_max DT + PROJ_ITERATIONS PROJ_MAX_SHIFT_RADII r_max)
    const float cell = d.hashCell;
    s.percRing = int(R_PERC / cell) + 1;
    const double reach = 2.0 * double(maxRadius) + 2.0 * (double(maxV0) * double(DT) + PROJ_ITERATIONS * double(PROJ_MAX_SHIFT_RADII) * double(maxRadius));
    s.projRing = std::max(1, int(std::ceil(reach / double(cell) - 1e-9)));

    // --- signed-distance grid of the walls, clamped ---
    s.sdfH = d.sdfSpacing;
    s.sdfInvH = float(1.0 / double(d.sdfSpacing));
    s.sdfNX = int(std::lround(double(d.xmax - d.xmin) / double(d.sdfSpacing)));
    s.sdfNY = int(std::lround(double(d.ymax - d.ymin) / double(d.sdfSpacing)));
    s.sdfOX = float(double(d.xmin) + 0.5 * double(d.sdfSpacing));
    s.sdfOY = float(double(d.ymin) + 0.5 * double(d.sdfSpacing));
    s.sdf.assign(size_t(s.sdfNX) * s.sdfNY, d.sdfClamp);
    {
        RectIndex idx;
        idx.build(d.walls, d.xmin, d.ymin, d.xmax, d.ymax, std::max(0.03f, 2.0f * d.sdfClamp), d.sdfClamp + 2.0f * d.sdfSpacing);
        parallelFor(threads, s.sdfNY, 8, [&](long long jl) {
            const int j = int(jl);
            for (int i = 0; i < s.sdfNX; ++i) {
                const double px = double(d.xmin) + (double(i) + 0.5) * double(d.sdfSpacing);
                const double py = double(d.ymin) + (double(j) + 0.5) * double(d.sdfSpacing);
                double v = double(d.sdfClamp);
                for (int r : idx.at(float(px), float(py))) v = std::min(v, sdBoxD(px, py, d.walls[size_t(r)]));
                s.sdf[size_t(j) * s.sdfNX + i] = float(v);
            }
        });
    }
    // min filter over [i - win, i + win + 1]^2: a point whose bilinear cell is i has every sample within
    // PLANE_PROBE_RADII r_max of it inside the window (one texel of slack each side)
    s.sdfMinWin = int(std::ceil(double(PLANE_PROBE_RADII) * double(maxRadius) / double(d.sdfSpacing))) + 1;
    {
        const int W = s.sdfMinWin, nx = s.sdfNX, ny = s.sdfNY;
        std::vector<float> rowMin(size_t(nx) * ny);
        parallelFor(threads, ny, 16, [&](long long jl) {
            const int j = int(jl);
            for (int i = 0; i < nx; ++i) {
                float m = std::numeric_limits<float>::infinity();
                for (int k = std::max(0, i - W); k <= std::min(nx - 1, i + W + 1); ++k) m = std::min(m, s.sdf[size_t(j) * nx + k]);
                rowMin[size_t(j) * nx + i] = m;
            }
        });
        s.sdfMin.assign(size_t(nx) * ny, 0.0f);
        parallelFor(threads, ny, 16, [&](long long jl) {
            const int j = int(jl);
            for (int i = 0; i < nx; ++i) {
                float m = std::numeric_limits<float>::infinity();
                for (int k = std::max(0, j - W); k <= std::min(ny - 1, j + W + 1); ++k) m = std::min(m, rowMin[size_t(k) * nx + i]);
                s.sdfMin[size_t(j) * nx + i] = m;
            }
        });
    }

    // --- navigation grid, per-exit fields, candidate table (exit_choice.cpp:143-317, one exit type) ---
    const NavField f = buildNavGrid(d, threads);
    s.nav = d.nav;
    const int E = int(d.exits.size());
    std::vector<double> widthM(size_t(E), 0.0);
    for (int e = 0; e < E; ++e) {   // an exit's clear width: given, else its rectangle's longer side
        double w = e < int(d.exitWidthM.size()) ? double(d.exitWidthM[size_t(e)]) : 0.0;
        if (w <= 0.0) {
            const Rect& r = d.exits[size_t(e)];
            w = double(std::max(r.hix - r.lox, r.hiy - r.loy));
        }
        widthM[size_t(e)] = w;
    }
    buildCandidateTable(f, widthM, std::vector<int>(size_t(E), 0), d.metresPerUnit, threads, s);
    const size_t n = d.nav.cells();
    // --- stair field: per nav cell, the factor of the first stair band holding its centre ---
    s.stair.assign(n, 1.0f);
    for (int iy = 0; iy < d.nav.resY; ++iy)
        for (int ix = 0; ix < d.nav.resX; ++ix) {
            float cx, cy;
            d.nav.centre(ix, iy, cx, cy);
            for (size_t k = 0; k < d.stairs.size(); ++k)
                if (d.stairs[k].contains(cx, cy)) { s.stair[size_t(d.nav.index(ix, iy))] = d.stairFactor[k]; break; }
        }
    s.exits = d.exits;
    buildWallBins(d, s);
    buildMetricGrid(d, s);
    s.buildSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return s;
}

// --- seeded generator ---
uint64_t Rng::next() {   // splitmix64 (Steele, Lea, Flood 2014)
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double Rng::uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
double Rng::normal() {
    double u1 = uniform();
    const double u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// --- the benchmark floors (metres, converted to model units once) ---
namespace {
inline float mu(double metres) { return float(metres / M_PER_UNIT); }
Rect rectM(double x0, double y0, double x1, double y1) { return Rect{ mu(x0), mu(y0), mu(x1), mu(y1) }; }

// A rectangular room with interior [x0, x1] x [y0, y1], walls of thickness t outside it, and gaps
// (side 0 bottom, 1 right, 2 top, 3 left; from, to along the side's axis), like nav_check::roomWalls
struct GapM { int side; double from, to; };
void roomWallsM(double x0, double y0, double x1, double y1, double t, const std::vector<GapM>& gaps, std::vector<Rect>& out) {
    auto addSide = [&](int side, double a0, double a1, auto make) {
        std::vector<std::pair<double, double>> cuts;
        for (const GapM& g : gaps)
            if (g.side == side) cuts.push_back({ g.from, g.to });
        std::sort(cuts.begin(), cuts.end());
        double a = a0;
        for (const auto& c : cuts) { if (c.first > a) out.push_back(make(a, c.first)); a = c.second; }
        if (a1 > a) out.push_back(make(a, a1));
    };
    addSide(0, x0 - t, x1 + t, [&](double a, double b) { return rectM(a, y0 - t, b, y0); });
    addSide(2, x0 - t, x1 + t, [&](double a, double b) { return rectM(a, y1, b, y1 + t); });
    addSide(3, y0, y1, [&](double a, double b) { return rectM(x0 - t, a, x0, b); });
    addSide(1, y0, y1, [&](double a, double b) { return rectM(x1, a, x1 + t, b); });
}

void setDomainM(SceneDesc& d, double w, double h) {
    d.xmin = 0.0f; d.ymin = 0.0f; d.xmax = mu(w); d.ymax = mu(h);
    d.hashCell = mu(0.5);
    d.sdfSpacing = mu(0.1);
    d.sdfClamp = mu(1.6);
    d.navBody = mu(0.25);
    d.nav.cell = mu(0.1);
    d.nav.invCell = float(1.0 / double(d.nav.cell));
    d.nav.originX = 0.0f; d.nav.originY = 0.0f;
    d.nav.resX = int(std::lround(w / 0.1));
    d.nav.resY = int(std::lround(h / 0.1));
    d.metresPerUnit = M_PER_UNIT;
}

// The seated bowl: blocks of 60 rows (0.1 m boxes at 0.8 m pitch along x, rows running along y over a 137.5 m
// span from y = 6.25 m, cut by 4 gangways of 1.2 m into 5 segments), a 0.3 m perimeter wall, per block two 2.4 m
// exits (bottom and top wall at the block's centre) with 6 m x 6 m stair bands (0.6x) in front of them.
constexpr double BOWL_ROW0 = 4.0, BOWL_PITCH = 0.8, BOWL_BOX = 0.1, BOWL_Y0 = 6.25, BOWL_LEN = 137.5, BOWL_GANG = 1.2;
constexpr int BOWL_SEGS = 5, BOWL_SEATS_PER_SEG = 53;
double bowlSegLen() { return (BOWL_LEN - (BOWL_SEGS - 1) * BOWL_GANG) / BOWL_SEGS; }

void bowlDesc(SceneDesc& d, int rows, int blocks, double width) {
    const double H = 150.0;
    setDomainM(d, width, H);
    const double t = 0.3;
    std::vector<double> xcs;
    const int rowsPerBlock = rows / blocks;
    for (int b = 0; b < blocks; ++b) xcs.push_back(BOWL_ROW0 + BOWL_PITCH * (rowsPerBlock * b) + 0.5 * BOWL_PITCH * rowsPerBlock);
    // perimeter: bottom and top with the exit gaps, left and right solid
    auto gapsAt = [&](std::vector<std::pair<double, double>>& g) { for (double xc : xcs) g.push_back({ xc - 1.2, xc + 1.2 }); };
    std::vector<std::pair<double, double>> gaps;
    gapsAt(gaps);
    std::sort(gaps.begin(), gaps.end());
    for (int side = 0; side < 2; ++side) {
        const double y0 = side == 0 ? 0.0 : H - t, y1 = side == 0 ? t : H;
        double a = 0.0;
        for (const auto& g : gaps) { d.walls.push_back(rectM(a, y0, g.first, y1)); a = g.second; }
        d.walls.push_back(rectM(a, y0, width, y1));
    }
    d.walls.push_back(rectM(0.0, t, t, H - t));
    d.walls.push_back(rectM(width - t, t, width, H - t));
    const double seg = bowlSegLen();
    for (int k = 0; k < rows; ++k) {
        const double x0 = BOWL_ROW0 + BOWL_PITCH * k;
        for (int j = 0; j < BOWL_SEGS; ++j) {
            const double ya = BOWL_Y0 + j * (seg + BOWL_GANG);
            d.walls.push_back(rectM(x0, ya, x0 + BOWL_BOX, ya + seg));
        }
    }
    for (double xc : xcs) {
        d.exits.push_back(rectM(xc - 1.2, 0.0, xc + 1.2, 0.2));
        d.exitWidthM.push_back(2.4f);
        d.exits.push_back(rectM(xc - 1.2, H - 0.2, xc + 1.2, H));
        d.exitWidthM.push_back(2.4f);
        d.stairs.push_back(rectM(xc - 3.0, t, xc + 3.0, t + 6.0));
        d.stairFactor.push_back(0.6f);
        d.stairs.push_back(rectM(xc - 3.0, H - t - 6.0, xc + 3.0, H - t));
        d.stairFactor.push_back(0.6f);
    }
}

void bowlSlots(int rows, std::vector<float>& sx, std::vector<float>& sy) {
    const double seg = bowlSegLen();
    for (int k = 0; k < rows; ++k) {
        const double x = BOWL_ROW0 + BOWL_PITCH * k + BOWL_BOX + 0.5 * (BOWL_PITCH - BOWL_BOX);
        for (int j = 0; j < BOWL_SEGS; ++j) {
            const double ya = BOWL_Y0 + j * (seg + BOWL_GANG);
            for (int s = 0; s < BOWL_SEATS_PER_SEG; ++s) {
                sx.push_back(float(x));
                sy.push_back(float(ya + 0.25 + 0.5 * s));
            }
        }
    }
}

// hexagonal lattice sites at pitch a inside [x0, x1] x [y0, y1] (metres), skipping a box
void hexSlots(double x0, double y0, double x1, double y1, double a, const Rect* hole, std::vector<float>& sx, std::vector<float>& sy) {
    const double dy = a * std::sqrt(3.0) / 2.0;
    int row = 0;
    for (double y = y0; y <= y1 + 1e-9; y += dy, ++row) {
        for (double x = x0 + ((row & 1) ? 0.5 * a : 0.0); x <= x1 + 1e-9; x += a) {
            if (hole && x >= hole->lox && x <= hole->hix && y >= hole->loy && y <= hole->hiy) continue;
            sx.push_back(float(x));
            sy.push_back(float(y));
        }
    }
}
constexpr double HEX_PITCH = 0.53, HEX_JITTER = 0.01, HEX_CLEAR = 0.27;
} // namespace

SceneDesc layoutDesc(int layout) {
    SceneDesc d;
    switch (layout) {
    case LAYOUT_ROOM: {   // 40 x 30 m room, 0.3 m walls, two 2 m exits in the bottom wall
        setDomainM(d, 41.0, 31.0);
        roomWallsM(0.5, 0.5, 40.5, 30.5, 0.3, { { 0, 10.5, 12.5 }, { 0, 28.5, 30.5 } }, d.walls);
        d.exits = { rectM(10.5, 0.0, 12.5, 0.35), rectM(28.5, 0.0, 30.5, 0.35) };
        d.exitWidthM = { 2.0f, 2.0f };
        break;
    }
    case LAYOUT_BOWL: bowlDesc(d, 240, 4, 200.0); break;
    case LAYOUT_BOWL_QUARTER: bowlDesc(d, 60, 1, 56.0); break;
    case LAYOUT_BOWL_EIGHTH: bowlDesc(d, 30, 1, 32.0); break;
    case LAYOUT_CORRIDOR: {   // 60 x 20 m corridor, one 1.5 m exit in its right end wall
        setDomainM(d, 61.0, 21.0);
        roomWallsM(0.5, 0.5, 60.5, 20.5, 0.3, { { 1, 9.75, 11.25 } }, d.walls);
        d.exits = { rectM(60.65, 9.75, 61.0, 11.25) };
        d.exitWidthM = { 1.5f };
        break;
    }
    case LAYOUT_PILLAR: {   // 30 x 30 m room, central 4 x 4 m pillar, two 2 m exits in the bottom wall
        setDomainM(d, 31.0, 31.0);
        roomWallsM(0.5, 0.5, 30.5, 30.5, 0.3, { { 0, 6.5, 8.5 }, { 0, 22.5, 24.5 } }, d.walls);
        d.walls.push_back(rectM(13.5, 13.5, 17.5, 17.5));
        d.exits = { rectM(6.5, 0.0, 8.5, 0.35), rectM(22.5, 0.0, 24.5, 0.35) };
        d.exitWidthM = { 2.0f, 2.0f };
        break;
    }
    default: throw std::invalid_argument("layoutDesc: unknown layout");
    }
    return d;
}

void layoutSlots(int layout, std::vector<float>& sx, std::vector<float>& sy) {
    sx.clear(); sy.clear();
    const double c = HEX_CLEAR;
    switch (layout) {
    case LAYOUT_ROOM: hexSlots(0.5 + c, 0.5 + c, 40.5 - c, 30.5 - c, HEX_PITCH, nullptr, sx, sy); break;
    case LAYOUT_CORRIDOR: hexSlots(0.5 + c, 0.5 + c, 60.5 - c, 20.5 - c, HEX_PITCH, nullptr, sx, sy); break;
    case LAYOUT_PILLAR: {
        const Rect hole{ float(13.5 - c), float(13.5 - c), float(17.5 + c), float(17.5 + c) };
        hexSlots(0.5 + c, 0.5 + c, 30.5 - c, 30.5 - c, HEX_PITCH, &hole, sx, sy);
        break;
    }
    case LAYOUT_BOWL: bowlSlots(240, sx, sy); break;
    case LAYOUT_BOWL_QUARTER: bowlSlots(60, sx, sy); break;
    case LAYOUT_BOWL_EIGHTH: bowlSlots(30, sx, sy); break;
    default: throw std::invalid_argument("layoutSlots: unknown layout");
    }
}

Agents placeAgents(int layout, size_t active, size_t total, uint64_t seed, double speedMean, double speedSd) {
    std::vector<float> sx, sy;   // metres
    layoutSlots(layout, sx, sy);
    if (active > sx.size()) throw std::invalid_argument("placeAgents: more agents than the layout's slots");
    if (total < active) throw std::invalid_argument("placeAgents: total < active");
    Rng rng(seed);
    const bool bowl = layout == LAYOUT_BOWL || layout == LAYOUT_BOWL_QUARTER || layout == LAYOUT_BOWL_EIGHTH;
    std::vector<size_t> pick(sx.size());
    for (size_t i = 0; i < pick.size(); ++i) pick[i] = i;
    if (!bowl)   // a seeded choice of lattice sites (Fisher-Yates)
        for (size_t i = pick.size(); i > 1; --i) std::swap(pick[i - 1], pick[size_t(rng.next() % i)]);
    const double jitter = bowl ? 0.05 : HEX_JITTER;
    Agents a;
    a.x.resize(total); a.y.resize(total); a.r.resize(total); a.v0.resize(total); a.w.resize(total);
    for (size_t i = 0; i < total; ++i) {
        const double r = 0.20 + 0.05 * rng.uniform();
        double v = speedMean + speedSd * rng.normal();
        v = std::min(std::max(v, 0.5), 2.0);
        const double jx = jitter * (2.0 * rng.uniform() - 1.0), jy = jitter * (2.0 * rng.uniform() - 1.0);
        a.r[i] = mu(r);
        a.v0[i] = float(v * S_PER_UNIT / M_PER_UNIT);
        if (i < active) {
            a.x[i] = mu(double(sx[pick[i]]) + jx);
            a.y[i] = mu(double(sy[pick[i]]) + jy);
            a.w[i] = 1.0f;
        } else {
            a.x[i] = PARK_COORD;
            a.y[i] = PARK_COORD;
            a.w[i] = 0.0f;
        }
    }
    return a;
}

} // namespace swarm
