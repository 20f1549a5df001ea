// The oracle's restatement of the reference's analytic seated venue (src/venue.cpp:411-935): TEST INFRASTRUCTURE ONLY.
// The product evaluates the venue with a cell generator indexed by (band, cell); this restatement first lists every band's
// cells in a table (restating seated::cell, src/venue.cpp:522-594), then answers each query by scanning that table:
//   floorHeight   src/venue.cpp:672-679      cellAt      src/venue.cpp:620-629
//   sdfGrad       src/venue.cpp:690-751 (sdfGradStats, with nearestInCell :456-477 and nearestRegion2 :434-449)
//   wall          src/venue.cpp:753-804      floorSlope  src/venue.cpp:640-650
//   exitByLat     src/venue.cpp:842-852      gangwayAt   src/venue.cpp:876-892
// FP32 in the reference's operation order; min / max / abs / clamp with glm's and std's tie and signed-zero rules.
// Pinned: the cell table against the reference's own cells, every query against the reference's answers at probe
// points (tests/golden/venue_kat.npz), the distance and the wall query against brute force over the unfolded cells.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "venue_oracle.h"

namespace {

enum { GROUND = 0, TREAD = 1, SEAT = 2, INCLINE = 3, BARRIER = 4, WALL = 5, CONC = 6, SIDE = 7 };
constexpr float BIG = 1e9f;
constexpr float NONE = 1e9f;

inline float gmin(float x, float y) { return y < x ? y : x; }   // glm::min, std::min
inline float gmax(float x, float y) { return x < y ? y : x; }   // glm::max, std::max
inline float gabs(float x) { return x >= 0.0f ? x : -x; }       // glm::abs
inline float sclamp(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }   // std::clamp

struct OCell { float l0, l1; int kind, row; };
struct OBand { float lo, hi; std::vector<OCell> cells; };

struct OVenue {
    // the layout (the order of the reference's dump: venue/<V>/layout_f, layout_i, the arrays)
    float standFront, row0, rowDepth, clearway, concFront, concBack, outerFace, landingEnd, avenueHalf, lat0, gangwayLo, gangwayHi,
        exitLo, exitHi, cornerInset, seatInset, faceMargin, wallAbove, wallClear, firstRow, rowRise, seatTop, barrierTop, concourseZ,
        wallTop, rake, seatWidth, domainHalf, cornerDoorLo;
    int rows, maxBandSteps, gangways, doors;
    float gwLo[8], gwHi[8], doorLo[8], doorHi[8];
    int gwRow[8], gwSeated[8];
    std::vector<OBand> bands;

    float rowFront(int k) const { return row0 + float(k) * rowDepth; }
    float rowTread(int k) const { return firstRow + float(k) * rowRise; }
    int nb() const { return 6 + 2 * rows; }
    bool corner() const { return cornerDoorLo < 1e8f; }
    int active(int r) const {
        int n = 0;
        for (int i = 0; i < gangways; ++i) n += gwRow[i] <= r ? 1 : 0;
        return n;
    }
    bool merged(int r) const {
        const int n = active(r);
        return n > 0 && r < gwSeated[n - 1];
    }

    void bandRange(int j, float& lo, float& hi) const {   // seated::band
        const int n = nb();
        if (j <= 0) { lo = 0.0f; hi = standFront; }
        else if (j == 1) { lo = standFront; hi = row0; }
        else if (j < n - 4) {
            const int k = (j - 2) >> 1;
            const float d = rowFront(k);
            if (((j - 2) & 1) == 0) { lo = d; hi = d + clearway; }
            else { lo = d + clearway; hi = rowFront(k + 1); }
        } else if (j == n - 4) { lo = concFront; hi = concBack; }
        else if (j == n - 3) { lo = concBack; hi = outerFace; }
        else if (j == n - 2) { lo = outerFace; hi = landingEnd; }
        else { lo = landingEnd; hi = BIG; }
    }

    void build() {
        bands.clear();
        const int n = nb();
        for (int j = 0; j < n; ++j) {
            OBand b;
            bandRange(j, b.lo, b.hi);
            auto push = [&b](float l0, float l1, int kind, int row) { b.cells.push_back(OCell{ l0, l1, kind, row }); };
            const float hi = b.hi;
            if (j == 0 || j == n - 1) push(0.0f, hi, GROUND, -1);
            else if (j == n - 2) {   // the landing
                for (int i = 0; i < doors; ++i) {
                    push(i == 0 ? 0.0f : doorHi[i - 1], doorLo[i], GROUND, -1);
                    push(doorLo[i], doorHi[i], CONC, -1);
                }
                push(doorHi[doors - 1], corner() ? cornerDoorLo : hi, GROUND, -1);
                if (corner()) push(cornerDoorLo, hi, CONC, -1);
            } else if (j == n - 3) {   // the outer wall
                push(0.0f, avenueHalf, GROUND, -1);
                for (int i = 0; i < doors; ++i) {
                    push(i == 0 ? avenueHalf : doorHi[i - 1], doorLo[i], WALL, -1);
                    push(doorLo[i], doorHi[i], CONC, -1);
                }
                push(doorHi[doors - 1], corner() ? cornerDoorLo : hi, WALL, -1);
                if (corner()) push(cornerDoorLo, hi, CONC, -1);
            } else if (j == n - 4) {   // the concourse
                push(0.0f, avenueHalf, GROUND, -1);
                push(avenueHalf, lat0, WALL, -1);
                push(lat0, hi, CONC, -1);
            } else {   // the barrier band and the row bands
                const bool rowBand = j >= 2;
                const int row = rowBand ? (j - 2) >> 1 : -1;
                const bool seatBand = !rowBand || ((j - 2) & 1) != 0;
                const int segKind = !rowBand ? BARRIER : (seatBand ? SEAT : TREAD);
                const int gapKind = rowBand ? INCLINE : GROUND;
                const float cornerLat = rowBand && !seatBand ? b.lo - cornerInset : hi - cornerInset - seatInset;
                const float inset = seatBand ? seatInset : 0.0f;
                const int r = (j - 1) >> 1;
                const int na = active(r);
                const bool m = merged(r);
                push(0.0f, avenueHalf, GROUND, -1);
                push(avenueHalf, lat0, SIDE, row);
                for (int i = 0; i < na; ++i) {
                    const float gl = seatBand ? gwLo[i] - inset : gwLo[i];
                    const float prevHi = i == 0 ? lat0 : (seatBand ? gwHi[i - 1] + inset : gwHi[i - 1]);
                    push(prevHi, gl, segKind, row);
                    push(gl, m && i == na - 1 ? hi : (seatBand ? gwHi[i] + inset : gwHi[i]), gapKind, row);
                }
                if (!m) {
                    const float lastHi = seatBand ? gwHi[na - 1] + inset : gwHi[na - 1];
                    push(lastHi, cornerLat, segKind, row);
                    push(cornerLat, hi, gapKind, row);
                }
            }
            bands.push_back(b);
        }
    }

    float top(int kind, int row, float dep) const {   // seated::cellTop
        switch (kind) {
        case TREAD: return rowTread(row);
        case SEAT: return rowTread(row) + seatTop;
        case INCLINE: return rake * (dep - row0);
        case BARRIER: return barrierTop;
        case WALL: return wallTop;
        case CONC: return concourseZ;
        case SIDE: return row < 0 ? barrierTop : rowTread(row) + seatTop;
        default: return 0.0f;
        }
    }

    int bandOf(float dep) const {
        const int n = nb();
        if (dep < standFront) return 0;
        if (dep < row0) return 1;
        if (dep < concFront) {
            int k = int(std::floor((dep - row0) / rowDepth));
            k = k < 0 ? 0 : (k > rows - 1 ? rows - 1 : k);
            if (dep < rowFront(k)) --k;
            else if (k + 1 < rows && dep >= rowFront(k + 1)) ++k;
            return 2 + 2 * k + (dep >= rowFront(k) + clearway ? 1 : 0);
        }
        if (dep < concBack) return n - 4;
        if (dep < outerFace) return n - 3;
        if (dep < landingEnd) return n - 2;
        return n - 1;
    }

    int own(int j, float lat) const {   // the index of the cell of band j holding lat
        const std::vector<OCell>& cs = bands[size_t(j)].cells;
        for (size_t i = 0; i < cs.size(); ++i)
            if (lat < cs[i].l1) return int(i);
        return int(cs.size()) - 1;
    }
};

void foldXY(float x, float y, float& lat, float& dep) {
    const float ax = gabs(x), ay = gabs(y);
    lat = gmin(ax, ay);
    dep = gmax(ax, ay);
}
void unfold(float x, float y, float dl, float dd, float& ox, float& oy) {
    const float sx = x < 0.0f ? -1.0f : 1.0f, sy = y < 0.0f ? -1.0f : 1.0f;
    if (gabs(y) >= gabs(x)) { ox = sx * dl; oy = sy * dd; }
    else { ox = sx * dd; oy = sy * dl; }
}

void region2(float u, float z, float u0, float u1, float s, float uRef, bool below, float& bu, float& bz) {
    const float uc = sclamp(u, u0, u1);
    const float tc = s * (uc - uRef);
    if (below ? z <= tc : z >= tc) { bu = uc; bz = z; return; }
    const float t0 = s * (u0 - uRef), t1 = s * (u1 - uRef);
    const float du = u1 - u0, dz = t1 - t0;
    const float len2 = du * du + dz * dz;
    const float t = len2 > 0.0f ? sclamp(((u - u0) * du + (z - t0) * dz) / len2, 0.0f, 1.0f) : 0.0f;
    bu = u0 + t * du;
    bz = t0 + t * dz;
    float bd = (u - bu) * (u - bu) + (z - bz) * (z - bz);
    const float r0z = below ? gmin(z, t0) : gmax(z, t0), r1z = below ? gmin(z, t1) : gmax(z, t1);
    const float d0 = (u - u0) * (u - u0) + (z - r0z) * (z - r0z), d1 = (u - u1) * (u - u1) + (z - r1z) * (z - r1z);
    if (d0 < bd) { bd = d0; bu = u0; bz = r0z; }
    if (d1 < bd) { bu = u1; bz = r1z; }
}

// nearest point of a cell's solid (air == false) or air column; returns the squared distance
float nearest(const OVenue& V, const OBand& b, const OCell& c, float px, float py, float pz, bool air, float q[3]) {
    const float latC = sclamp(px, c.l0, c.l1);
    if (c.kind == INCLINE) {
        float u, z;
        region2(py, pz, b.lo, b.hi, V.rake, V.row0, !air, u, z);
        q[0] = latC; q[1] = u; q[2] = z;
        if (air && q[0] > q[1]) {
            const float s0 = gmax(c.l0, b.lo), s1 = gmin(c.l1, b.hi);
            if (s0 > s1) return BIG;
            const float r2 = 1.41421356f;
            const float m = 0.5f * (px + py);
            float wu, wz;
            region2(m * r2, pz, s0 * r2, s1 * r2, V.rake / r2, V.row0 * r2, false, wu, wz);
            const float s = wu / r2;
            q[0] = s; q[1] = s; q[2] = wz;
        }
    } else {
        const float t = V.top(c.kind, c.row, 0.0f);
        q[0] = latC; q[1] = sclamp(py, b.lo, b.hi); q[2] = air ? gmax(pz, t) : gmin(pz, t);
    }
    const float dx = px - q[0], dy = py - q[1], dz = pz - q[2];
    return dx * dx + dy * dy + dz * dz;
}

// The outward band search of sdfGradStats and wall: visit(jp), then s = 1 .. maxBandSteps below and above while the
// band's depth gap is under the best distance
template <class Visit> void search(const OVenue& V, int jp, float dep, const float& best, Visit visit) {
    visit(jp, 0.0f);
    bool open[2] = { true, true };
    for (int s = 1; s <= V.maxBandSteps && (open[0] || open[1]); ++s)
        for (int side = 0; side < 2; ++side) {
            if (!open[side]) continue;
            const int j = side == 0 ? jp - s : jp + s;
            if (j < 0 || j >= V.nb()) { open[side] = false; continue; }
            const OBand& b = V.bands[size_t(j)];
            const float g = gmax(gmax(b.lo - dep, dep - b.hi), 0.0f);
            if (g * g >= best) { open[side] = false; continue; }
            visit(j, g * g);
        }
}

void sdfGrad(const OVenue& V, float x, float y, float z, float out[4]) {
    float lat, dep;
    foldXY(x, y, lat, dep);
    const int jp = V.bandOf(dep);
    const OCell& oc = V.bands[size_t(jp)].cells[size_t(V.own(jp, lat))];
    const bool inAir = z > V.top(oc.kind, oc.row, dep);
    float best = BIG;
    float bq[3] = { lat, dep, z };
    float rect[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    int bestKind = oc.kind;
    search(V, jp, dep, best, [&](int j, float g2) {
        const OBand& b = V.bands[size_t(j)];
        for (const OCell& c : b.cells) {
            const float lg = gmax(gmax(c.l0 - lat, lat - c.l1), 0.0f);
            if (g2 + lg * lg >= best) {
                if (c.l0 > lat) break;
                continue;
            }
            float q[3];
            const float d2 = nearest(V, b, c, lat, dep, z, !inAir, q);
            if (d2 < best) {
                best = d2;
                bq[0] = q[0]; bq[1] = q[1]; bq[2] = q[2];
                rect[0] = c.l0; rect[1] = c.l1; rect[2] = b.lo; rect[3] = b.hi;
                bestKind = c.kind;
            }
        }
    });
    const float d = std::sqrt(best);
    float g[3];
    if (d > 0.0f) {
        for (int k = 0; k < 3; ++k) {
            const float pk = k == 0 ? lat : (k == 1 ? dep : z);
            g[k] = inAir ? (pk - bq[k]) / d : (bq[k] - pk) / d;
        }
    } else {
        const float sgn = inAir ? -1.0f : 1.0f;
        const float hx = lat <= rect[0] ? sgn : (lat >= rect[1] ? -sgn : 0.0f);
        const float hy = dep <= rect[2] ? sgn : (dep >= rect[3] ? -sgn : 0.0f);
        if (hx != 0.0f || hy != 0.0f) {
            const float inv = 1.0f / std::sqrt(hx * hx + hy * hy);
            g[0] = hx * inv; g[1] = hy * inv; g[2] = 0.0f;
        } else if (bestKind == INCLINE) {
            const float inv = 1.0f / std::sqrt((0.0f * 0.0f + V.rake * V.rake) + 1.0f);
            g[0] = 0.0f * inv; g[1] = -V.rake * inv; g[2] = 1.0f * inv;
        } else {
            g[0] = 0.0f; g[1] = 0.0f; g[2] = 1.0f;
        }
    }
    unfold(x, y, g[0], g[1], out[0], out[1]);
    out[2] = g[2];
    out[3] = inAir ? d : -d;
}

void wall(const OVenue& V, float x, float y, float z, float reach, float out[3]) {
    float lat, dep;
    foldXY(x, y, lat, dep);
    const int jp = V.bandOf(dep);
    const OCell& oc = V.bands[size_t(jp)].cells[size_t(V.own(jp, lat))];
    const bool inWall = oc.kind != INCLINE && V.top(oc.kind, oc.row, dep) > z;
    float best = inWall ? BIG : reach * reach;
    float bq[2] = { lat, dep };
    float rect[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    bool found = false;
    search(V, jp, dep, best, [&](int j, float) {
        const OBand& b = V.bands[size_t(j)];
        for (const OCell& c : b.cells) {
            const bool isWall = c.kind != INCLINE && V.top(c.kind, c.row, b.lo) > z;
            if (isWall == inWall) continue;
            const float qx = gmin(gmax(lat, c.l0), c.l1), qy = gmin(gmax(dep, b.lo), b.hi);
            const float dx = lat - qx, dy = dep - qy;
            const float d2 = dx * dx + dy * dy;
            if (d2 < best) {
                best = d2;
                bq[0] = qx; bq[1] = qy;
                rect[0] = c.l0; rect[1] = c.l1; rect[2] = b.lo; rect[3] = b.hi;
                found = true;
            }
        }
    });
    if (!found) { out[0] = 0.0f; out[1] = 0.0f; out[2] = NONE; return; }
    const float d = std::sqrt(best);
    float dl, dd;
    if (d > 0.0f) {
        dl = inWall ? (bq[0] - lat) / d : (lat - bq[0]) / d;
        dd = inWall ? (bq[1] - dep) / d : (dep - bq[1]) / d;
    } else {
        const float sgn = inWall ? 1.0f : -1.0f;
        const float hx = lat <= rect[0] ? sgn : (lat >= rect[1] ? -sgn : 0.0f);
        const float hy = dep <= rect[2] ? sgn : (dep >= rect[3] ? -sgn : 0.0f);
        const float inv = 1.0f / std::sqrt(hx * hx + hy * hy);
        dl = hx * inv; dd = hy * inv;
    }
    unfold(x, y, dl, dd, out[0], out[1]);
    out[2] = inWall ? -d : d;
}

int quadrant(float x, float y) { return x >= 0.0f ? (y >= 0.0f ? 0 : 3) : (y >= 0.0f ? 1 : 2); }

int exitByLat(const OVenue& V, float x, float y) {
    const float ax = gabs(x), ay = gabs(y);
    const bool alongY = ay >= ax;
    const float lat = alongY ? ax : ay;
    if (lat < V.avenueHalf) return alongY ? (y > 0.0f ? 2 : 3) : (x > 0.0f ? 0 : 1);
    const int dpq = 2 * V.doors + (V.corner() ? 1 : 0);
    const int base = 4 + quadrant(x, y) * dpq;
    for (int i = 0; i < V.doors; ++i)
        if (lat >= V.doorLo[i] && lat <= V.doorHi[i]) return base + (alongY ? 0 : V.doors) + i;
    if (V.corner() && lat >= V.cornerDoorLo) return base + 2 * V.doors;
    return -1;
}

int gangwayAt(const OVenue& V, float x, float y) {
    float lat, dep;
    foldXY(x, y, lat, dep);
    const int j = V.bandOf(dep);
    if (j < 1 || j >= V.nb() - 4) return -1;
    const int c = V.own(j, lat);
    const int kind = V.bands[size_t(j)].cells[size_t(c)].kind;
    if (!(kind == INCLINE || (j == 1 && kind == GROUND && c >= 2))) return -1;
    const int r = (j - 1) >> 1;
    const int n = V.active(r);
    const bool m = V.merged(r);
    const int i = (c - 3) >> 1;
    const bool straight = (c & 1) == 1 && i < n && !(m && i == n - 1);
    const int base = quadrant(x, y) * (2 * V.gangways + 1);
    return straight ? base + (gabs(y) >= gabs(x) ? 0 : V.gangways) + i : base + 2 * V.gangways;
}

float floorSlope(const OVenue& V, float x, float y, float ex, float ey) {   // seated::floorSlope
    float lat, dep;
    foldXY(x, y, lat, dep);
    const int j = V.bandOf(dep);
    const OCell& c = V.bands[size_t(j)].cells[size_t(V.own(j, lat))];
    if (c.kind != INCLINE) return 0.0f;
    float gx, gy;
    unfold(x, y, 0.0f, V.rake, gx, gy);
    const float a = gx * ex, b = gy * ey;
    return a + b;
}

OVenue make(const float* lf, const int* li, const float* gwLo, const float* gwHi, const int* gwRow, const int* gwSeated,
            const float* doorLo, const float* doorHi) {
    OVenue V{};
    float* dst[] = { &V.standFront, &V.row0, &V.rowDepth, &V.clearway, &V.concFront, &V.concBack, &V.outerFace, &V.landingEnd,
                     &V.avenueHalf, &V.lat0, &V.gangwayLo, &V.gangwayHi, &V.exitLo, &V.exitHi, &V.cornerInset, &V.seatInset,
                     &V.faceMargin, &V.wallAbove, &V.wallClear, &V.firstRow, &V.rowRise, &V.seatTop, &V.barrierTop, &V.concourseZ,
                     &V.wallTop, &V.rake, &V.seatWidth, &V.domainHalf, &V.cornerDoorLo };
    for (int i = 0; i < 29; ++i) *dst[i] = lf[i];
    V.rows = li[0]; V.maxBandSteps = li[1]; V.gangways = li[2]; V.doors = li[3];
    for (int i = 0; i < 8; ++i) {
        V.gwLo[i] = gwLo[i]; V.gwHi[i] = gwHi[i]; V.gwRow[i] = gwRow[i]; V.gwSeated[i] = gwSeated[i];
        V.doorLo[i] = doorLo[i]; V.doorHi[i] = doorHi[i];
    }
    V.build();
    return V;
}

}  // namespace

namespace vo {

namespace {
const OVenue& ven(const void* v) { return *static_cast<const OVenue*>(v); }
const OCell& ownCellOf(const OVenue& V, float x, float y, int& band, float& dep) {
    float lat;
    foldXY(x, y, lat, dep);
    band = V.bandOf(dep);
    return V.bands[size_t(band)].cells[size_t(V.own(band, lat))];
}
constexpr float STEP_UP_SAMPLE = 0x1.a36e2ep-11f;   // STEP_UP_SAMPLE_RADII * BODY_RADIUS folded (0.125 x 0.0064 = 0.0008)
constexpr float STEP_UP_TOL = 1e-6f;
constexpr int STEP_UP_MAX_SAMPLES = 32, SCAN_CROSSINGS = 4, SCAN_BISECTIONS = 12;
inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }   // glm::mix
}  // namespace

void sdfGrad(const void* v, float x, float y, float z, float out[4]) { ::sdfGrad(ven(v), x, y, z, out); }
void wall(const void* v, float x, float y, float z, float reach, float out[3]) { ::wall(ven(v), x, y, z, reach, out); }
float floorHeight(const void* v, float x, float y) {
    int j;
    float dep;
    const OCell& c = ownCellOf(ven(v), x, y, j, dep);
    return ven(v).top(c.kind, c.row, dep);
}
int cellAt(const void* v, float x, float y, int* row) {
    int j;
    float dep;
    const OCell& c = ownCellOf(ven(v), x, y, j, dep);
    if (row) *row = c.row;
    return c.kind;
}
void cellKey(const void* v, float x, float y, float out[2]) {
    int j;
    float dep;
    const OCell& c = ownCellOf(ven(v), x, y, j, dep);
    out[0] = float(j);
    out[1] = c.l0;
}
float floorSlope(const void* v, float x, float y, float ex, float ey) { return ::floorSlope(ven(v), x, y, ex, ey); }
int exitByLat(const void* v, float x, float y) { return ::exitByLat(ven(v), x, y); }
float rake(const void* v) { return ven(v).rake; }
float wallAbove(const void* v) { return ven(v).wallAbove; }
float wallClear(const void* v) { return ven(v).wallClear; }
float faceMargin(const void* v) { return ven(v).faceMargin; }

float stepUpScan(const void* v, float p0x, float p0y, float p1x, float p1y, float f0, float n[2]) {
    n[0] = 0.0f;
    n[1] = 0.0f;
    const float dx = p1x - p0x, dy = p1y - p0y;
    const float len = std::sqrt(dx * dx + dy * dy);
    const int ns = std::min(std::max(int(std::ceil(len / STEP_UP_SAMPLE)), 1), STEP_UP_MAX_SAMPLES);
    float lo = 0.0f;
    float key[2], k2[2];
    cellKey(v, p0x, p0y, key);
    auto at = [&](float t, float& x, float& y) { x = mixf(p0x, p1x, t); y = mixf(p0y, p1y, t); };
    auto same = [&](float t) {
        float x, y;
        at(t, x, y);
        cellKey(v, x, y, k2);
        return k2[0] == key[0] && k2[1] == key[1];
    };
    for (int i = 1; i <= ns; ++i) {
        const float t = float(i) / float(ns);
        for (int k = 0; k < SCAN_CROSSINGS; ++k) {
            if (same(t)) break;
            float a = lo, b = t;
            for (int m = 0; m < SCAN_BISECTIONS; ++m) {
                const float tm = 0.5f * (a + b);
                if (same(tm)) a = tm;
                else b = tm;
            }
            float ax, ay, bx, by;
            at(a, ax, ay);
            at(b, bx, by);
            const float ha = floorHeight(v, ax, ay), hb = floorHeight(v, bx, by);
            if (hb > f0 + STEP_UP_TOL && hb - ha > rake(v) * len * (b - a) + STEP_UP_TOL) {
                float w[3];
                ::wall(ven(v), ax, ay, ha + STEP_UP_TOL, STEP_UP_SAMPLE, w);
                if (w[2] < NONE) { n[0] = w[0]; n[1] = w[1]; }
                else {
                    const float l = std::max(len, 1e-30f);
                    n[0] = (p0x - p1x) / l;
                    n[1] = (p0y - p1y) / l;
                }
                return a;
            }
            lo = b;
            cellKey(v, bx, by, key);
        }
        lo = t;
    }
    return 1.0f;
}

}  // namespace vo

extern "C" {

// The cell table: cells_i (band, index, kind, row) and cells_f (l0, l1, lo, hi), at most cap rows; returns the row count
int oracle_venue_cells(const float* lf, const int* li, const float* gwLo, const float* gwHi, const int* gwRow, const int* gwSeated,
                       const float* doorLo, const float* doorHi, int cap, int* cellsI, float* cellsF) {
    const OVenue V = make(lf, li, gwLo, gwHi, gwRow, gwSeated, doorLo, doorHi);
    int n = 0;
    for (size_t j = 0; j < V.bands.size(); ++j)
        for (size_t c = 0; c < V.bands[j].cells.size(); ++c, ++n) {
            if (n >= cap) continue;
            const OCell& k = V.bands[j].cells[c];
            cellsI[4 * n] = int(j); cellsI[4 * n + 1] = int(c); cellsI[4 * n + 2] = k.kind; cellsI[4 * n + 3] = k.row;
            cellsF[4 * n] = k.l0; cellsF[4 * n + 1] = k.l1; cellsF[4 * n + 2] = V.bands[j].lo; cellsF[4 * n + 3] = V.bands[j].hi;
        }
    return n;
}

// Every query at n points p (n x 3) with directions e (n x 2): grad n x 4, floor n, cell n x 3, wall n x 9 (three
// reaches), slope n, exit n, gangway n
void oracle_venue_eval(const float* lf, const int* li, const float* gwLo, const float* gwHi, const int* gwRow, const int* gwSeated,
                       const float* doorLo, const float* doorHi, int n, const float* p, const float* e, const float* reach3,
                       float* grad, float* floorH, int* cell, float* wallOut, float* slope, int* exitId, int* gangway) {
    const OVenue V = make(lf, li, gwLo, gwHi, gwRow, gwSeated, doorLo, doorHi);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        const float x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
        sdfGrad(V, x, y, z, grad + 4 * i);
        float lat, dep;
        foldXY(x, y, lat, dep);
        const int j = V.bandOf(dep);
        const OCell& c = V.bands[size_t(j)].cells[size_t(V.own(j, lat))];
        floorH[i] = V.top(c.kind, c.row, dep);
        cell[3 * i] = c.kind; cell[3 * i + 1] = c.row; cell[3 * i + 2] = j;
        for (int k = 0; k < 3; ++k) wall(V, x, y, z, reach3[k], wallOut + 9 * i + 3 * k);
        slope[i] = floorSlope(V, x, y, e[2 * i], e[2 * i + 1]);
        exitId[i] = exitByLat(V, x, y);
        gangway[i] = gangwayAt(V, x, y);
    }
}

// A venue for the step's incline rules (oracle_scene.venue): made once, read by oracle_venue_slope, freed by
// oracle_venue_free
void* oracle_venue_new(const float* lf, const int* li, const float* gwLo, const float* gwHi, const int* gwRow, const int* gwSeated,
                       const float* doorLo, const float* doorHi) {
    return new OVenue(make(lf, li, gwLo, gwHi, gwRow, gwSeated, doorLo, doorHi));
}
void oracle_venue_free(void* v) { delete static_cast<OVenue*>(v); }
float oracle_venue_slope(const void* v, float x, float y, float ex, float ey) {
    return floorSlope(*static_cast<const OVenue*>(v), x, y, ex, ey);
}

}  // extern "C"
