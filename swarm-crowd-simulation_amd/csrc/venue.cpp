// The reference's analytic seated venues S, M and L on the host (src/venue.cpp:20-340, 1019-1083): the parameter sets,
// the layout generator with its design rules, the seats, and the venue's plan as a rectangle world.
//
// The generator's arithmetic is the reference's (metres converted to model units in double and rounded to float once,
// the row fronts accumulated in float as the shader computes them, the fence's face and the wall clearance folded
// from their decimal literals): the layout's floats decide every cell boundary, so they are reproduced bit for bit.
#include "venue.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "swarm_params.h"

namespace swarm {

namespace {

constexpr double SQRT2 = 1.4142135623730951;
constexpr float FENCE_HALF_OUTER = 0.335432f;   // the pitch fence's measured outer face (sim_params.hpp:316)
constexpr float BODY_RADIUS_M = 0.20f;
constexpr float MAX_STEP_UP_M = 0.30f;
constexpr float WALL_MAX_NZ = 0.6f;
constexpr int TERRACE_GRID_RES = 40;            // the terrace's neighbour grid: 40 cells over [-1, 1]
constexpr float TERRACE_DOMAIN_MIN = -1.0f, TERRACE_DOMAIN_SPAN = 2.0f;

// The double of f's shortest round-trip decimal: the value the reference's shader compiler folds a literal to
double decimalOf(float f) {
    if (f == 0.0f) return 0.0;
    char buf[48];
    for (int digits = 1; digits <= 9; ++digits) {
        std::snprintf(buf, sizeof buf, "%.*e", digits - 1, double(f));
        const double d = std::strtod(buf, nullptr);
        if (float(d) == f) return d;
    }
    return double(f);
}

double mu(double metres) { return metres / double(WORLD_SCALE_M_PER_UNIT); }
double metresOf(double units) { return units * double(WORLD_SCALE_M_PER_UNIT); }

VenueParams paramsS() {
    VenueParams p;
    p.name = "S";
    p.walkwayWidthM = 2.63; p.barrierThicknessM = 0.30; p.avenueWidthM = 5.66; p.sideWallThicknessM = 0.30;
    p.rows = 17;
    p.rowDepthM = 0.80; p.clearwayM = 0.50; p.seatWidthM = 0.50; p.gangwayWidthM = 1.20; p.seatBlockInsetM = 0.20;
    p.maxSeatsBetweenGangways = 28; p.maxSeatsDeadEnd = 14;
    p.concourseWidthM = 3.00; p.outerWallThicknessM = 0.30; p.landingDepthM = 0.50;
    p.rowRiseM = 0.40; p.seatBlockTopM = 0.85; p.barrierTopM = 1.10; p.wallTopAboveConcourseM = 1.10;
    p.exitsPerQuadrant = 2; p.exitRatePerMMin = 66.0; p.egressTimeMin = 8.0; p.minExitWidthM = 1.20;
    p.doorPerGangwayHead = false;
    p.gridRes = 0;
    p.metricGridRes = 0;
    return p;
}

// M and L: S's plan and rows, more rows, a door at every gangway head, the grids sized to the venue
VenueParams generated(const char* name, int rows, int gridRes, int metricGridRes) {
    VenueParams p = paramsS();
    p.name = name;
    p.rows = rows;
    p.exitsPerQuadrant = 0;
    p.doorPerGangwayHead = true;
    p.gridRes = gridRes;
    p.metricGridRes = metricGridRes;
    return p;
}

struct Segment { float a, b; };

// Row k's seat segment seg: in front of the seat blocks of the band before it, which end seatInset short of the gangways
Segment segmentOf(const seated::Layout& L, int k, int seg) {
    const int n = seated::activeGangways(L, k);
    if (seg == 0) return { L.lat0, L.gangwayLo - L.seatInset };
    if (seg < n) return { L.gwHi[seg - 1] + L.seatInset, L.gwLo[seg] - L.seatInset };
    return { L.gwHi[n - 1] + L.seatInset, seated::rowFront(L, k) - L.cornerInset - L.seatInset };
}
int segmentsOf(const seated::Layout& L, int k) { return seated::activeGangways(L, k) + (seated::mergedCorner(L, k) ? 0 : 1); }

struct RowCount { int deadEnd, between, total; };
RowCount rowCount(const seated::Layout& L, int k) {
    RowCount rc{ 0, 0, 0 };
    for (int seg = 0; seg < segmentsOf(L, k); ++seg) {
        const Segment s = segmentOf(L, k, seg);
        const int n = seatCountIn(s.a, s.b, L.seatWidth);
        if (seg == 0) rc.deadEnd = n;
        else rc.between = std::max(rc.between, n);
        rc.total += n;
    }
    return rc;
}

void fail(const VenueParams& p, const std::string& what) { throw std::runtime_error("seated venue " + p.name + ": " + what); }

void buildLayout(const VenueParams& p, float domainHalf, Venue& v) {
    seated::Layout& L = v.layout;
    L = seated::Layout{};
    // depth breakpoints; the row fronts and the concourse front in float, row0 + float(k) rowDepth
    L.standFront = float(decimalOf(FENCE_HALF_OUTER) + mu(p.walkwayWidthM));
    L.row0 = float(double(L.standFront) + mu(p.barrierThicknessM));
    L.rowDepth = float(mu(p.rowDepthM));
    L.clearway = float(mu(p.clearwayM));
    L.rows = p.rows;
    L.concFront = L.row0 + float(L.rows) * L.rowDepth;
    L.concBack = float(double(L.concFront) + mu(p.concourseWidthM));
    L.outerFace = float(double(L.concBack) + mu(p.outerWallThicknessM));
    L.landingEnd = float(double(L.outerFace) + mu(p.landingDepthM));
    // lateral breakpoints: the straight gangway after a dead-end segment of exactly maxSeatsDeadEnd seats; the corner
    // gangway's narrowest clear width (between two mirrored seat blocks' front corners) is the gangway width
    L.avenueHalf = float(mu(0.5 * p.avenueWidthM));
    L.lat0 = float(double(L.avenueHalf) + mu(p.sideWallThicknessM));
    L.seatInset = float(mu(p.seatBlockInsetM));
    L.faceMargin = float(mu(0.3e-3));
    L.wallAbove = BODY_RADIUS;
    L.wallClear = float(decimalOf(BODY_RADIUS) + decimalOf(L.faceMargin));
    const double gangwayCentre = double(L.lat0) + mu(p.maxSeatsDeadEnd * p.seatWidthM + p.seatBlockInsetM + 0.5 * p.gangwayWidthM);
    const double halfGang = mu(0.5 * p.gangwayWidthM);
    L.gangwayLo = float(gangwayCentre - halfGang);
    L.gangwayHi = float(gangwayCentre + halfGang);
    L.cornerInset = float(mu(p.rowDepthM - p.clearwayM + p.gangwayWidthM / SQRT2 - p.seatBlockInsetM));
    // heights: the incline through every clearway's back edge meets the walkway at row 0's front
    L.rowRise = float(mu(p.rowRiseM));
    L.rake = float(p.rowRiseM / p.rowDepthM);
    L.firstRow = float(mu(p.rowRiseM / p.rowDepthM * p.clearwayM));
    L.seatTop = float(mu(p.seatBlockTopM));
    L.barrierTop = float(mu(p.barrierTopM));
    L.concourseZ = L.rake * (L.concFront - L.row0);
    L.wallTop = float(double(L.concourseZ) + mu(p.wallTopAboveConcourseM));
    L.seatWidth = float(mu(p.seatWidthM));
    L.maxBandSteps = 4;   // 4 bands either side reach 2 clearways + 2 seat blocks (1.6 m)
    L.domainHalf = domainHalf;

    // straight gangways: gangway i + 1 exactly maxSeatsBetweenGangways seats beyond gangway i, from the first row front
    // before the corner gangway's start; one no row reaches is not built
    const double pitch = mu(p.maxSeatsBetweenGangways * p.seatWidthM + 2.0 * p.seatBlockInsetM + p.gangwayWidthM);
    auto frontCorner = [&L](int r) { return seated::rowFront(L, r) - L.cornerInset; };
    L.gangways = 0;
    for (int i = 0; i < seated::MAX_GANGWAYS; ++i) {
        const double c = gangwayCentre + double(i) * pitch;
        const float lo = i == 0 ? L.gangwayLo : float(c - halfGang), hi = i == 0 ? L.gangwayHi : float(c + halfGang);
        int r = 0;
        while (r <= L.rows && !(lo < frontCorner(r))) ++r;
        if (r >= L.rows) break;
        L.gwLo[i] = lo;
        L.gwHi[i] = hi;
        L.gwRow[i] = r;
        ++L.gangways;
    }
    if (L.gangways == 0 || L.gwRow[0] != 0) fail(p, "the first straight gangway must cross every row (reach the walkway)");
    if (L.gangways == seated::MAX_GANGWAYS) {
        const double c = gangwayCentre + double(seated::MAX_GANGWAYS) * pitch;
        if (float(c - halfGang) < frontCorner(L.rows - 1)) fail(p, "more straight gangways than the layout holds");
    }
    // while gangway i is a row front's last, the first row front whose segment after it holds a seat
    for (int i = 0; i < L.gangways; ++i) {
        const int rEnd = i + 1 < L.gangways ? L.gwRow[i + 1] : L.rows + 1;
        int r = L.gwRow[i];
        while (r < rEnd && seatCountIn(L.gwHi[i] + L.seatInset, frontCorner(r) - L.seatInset, L.seatWidth) < 1) ++r;
        L.gwSeated[i] = r;
    }

    // concourse exits, sized by the capacity rule: the total exit width clears every seat within the egress time
    long total = 0;
    for (int k = 0; k < L.rows; ++k) total += 8L * rowCount(L, k).total;
    v.seatTotal = int(total);
    const double avenuesM = 4.0 * p.avenueWidthM;
    v.requiredExitWidthM = double(total) / (p.exitRatePerMMin * p.egressTimeMin);
    if (!p.doorPerGangwayHead) {
        // one door per wing on the straight gangway; per quadrant the doors alone clear the quadrant's seats, rounded up
        // to 0.1 m, at least minExitWidthM each
        if (p.exitsPerQuadrant != 2) fail(p, "one concourse exit per wing (exitsPerQuadrant = 2) is the only door-per-wing layout");
        const double perQuadrant = double(total) / 4.0;
        const double needM = perQuadrant / (p.exitRatePerMMin * p.egressTimeMin) / p.exitsPerQuadrant;
        const double exitM = std::max(p.minExitWidthM, std::ceil(needM * 10.0 - 1e-9) / 10.0);
        L.exitLo = float(gangwayCentre - mu(0.5 * exitM));
        L.exitHi = float(gangwayCentre + mu(0.5 * exitM));
        L.doors = 1;
        L.doorLo[0] = L.exitLo;
        L.doorHi[0] = L.exitHi;
        L.cornerDoorLo = 1e9f;
        v.doorWidthM = exitM;
        v.cornerDoorWidthM = 0.0;
        v.totalExitWidthM = 8.0 * exitM + avenuesM;
    } else {
        // a door at every gangway head, all of one width: the width the avenues leave, split evenly, rounded up to 0.1 m,
        // at least the gangway's; the corner door through the outer wall's corner
        if (L.gangways > seated::MAX_DOORS) fail(p, "more straight gangways than doors per wing");
        const int doorCount = 4 * (2 * L.gangways + 1);
        const double needM = std::max(0.0, v.requiredExitWidthM - avenuesM) / doorCount;
        const double w = std::max(std::max(p.minExitWidthM, p.gangwayWidthM), std::ceil(needM * 10.0 - 1e-9) / 10.0);
        L.doors = L.gangways;
        for (int i = 0; i < L.doors; ++i) {
            const double c = gangwayCentre + double(i) * pitch;
            L.doorLo[i] = float(c - mu(0.5 * w));
            L.doorHi[i] = float(c + mu(0.5 * w));
        }
        L.exitLo = L.doorLo[0];
        L.exitHi = L.doorHi[0];
        L.cornerDoorLo = float(double(L.concBack) - mu(w) / SQRT2);
        v.doorWidthM = w;
        v.cornerDoorWidthM = SQRT2 * metresOf(double(L.concBack) - double(L.cornerDoorLo));
        v.totalExitWidthM = 8.0 * L.doors * w + 4.0 * v.cornerDoorWidthM + avenuesM;
        if (L.doorLo[0] <= L.lat0) fail(p, "the first door must lie beside the stand (past the side wall)");
        if (L.doorHi[L.doors - 1] >= L.cornerDoorLo) fail(p, "the last straight door overlaps the corner door");
    }

    // design rules
    if (p.rows < 1 || p.rowDepthM <= p.clearwayM) fail(p, "rows and a seat block behind the clearway are required");
    if (p.clearwayM < 2.0 * BODY_RADIUS_M || p.seatWidthM < 2.0 * BODY_RADIUS_M) fail(p, "clearway and seat pitch must be at least a body wide");
    if (p.seatBlockTopM - p.rowRiseM <= MAX_STEP_UP_M) fail(p, "a seat block must stand more than the step-up limit above the next row's tread");
    if (p.barrierTopM - p.rowRiseM / p.rowDepthM * p.clearwayM <= MAX_STEP_UP_M) fail(p, "the front barrier must stand more than the step-up limit above row 0");
    if (p.seatBlockTopM - p.rowRiseM <= 2.0 * BODY_RADIUS_M || p.barrierTopM <= 2.0 * BODY_RADIUS_M + p.rowRiseM / p.rowDepthM * p.clearwayM)
        fail(p, "seat blocks and the barrier must reach a body diameter above every floor beside them");
    if (p.rowRiseM / p.rowDepthM * (p.clearwayM + BODY_RADIUS_M) > 2.0 * BODY_RADIUS_M)
        fail(p, "a tread face must not reach a body diameter within a body radius of its corner");
    if (1.0 / std::sqrt(1.0 + double(L.rake) * L.rake) <= WALL_MAX_NZ) fail(p, "the gangway rake must be walkable");
    if (L.gangwayHi + 2.0f * L.seatInset > L.row0 - L.cornerInset) fail(p, "the straight gangway must cross every row (reach the walkway)");
    if (float(p.seatBlockInsetM) < BODY_RADIUS_M) fail(p, "seat blocks must end at least a body radius short of the gangways");
    if (p.maxSeatsDeadEnd * p.seatWidthM < p.seatWidthM) fail(p, "empty dead-end segment");
    for (int k = 0; k < L.rows; ++k) {
        const RowCount rc = rowCount(L, k);
        if (rc.deadEnd > p.maxSeatsDeadEnd) fail(p, "a dead-end segment exceeds maxSeatsDeadEnd");
        if (rc.between > p.maxSeatsBetweenGangways) fail(p, "a segment between gangways exceeds maxSeatsBetweenGangways");
        if (rc.between < 1) fail(p, "a row without seats between its gangways");
    }
    for (int i = 1; i < L.gangways; ++i)
        if (L.gwSeated[i - 1] > L.gwRow[i]) fail(p, "a straight gangway starts before the previous one's last segment holds a seat");
    if (L.landingEnd >= domainHalf) fail(p, "the venue does not fit the domain");
}

// A folded rectangle [l0, l1] x [lo, hi] (lat x dep, lat <= dep) as plan rectangles in all eight octants. A cell that
// reaches the diagonal (l1 >= hi) covers, with its mirror, [l0, hi] x [lo, hi] in each wing; one that stops short of
// it (l1 <= lo) is a plain rectangle; one that crosses it inside the band is not a rectangle and is refused.
void unfold(float l0, float l1, float lo, float hi, std::vector<Rect>& out) {
    if (l1 > lo && l1 < hi) throw std::logic_error("venuePlan: a cell crosses the diagonal inside its band");
    const float b = l1 >= hi ? hi : l1;
    if (!(b > l0) || !(hi > lo)) return;
    for (int q = 0; q < 4; ++q) {
        const bool nx = q == 1 || q == 2, ny = q >= 2;
        auto span = [](bool neg, float a, float c, float& r0, float& r1) { if (neg) { r0 = -c; r1 = -a; } else { r0 = a; r1 = c; } };
        Rect ry{}, rx{};
        span(nx, l0, b, ry.lox, ry.hix);   // the y wing: lat along x
        span(ny, lo, hi, ry.loy, ry.hiy);
        span(nx, lo, hi, rx.lox, rx.hix);  // the x wing: lat along y
        span(ny, l0, b, rx.loy, rx.hiy);
        out.push_back(ry);
        out.push_back(rx);
    }
}

}  // namespace

int seatCountIn(float a, float b, float w) { return b > a ? int(std::floor(double(b - a) / double(w) + 1e-4)) : 0; }

Venue makeVenue(const VenueParams& p) {
    Venue v{};
    v.params = p;
    v.gridRes = p.gridRes > 0 ? p.gridRes : TERRACE_GRID_RES;
    if (p.gridRes > 0) {
        v.domainSpan = float(double(p.gridRes) * (double(TERRACE_DOMAIN_SPAN) / double(TERRACE_GRID_RES)));
        v.domainMin = -0.5f * v.domainSpan;
    } else {
        v.domainMin = TERRACE_DOMAIN_MIN;
        v.domainSpan = TERRACE_DOMAIN_SPAN;
    }
    v.metricGridRes = p.metricGridRes > 0 ? p.metricGridRes : 64;
    buildLayout(p, v.domainMin + v.domainSpan, v);
    v.exitDepth = v.layout.outerFace;
    v.exitMaxZ = v.layout.wallTop < 0.5f ? 1.0f : float(std::ceil(double(v.layout.wallTop)) + 1.0);
    return v;
}

const Venue& venueByName(const std::string& name) {
    static const Venue S = makeVenue(paramsS());
    static const Venue M = makeVenue(generated("M", 43, 76, 128));   // 15,832 seats
    static const Venue L = makeVenue(generated("L", 95, 128, 208));  // 63,424 seats
    if (name == "S" || name == "s") return S;
    if (name == "M" || name == "m") return M;
    if (name == "L" || name == "l") return L;
    throw std::invalid_argument("unknown venue '" + name + "' (S | M | L)");
}

std::vector<VenueSeat> venueSeats(const seated::Layout& L) {
    std::vector<VenueSeat> out;
    for (int q = 0; q < 4; ++q) {
        const float sx = (q == 1 || q == 2) ? -1.0f : 1.0f, sy = q >= 2 ? -1.0f : 1.0f;
        for (int wing = 0; wing < 2; ++wing)
            for (int k = 0; k < L.rows; ++k)
                for (int seg = 0; seg < segmentsOf(L, k); ++seg) {
                    const Segment s = segmentOf(L, k, seg);
                    const int n = seatCountIn(s.a, s.b, L.seatWidth);
                    const double off = 0.5 * (double(s.b - s.a) - n * double(L.seatWidth));   // centred in the segment
                    const float dep = seated::rowFront(L, k) + 0.5f * L.clearway;
                    const float z = seated::rowTread(L, k) + BODY_RADIUS;
                    for (int i = 0; i < n; ++i) {
                        const float lat = float(double(s.a) + off + (i + 0.5) * double(L.seatWidth));
                        VenueSeat st;
                        st.x = wing == 0 ? sx * lat : sx * dep;
                        st.y = wing == 0 ? sy * dep : sy * lat;
                        st.z = z;
                        st.lat = lat;
                        st.quadrant = q;
                        st.wing = wing;
                        st.row = k;
                        st.segment = seg;
                        st.index = i;
                        out.push_back(st);
                    }
                }
    }
    return out;
}

double venueExitWidthM(const Venue& v, int e) {
    const seated::Layout& L = v.layout;
    if (e >= 0 && e < 4) return metresOf(2.0 * double(L.avenueHalf));
    if (e < 4 || e >= seated::exitCount(L)) return 0.0;
    const int d = (e - 4) % seated::doorsPerQuadrant(L);
    if (d == 2 * L.doors) return v.cornerDoorWidthM;
    const int i = d % L.doors;
    return metresOf(double(L.doorHi[i]) - double(L.doorLo[i]));
}

VenuePlan venuePlan(const Venue& v) {
    const seated::Layout& L = v.layout;
    VenuePlan plan;
    SceneDesc& d = plan.desc;
    const float of = L.outerFace, le = L.landingEnd;
    d.xmin = -of; d.ymin = -of; d.xmax = of; d.ymax = of;
    d.metresPerUnit = M_PER_UNIT;
    d.navBody = BODY_RADIUS;
    d.hasVenue = true;   // the step's incline rules read the layout
    d.venue = L;
    // the navigation grid: 0.1 m cells, symmetric about 0, reaching the landings
    d.nav.cell = 0.0032f;
    d.nav.invCell = float(1.0 / double(d.nav.cell));
    const int half = int(std::ceil(double(le) / double(d.nav.cell)));
    d.nav.resX = d.nav.resY = 2 * half;
    d.nav.originX = d.nav.originY = float(-double(half) * double(d.nav.cell));

    // walls: the seat blocks, the barrier, the side and outer walls; the fenced pitch
    const int nb = seated::bandCount(L);
    for (int j = 1; j < nb - 1; ++j) {
        float lo, hi;
        seated::band(L, j, lo, hi);
        seated::Cell c{};
        for (int i = 0; seated::cellAt(L, j, i, c); ++i) {
            if (c.kind == seated::SEAT || c.kind == seated::BARRIER || c.kind == seated::WALL || c.kind == seated::SIDE)
                unfold(c.l0, c.l1, lo, hi, d.walls);
        }
    }
    d.walls.push_back(Rect{ -FENCE_HALF_OUTER, -FENCE_HALF_OUTER, FENCE_HALF_OUTER, FENCE_HALF_OUTER });

    // exits beyond the outer face (agents leave on crossing it: the domain's edge), in the reference's order
    const float past = std::nextafter(of, 2.0f * le);
    auto add = [&](Rect r, int ref, double widthM) {
        d.exits.push_back(r);
        d.exitWidthM.push_back(float(widthM));
        plan.refExit.push_back(ref);
    };
    const float ah = L.avenueHalf;
    add(Rect{ past, -ah, le, ah }, 0, venueExitWidthM(v, 0));
    add(Rect{ -le, -ah, -past, ah }, 1, venueExitWidthM(v, 1));
    add(Rect{ -ah, past, ah, le }, 2, venueExitWidthM(v, 2));
    add(Rect{ -ah, -le, ah, -past }, 3, venueExitWidthM(v, 3));
    auto signedSpan = [](bool neg, float a, float b, float& r0, float& r1) { if (neg) { r0 = -b; r1 = -a; } else { r0 = a; r1 = b; } };
    const int dpq = seated::doorsPerQuadrant(L);
    for (int q = 0; q < 4; ++q) {
        const bool nx = q == 1 || q == 2, ny = q >= 2;
        const int base = 4 + q * dpq;
        for (int wing = 0; wing < 2; ++wing)
            for (int i = 0; i < L.doors; ++i) {
                Rect r{};
                if (wing == 0) {   // the y wall: the door's lat along x
                    signedSpan(nx, L.doorLo[i], L.doorHi[i], r.lox, r.hix);
                    signedSpan(ny, past, le, r.loy, r.hiy);
                } else {
                    signedSpan(nx, past, le, r.lox, r.hix);
                    signedSpan(ny, L.doorLo[i], L.doorHi[i], r.loy, r.hiy);
                }
                const int ref = base + wing * L.doors + i;
                add(r, ref, venueExitWidthM(v, ref));
            }
        if (seated::hasCornerDoor(L)) {   // its y-wall half (the other half follows the reference's exits)
            Rect r{};
            signedSpan(nx, L.cornerDoorLo, le, r.lox, r.hix);
            signedSpan(ny, past, le, r.loy, r.hiy);
            add(r, base + 2 * L.doors, 0.5 * v.cornerDoorWidthM);
        }
    }
    if (seated::hasCornerDoor(L))
        for (int q = 0; q < 4; ++q) {
            const bool nx = q == 1 || q == 2, ny = q >= 2;
            Rect r{};
            signedSpan(nx, past, le, r.lox, r.hix);
            signedSpan(ny, L.cornerDoorLo, le, r.loy, r.hiy);
            add(r, 4 + q * dpq + 2 * L.doors, 0.5 * v.cornerDoorWidthM);
        }
    return plan;
}

CellTableData makeCellTable(const seated::Layout& L) {
    CellTableData t;
    const int nb = seated::bandCount(L);
    for (int j = 0; j < nb; ++j) {
        const int a = int(t.l1.size());
        t.start.push_back(a);
        seated::Cell c{};
        for (int i = 0; seated::cellAt(L, j, i, c); ++i) {
            if (!(c.l0 <= c.l1) || (i > 0 && !(c.l0 >= t.l1.back())))
                throw std::logic_error("makeCellTable: band " + std::to_string(j) + " is not sorted at cell " + std::to_string(i));
            t.l1.push_back(c.l1);
            t.cell.push_back(make_int4(__builtin_bit_cast(int, c.l0), __builtin_bit_cast(int, c.l1), c.kind, c.row));
        }
        const int last = int(t.l1.size()) - 1;
        if (last < a) throw std::logic_error("makeCellTable: band " + std::to_string(j) + " has no cell");
        // buckets over [0, the last cell's l1): each starts at the first cell reaching past its left edge
        const float w = t.l1[size_t(last)];
        const float invW = w > 0.0f && std::isfinite(w) ? float(seated::CELL_BUCKETS) / w : 0.0f;
        t.invW.push_back(invW);
        for (int b = 0; b < seated::CELL_BUCKETS; ++b) {
            const float left = invW > 0.0f ? float(b) / invW : 0.0f;
            int i = a;
            while (i < last && !(left < t.l1[size_t(i)])) ++i;
            t.bucket.push_back(uint16_t(i));
        }
    }
    t.start.push_back(int(t.l1.size()));
    if (t.l1.size() >= 65536) throw std::logic_error("makeCellTable: 65,536 cells or more");
    // the blob: start, invW, l1, bucket, cell; each array at a 16-byte boundary
    auto add = [&t](const void* p, size_t bytes) {
        const int o = int(t.blob.size());
        t.blob.insert(t.blob.end(), static_cast<const char*>(p), static_cast<const char*>(p) + bytes);
        t.blob.resize((t.blob.size() + 15) & ~size_t(15), 0);
        return o;
    };
    t.blobLayout.oStart = add(t.start.data(), t.start.size() * sizeof(int));
    t.blobLayout.oInvW = add(t.invW.data(), t.invW.size() * sizeof(float));
    t.blobLayout.oL1 = add(t.l1.data(), t.l1.size() * sizeof(float));
    t.blobLayout.oBucket = add(t.bucket.data(), t.bucket.size() * sizeof(uint16_t));
    t.blobLayout.oCell = add(t.cell.data(), t.cell.size() * sizeof(int4));
    t.blobLayout.bytes = int(t.blob.size());
    return t;
}

NavGrid venueNavGrid(const Venue& v) {   // nav::venueGrid (src/nav_field.cpp:41-52): 0.05 m cells over the domain
    const double cell = 0.05 / M_PER_UNIT;
    NavGrid g;
    g.resX = g.resY = int(std::lround(double(v.domainSpan) / cell));
    if (float(double(g.resX) * cell) != v.domainSpan) throw std::logic_error("venueNavGrid: the domain is not a whole number of cells");
    g.originX = g.originY = v.domainMin;
    g.cell = float(double(v.domainSpan) / g.resX);
    g.invCell = float(double(g.resX) / double(v.domainSpan));
    return g;
}

Scene buildSeatedScene(const Venue& v, int threads) {
    const auto t0 = std::chrono::steady_clock::now();
    const seated::Layout& L = v.layout;
    Scene s;
    s.xmin = s.ymin = v.domainMin;
    s.xmax = s.ymax = v.domainMin + v.domainSpan;
    s.hashSpanX = s.hashSpanY = v.domainSpan;
    s.hashNX = s.hashNY = v.gridRes;
    // the predict ring covers R_PERC (sim_params.hpp:516); the projection's 3 x 3 ring is the reference's (its grid
    // cell bounds a step's moves, sim_params.hpp:533-534)
    s.percRing = int(R_PERC / (v.domainSpan / float(v.gridRes))) + 1;
    s.projRing = 1;
    s.maxRadius = BODY_RADIUS;
    s.maxV0 = V_MAX;
    // no distance grid: the heightfield is evaluated exactly
    s.sdfNX = s.sdfNY = 2;
    s.sdfOX = s.sdfOY = 0.0f;
    s.sdfH = s.sdfInvH = 1.0f;
    s.sdf.assign(4, 1.0f);
    s.sdfMin.assign(4, 1.0f);
    s.sdfMinWin = 0;
    s.nav = venueNavGrid(v);
    const NavField f = buildSeatedNavGrid(L, s.nav, threads);
    const int E = seated::exitCount(L);
    std::vector<double> widthM(static_cast<size_t>(E));
    std::vector<int> type(static_cast<size_t>(E));
    for (int e = 0; e < E; ++e) {
        widthM[size_t(e)] = venueExitWidthM(v, e);
        type[size_t(e)] = e < 4 ? 0 : 1;   // the avenues, the doors (SeatedGeometry::exitType)
    }
    buildCandidateTable(f, widthM, type, M_PER_UNIT, threads, s);
    s.stair.assign(s.nav.cells(), 1.0f);
    s.wallBinStart.assign(size_t(s.hashNX) * size_t(s.hashNY) + 1, 0u);
    s.hasVenue = 1;
    s.venue = L;
    s.venue3d = 1;
    s.exitDepth = v.exitDepth;
    s.exitMaxZ = v.exitMaxZ;
    s.domainMax = v.domainMin + v.domainSpan;
    s.buildSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return s;
}

}  // namespace swarm
