// The analytic seated venue (the reference's venues S, M and L): a square stand folded eight ways, whose floor is a
// heightfield of columns ("cells") over bands of depth. This header is its evaluation, on the GPU and on the host:
// the cell enumeration of the folded frame, the floor height, the exact signed distance to the heightfield and its
// gradient, the nearest wall in a horizontal plane, the incline slope, and the exit and gangway of a plan point.
//
// Reference: src/venue.cpp:411-935 (seated::band / bandOf / cell / cellTop / fold / unfoldDir / floorHeight /
// sdfGradStats / wall / floorSlope / exitByLat / gangwayAt) and shaders/venue_seated.glsl, which evaluates the same
// functions inside physics.comp. Every value is FP32 in the reference's operation order (its results are reproduced bit
// for bit); min / max / abs / clamp follow the reference's semantics for ties and signed zeros (glm::min(x, y) =
// y < x ? y : x, glm::max(x, y) = x < y ? y : x, glm::abs(x) = x >= 0 ? x : -x).
//
// The folded frame: lat = min(|x|, |y|), dep = max(|x|, |y|). Bands along dep: 0 the walkway (and the pitch), 1 the
// barrier, 2 + 2k the clearway of row k, 3 + 2k the seat block of row k, then the concourse, the outer wall, the
// landing and the outside. Each band is a row of cells ascending in lat; the last cell of a band reaches the band's
// upper depth (the diagonal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace swarm {
namespace seated {

constexpr int MAX_GANGWAYS = 8;   // straight gangways per wing
constexpr int MAX_DOORS = 8;      // straight-gangway doors per wing
constexpr float BIG = 1e9f;
constexpr float NO_WALL = 1e9f;   // wall(): nothing within reach

enum CellKind : int { GROUND = 0, TREAD = 1, SEAT = 2, INCLINE = 3, BARRIER = 4, WALL = 5, CONC = 6, SIDE = 7 };

// The layout (model units), as the generator makes it; plain data, passed to kernels by value
struct Layout {
    float standFront, row0, rowDepth, clearway;
    int rows;
    float concFront, concBack, outerFace, landingEnd;
    float avenueHalf, lat0, gangwayLo, gangwayHi, exitLo, exitHi, cornerInset;
    float seatInset, faceMargin, wallAbove, wallClear;
    float firstRow, rowRise, seatTop, barrierTop, concourseZ, wallTop, rake;
    int maxBandSteps;
    float seatWidth, domainHalf;
    int gangways;
    float gwLo[MAX_GANGWAYS], gwHi[MAX_GANGWAYS];
    int gwRow[MAX_GANGWAYS], gwSeated[MAX_GANGWAYS];
    int doors;
    float doorLo[MAX_DOORS], doorHi[MAX_DOORS];
    float cornerDoorLo;   // >= 1e8: no corner door
};

struct F2 { float x, y; };
struct F3 { float x, y, z; };
struct F4 { float x, y, z, w; };

__host__ __device__ inline float mn(float a, float b) { return b < a ? b : a; }
__host__ __device__ inline float mx(float a, float b) { return a < b ? b : a; }
__host__ __device__ inline float ab(float a) { return a >= 0.0f ? a : -a; }
__host__ __device__ inline float cl(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }   // std::clamp
__host__ __device__ inline int cli(int v, int lo, int hi) { return v < lo ? lo : (hi < v ? hi : v); }

__host__ __device__ inline float rowFront(const Layout& L, int k) { return L.row0 + float(k) * L.rowDepth; }
__host__ __device__ inline float rowTread(const Layout& L, int k) { return L.firstRow + float(k) * L.rowRise; }
__host__ __device__ inline int bandCount(const Layout& L) { return 6 + 2 * L.rows; }
__host__ __device__ inline bool hasCornerDoor(const Layout& L) { return L.cornerDoorLo < 1e8f; }
__host__ __device__ inline int doorsPerQuadrant(const Layout& L) { return 2 * L.doors + (hasCornerDoor(L) ? 1 : 0); }
__host__ __device__ inline int gangwaysPerQuadrant(const Layout& L) { return 2 * L.gangways + 1; }
__host__ __device__ inline int exitCount(const Layout& L) { return 4 + 4 * doorsPerQuadrant(L); }
__host__ __device__ inline int gangwayCount(const Layout& L) { return 4 * gangwaysPerQuadrant(L); }

// Straight gangways present at row front r, and whether the last one merges with the corner gangway (no seat after it)
__host__ __device__ inline int activeGangways(const Layout& L, int r) {
    int n = 0;
    for (int i = 0; i < L.gangways; ++i) n += L.gwRow[i] <= r ? 1 : 0;
    return n;
}
__host__ __device__ inline bool mergedCorner(const Layout& L, int r) {
    const int n = activeGangways(L, r);
    return n > 0 && r < L.gwSeated[n - 1];
}

__host__ __device__ inline F2 fold(F2 p) {
    const float ax = ab(p.x), ay = ab(p.y);
    return F2{ mn(ax, ay), mx(ax, ay) };
}
// A folded direction (d_lat, d_dep) at plan point p, back in the plan
__host__ __device__ inline F2 unfoldDir(F2 p, F2 d) {
    const float sx = p.x < 0.0f ? -1.0f : 1.0f, sy = p.y < 0.0f ? -1.0f : 1.0f;
    return ab(p.y) >= ab(p.x) ? F2{ sx * d.x, sy * d.y } : F2{ sx * d.y, sy * d.x };
}

// Depth extent of band j
__host__ __device__ inline void band(const Layout& L, int j, float& lo, float& hi) {
    const int nb = bandCount(L);
    if (j <= 0) { lo = 0.0f; hi = L.standFront; return; }
    if (j == 1) { lo = L.standFront; hi = L.row0; return; }
    if (j < nb - 4) {
        const int k = (j - 2) >> 1;
        const float d = rowFront(L, k);
        if (((j - 2) & 1) == 0) { lo = d; hi = d + L.clearway; }
        else { lo = d + L.clearway; hi = rowFront(L, k + 1); }
        return;
    }
    if (j == nb - 4) { lo = L.concFront; hi = L.concBack; return; }
    if (j == nb - 3) { lo = L.concBack; hi = L.outerFace; return; }
    if (j == nb - 2) { lo = L.outerFace; hi = L.landingEnd; return; }
    lo = L.landingEnd;
    hi = BIG;
}

__host__ __device__ inline int bandOf(const Layout& L, float dep) {
    const int nb = bandCount(L);
    if (dep < L.standFront) return 0;
    if (dep < L.row0) return 1;
    if (dep < L.concFront) {
        // the row from an estimate within one of it, then the row fronts themselves decide: the reference divides
        // ((dep - row0) / rowDepth), a product with the reciprocal (a few ulps off the quotient; hoisted out of the
        // callers' loops) gives the same row after the same correction
        int k = cli(int(floorf((dep - L.row0) * (1.0f / L.rowDepth))), 0, L.rows - 1);
        if (dep < rowFront(L, k)) --k;   // the estimate rounded across a row front
        else if (k + 1 < L.rows && dep >= rowFront(L, k + 1)) ++k;
        return 2 + 2 * k + (dep >= rowFront(L, k) + L.clearway ? 1 : 0);
    }
    if (dep < L.concBack) return nb - 4;
    if (dep < L.outerFace) return nb - 3;
    if (dep < L.landingEnd) return nb - 2;
    return nb - 1;
}

struct Cell {
    float l0, l1;   // lateral extent [l0, l1)
    int kind, row;  // row: the tread row of TREAD / SEAT / SIDE cells, else -1
};

// Cell c of band j (false past the band's last cell). Row bands and the barrier band: the dead-end segment, then per
// straight gangway of the band's row front the gangway and the segment after it, then the corner gangway (merged with
// the last straight gangway while no seat follows it). Seat blocks and the barrier end seatInset short of the gangways.
__host__ __device__ inline bool cellAt(const Layout& L, int j, int c, Cell& out) {
    const int nb = bandCount(L);
    float lo, hi;
    band(L, j, lo, hi);
    out.row = -1;
    if (j == 0 || j == nb - 1) {   // the walkway (with the pitch), the outside
        if (c > 0) return false;
        out.l0 = 0.0f; out.l1 = hi; out.kind = GROUND;
        return true;
    }
    const bool corner = hasCornerDoor(L);
    if (j == nb - 2) {   // the landing: ground between the doors' landings
        const int i = c >> 1;
        if (i < L.doors) {
            if ((c & 1) == 0) { out.l0 = i == 0 ? 0.0f : L.doorHi[i - 1]; out.l1 = L.doorLo[i]; out.kind = GROUND; }
            else { out.l0 = L.doorLo[i]; out.l1 = L.doorHi[i]; out.kind = CONC; }
            return true;
        }
        if (c == 2 * L.doors) { out.l0 = L.doorHi[L.doors - 1]; out.l1 = corner ? L.cornerDoorLo : hi; out.kind = GROUND; return true; }
        if (c == 2 * L.doors + 1 && corner) { out.l0 = L.cornerDoorLo; out.l1 = hi; out.kind = CONC; return true; }
        return false;
    }
    if (c == 0) { out.l0 = 0.0f; out.l1 = L.avenueHalf; out.kind = GROUND; return true; }   // the avenue
    if (j == nb - 3) {   // the outer wall, cut by the doors
        const int i = (c - 1) >> 1;
        if (i < L.doors) {
            if ((c & 1) == 1) { out.l0 = i == 0 ? L.avenueHalf : L.doorHi[i - 1]; out.l1 = L.doorLo[i]; out.kind = WALL; }
            else { out.l0 = L.doorLo[i]; out.l1 = L.doorHi[i]; out.kind = CONC; }
            return true;
        }
        if (c == 2 * L.doors + 1) { out.l0 = L.doorHi[L.doors - 1]; out.l1 = corner ? L.cornerDoorLo : hi; out.kind = WALL; return true; }
        if (c == 2 * L.doors + 2 && corner) { out.l0 = L.cornerDoorLo; out.l1 = hi; out.kind = CONC; return true; }
        return false;
    }
    if (j == nb - 4) {   // the concourse, the side wall's end
        if (c == 1) { out.l0 = L.avenueHalf; out.l1 = L.lat0; out.kind = WALL; return true; }
        if (c > 2) return false;
        out.l0 = L.lat0; out.l1 = hi; out.kind = CONC;
        return true;
    }
    int segKind = BARRIER, gapKind = GROUND;
    bool seatBand = true;
    float cornerLat = hi - L.cornerInset - L.seatInset;
    if (j >= 2) {
        out.row = (j - 2) >> 1;
        seatBand = ((j - 2) & 1) != 0;
        segKind = seatBand ? SEAT : TREAD;
        gapKind = INCLINE;
        if (!seatBand) cornerLat = lo - L.cornerInset;
    }
    if (c == 1) { out.l0 = L.avenueHalf; out.l1 = L.lat0; out.kind = SIDE; return true; }   // the side wall (a parapet)
    const int r = (j - 1) >> 1;
    const int n = activeGangways(L, r);
    const bool merged = mergedCorner(L, r);
    const int i = (c - 2) >> 1;
    if (i < n) {
        const float gLo = seatBand ? L.gwLo[i] - L.seatInset : L.gwLo[i];
        if ((c & 1) == 0) {
            out.l0 = i == 0 ? L.lat0 : (seatBand ? L.gwHi[i - 1] + L.seatInset : L.gwHi[i - 1]);
            out.l1 = gLo;
            out.kind = segKind;
        } else {
            out.l0 = gLo;
            out.l1 = merged && i == n - 1 ? hi : (seatBand ? L.gwHi[i] + L.seatInset : L.gwHi[i]);
            out.kind = gapKind;
        }
        return true;
    }
    if (merged) return false;
    const float lastHi = seatBand ? L.gwHi[n - 1] + L.seatInset : L.gwHi[n - 1];
    if (c == 2 + 2 * n) { out.l0 = lastHi; out.l1 = cornerLat; out.kind = segKind; return true; }
    if (c == 3 + 2 * n) { out.l0 = cornerLat; out.l1 = hi; out.kind = gapKind; return true; }
    return false;
}

// Top height of a cell at depth dep
__host__ __device__ inline float cellTop(const Layout& L, int kind, int row, float dep) {
    switch (kind) {
    case TREAD: return rowTread(L, row);
    case SEAT: return rowTread(L, row) + L.seatTop;
    case INCLINE: return L.rake * (dep - L.row0);
    case BARRIER: return L.barrierTop;
    case WALL: return L.wallTop;
    case CONC: return L.concourseZ;
    case SIDE: return row < 0 ? L.barrierTop : rowTread(L, row) + L.seatTop;
    default: return 0.0f;
    }
}

// The cell of band j holding lateral position lat (its index; the band's last cell past every l1)
__host__ __device__ inline int ownCell(const Layout& L, int j, float lat, Cell& out) {
    Cell c{};
    int i = 0;
    for (; cellAt(L, j, i, c); ++i) {
        out = c;
        if (lat < c.l1) return i;
    }
    return i - 1;
}

// Where the queries below take their cells from: the generator above (host and device), or a table of every band's cells
// made from it once (CellTable: the same values, read instead of generated; the device step uses it)
struct GenCells {
    static constexpr bool sorted = false;   // the queries scan every cell of a band, as the reference does
    const Layout& L;
    __host__ __device__ bool at(int j, int c, Cell& out) const { return cellAt(L, j, c, out); }
    __host__ __device__ int own(int j, float lat, Cell& out) const { return ownCell(L, j, lat, out); }
};
constexpr int CELL_BUCKETS = 32;   // lateral buckets per band of a cell table
// Every band's cells in order: band j's are [start[j], start[j + 1]); per cell its l1 alone (the lateral search) and its
// whole record (l0, l1 as float bits, kind, row). The table is built only from layouts whose bands are sorted (l0 <= l1,
// each cell's l0 at or past the previous cell's l1; checked by the builder), so a band's l1 never decreases and the
// queries may search it (sorted): bucket[j * CELL_BUCKETS + b] is a starting index for the lateral search of band j at
// lat in [b w, (b + 1) w), w = 1 / invW[j]. Any start gives the same cell (the search walks both ways to the first
// l1 above lat); a good one saves the walk.
struct CellTable {
    static constexpr bool sorted = true;
    const int* start;
    const float* l1;
    const int4* cell;
    const uint16_t* bucket;
    const float* invW;
    __host__ __device__ bool at(int j, int c, Cell& out) const {
        const int i = start[j] + c;
        if (i >= start[j + 1]) return false;
        const int4 v = cell[i];
        out.l0 = __builtin_bit_cast(float, v.x);
        out.l1 = __builtin_bit_cast(float, v.y);
        out.kind = v.z;
        out.row = v.w;
        return true;
    }
    // ownCell's answer: the first cell with lat < l1, else the band's last (a NaN lat: the last)
    __host__ __device__ int own(int j, float lat, Cell& out) const {
        const int a = start[j], last = start[j + 1] - 1;
        const float t = lat * invW[j];
        const int b = t >= 0.0f ? (t < float(CELL_BUCKETS) ? int(t) : CELL_BUCKETS - 1) : 0;
        int i = bucket[j * CELL_BUCKETS + b];
        while (i < last && !(lat < l1[i])) ++i;
        while (i > a && lat < l1[i - 1]) --i;
        at(j, i - a, out);
        return i - a;
    }
};

// A cell table stored as one blob (byte offsets from its start, each array 16-byte aligned, the size a multiple of 16),
// so that a kernel can copy it whole into LDS and search it there
struct CellBlob {
    int bytes, oStart, oInvW, oL1, oBucket, oCell;
};
__host__ __device__ inline CellTable cellTableAt(const char* base, const CellBlob& b) {
    return CellTable{ reinterpret_cast<const int*>(base + b.oStart), reinterpret_cast<const float*>(base + b.oL1),
                      reinterpret_cast<const int4*>(base + b.oCell), reinterpret_cast<const uint16_t*>(base + b.oBucket),
                      reinterpret_cast<const float*>(base + b.oInvW) };
}

template <class C> __host__ __device__ inline float floorHeightC(const Layout& L, const C& cs, F2 p) {
    const F2 f = fold(p);
    const int j = bandOf(L, f.y);
    Cell c{};
    cs.own(j, f.x, c);
    return cellTop(L, c.kind, c.row, f.y);
}

// Kind, row and band of the cell under p
template <class C> __host__ __device__ inline int kindAtC(const Layout& L, const C& cs, F2 p, int* row, int* bandIndex) {
    const F2 f = fold(p);
    const int j = bandOf(L, f.y);
    Cell c{};
    cs.own(j, f.x, c);
    *row = c.row;
    *bandIndex = j;
    return c.kind;
}

// The nearest point (u, z) of {u in [u0, u1]} x (below ? z <= t(u) : z >= t(u)), t(u) = s (u - uRef)
__host__ __device__ inline F2 nearestRegion2(float u, float z, float u0, float u1, float s, float uRef, bool below) {
    const float uc = cl(u, u0, u1);
    const float tc = s * (uc - uRef);
    if (below ? z <= tc : z >= tc) return F2{ uc, z };
    const float t0 = s * (u0 - uRef), t1 = s * (u1 - uRef);
    const float du = u1 - u0, dz = t1 - t0;
    const float len2 = du * du + dz * dz;
    const float t = len2 > 0.0f ? cl(((u - u0) * du + (z - t0) * dz) / len2, 0.0f, 1.0f) : 0.0f;
    F2 best{ u0 + t * du, t0 + t * dz };
    float bd = (u - best.x) * (u - best.x) + (z - best.y) * (z - best.y);
    const F2 r0{ u0, below ? mn(z, t0) : mx(z, t0) }, r1{ u1, below ? mn(z, t1) : mx(z, t1) };
    const float d0 = (u - r0.x) * (u - r0.x) + (z - r0.y) * (z - r0.y), d1 = (u - r1.x) * (u - r1.x) + (z - r1.y) * (z - r1.y);
    if (d0 < bd) { bd = d0; best = r0; }
    if (d1 < bd) best = r1;
    return best;
}

// The nearest point to pf (folded) of a cell's solid column (below its top) or air column (above it); .w its squared
// distance. The air above an incline that crosses the diagonal is clipped to the fold (beyond it lies the mirrored,
// higher incline): its nearest point is then on the fold plane.
__host__ __device__ inline F4 nearestInCell(const Layout& L, float lo, float hi, const Cell& c, F3 pf, bool airColumn) {
    const float latC = cl(pf.x, c.l0, c.l1);
    F3 q;
    if (c.kind == INCLINE) {
        const F2 n2 = nearestRegion2(pf.y, pf.z, lo, hi, L.rake, L.row0, !airColumn);
        q = F3{ latC, n2.x, n2.y };
        if (airColumn && q.x > q.y) {
            const float s0 = mx(c.l0, lo), s1 = mn(c.l1, hi);
            if (s0 > s1) return F4{ q.x, q.y, q.z, BIG };
            const float r2 = 1.41421356f;
            const float m = 0.5f * (pf.x + pf.y);
            const F2 w = nearestRegion2(m * r2, pf.z, s0 * r2, s1 * r2, L.rake / r2, L.row0 * r2, false);
            const float s = w.x / r2;
            q = F3{ s, s, w.y };
        }
    } else {
        const float top = cellTop(L, c.kind, c.row, 0.0f);
        q = F3{ latC, cl(pf.y, lo, hi), airColumn ? mx(pf.z, top) : mn(pf.z, top) };
    }
    const float dx = pf.x - q.x, dy = pf.y - q.y, dz = pf.z - q.z;
    return F4{ q.x, q.y, q.z, dx * dx + dy * dy + dz * dz };
}

// The exact signed distance to the heightfield {z <= floorHeight(xy)} and its unit gradient: (gx, gy, gz, d).
// Bands are searched outward from p's own band until the band gap exceeds the best distance, at most maxBandSteps on
// either side; the cells of a band are scanned in lat order, stopping at the first one beyond the best distance (a
// sorted cell table starts the scan at the first cell not skipped: the same cells are visited in the same order).
template <class C> __host__ __device__ inline F4 sdfGradC(const Layout& L, const C& cs, F3 p) {
    const F2 f = fold(F2{ p.x, p.y });
    const F3 pf{ f.x, f.y, p.z };
    const int nb = bandCount(L);
    const int jp = bandOf(L, pf.y);
    Cell own{};
    cs.own(jp, pf.x, own);
    const bool inAir = pf.z > cellTop(L, own.kind, own.row, pf.y);
    float best = BIG;
    F3 bestQ = pf;
    float rl0 = 0.0f, rl1 = 0.0f, rlo = 0.0f, rhi = 0.0f;   // the nearest cell's rectangle
    int bestKind = own.kind;
    auto visit = [&](int j, float g2) {
        float lo, hi;
        band(L, j, lo, hi);
        Cell c{};
        int i = 0;
        if constexpr (C::sorted) {
            // the cells before this run are all skipped below (their lateral gap only grows away from p's own cell)
            const float b0 = best;
            i = cs.own(j, pf.x, c);
            Cell t{};
            while (i > 0 && cs.at(j, i - 1, t)) {
                const float lg = mx(mx(t.l0 - pf.x, pf.x - t.l1), 0.0f);
                if (g2 + lg * lg >= b0) break;
                --i;
            }
        }
        for (; cs.at(j, i, c); ++i) {
            const float lg = mx(mx(c.l0 - pf.x, pf.x - c.l1), 0.0f);
            if (g2 + lg * lg >= best) {
                if (c.l0 > pf.x) break;
                continue;
            }
            const F4 n = nearestInCell(L, lo, hi, c, pf, !inAir);
            if (n.w < best) {
                best = n.w;
                bestQ = F3{ n.x, n.y, n.z };
                rl0 = c.l0; rl1 = c.l1; rlo = lo; rhi = hi;
                bestKind = c.kind;
            }
        }
    };
    visit(jp, 0.0f);
    bool down = true, up = true;
    for (int s = 1; s <= L.maxBandSteps && (down || up); ++s) {
        for (int side = 0; side < 2; ++side) {
            bool& open = side == 0 ? down : up;
            if (!open) continue;
            const int j = side == 0 ? jp - s : jp + s;
            if (j < 0 || j >= nb) { open = false; continue; }
            float lo, hi;
            band(L, j, lo, hi);
            const float g = mx(mx(lo - pf.y, pf.y - hi), 0.0f);
            if (g * g >= best) { open = false; continue; }
            visit(j, g * g);
        }
    }
    const float d = sqrtf(best);
    F3 gf;
    if (d > 0.0f) {
        gf = inAir ? F3{ (pf.x - bestQ.x) / d, (pf.y - bestQ.y) / d, (pf.z - bestQ.z) / d }
                   : F3{ (bestQ.x - pf.x) / d, (bestQ.y - pf.y) / d, (bestQ.z - pf.z) / d };
    } else {
        // on the surface: the normal of the face p lies on (a side of the nearest cell's footprint: a vertical face;
        // inside the footprint: the cell's top)
        const float sgn = inAir ? -1.0f : 1.0f;
        const float hx = pf.x <= rl0 ? sgn : (pf.x >= rl1 ? -sgn : 0.0f);
        const float hy = pf.y <= rlo ? sgn : (pf.y >= rhi ? -sgn : 0.0f);
        if (hx != 0.0f || hy != 0.0f) {
            const float inv = 1.0f / sqrtf(hx * hx + hy * hy);
            gf = F3{ hx * inv, hy * inv, 0.0f };
        } else if (bestKind == INCLINE) {
            const float nz = -L.rake;
            const float inv = 1.0f / sqrtf((0.0f * 0.0f + nz * nz) + 1.0f * 1.0f);
            gf = F3{ 0.0f * inv, nz * inv, 1.0f * inv };
        } else {
            gf = F3{ 0.0f, 0.0f, 1.0f };
        }
    }
    const F2 g = unfoldDir(F2{ p.x, p.y }, F2{ gf.x, gf.y });
    return F4{ g.x, g.y, gf.z, inAir ? d : -d };
}

// The nearest wall in the horizontal plane z = c.z: walls are the footprints of the cells whose top lies above the
// plane (inclines are walkable slopes, never walls). (unit direction away from the wall, clearance); clearance < 0:
// c is inside a wall; NO_WALL: none within reach.
template <class C> __host__ __device__ inline F3 wallC(const Layout& L, const C& cs, F3 cpt, float reach) {
    const F2 f = fold(F2{ cpt.x, cpt.y });
    const int nb = bandCount(L);
    const int jp = bandOf(L, f.y);
    Cell own{};
    cs.own(jp, f.x, own);
    const bool inWall = own.kind != INCLINE && cellTop(L, own.kind, own.row, f.y) > cpt.z;
    float best = inWall ? BIG : reach * reach;
    F2 bestQ = f;
    float rl0 = 0.0f, rl1 = 0.0f, rlo = 0.0f, rhi = 0.0f;
    bool found = false;
    auto planeD2 = [&](const Cell& c, float lo, float hi) {   // the squared distance to a cell's footprint
        const F2 q{ mn(mx(f.x, c.l0), c.l1), mn(mx(f.y, lo), hi) };
        const float dx = f.x - q.x, dy = f.y - q.y;
        return dx * dx + dy * dy;
    };
    auto consider = [&](const Cell& c, float lo, float hi) {
        const bool isWall = c.kind != INCLINE && cellTop(L, c.kind, c.row, lo) > cpt.z;
        if (isWall == inWall) return;
        const F2 q{ mn(mx(f.x, c.l0), c.l1), mn(mx(f.y, lo), hi) };
        const float dx = f.x - q.x, dy = f.y - q.y;
        const float d2 = dx * dx + dy * dy;
        if (d2 < best) {
            best = d2;
            bestQ = q;
            rl0 = c.l0; rl1 = c.l1; rlo = lo; rhi = hi;
            found = true;
        }
    };
    auto visit = [&](int j) {
        float lo, hi;
        band(L, j, lo, hi);
        Cell c{};
        if constexpr (C::sorted) {
            // In a sorted band a footprint's distance grows with the cells' index distance from p's own cell, so the
            // cells closer than the best so far are one run about it: visited in order, the others could not win
            const float b0 = best;
            const int io = cs.own(j, f.x, c);
            int i = io;
            Cell t{};
            while (i > 0 && cs.at(j, i - 1, t) && planeD2(t, lo, hi) < b0) --i;
            for (; cs.at(j, i, c); ++i) {
                if (i > io && !(planeD2(c, lo, hi) < b0)) break;
                consider(c, lo, hi);
            }
        } else {
            for (int i = 0; cs.at(j, i, c); ++i) consider(c, lo, hi);
        }
    };
    visit(jp);
    bool down = true, up = true;
    for (int s = 1; s <= L.maxBandSteps && (down || up); ++s) {
        for (int side = 0; side < 2; ++side) {
            bool& open = side == 0 ? down : up;
            if (!open) continue;
            const int j = side == 0 ? jp - s : jp + s;
            if (j < 0 || j >= nb) { open = false; continue; }
            float lo, hi;
            band(L, j, lo, hi);
            const float g = mx(mx(lo - f.y, f.y - hi), 0.0f);
            if (g * g >= best) { open = false; continue; }
            visit(j);
        }
    }
    if (!found) return F3{ 0.0f, 0.0f, NO_WALL };
    const float d = sqrtf(best);
    F2 dir;
    if (d > 0.0f) {
        dir = inWall ? F2{ (bestQ.x - f.x) / d, (bestQ.y - f.y) / d } : F2{ (f.x - bestQ.x) / d, (f.y - bestQ.y) / d };
    } else {
        const float sgn = inWall ? 1.0f : -1.0f;
        const float hx = f.x <= rl0 ? sgn : (f.x >= rl1 ? -sgn : 0.0f);
        const float hy = f.y <= rlo ? sgn : (f.y >= rhi ? -sgn : 0.0f);
        const float inv = 1.0f / sqrtf(hx * hx + hy * hy);
        dir = F2{ hx * inv, hy * inv };
    }
    const F2 u = unfoldDir(F2{ cpt.x, cpt.y }, dir);
    return F3{ u.x, u.y, inWall ? -d : d };
}

// The floor's rise per unit of horizontal travel along the plan direction e at p: on an incline its gradient (rake
// along dep) projected on e, 0 on every level floor
template <class C> __host__ __device__ inline float floorSlopeC(const Layout& L, const C& cs, F2 p, F2 e) {
    const F2 f = fold(p);
    const int j = bandOf(L, f.y);
    Cell c{};
    cs.own(j, f.x, c);
    if (c.kind != INCLINE) return 0.0f;
    const F2 g = unfoldDir(p, F2{ 0.0f, L.rake });
    const float a = g.x * e.x, b = g.y * e.y;
    return a + b;
}

__host__ __device__ inline int quadrantOf(F2 p) { return p.x >= 0.0f ? (p.y >= 0.0f ? 0 : 3) : (p.y >= 0.0f ? 1 : 2); }

// The exit whose opening contains p's folded lateral position: 0-3 the avenues (+x, -x, +y, -y), then per quadrant
// (+x+y, -x+y, -x-y, +x-y) the doors of the y wall, of the x wall, then the corner door; -1: none
__host__ __device__ inline int exitByLat(const Layout& L, F2 p) {
    const float ax = ab(p.x), ay = ab(p.y);
    const bool alongY = ay >= ax;
    const float lat = alongY ? ax : ay;
    if (lat < L.avenueHalf) return alongY ? (p.y > 0.0f ? 2 : 3) : (p.x > 0.0f ? 0 : 1);
    const int base = 4 + quadrantOf(p) * doorsPerQuadrant(L);
    for (int i = 0; i < L.doors; ++i)
        if (lat >= L.doorLo[i] && lat <= L.doorHi[i]) return base + (alongY ? 0 : L.doors) + i;
    if (hasCornerDoor(L) && lat >= L.cornerDoorLo) return base + 2 * L.doors;
    return -1;
}

// The gangway under p (an incline, or a barrier gap at a gangway's foot): per quadrant the y wing's straight gangways,
// the x wing's, then the corner gangway (a straight gangway merged with it counts as the corner gangway); -1: none
template <class C> __host__ __device__ inline int gangwayAtC(const Layout& L, const C& cs, F2 p) {
    const F2 f = fold(p);
    const int j = bandOf(L, f.y);
    if (j < 1 || j >= bandCount(L) - 4) return -1;
    Cell c{};
    const int ci = cs.own(j, f.x, c);
    if (!(c.kind == INCLINE || (j == 1 && c.kind == GROUND && ci >= 2))) return -1;
    const int r = (j - 1) >> 1;
    const int n = activeGangways(L, r);
    const bool merged = mergedCorner(L, r);
    const int i = (ci - 3) >> 1;
    const bool straight = (ci & 1) == 1 && i < n && !(merged && i == n - 1);
    const int base = quadrantOf(p) * gangwaysPerQuadrant(L);
    return straight ? base + (ab(p.y) >= ab(p.x) ? 0 : L.gangways) + i : base + 2 * L.gangways;
}

// ---------------- the venue's own 3D step: cell keys, the step-up scan, the pitch fence ----------------

// The cell under p as a key (its band and its lower lateral bound): two points share a cell iff their keys are equal
// (seated::cellKey, src/venue.cpp:631-638)
template <class C> __host__ __device__ inline F2 cellKeyC(const Layout& L, const C& cs, F2 p) {
    const F2 f = fold(p);
    const int j = bandOf(L, f.y);
    Cell c{};
    cs.own(j, f.x, c);
    return F2{ float(j), c.l0 };
}

constexpr float STEP_UP_SAMPLE = 0x1.a36e2ep-11f;   // 0.125 BODY_RADIUS folded (0.0008): the scan's sample spacing
constexpr float STEP_UP_TOL = 1e-6f;
constexpr int STEP_UP_MAX_SAMPLES = 32;
constexpr int SCAN_CROSSINGS = 4;    // cell boundaries located per sample
constexpr int SCAN_BISECTIONS = 12;  // bisections per boundary

__host__ __device__ inline F2 lerp2(F2 a, F2 b, float t) {   // glm::mix: a (1 - t) + b t
    const float u = 1.0f - t;
    return F2{ a.x * u + b.x * t, a.y * u + b.y * t };
}
__host__ __device__ inline bool sameKey(F2 a, F2 b) { return a.x == b.x && a.y == b.y; }

// "Is q's cell key this key?", asked of many points against one key. The generator compares the keys; a sorted cell
// table answers from the key's lateral run in its band: the cells sharing the key's l0 are adjacent, and the lateral
// search gives them exactly the lat in [l1 of the cell before them, l1 of their last cell), every lat beyond when their
// last is the band's last (NaN included), every lat below when their first is the band's first.
template <class C, bool SORTED = C::sorted> struct KeyMatch;
template <class C> struct KeyMatch<C, false> {
    F2 key;
    __host__ __device__ void set(const Layout& L, const C& cs, F2 p) { key = cellKeyC(L, cs, p); }
    __host__ __device__ bool test(const Layout& L, const C& cs, F2 q) const { return sameKey(cellKeyC(L, cs, q), key); }
};
template <class C> struct KeyMatch<C, true> {
    int j;
    float lo, hi;
    bool toEnd, fromStart;
    __host__ __device__ void set(const Layout& L, const C& cs, F2 p) {
        const F2 f = fold(p);
        j = bandOf(L, f.y);
        Cell c{}, t{};
        const int io = cs.own(j, f.x, c);
        int i0 = io, i1 = io;
        while (i0 > 0 && cs.at(j, i0 - 1, t) && t.l0 == c.l0) --i0;
        while (cs.at(j, i1 + 1, t) && t.l0 == c.l0) ++i1;
        fromStart = i0 == 0;
        lo = 0.0f;
        if (!fromStart) { cs.at(j, i0 - 1, t); lo = t.l1; }
        cs.at(j, i1, t);
        hi = t.l1;
        toEnd = !cs.at(j, i1 + 1, t);
    }
    __host__ __device__ bool test(const Layout& L, const C&, F2 q) const {
        const F2 f = fold(q);
        return bandOf(L, f.y) == j && (fromStart || !(f.x < lo)) && (toEnd || f.x < hi);
    }
};

// The step-up limit's scan of a horizontal move p0 -> p1 from the start floor f0: walking the cells the move crosses
// (every STEP_UP_SAMPLE, each boundary located by bisection), the move is blocked at the first boundary into a floor
// above f0 + STEP_UP_TOL that rises faster than the steepest incline along the move; returns the allowed fraction (1: not
// blocked) and n, the blocking face's normal away from it (seated::stepUpScan, src/venue.cpp:807-835)
template <class C> __host__ __device__ inline float stepUpScanC(const Layout& L, const C& cs, F2 p0, F2 p1, float f0, F2& n) {
    n = F2{ 0.0f, 0.0f };
    const float dx = p1.x - p0.x, dy = p1.y - p0.y;
    const float len = sqrtf(dx * dx + dy * dy);
    const int ns = cli(int(ceilf(len / STEP_UP_SAMPLE)), 1, STEP_UP_MAX_SAMPLES);
    float lo = 0.0f;
    KeyMatch<C> key;
    key.set(L, cs, p0);
    for (int i = 1; i <= ns; ++i) {
        const float t = float(i) / float(ns);
        for (int k = 0; k < SCAN_CROSSINGS; ++k) {
            if (key.test(L, cs, lerp2(p0, p1, t))) break;
            float a = lo, b = t;
            for (int m = 0; m < SCAN_BISECTIONS; ++m) {
                const float tm = 0.5f * (a + b);
                if (key.test(L, cs, lerp2(p0, p1, tm))) a = tm;
                else b = tm;
            }
            const F2 pa = lerp2(p0, p1, a), pb = lerp2(p0, p1, b);
            const float ha = floorHeightC(L, cs, pa), hb = floorHeightC(L, cs, pb);
            if (hb > f0 + STEP_UP_TOL && hb - ha > L.rake * len * (b - a) + STEP_UP_TOL) {
                const F3 w = wallC(L, cs, F3{ pa.x, pa.y, ha + STEP_UP_TOL }, STEP_UP_SAMPLE);
                if (w.z < NO_WALL) n = F2{ w.x, w.y };
                else {
                    const float l = mx(len, 1e-30f);
                    n = F2{ (p0.x - p1.x) / l, (p0.y - p1.y) / l };
                }
                return a;
            }
            lo = b;
            key.set(L, cs, pb);
        }
        lo = t;
    }
    return 1.0f;
}

// The pitch fence: a closed square ring about the origin at the measured faces, inflated by the body radius for
// collision (sim_params.hpp:315-327; sdBox2 / sdBox2Grad / sdFence / sdFenceGradient, src/sdf.cpp:45-82)
constexpr float FENCE_HALF_MID = 0.5f * (0.328395f + 0.335432f);
constexpr float FENCE_HALF_THK_COLL = 0.5f * (0.335432f - 0.328395f) + 0.0064f;
__host__ __device__ inline float sdBox2(F2 p, float h) {
    const float qx = ab(p.x) - h, qy = ab(p.y) - h;
    const float ox = mx(qx, 0.0f), oy = mx(qy, 0.0f);
    return sqrtf(ox * ox + oy * oy) + mn(mx(qx, qy), 0.0f);
}
__host__ __device__ inline F2 sdBox2Grad(F2 p, float h) {
    const float qx = ab(p.x) - h, qy = ab(p.y) - h;
    const float sx = p.x < 0.0f ? -1.0f : 1.0f, sy = p.y < 0.0f ? -1.0f : 1.0f;
    F2 g;
    if (mx(qx, qy) > 0.0f) {
        const float ox = mx(qx, 0.0f), oy = mx(qy, 0.0f);
        const float l = sqrtf(ox * ox + oy * oy);
        g = F2{ ox / l, oy / l };
    } else {
        g = qx > qy ? F2{ 1.0f, 0.0f } : F2{ 0.0f, 1.0f };
    }
    return F2{ sx * g.x, sy * g.y };
}
__host__ __device__ inline float sdFence(F2 p) { return ab(sdBox2(F2{ p.x - 0.0f, p.y - 0.0f }, FENCE_HALF_MID)) - FENCE_HALF_THK_COLL; }
__host__ __device__ inline F2 sdFenceGradient(F2 p) {
    const F2 rel{ p.x - 0.0f, p.y - 0.0f };
    const float s = sdBox2(rel, FENCE_HALF_MID) < 0.0f ? -1.0f : 1.0f;
    const F2 g = sdBox2Grad(rel, FENCE_HALF_MID);
    return F2{ s * g.x, s * g.y };
}

// The queries on the generator (the host, the probe kernel)
__host__ __device__ inline float floorHeight(const Layout& L, F2 p) { return floorHeightC(L, GenCells{ L }, p); }
__host__ __device__ inline int kindAt(const Layout& L, F2 p, int* row, int* bandIndex) { return kindAtC(L, GenCells{ L }, p, row, bandIndex); }
__host__ __device__ inline F4 sdfGrad(const Layout& L, F3 p) { return sdfGradC(L, GenCells{ L }, p); }
__host__ __device__ inline F3 wall(const Layout& L, F3 cpt, float reach) { return wallC(L, GenCells{ L }, cpt, reach); }
__host__ __device__ inline float floorSlope(const Layout& L, F2 p, F2 e) { return floorSlopeC(L, GenCells{ L }, p, e); }
__host__ __device__ inline int gangwayAt(const Layout& L, F2 p) { return gangwayAtC(L, GenCells{ L }, p); }
__host__ __device__ inline F2 cellKey(const Layout& L, F2 p) { return cellKeyC(L, GenCells{ L }, p); }
__host__ __device__ inline float stepUpScan(const Layout& L, F2 p0, F2 p1, float f0, F2& n) { return stepUpScanC(L, GenCells{ L }, p0, p1, f0, n); }

}  // namespace seated
}  // namespace swarm
