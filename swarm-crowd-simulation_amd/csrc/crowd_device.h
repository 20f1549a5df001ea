// Device-side model functions of the per-step pipeline (gfx950). Each follows the reference's
// physics.comp / cpu_crowd_sim.cpp expression by expression on a flat 2D floor, in FP32 with no FMA
// contraction (the library is built with -ffp-contract=off and correctly rounded divide and sqrt), so the
// results are bit-identical to the reference's CPU arithmetic.
#pragma once
#include <hip/hip_runtime.h>

#include "seated_venue.h"
#include "swarm_params.h"

namespace swarm {

// The read-only tables of one scene as the kernels see them
struct SceneView {
    float xmin, ymin, xmax, ymax;
    float hashSpanX, hashSpanY;
    int hashNX, hashNY;
    int percRing, projRing;
    int sdfNX, sdfNY;
    float sdfOX, sdfOY, sdfInvH;
    const float* sdf;
    const float* sdfMin;
    int sdfMinWin;
    int navRX, navRY;
    float navOX, navOY, navCell, navInvCell;
    int K, E;
    const uint32_t* navInfo;           // per nav cell: its class bits (the fallback search reads its neighbours here)
    // per nav cell one record of navStride words: the class bits, the K candidate slots (anchor | exit << 25), the K
    // 16-bit distance quanta; an agent's choice reads one record (one cache line where K <= 8) instead of three arrays
    const uint32_t* navRec;
    int navStride;
    const float* invCapS;
    float secondsPerUnit, distQuantum;
    const float* stair;
    const float* exitRects;   // E x 4
    // per hash cell: its wall boxes within reach, in ascending box order, as the boxes themselves (lox, loy, hix, hiy):
    // wallBinBox[wallBinStart[c] .. wallBinStart[c + 1])
    const uint32_t* wallBinStart;
    const float4* wallBinBox;
    const seated::Layout* venue;       // a seated venue: its layout (a plan's incline rules, or the 3D step), else null
    int venue3d;                       // the venue's own 3D step: the surface rule in choiceCell, the venue's wall term
    float exitDepth, exitMaxZ, domainMax;   // its exit test
    seated::CellTable venueCells;      // the venue's cells, every band's (a venue scene)
    const char* venueBlob;             // the same table as one blob (venueCells points into it)
    seated::CellBlob venueBlobLayout;
};

// The largest venue cell table a block copies into its LDS (the rest read it from global memory)
constexpr int VENUE_LDS_MAX = 64 * 1024;
__host__ __device__ inline int venueLdsBytes(const SceneView& s) {
    return s.venueBlob && s.venueBlobLayout.bytes <= VENUE_LDS_MAX ? s.venueBlobLayout.bytes : 0;
}
// The venue's cell table copied into the block's LDS (dynamic shared memory of venueLdsBytes(s) bytes, launched so):
// the searches then read LDS instead of global memory. Every thread of the block must call it.
__device__ inline seated::CellTable venueCellsInLds(const SceneView& s) {
    extern __shared__ int4 venueLds[];
    const int n16 = s.venueBlobLayout.bytes / 16;
    const int4* src = reinterpret_cast<const int4*>(s.venueBlob);
    for (int i = int(threadIdx.x); i < n16; i += int(blockDim.x)) venueLds[i] = src[i];
    __syncthreads();
    return seated::cellTableAt(reinterpret_cast<const char*>(venueLds), s.venueBlobLayout);
}

struct ModelParams {
    float relaxStep;      // exp(-DT / tau)
    float tau;            // relaxation time, model time
    float obstacleGain;   // OBSTACLE_GAIN
    int projIters;        // PROJ_ITERATIONS
};

struct V2 { float x, y; };

// a nav cell's record: its class bits, its candidate slots, its distance quanta
__device__ inline const uint32_t* navRecOf(const SceneView& s, int cell) { return s.navRec + size_t(cell) * size_t(s.navStride); }
__device__ inline const uint32_t* navSlots(const SceneView& s, int cell) { return navRecOf(s, cell) + 1; }
__device__ inline const uint16_t* navDists(const SceneView& s, int cell) {
    return reinterpret_cast<const uint16_t*>(navRecOf(s, cell) + 1 + s.K);
}
__device__ inline V2 operator+(V2 a, V2 b) { return { a.x + b.x, a.y + b.y }; }
__device__ inline V2 operator-(V2 a, V2 b) { return { a.x - b.x, a.y - b.y }; }
__device__ inline V2 operator-(V2 a) { return { -a.x, -a.y }; }
__device__ inline V2 operator*(V2 a, float s) { return { a.x * s, a.y * s }; }
__device__ inline V2 operator*(float s, V2 a) { return { s * a.x, s * a.y }; }
__device__ inline V2 operator/(V2 a, float s) { return { a.x / s, a.y / s }; }
__device__ inline float dot2(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
__device__ inline float len2(V2 a) { return sqrtf(dot2(a, a)); }
__device__ inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ inline V2 mix2(V2 a, V2 b, float t) { return { mixf(a.x, b.x, t), mixf(a.y, b.y, t) }; }
__device__ inline bool nonzero(V2 a) { return a.x != 0.0f || a.y != 0.0f; }
__device__ inline V2 safeNormalize(V2 v) { const float l = len2(v); return l > SAFE_NORMALIZE_EPS ? v / l : V2{ 0.0f, 0.0f }; }
__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

struct G3 { float gx, gy, w; };

// calc_grid.comp cellID on a rectangular grid
__device__ inline uint32_t hashCell(const SceneView& s, float px, float py) {
    const int gx = clampi(int((px - s.xmin) / s.hashSpanX * float(s.hashNX)), 0, s.hashNX - 1);
    const int gy = clampi(int((py - s.ymin) / s.hashSpanY * float(s.hashNY)), 0, s.hashNY - 1);
    return uint32_t(gx + gy * s.hashNX);
}

// mapExactGrad (cpu_crowd_sim.cpp:239-262) on the 2D grid; map() is its value (exact weights, no filtering)
__device__ inline G3 mapExactGrad(const SceneView& s, V2 p) {
    const float fx = (p.x - s.sdfOX) * s.sdfInvH, fy = (p.y - s.sdfOY) * s.sdfInvH;
    const bool outside = fx < -0.5f || fy < -0.5f || fx > float(s.sdfNX) - 0.5f || fy > float(s.sdfNY) - 0.5f;
    const int ix = clampi(int(floorf(fx)), 0, s.sdfNX - 2), iy = clampi(int(floorf(fy)), 0, s.sdfNY - 2);
    const float wx = fx - float(ix), wy = fy - float(iy);
    const float* row0 = s.sdf + size_t(iy) * size_t(s.sdfNX);
    const float* row1 = row0 + s.sdfNX;
    const float c00 = row0[ix], c10 = row0[ix + 1], c01 = row1[ix], c11 = row1[ix + 1];
    const float x0 = mixf(c00, c10, wx), x1 = mixf(c01, c11, wx);
    if (outside) return G3{ 0.0f, 0.0f, 1.0f };
    return G3{ mixf(c10 - c00, c11 - c01, wy) * s.sdfInvH, (x1 - x0) * s.sdfInvH, mixf(x0, x1, wy) };
}
__device__ inline float mapSdf(const SceneView& s, V2 p) { return mapExactGrad(s, p).w; }

// True when no wall sample within PLANE_PROBE_RADII r_max of p can be <= 0 (the min-filtered grid is > 0 over
// the window of p's bilinear cell), so the wall passes would find nothing: the exact results, without the samples
__device__ inline bool wallFree(const SceneView& s, V2 p) {
    const float fx = (p.x - s.sdfOX) * s.sdfInvH, fy = (p.y - s.sdfOY) * s.sdfInvH;
    const float lo = float(s.sdfMinWin);
    if (!(fx >= lo && fy >= lo && fx <= float(s.sdfNX - 2 - s.sdfMinWin) && fy <= float(s.sdfNY - 2 - s.sdfMinWin))) return false;
    const int ix = int(floorf(fx)), iy = int(floorf(fy));
    return s.sdfMin[size_t(iy) * size_t(s.sdfNX) + size_t(ix)] > 0.0f;
}

// calcGradient's xy rows (cpu_crowd_sim.cpp:297-303)
__device__ inline V2 calcGradient(const SceneView& s, V2 p) {
    const float H = OBSTACLE_GRAD_H;
    const float m1 = mapSdf(s, V2{ p.x + H, p.y + -H }), m2 = mapSdf(s, V2{ p.x + -H, p.y + -H });
    const float m3 = mapSdf(s, V2{ p.x + -H, p.y + H }), m4 = mapSdf(s, V2{ p.x + H, p.y + H });
    const V2 g{ ((m1 + -m2) + -m3) + m4, ((-m1 + -m2) + m3) + m4 };
    return g * OBSTACLE_GRAD_SCALE;
}

__device__ inline V2 pairDirection(uint32_t self, uint32_t other) {   // cpu_crowd_sim.cpp:123-129
    const uint32_t lo = self < other ? self : other, hi = self < other ? other : self;
    const float angle = float(hash32(hash32(lo) ^ hi)) * HASH_TO_RADIANS;
    float sn, cs;
    sincos_cephes(angle, &sn, &cs);
    const V2 dir{ cs, sn };
    return self == lo ? dir : -dir;
}

// squared separation support / contact distance: the reference's folded constants at its distances
__device__ inline float sepSupport2(float S) { return S == R_SEP ? R_SEP2 : S * S; }
__device__ inline float contact2(float D) { return D == BODY_DIAM ? BODY_DIAM2 : D * D; }

__device__ inline V2 separationForce(V2 offset, uint32_t self, uint32_t other, float S) {   // :131-139
    const float d2 = dot2(offset, offset);
    if (d2 >= sepSupport2(S)) return V2{ 0.0f, 0.0f };
    const float w = 1.0f - sqrtf(d2) / S;
    const float magnitude = K_SEP * w * w;
    const float dxy2 = dot2(offset, offset);
    if (dxy2 <= SEP_STACKED_XY2) return magnitude * pairDirection(self, other);
    return magnitude * offset * (1.0f / sqrtf(dxy2));
}

__device__ inline V2 contactCorrection(V2 offset, uint32_t self, uint32_t other, float D) {   // :141-148
    const float dxy2 = dot2(offset, offset);
    if (dxy2 >= contact2(D)) return V2{ 0.0f, 0.0f };
    const float dxy = sqrtf(dxy2);
    const float shift = 0.5f * (D - dxy);
    if (dxy2 <= SEP_STACKED_XY2) return shift * pairDirection(self, other);
    return shift * (offset / dxy);
}

// ---- navigation / exit choice (exit_choice.cpp:512-603) ----
// On a seated venue's own step: whether the cell centre of nav cell a is reachable on the venue's surface from p (the
// same venue cell, or a move the step-up scan allows; surfaceReaches, src/nav_field.cpp:375-385)
__device__ inline bool surfaceReaches(const SceneView& s, V2 p, uint32_t a) {
    const uint32_t res = uint32_t(s.navRX);
    const float cx = float(a % res) + 0.5f, cy = float(a / res) + 0.5f;
    const float px = cx * s.navCell, py = cy * s.navCell;
    const seated::F2 c{ s.navOX + px, s.navOY + py }, x{ p.x, p.y };
    const seated::Layout& L = *s.venue;
    if (seated::sameKey(seated::cellKeyC(L, s.venueCells, c), seated::cellKeyC(L, s.venueCells, x))) return true;
    seated::F2 n;
    return seated::stepUpScanC(L, s.venueCells, x, c, seated::floorHeightC(L, s.venueCells, x), n) >= 1.0f;
}
__device__ inline int choiceCell(const SceneView& s, V2 p) {
    const float gx = (p.x - s.navOX) * s.navInvCell, gy = (p.y - s.navOY) * s.navInvCell;
    const int cx = clampi(int(floorf(gx)), 0, s.navRX - 1), cy = clampi(int(floorf(gy)), 0, s.navRY - 1);
    int i = cx + cy * s.navRX;
    if (!(navRecOf(s, i)[0] & NAV_FIT)) {   // the record's own copy: the slots the caller reads next share its line
        float best = 1e9f;
        for (int r = 1; r <= 2 && best == 1e9f; r++)
            for (int dy = -r; dy <= r; dy++)
                for (int dx = -r; dx <= r; dx++) {
                    if ((dx < 0 ? -dx : dx) != r && (dy < 0 ? -dy : dy) != r) continue;
                    const int mx = cx + dx, my = cy + dy;
                    if (mx < 0 || my < 0 || mx >= s.navRX || my >= s.navRY) continue;
                    const int j = mx + my * s.navRX;
                    const uint32_t info = s.navInfo[j];
                    if (!(info & NAV_FIT) || !(info & NAV_HAS)) continue;
                    const float qx = gx - (float(mx) + 0.5f), qy = gy - (float(my) + 0.5f);
                    const float d2 = qx * qx + qy * qy;
                    if (d2 < best) { best = d2; i = j; }
                }
    }
    if (s.venue3d && !surfaceReaches(s, p, uint32_t(i))) {   // the surface rule (exit_choice.cpp:533-546)
        float best = 1e9f;
        int pick = i;
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int mx = cx + dx, my = cy + dy;
                if (mx < 0 || my < 0 || mx >= s.navRX || my >= s.navRY) continue;
                const int j = mx + my * s.navRX;
                const uint32_t info = s.navInfo[j];
                if (j == i || !(info & NAV_FIT) || !(info & NAV_HAS)) continue;
                const float qx = gx - (float(mx) + 0.5f), qy = gy - (float(my) + 0.5f);
                const float d2 = qx * qx + qy * qy;
                if (d2 < best && surfaceReaches(s, p, uint32_t(j))) { best = d2; pick = j; }
            }
        i = pick;
    }
    return i;
}
__device__ inline uint32_t slotExitOf(uint32_t w) { return (w >> CHOICE_EXIT_SHIFT) & 63u; }
__device__ inline float slotDistOf(const SceneView& s, uint32_t q) {
    return q == CHOICE_DIST_EMPTY ? __builtin_inff() : float(q) * s.distQuantum;
}
__device__ inline V2 anchorCentre(const SceneView& s, uint32_t a) {
    const uint32_t res = uint32_t(s.navRX);
    const float cx = float(a % res) + 0.5f, cy = float(a / res) + 0.5f;
    const float px = cx * s.navCell, py = cy * s.navCell;
    return V2{ s.navOX + px, s.navOY + py };
}
__device__ inline V2 choiceDirection(const SceneView& s, int cell, V2 p, uint32_t exit) {
    const uint32_t* sl = navSlots(s, cell);
    for (int k = 0; k < s.K; ++k) {
        const uint32_t w = sl[k];
        if (slotExitOf(w) != exit) continue;
        const V2 d = anchorCentre(s, w & CHOICE_ANCHOR_MASK) - p;
        const float l = len2(d);
        return l > GOAL_DEADZONE ? d / l : V2{ 0.0f, 0.0f };
    }
    return V2{ 0.0f, 0.0f };
}
// The exact wall query (seated::wall, src/venue.cpp:763-813, on a rectangle world): the nearest point of the wall
// boxes within 'reach' of p, as (unit direction away from the wall, clearance); w = WALL_NONE when none lies closer
// than reach. Candidates in ascending box order, the first of equal distances kept. On a box's boundary the
// direction is the normal of the face p lies on; strictly inside a box, the normal of its nearest face and the
// clearance the negative depth.
__device__ inline G3 wallQuery(const SceneView& s, V2 p, float reach) {
    const uint32_t c = hashCell(s, p.x, p.y);
    const uint32_t k0 = s.wallBinStart[c], k1 = s.wallBinStart[c + 1];
    float best = reach * reach;
    bool found = false;
    V2 bq{ 0.0f, 0.0f };
    float4 r{ 0.0f, 0.0f, 0.0f, 0.0f };   // the nearest box
    for (uint32_t k = k0; k < k1; ++k) {
        const float4 b = s.wallBinBox[k];
        const V2 q{ fminf(fmaxf(p.x, b.x), b.z), fminf(fmaxf(p.y, b.y), b.w) };
        const V2 dd = p - q;
        const float d2 = dot2(dd, dd);
        if (d2 < best) { best = d2; bq = q; r = b; found = true; }
    }
    if (!found) return G3{ 0.0f, 0.0f, WALL_NONE };
    const float d = sqrtf(best);
    if (d > 0.0f) { const V2 u = (p - bq) / d; return G3{ u.x, u.y, d }; }
    const V2 h{ p.x <= r.x ? -1.0f : (p.x >= r.z ? 1.0f : 0.0f), p.y <= r.y ? -1.0f : (p.y >= r.w ? 1.0f : 0.0f) };
    if (nonzero(h)) { const V2 u = h / len2(h); return G3{ u.x, u.y, 0.0f }; }
    // strictly inside: the nearest face (left, right, bottom, top; the first of equal depths)
    float dep = p.x - r.x;
    V2 u{ -1.0f, 0.0f };
    if (r.z - p.x < dep) { dep = r.z - p.x; u = V2{ 1.0f, 0.0f }; }
    if (p.y - r.y < dep) { dep = p.y - r.y; u = V2{ 0.0f, -1.0f }; }
    if (r.w - p.y < dep) { dep = r.w - p.y; u = V2{ 0.0f, 1.0f }; }
    return G3{ u.x, u.y, -dep };
}
// e_goal's wall rule (physics.comp navChoiceGoal :719-729, CPU nav::choiceGoal, src/exit_choice.cpp:568-578): the
// component of e into the nearest wall within reach removed, the rest renormalized (0 when nothing is left).
// *applied: the rule changed e; *wn: the wall direction it used.
__device__ inline V2 goalWallRule(const SceneView& s, V2 p, float reach, V2 e, bool* applied = nullptr, V2* wn = nullptr) {
    const G3 w = wallQuery(s, p, reach);
    const V2 n{ w.gx, w.gy };
    const float into = dot2(e, n);
    if (w.w < reach && into < 0.0f) {
        e = e - n * into;
        const float l = len2(e);
        e = l > SAFE_NORMALIZE_EPS ? e / l : V2{ 0.0f, 0.0f };
        if (applied) *applied = true;
        if (wn) *wn = n;
    }
    return e;
}
__device__ inline float estimatedTime(const SceneView& s, float dist, uint32_t exit, uint32_t queued) {
    const float walk = dist * s.secondsPerUnit;
    const float wait = float(queued) * s.invCapS[exit];
    return walk + wait;
}
__device__ inline uint32_t aheadBin(uint32_t q) {
    const uint32_t b = q >> CHOICE_AHEAD_SHIFT;
    return b < uint32_t(CHOICE_AHEAD_BINS - 1) ? b : uint32_t(CHOICE_AHEAD_BINS - 1);
}
__device__ inline int slotOfExit(const SceneView& s, int cell, uint32_t exit) {
    int k = -1;
    const uint32_t* sl = navSlots(s, cell);
    for (int i = 0; i < s.K; ++i)
        if (slotExitOf(sl[i]) == exit) k = i;
    return k;
}
__device__ inline int bestCandidate(const SceneView& s, int cell, const uint32_t* aRead, float& eBest) {
    int best = -1;
    eBest = 0.0f;
    const uint32_t* sl = navSlots(s, cell);
    const uint16_t* dq = navDists(s, cell);
    for (int k = 0; k < s.K; ++k) {
        const uint32_t ex = slotExitOf(sl[k]);
        if (ex == CHOICE_EMPTY) break;
        const uint32_t q = dq[k];
        const float e = estimatedTime(s, slotDistOf(s, q), ex, aRead[size_t(ex) * CHOICE_AHEAD_BINS + aheadBin(q)]);
        if (best < 0 || e < eBest) { best = k; eBest = e; }
    }
    return best;
}

__device__ inline int exitOf(const SceneView& s, V2 p) {
    for (int e = 0; e < s.E; ++e) {
        const float* r = s.exitRects + 4 * e;
        if (p.x >= r[0] && p.x <= r[2] && p.y >= r[1] && p.y <= r[3]) return e;
    }
    return -1;
}
__device__ inline bool parkTest(const SceneView& s, V2 p) {
    return exitOf(s, p) >= 0 || p.x < s.xmin || p.x > s.xmax || p.y < s.ymin || p.y > s.ymax;
}
// The seated venues' incline rules (physics.comp inclineFreeSpeed / inclineHeadwayTime, src/venue.cpp:652-670), with
// the agent's own free speed v0 in place of V_MAX (v0 R_STAIR is V_STAIR bit for bit at v0 = V_MAX): the free speed and
// the headway rule's time gap for the floor's slope g along the travel direction, blended by the smoothstep of
// |g| / INCLINE_REF_SLOPE
__device__ inline float inclineFreeSpeed(float v0, float g) {
    const float a = fabsf(g) * INCLINE_INV_REF_SLOPE;
    const float u = a < 0.0f ? 0.0f : (1.0f < a ? 1.0f : a);
    const float uu = u * u, w = 3.0f - 2.0f * u;
    const float t = uu * w;
    const float vStair = v0 * (g > 0.0f ? R_STAIR_UP : R_STAIR_DOWN);
    const float dv = vStair - v0;
    const float k = dv * t;
    return v0 + k;
}
__device__ inline float inclineHeadwayTime(float g) {
    const float a = fabsf(g) * INCLINE_INV_REF_SLOPE;
    const float u = a < 0.0f ? 0.0f : (1.0f < a ? 1.0f : a);
    const float uu = u * u, w = 3.0f - 2.0f * u;
    const float t = uu * w;
    const float tStair = g > 0.0f ? T_HEADWAY_STAIR_UP : T_HEADWAY_STAIR_DOWN;
    const float dt = tStair - T_HEADWAY;
    const float k = dt * t;
    return T_HEADWAY + k;
}
__device__ inline float stairFactor(const SceneView& s, V2 p) {
    const float gx = (p.x - s.navOX) * s.navInvCell, gy = (p.y - s.navOY) * s.navInvCell;
    const int cx = clampi(int(floorf(gx)), 0, s.navRX - 1), cy = clampi(int(floorf(gy)), 0, s.navRY - 1);
    return s.stair[size_t(cx) + size_t(cy) * size_t(s.navRX)];
}

// ---- static wall contacts (cpu_crowd_sim.cpp:173-500, 2D) ----
__device__ inline float rootStep(float br[4], float t, G3 g, V2 u) {
    if (g.w > 0.0f) { br[0] = t; br[1] = g.w; } else { br[2] = t; br[3] = g.w; }
    const float du = g.gx * u.x + g.gy * u.y;
    const float tn = du < 0.0f ? t - g.w / du : -1.0f;
    return tn > br[0] && tn < br[2] ? tn : (br[0] * br[3] - br[2] * br[1]) / (br[3] - br[1]);
}
__device__ inline float newtonStep(float t, G3 g, V2 u) {
    const float du = g.gx * u.x + g.gy * u.y;
    return du < 0.0f ? t - g.w / du : t;
}
// ---- wall contacts, with an optional recorder ----
// A recorder sees every push of resolveStatic / clearWalls and the wall plane's source, so that a taped forward
// (the adjoint's) is the production step bit for bit: NoRec compiles to nothing. A clearance measurement is kept
// as its local linear model: the clearance t0 measured from x0 along the ray u, and the distance field's gradient
// g at the hit point, so that t(x) = t0 - g.(x - x0) / (g.u) (exact for a flat wall).
struct ContactRec {
    V2 u, g, x0;
    float t0;
};
struct PlaneRec {
    int kind;          // 0 none, 1 a push (offset n.np), 2 a contact beyond the radius (offset n.np - (t - r)), 3 a probe
    V2 n;
    ContactRec c;
};
struct NoRec {
    static constexpr bool on = false;
    __device__ void push(V2, const ContactRec&, bool) {}
    __device__ void sweep(float, V2, bool) {}
};

template <bool REC>
__device__ inline float wallContact(const SceneView& s, V2 np, float fc, V2 u, float fr, float reach, float recastTol, V2& away,
                                    ContactRec* info = nullptr) {
    float br[4] = { 0.0f, fc, reach, fr };
    float t = reach * fc / (fc - fr);
    G3 g{ 0.0f, 0.0f, 0.0f };
    for (int i = 0; i < WALL_ROOT_STEPS; ++i) {
        g = mapExactGrad(s, np + u * t);
        t = rootStep(br, t, g, u);
    }
    float e = t;
    away = -u;
    const V2 nn = safeNormalize(V2{ g.gx, g.gy });
    const V2 u2 = -nn;
    float t2 = 0.0f, t3 = t * dot2(u, u2);
    G3 g2{ 0.0f, 0.0f, 0.0f };
    for (int i = 0; i < WALL_RECAST_STEPS; ++i) {
        t2 = t3;
        g2 = mapExactGrad(s, np + u2 * t2);
        t3 = newtonStep(t2, g2, u2);
    }
    const bool ok = (g2.gx * u2.x + g2.gy * u2.y) < 0.0f && fabsf(t3 - t2) < recastTol && t3 > 0.0f;
    const bool recast = ok && t3 < e;
    if (recast) { e = t3; away = nn; }
    if constexpr (REC) {
        const V2 ray = recast ? u2 : u;
        const G3 gh = mapExactGrad(s, np + ray * e);
        *info = ContactRec{ ray, V2{ gh.gx, gh.gy }, np, e };
    }
    return e;
}

template <class R>
__device__ inline bool clearWalls(const SceneView& s, V2& np, V2& nv, float rad, float wall[3], R& rec, PlaneRec* plane = nullptr) {
    bool moved = false;
    wall[0] = wall[1] = wall[2] = 0.0f;
    if constexpr (R::on) plane->kind = 0;
    const float reach = FEELER_REACH_RADII * rad;
    const float recastTol = WALL_RECAST_TOL_RADII * rad;
    for (int pass = 0; pass < WALL_PASSES; ++pass) {
        const float fc = mapSdf(s, np);
        float fr1 = 0.0f, fr2 = 0.0f;
        int i1 = -1, i2 = -1;
        for (int i = 0; i < FEELER_COUNT; ++i) {
            const V2 f = np + V2{ FEELER_DIRS[i].x * reach, FEELER_DIRS[i].y * reach };
            const float v = mapSdf(s, f);
            if (v < fr1) { fr2 = fr1; i2 = i1; fr1 = v; i1 = i; }
            else if (v < fr2) { fr2 = v; i2 = i; }
        }
        if (i1 < 0) break;
        V2 hNorm;
        float clearance;
        ContactRec ci{};
        if (fc > 0.0f) {
            clearance = wallContact<R::on>(s, np, fc, V2{ FEELER_DIRS[i1].x, FEELER_DIRS[i1].y }, fr1, reach, recastTol, hNorm, &ci);
            if (i2 >= 0) {
                V2 h2;
                ContactRec ci2{};
                const float c2 = wallContact<R::on>(s, np, fc, V2{ FEELER_DIRS[i2].x, FEELER_DIRS[i2].y }, fr2, reach, recastTol, h2, &ci2);
                if (c2 < clearance) { clearance = c2; hNorm = h2; ci = ci2; }
            }
        } else {
            const G3 g = mapExactGrad(s, np);
            hNorm = safeNormalize(V2{ g.gx, g.gy });
            const float gl = len2(V2{ g.gx, g.gy });
            const float L = gl > CENTRE_GRAD_MIN ? gl : CENTRE_GRAD_MIN;
            clearance = g.w / L;
            if constexpr (R::on) {   // the centre's clearance d / L: the ray u with g.u = -L
                const float g2n = g.gx * g.gx + g.gy * g.gy;
                const V2 u = g2n > 0.0f ? V2{ g.gx, g.gy } * (-L / g2n) : V2{ 0.0f, 0.0f };
                ci = ContactRec{ u, V2{ g.gx, g.gy }, np, clearance };
            }
        }
        if (clearance >= rad) {
            if (clearance < NO_CONTACT) {
                wall[0] = hNorm.x; wall[1] = hNorm.y; wall[2] = clearance - rad;
                if constexpr (R::on) *plane = PlaneRec{ 2, hNorm, ci };
            }
            break;
        }
        np = np + hNorm * (rad - clearance);
        moved = true;
        wall[0] = hNorm.x; wall[1] = hNorm.y; wall[2] = 0.0f;
        const float vDotN = dot2(nv, hNorm);
        if (vDotN < 0.0f) nv = nv - hNorm * vDotN;
        if constexpr (R::on) {   // the clearance's model is at the pre-push position
            rec.push(hNorm, ci, vDotN < 0.0f);
            *plane = PlaneRec{ 1, hNorm, ci };
        }
    }
    return moved;
}

template <class R>
__device__ inline void wallPlaneProbe(const SceneView& s, V2 np, float rad, float wall[3], R& rec, PlaneRec* plane = nullptr) {
    wall[0] = wall[1] = wall[2] = 0.0f;
    const float reach = PLANE_PROBE_RADII * rad;
    const float fc = mapSdf(s, np);
    float fr = 0.0f;
    int i1 = -1;
    for (int i = 0; i < FEELER_COUNT; ++i) {
        const V2 f = np + V2{ FEELER_DIRS[i].x * reach, FEELER_DIRS[i].y * reach };
        const float v = mapSdf(s, f);
        if (v < fr) { fr = v; i1 = i; }
    }
    if (i1 < 0 || fc <= 0.0f) return;
    V2 away;
    ContactRec ci{};
    const float c = wallContact<R::on>(s, np, fc, V2{ FEELER_DIRS[i1].x, FEELER_DIRS[i1].y }, fr, reach, WALL_RECAST_TOL_RADII * rad, away, &ci);
    if (c < NO_CONTACT) {
        wall[0] = away.x; wall[1] = away.y;
        const float m = c - rad;
        wall[2] = m > 0.0f ? m : 0.0f;
        if constexpr (R::on) *plane = PlaneRec{ m > 0.0f ? 3 : 1, away, ci };
    }
    (void)rec;
}

__device__ inline bool solidAt(const SceneView& s, V2 p) { return mapExactGrad(s, p).w <= 0.0f; }

template <class R>
__device__ inline bool sweepBack(const SceneView& s, V2 pOld, V2& np, V2& nv, R& rec) {
    if (!solidAt(s, np)) return false;
    float lo = 0.0f, hi = 1.0f;
    for (int i = 0; i < STATIC_SWEEP_BISECTIONS; ++i) {
        const float t = 0.5f * (lo + hi);
        if (solidAt(s, mix2(pOld, np, t))) hi = t;
        else lo = t;
    }
    const V2 d = safeNormalize(np - pOld);
    np = mix2(pOld, np, lo);
    const float vd = dot2(nv, d);
    if (vd > 0.0f) nv = nv - d * vd;
    if constexpr (R::on) rec.sweep(lo, d, vd > 0.0f);
    return true;
}

// resolveStatic on a flat floor: grounded unless the centre lies in a wall
template <class R>
__device__ inline bool resolveStatic(const SceneView& s, V2 pOld, V2& np, V2& nv, float rad, float wall[3], R& rec,
                                     PlaneRec* plane = nullptr) {
    if (wallFree(s, np)) {
        wall[0] = wall[1] = wall[2] = 0.0f;
        if constexpr (R::on) plane->kind = 0;
        return true;
    }
    bool grounded = !solidAt(s, np);
    if (!grounded && sweepBack(s, pOld, np, nv, rec)) grounded = !solidAt(s, np);
    const bool walls = clearWalls(s, np, nv, rad, wall, rec, plane);
    if (walls && grounded) {
        float w2[3];
        PlaneRec p2{};
        clearWalls(s, np, nv, rad, w2, rec, &p2);
        if (w2[0] != 0.0f || w2[1] != 0.0f) {
            wall[0] = w2[0]; wall[1] = w2[1]; wall[2] = w2[2];
            if constexpr (R::on) *plane = p2;
        }
    }
    return grounded;
}
__device__ inline bool resolveStatic(const SceneView& s, V2 pOld, V2& np, V2& nv, float rad, float wall[3]) {
    NoRec r;
    return resolveStatic(s, pOld, np, nv, rad, wall, r);
}

// ---- the same wall contacts shared by a slot's Q adjacent lanes ----
// Every lane of the slot calls these with the same arguments and gets the same results (the state stays identical
// across the lanes); only the work splits. The feelers: lane l samples feelers l, l + Q, ...; the clearing pass's
// choice (the deepest feeler below zero, then the next, a tie going to the lower feeler index: the sequential scan's
// rule) is the two smallest (value, index) pairs, merged across the lanes. The two contacts of a pass are measured
// on lanes 0 and 1 and exchanged.
struct Deep { float v; int i; };   // i < 0: none
__device__ inline bool deeper(Deep a, Deep b) { return a.i >= 0 && (b.i < 0 || a.v < b.v || (a.v == b.v && a.i < b.i)); }

template <int Q>
__device__ inline void deepestFeelers(const SceneView& s, V2 np, float reach, int lane, Deep& d1, Deep& d2) {
    d1 = Deep{ 0.0f, -1 };
    d2 = Deep{ 0.0f, -1 };
    for (int i = lane; i < FEELER_COUNT; i += Q) {   // ascending within the lane: the sequential rule on its share
        const V2 f = np + V2{ FEELER_DIRS[i].x * reach, FEELER_DIRS[i].y * reach };
        const float v = mapSdf(s, f);
        if (!(v < 0.0f)) continue;
        const Deep c{ v, i };
        if (deeper(c, d1)) { d2 = d1; d1 = c; }
        else if (deeper(c, d2)) d2 = c;
    }
#pragma unroll
    for (int m = 1; m < Q; m <<= 1) {
        const Deep o1{ __shfl_xor(d1.v, m, Q), __shfl_xor(d1.i, m, Q) };
        const Deep o2{ __shfl_xor(d2.v, m, Q), __shfl_xor(d2.i, m, Q) };
        // the two smallest of d1 <= d2 and o1 <= o2
        if (deeper(o1, d1)) { d2 = deeper(d1, o2) ? d1 : o2; d1 = o1; }
        else if (deeper(o1, d2)) d2 = o1;
    }
}

template <int Q>
__device__ inline bool clearWallsQ(const SceneView& s, V2& np, V2& nv, float rad, float wall[3], int lane) {
    bool moved = false;
    wall[0] = wall[1] = wall[2] = 0.0f;
    const float reach = FEELER_REACH_RADII * rad;
    const float recastTol = WALL_RECAST_TOL_RADII * rad;
    for (int pass = 0; pass < WALL_PASSES; ++pass) {
        const float fc = mapSdf(s, np);
        Deep d1, d2;
        deepestFeelers<Q>(s, np, reach, lane, d1, d2);
        if (d1.i < 0) break;
        V2 hNorm;
        float clearance;
        if (fc > 0.0f) {
            // lane 0 the deepest feeler's contact, lane 1 the next one's (the other lanes repeat lane 0's)
            const Deep d = (lane == 1 && d2.i >= 0) ? d2 : d1;
            V2 h;
            const float c = wallContact<false>(s, np, fc, V2{ FEELER_DIRS[d.i].x, FEELER_DIRS[d.i].y }, d.v, reach, recastTol, h);
            clearance = __shfl(c, 0, Q);
            hNorm = V2{ __shfl(h.x, 0, Q), __shfl(h.y, 0, Q) };
            if (d2.i >= 0) {
                const float c2 = __shfl(c, 1, Q);
                const V2 h2{ __shfl(h.x, 1, Q), __shfl(h.y, 1, Q) };
                if (c2 < clearance) { clearance = c2; hNorm = h2; }
            }
        } else {
            const G3 g = mapExactGrad(s, np);
            hNorm = safeNormalize(V2{ g.gx, g.gy });
            const float gl = len2(V2{ g.gx, g.gy });
            const float L = gl > CENTRE_GRAD_MIN ? gl : CENTRE_GRAD_MIN;
            clearance = g.w / L;
        }
        if (clearance >= rad) {
            if (clearance < NO_CONTACT) { wall[0] = hNorm.x; wall[1] = hNorm.y; wall[2] = clearance - rad; }
            break;
        }
        np = np + hNorm * (rad - clearance);
        moved = true;
        wall[0] = hNorm.x; wall[1] = hNorm.y; wall[2] = 0.0f;
        const float vDotN = dot2(nv, hNorm);
        if (vDotN < 0.0f) nv = nv - hNorm * vDotN;
    }
    return moved;
}

template <int Q>
__device__ inline bool resolveStaticQ(const SceneView& s, V2 pOld, V2& np, V2& nv, float rad, float wall[3], int lane) {
    if (wallFree(s, np)) {
        wall[0] = wall[1] = wall[2] = 0.0f;
        return true;
    }
    NoRec r;
    bool grounded = !solidAt(s, np);
    if (!grounded && sweepBack(s, pOld, np, nv, r)) grounded = !solidAt(s, np);
    const bool walls = clearWallsQ<Q>(s, np, nv, rad, wall, lane);
    if (walls && grounded) {
        float w2[3];
        clearWallsQ<Q>(s, np, nv, rad, w2, lane);
        if (w2[0] != 0.0f || w2[1] != 0.0f) { wall[0] = w2[0]; wall[1] = w2[1]; wall[2] = w2[2]; }
    }
    return grounded;
}

template <int Q>
__device__ inline void wallPlaneProbeQ(const SceneView& s, V2 np, float rad, float wall[3], int lane) {
    wall[0] = wall[1] = wall[2] = 0.0f;
    const float reach = PLANE_PROBE_RADII * rad;
    const float fc = mapSdf(s, np);
    Deep d1, d2;
    deepestFeelers<Q>(s, np, reach, lane, d1, d2);
    if (d1.i < 0 || fc <= 0.0f) return;
    V2 away;
    const float c = wallContact<false>(s, np, fc, V2{ FEELER_DIRS[d1.i].x, FEELER_DIRS[d1.i].y }, d1.v, reach, WALL_RECAST_TOL_RADII * rad, away);
    if (c < NO_CONTACT) {
        wall[0] = away.x; wall[1] = away.y;
        const float m = c - rad;
        wall[2] = m > 0.0f ? m : 0.0f;
    }
}
__device__ inline void wallPlaneProbe(const SceneView& s, V2 np, float rad, float wall[3]) {
    NoRec r;
    wallPlaneProbe(s, np, rad, wall, r);
}

} // namespace swarm
