// Device functions of a seated venue's own 3D step (gfx950): the reference's CpuCrowdSim / physics.comp on an analytic
// venue (VENUE_ANALYTIC), expression by expression in FP32 with the reference's operation order and its glm / std tie and
// signed-zero rules, so that the results are bit-identical to the reference's CPU step. Agents carry a height z: they
// rest on the venue's floor (gravity pulls, the floor holds), collide with the pitch fence, the exact walls of the
// heightfield at their body's top, and each other only when their heights overlap; a move that would climb a face higher
// than a step is cut at the face (the step-up limit).
//
// Reference: src/cpu_crowd_sim.cpp:94-112 (fenceCCD), :131-148 (separation, contact), :265-336 (floors), :367-424 (walls,
// wall plane), :445-500 (sweep-back, static contacts), :555-636 (step-up limit, blocked wall, face margin), :655-777
// (predict), :780-820 (projection), :835-866 (final); src/venue.cpp (the heightfield queries, seated_venue.h).
#pragma once
#include <hip/hip_runtime.h>

#include "crowd_device.h"
#include "seated_venue.h"

namespace swarm {
namespace seatedstep {

using seated::F2;
using seated::F3;
using seated::Layout;

// the venue as the step reads it: its layout and its cell table
struct Venue3 {
    const Layout& L;
    seated::CellTable cs;
};

struct V3 { float x, y, z; };
__device__ inline V2 xy(V3 a) { return V2{ a.x, a.y }; }
__device__ inline V3 v3(V2 a, float z) { return V3{ a.x, a.y, z }; }
__device__ inline void setXY(V3& v, V2 a) { v.x = a.x; v.y = a.y; }
__device__ inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }   // glm: (x x' + y y') + z z'
__device__ inline float len3(V3 a) { return sqrtf(dot3(a, a)); }
__device__ inline float mnf(float a, float b) { return b < a ? b : a; }   // std::min
__device__ inline float mxf(float a, float b) { return a < b ? b : a; }   // std::max

// contact bits (CpuCrowdSim::ContactBits)
enum : uint32_t { C_EXIT = 1, C_FENCE_CCD = 2, C_FENCE = 4, C_FLOOR = 8, C_WALL = 16, C_AGENT = 32, C_SWEEP = 64, C_STEP_UP = 128 };

__device__ inline float fence(V3 p) { return seated::sdFence(F2{ p.x, p.y }); }
__device__ inline V2 fenceGrad(V3 p) { const F2 g = seated::sdFenceGradient(F2{ p.x, p.y }); return V2{ g.x, g.y }; }
__device__ inline float box(V2 p) { return seated::sdBox2(F2{ p.x, p.y }, seated::FENCE_HALF_MID); }
__device__ inline float floorAt(const Venue3& V, float x, float y) { return seated::floorHeightC(V.L, V.cs, F2{ x, y }); }
__device__ inline seated::F3 wallAt(const Venue3& V, float x, float y, float z, float reach) {
    return seated::wallC(V.L, V.cs, F3{ x, y, z }, reach);
}
__device__ inline float venueMap(const Venue3& V, V3 p) { return seated::sdfGradC(V.L, V.cs, F3{ p.x, p.y, p.z }).w; }

// the step crossing the fence's midline is cut where it enters the fence's collision band (cpu_crowd_sim.cpp:94-112)
__device__ inline bool fenceCCD(V3 pOld, V3& pNew, V3& vel) {
    const float sOld = box(xy(pOld)), sNew = box(xy(pNew));
    if (sOld * sNew >= 0.0f) return false;
    const float side = sOld > 0.0f ? 1.0f : -1.0f;
    float lo = 0.0f;
    if (side * sOld > seated::FENCE_HALF_THK_COLL) {
        float hi = 1.0f;
        for (int i = 0; i < 8; ++i) {
            const float t = 0.5f * (lo + hi);
            if (side * box(mix2(xy(pOld), xy(pNew), t)) > seated::FENCE_HALF_THK_COLL) lo = t;
            else hi = t;
        }
    }
    setXY(pNew, mix2(xy(pOld), xy(pNew), lo));
    const V2 n = fenceGrad(pNew);
    const float vn = dot2(xy(vel), n);
    if (vn < 0.0f) setXY(vel, xy(vel) - n * vn);
    return true;
}

// the floor under a centre, the venue's branch of floorUnderReach (cpu_crowd_sim.cpp:265-271)
__device__ inline bool floorUnderReach(const Venue3& V, V3 c, float zFrom, float down, float& zFloor) {
    const float h = floorAt(V, c.x, c.y);
    zFloor = h;
    if (c.z <= h) return h <= mxf(zFrom, c.z) + BODY_RADIUS;
    return h >= c.z - down;
}
__device__ inline bool restOnFloor(const Venue3& V, V3& np, V3& nv, float zFrom) {
    float zF;
    if (floorUnderReach(V, np, zFrom, BODY_RADIUS, zF) && np.z < zF + BODY_RADIUS) { np.z = zF + BODY_RADIUS; nv.z = 0.0f; return true; }
    return false;
}
__device__ inline bool settleOnFloor(const Venue3& V, V3& np, V3& nv, float zFrom) {
    float zF;
    if (floorUnderReach(V, np, zFrom, FLOOR_SETTLE_REACH, zF)) { np.z = zF + BODY_RADIUS; nv.z = 0.0f; return true; }
    return false;
}
__device__ inline bool bakedSolid(const Venue3& V, V3 p) { return fence(p) >= 0.0f && venueMap(V, p) <= 0.0f; }
// a centre left inside the venue is taken back along its step at its start height (cpu_crowd_sim.cpp:447-463)
__device__ inline bool sweepBack(const Venue3& V, V3 pOld, V3& np, V3& nv) {
    if (!bakedSolid(V, np)) return false;
    float lo = 0.0f, hi = 1.0f;
    if (!bakedSolid(V, V3{ np.x, np.y, pOld.z })) lo = 1.0f;
    for (int i = 0; i < STATIC_SWEEP_BISECTIONS; ++i) {
        const float t = 0.5f * (lo + hi);
        if (bakedSolid(V, v3(mix2(xy(pOld), xy(np), t), pOld.z))) hi = t;
        else lo = t;
    }
    const V2 d = safeNormalize(xy(np) - xy(pOld));
    np = v3(mix2(xy(pOld), xy(np), lo), pOld.z);
    const float vd = dot2(xy(nv), d);
    if (vd > 0.0f) setXY(nv, xy(nv) - d * vd);
    nv.z = 0.0f;
    return true;
}
// the exact walls at the body's top pushed out to the wall clearance, at most WALL_PASSES times (cpu_crowd_sim.cpp:367-381)
__device__ inline bool clearWalls(const Venue3& V, V3& np, V3& nv, V3& wall) {
    bool moved = false;
    wall = V3{ 0.0f, 0.0f, 0.0f };
    for (int pass = 0; pass < WALL_PASSES; ++pass) {
        const F3 w = wallAt(V, np.x, np.y, np.z + V.L.wallAbove, FEELER_REACH);
        const V2 n{ w.x, w.y };
        if (w.z >= V.L.wallClear) {
            if (w.z < WALL_NONE) wall = v3(n, w.z - V.L.wallClear);
            break;
        }
        setXY(np, xy(np) + n * (V.L.wallClear - w.z));
        moved = true;
        wall = v3(n, 0.0f);
        const float vDotN = dot2(xy(nv), n);
        if (vDotN < 0.0f) setXY(nv, xy(nv) - n * vDotN);
    }
    return moved;
}
__device__ inline V3 wallPlaneProbe(const Venue3& V, V3 np) {   // cpu_crowd_sim.cpp:420-424
    const F3 w = wallAt(V, np.x, np.y, np.z + V.L.wallAbove, PLANE_PROBE_REACH);
    return w.z >= 0.0f && w.z < WALL_NONE ? V3{ w.x, w.y, mxf(w.z - V.L.wallClear, 0.0f) } : V3{ 0.0f, 0.0f, 0.0f };
}
// the static contacts of a centre that moved from pOld to np this step; returns whether it stands on a floor
// (cpu_crowd_sim.cpp:467-500)
__device__ inline bool resolveStatic(const Venue3& V, V3 pOld, V3& np, V3& nv, bool settle, V3& wall) {
    fenceCCD(pOld, np, nv);
    const float dFence = fence(np);
    if (dFence < 0.0f) {
        const V2 n = fenceGrad(np);
        setXY(np, xy(np) - n * dFence);
        const float vDotN = dot2(xy(nv), n);
        if (vDotN < 0.0f) setXY(nv, xy(nv) - n * vDotN);
    }
    bool grounded = settle ? settleOnFloor(V, np, nv, pOld.z) : restOnFloor(V, np, nv, pOld.z);
    if (!grounded && sweepBack(V, pOld, np, nv)) grounded = restOnFloor(V, np, nv, pOld.z);
    if (clearWalls(V, np, nv, wall)) {
        if (grounded) {
            settleOnFloor(V, np, nv, pOld.z);
            V3 w2;
            const bool again = clearWalls(V, np, nv, w2);
            if (w2.x != 0.0f || w2.y != 0.0f) wall = w2;
            if (again) restOnFloor(V, np, nv, pOld.z);
        } else {
            restOnFloor(V, np, nv, pOld.z);
        }
    }
    return grounded;
}
__device__ inline float scan(const Venue3& V, V2 a, V2 b, float f0, V2& n) {
    F2 nn;
    const float t = seated::stepUpScanC(V.L, V.cs, F2{ a.x, a.y }, F2{ b.x, b.y }, f0, nn);
    n = V2{ nn.x, nn.y };
    return t;
}
// a blocked move's end in a wall's clearance band: pushed out unless the push climbs a face (cpu_crowd_sim.cpp:591-605)
__device__ inline void venueBlockedWall(const Venue3& V, float f0, V3 pOld, V3& np, V3& nv) {
    const F3 w = wallAt(V, np.x, np.y, np.z + V.L.wallAbove, FEELER_REACH);
    if (w.z >= V.L.wallClear) return;
    const V2 n{ w.x, w.y };
    const V2 pushed = xy(np) + n * (V.L.wallClear - w.z);
    V2 nn;
    if (scan(V, xy(np), pushed, f0, nn) < 1.0f) return;
    const float vDotN = dot2(xy(nv), n);
    if (vDotN < 0.0f) setXY(nv, xy(nv) - n * vDotN);
    float zF;
    if (floorUnderReach(V, V3{ pushed.x, pushed.y, pOld.z }, pOld.z, FLOOR_SETTLE_REACH, zF)) {
        np = v3(pushed, zF + BODY_RADIUS);
        nv.z = 0.0f;
    } else {
        setXY(np, pushed);
    }
}
// every centre a face margin off the side face of a column higher than its feet (cpu_crowd_sim.cpp:610-636)
__device__ inline void venueEdgeMargin(const Venue3& V, V3 pOld, float f0, V3& np, V3& nv) {
    const float margin = V.L.faceMargin;
    float feet = np.z - BODY_RADIUS;
    const float fl = floorAt(V, np.x, np.y);
    if (feet > fl + margin) {
        const F3 w = wallAt(V, np.x, np.y, fl + seated::STEP_UP_TOL, margin);
        if (w.z < margin) {
            const V2 into = xy(np) - V2{ w.x, w.y } * (w.z + margin);
            const float top = floorAt(V, into.x, into.y);
            if (fabsf(top - feet) <= margin && top <= f0 + seated::STEP_UP_TOL) { setXY(np, into); return; }
        }
        if (np.z <= pOld.z) {
            const F3 wa = wallAt(V, np.x, np.y, feet + seated::STEP_UP_TOL, margin);
            if (wa.z < margin) setXY(np, xy(np) + V2{ wa.x, wa.y } * (margin - wa.z));
            return;
        }
        float zF;
        if (!floorUnderReach(V, V3{ np.x, np.y, pOld.z }, pOld.z, FLOOR_SETTLE_REACH, zF)) { np.z = pOld.z; return; }
        np.z = zF + BODY_RADIUS;
        nv.z = 0.0f;
        feet = zF;
    }
    const F3 w = wallAt(V, np.x, np.y, feet + seated::STEP_UP_TOL, margin);
    if (w.z >= margin) return;
    setXY(np, xy(np) + V2{ w.x, w.y } * (margin - w.z));
    np.z = floorAt(V, np.x, np.y) + BODY_RADIUS;
}
// the step-up limit with the venue's scan: a move that climbs a face is cut at the face and slid along it (0, 30 or
// 60 degrees off it, the first slide the scan allows); returns whether it cut the move (cpu_crowd_sim.cpp:555-587)
__device__ inline bool limitStepUp(const Venue3& V, V3 pOld, V3& np, V3& nv) {
    float f0;
    if (!floorUnderReach(V, pOld, pOld.z, STEP_UP_FLOOR_REACH, f0)) f0 = pOld.z - STEP_UP_FLOOR_REACH;
    V2 n;
    const float t = scan(V, xy(pOld), xy(np), f0, n);
    if (t >= 1.0f) {
        venueEdgeMargin(V, pOld, f0, np, nv);
        return false;
    }
    V2 q = mix2(xy(pOld), xy(np), t);
    const V2 rest = xy(np) - q;
    const V2 tangent = rest - n * dot2(rest, n);
    const float lt = len2(tangent);
    if (lt > 0.0f) {
        const float dirs[3][2] = { { 1.0f, 0.0f }, { 0.86602540f, 0.5f }, { 0.5f, 0.86602540f } };
        for (int k = 0; k < STEP_UP_SLIDE_TRIES; ++k) {
            const V2 slide = (q + tangent * dirs[k][0]) + n * (lt * dirs[k][1]);
            V2 n2;
            if (seated::sdFence(F2{ slide.x, slide.y }) >= 0.0f && scan(V, q, slide, f0, n2) >= 1.0f) { q = slide; break; }
        }
    }
    const float vn = dot2(xy(nv), n);
    if (vn < 0.0f) setXY(nv, xy(nv) - n * vn);
    float zF;
    if (floorUnderReach(V, V3{ q.x, q.y, pOld.z }, pOld.z, FLOOR_SETTLE_REACH, zF)) {
        np = v3(q, zF + BODY_RADIUS);
        nv.z = 0.0f;
    } else {
        np = v3(q, mnf(np.z, pOld.z));
    }
    venueBlockedWall(V, f0, pOld, np, nv);
    venueEdgeMargin(V, pOld, f0, np, nv);
    return true;
}
// separation and contact over 3D offsets at the reference's global radius (cpu_crowd_sim.cpp:131-148)
__device__ inline V2 separation3(V3 offset, uint32_t self, uint32_t other) {
    const float d2 = dot3(offset, offset);
    if (d2 >= R_SEP2) return V2{ 0.0f, 0.0f };
    const float w = 1.0f - sqrtf(d2) / R_SEP;
    const float magnitude = K_SEP * w * w;
    const float dxy2 = dot2(xy(offset), xy(offset));
    if (dxy2 <= SEP_STACKED_XY2) return magnitude * pairDirection(self, other);
    return magnitude * xy(offset) * (1.0f / sqrtf(dxy2));
}
__device__ inline V2 contact3(V3 offset, uint32_t self, uint32_t other) {
    const float dxy2 = dot2(xy(offset), xy(offset));
    if (fabsf(offset.z) >= BODY_DIAM || dxy2 >= BODY_DIAM2) return V2{ 0.0f, 0.0f };
    const float dxy = sqrtf(dxy2);
    const float shift = 0.5f * (BODY_DIAM - dxy);
    if (dxy2 <= SEP_STACKED_XY2) return shift * pairDirection(self, other);
    return shift * (xy(offset) / dxy);
}
// the obstacle avoidance's gradient at torso height: the tetrahedral stencil of the venue's distance in 3D
// (calcGradient, cpu_crowd_sim.cpp:297-303), normalized (calcNormal)
__device__ inline V2 venueNormalXY(const Venue3& V, V3 p) {
    const float H = OBSTACLE_GRAD_H;
    const float m1 = venueMap(V, V3{ p.x + H, p.y + -H, p.z + -H }), m2 = venueMap(V, V3{ p.x + -H, p.y + -H, p.z + H });
    const float m3 = venueMap(V, V3{ p.x + -H, p.y + H, p.z + -H }), m4 = venueMap(V, V3{ p.x + H, p.y + H, p.z + H });
    const V3 g{ (((m1 + -m2) + -m3) + m4) * OBSTACLE_GRAD_SCALE, (((-m1 + -m2) + m3) + m4) * OBSTACLE_GRAD_SCALE,
                (((-m1 + m2) + -m3) + m4) * OBSTACLE_GRAD_SCALE };
    const float l = len3(g);
    return l > SAFE_NORMALIZE_EPS ? V2{ g.x / l, g.y / l } : V2{ 0.0f, 0.0f };
}
// e_goal's wall term on the venue: its exact wall at the centre within FEELER_REACH (exit_choice.cpp:568-578)
__device__ inline V2 venueGoalWallRule(const Venue3& V, V3 p, V2 e) {
    const F3 w = wallAt(V, p.x, p.y, p.z, FEELER_REACH);
    const V2 n{ w.x, w.y };
    const float into = dot2(e, n);
    if (w.z < FEELER_REACH && into < 0.0f) {
        e = e - n * into;
        const float l = len2(e);
        e = l > SAFE_NORMALIZE_EPS ? e / l : V2{ 0.0f, 0.0f };
    }
    return e;
}

}  // namespace seatedstep
}  // namespace swarm
