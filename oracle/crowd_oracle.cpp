// Oracle: CPU restatement of the reference's per-step pipeline (CpuCrowdSim::step,
// src/cpu_crowd_sim.cpp:967-1088) on flat 2D floors.
//
// TEST INFRASTRUCTURE ONLY: the checker of the GPU path (tests, smoke check, CPU baseline).
//
// Operation order follows the reference expression by expression (float, no FMA contraction:
// build with -ffp-contract=off). The 2D restriction of the reference's 3D model: every agent stands on
// one flat floor, so the floor rest, gravity, pitch fence, fence CCD, sweep height test and step-up
// limit are identities, and wall normals have no vertical component. Generalisations: a rectangular
// domain and hash grid, per-agent radius r_i and free speed v0_i, a 2D wall-distance grid sampled
// bilinearly with exact weights, exits as rectangles, a stair speed-factor field, and a configurable
// projection ring (the reference searches 3x3: its cell is large enough, sim_params.hpp:533).
#include "crowd_oracle.h"
#include "venue_oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace {

// ---- model constants (include/swarm/sim_params.hpp), folded values bit-identical ----
constexpr float DT = 0.01f;
constexpr float K_SEP = 100.0f;
constexpr float R_SEP = 0.016f, BODY_DIAM = 0.0128f;
constexpr float SEP_MARGIN = R_SEP - BODY_DIAM;   // exact (Sterbenz)
constexpr float HEADWAY_RATE = 0x1.d26afcp+5f;     // 1.0 / T_HEADWAY (sim_params.hpp:511)
constexpr float R_PERC2 = 0x1.096bbap-9f;          // R_PERC * R_PERC (sim_params.hpp:503)
// the seated venues' incline rules (sim_params.hpp:441-481; EMPIRICAL's stair time gaps, exit_choice.cpp:43-48)
constexpr float V_MAX = 0.75f;
constexpr float T_HEADWAY = 0x1.1904b4p-6f;        // 0.3 s / TIME_SCALE_S_PER_UNIT
constexpr float V_STAIR_UP = 0x1.4c6afcp-2f;       // 0.58 m/s in model units
constexpr float V_STAIR_DOWN = 0x1.77672ap-2f;     // 0.655 m/s
constexpr float INCLINE_INV_REF_SLOPE = 0x1.f66da2p+0f;   // 1 / tan(27 deg)
constexpr float T_HEADWAY_STAIR_UP = 0x1.24ba3cp-5f;      // 0.625 s
constexpr float T_HEADWAY_STAIR_DOWN = 0x1.306fc4p-5f;    // 0.65 s
constexpr float PROJ_RELAX = 1.0f;
constexpr float PROJ_MAX_SHIFT_RADII = 0.05f;
constexpr float W_OBS = 10.0f;
constexpr float GOAL_DEADZONE = 1e-4f;
constexpr float WALL_NONE = 1e9f;   // venue.hpp:251
constexpr float OBSTACLE_RANGE = 0.025f;
constexpr float OBSTACLE_MIN_NORMAL = 0.01f;
constexpr float OBSTACLE_GRAD_H = 0.0015f;
constexpr float OBSTACLE_GRAD_SCALE = 0x1.4d5556p+7f;   // 0.25 / OBSTACLE_GRAD_H (sim_params.hpp:505)
constexpr float SAFE_NORMALIZE_EPS = 1e-8f;
constexpr float SEP_STACKED_XY2 = 1e-12f;
constexpr float FEELER_REACH_RADII = 1.25f;
constexpr int WALL_ROOT_STEPS = 3, WALL_RECAST_STEPS = 2, WALL_PASSES = 3, STATIC_SWEEP_BISECTIONS = 8;
constexpr float WALL_RECAST_TOL_RADII = 0.02f;
constexpr float CENTRE_GRAD_MIN = 1e-3f;
constexpr float PLANE_PROBE_RADII = 2.5f;
constexpr float PARK_COORD = 1000.0f;
constexpr uint32_t INACTIVE_CELL = 0xFFFFFFFFu;
constexpr float HASH_TO_RADIANS = 0x1.921fb6p-30f;
constexpr int CHOICE_STEPS = 17;
constexpr float CHOICE_REL_MARGIN = 0.10f, CHOICE_ABS_MARGIN_S = 5.0f, CHOICE_SWITCH_SHARE = 0.10f;
constexpr uint32_t CHOICE_MIN_SWITCHES = 1u;
constexpr int CHOICE_GAIN_BINS = 256;
constexpr int CHOICE_EXIT_SHIFT = 25;
constexpr uint32_t CHOICE_ANCHOR_MASK = (1u << CHOICE_EXIT_SHIFT) - 1u;
constexpr uint32_t CHOICE_EMPTY = 63u, CHOICE_DIST_EMPTY = 0xFFFFu, CHOICE_NONE = 0xFFFFFFFFu;
constexpr int CHOICE_MAX_EXITS = 63, CHOICE_QUEUE_STRIDE = 64, CHOICE_AHEAD_BINS = 1024, CHOICE_AHEAD_SHIFT = 6;
constexpr uint32_t NAV_FIT = 0x80000000u, NAV_HAS = 0x40000000u;
constexpr float PLANE_WALL = 1.0f, PLANE_GROUNDED = 2.0f;

const float FEELER_DIRS[16][2] = {
    { 1.0f, 0.0f }, { 0.92387953f, 0.38268343f }, { 0.70710678f, 0.70710678f }, { 0.38268343f, 0.92387953f },
    { 0.0f, 1.0f }, { -0.38268343f, 0.92387953f }, { -0.70710678f, 0.70710678f }, { -0.92387953f, 0.38268343f },
    { -1.0f, 0.0f }, { -0.92387953f, -0.38268343f }, { -0.70710678f, -0.70710678f }, { -0.38268343f, -0.92387953f },
    { 0.0f, -1.0f }, { 0.38268343f, -0.92387953f }, { 0.70710678f, -0.70710678f }, { 0.92387953f, -0.38268343f } };

// ---- glm semantics written out (external/glm: dot = x*x' + y*y', length = sqrt(dot), mix = x(1-a) + y a) ----
struct V2 { float x, y; };
inline V2 v2(float x, float y) { return V2{ x, y }; }
inline V2 operator+(V2 a, V2 b) { return { a.x + b.x, a.y + b.y }; }
inline V2 operator-(V2 a, V2 b) { return { a.x - b.x, a.y - b.y }; }
inline V2 operator-(V2 a) { return { -a.x, -a.y }; }
inline V2 operator*(V2 a, float s) { return { a.x * s, a.y * s }; }
inline V2 operator*(float s, V2 a) { return { s * a.x, s * a.y }; }
inline V2 operator/(V2 a, float s) { return { a.x / s, a.y / s }; }
inline float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline float length(V2 a) { return std::sqrt(dot(a, a)); }
inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
inline V2 mix2(V2 a, V2 b, float t) { return { mixf(a.x, b.x, t), mixf(a.y, b.y, t) }; }
inline bool nonzero(V2 a) { return a.x != 0.0f || a.y != 0.0f; }
V2 safeNormalize(V2 v) { float l = length(v); return l > SAFE_NORMALIZE_EPS ? v / l : V2{ 0.0f, 0.0f }; }   // physics.comp / cpu_crowd_sim.cpp:89
struct G3 { float gx, gy, w; };   // mapExactGrad's vec4 (gradient xy, value)

uint32_t hash32(uint32_t x) {   // lowbias32 (cpu_crowd_sim.cpp:116-121)
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// The stacked-pair direction's sine and cosine: Cody-Waite reduction by pi/2 and the Cephes single
// precision polynomials (S. L. Moshier, sinf.c / cosf.c), the same float operations as the GPU's.
void sincosCephes(float x, float* s, float* c) {
    const float FOPI = 1.27323954473516f;
    const float DP1 = 0.78515625f, DP2 = 2.4187564849853515625e-4f, DP3 = 3.77489497744594108e-8f;
    int j = (int)(x * FOPI);
    float y = (float)j;
    if (j & 1) { j += 1; y += 1.0f; }
    j &= 7;
    const float z = ((x - y * DP1) - y * DP2) - y * DP3;
    const float zz = z * z;
    const float ps = ((-1.9515295891e-4f * zz + 8.3321608736e-3f) * zz - 1.6666654611e-1f) * zz * z + z;
    const float pc = ((2.443315711809948e-5f * zz - 1.388731625493765e-3f) * zz + 4.166664568298827e-2f) * zz * zz - 0.5f * zz + 1.0f;
    float sn, cs;
    switch (j) {
    case 0: sn = ps; cs = pc; break;
    case 2: sn = pc; cs = -ps; break;
    case 4: sn = -ps; cs = -pc; break;
    default: sn = -pc; cs = ps; break;
    }
    *s = sn; *c = cs;
}

V2 pairDirection(uint32_t self, uint32_t other) {   // cpu_crowd_sim.cpp:123-129
    uint32_t lo = std::min(self, other), hi = std::max(self, other);
    float angle = float(hash32(hash32(lo) ^ hi)) * HASH_TO_RADIANS;
    float s, c;
    sincosCephes(angle, &s, &c);
    V2 dir{ c, s };
    return self == lo ? dir : -dir;
}

// Squared support and contact distances: the reference's folded constants R_SEP * R_SEP and BODY_DIAM * BODY_DIAM
// (sim_params.hpp:502,512; the squares of the decimals, one float step off the float squares) where the pair has the
// reference's distances, else the float square
constexpr float R_SEP2 = 0x1.0c6f7ap-12f, BODY_DIAM2 = 0x1.5798eep-13f;
inline float sepSupport2(float S) { return S == R_SEP ? R_SEP2 : S * S; }
inline float contact2(float D) { return D == BODY_DIAM ? BODY_DIAM2 : D * D; }

// separationForce (cpu_crowd_sim.cpp:131-139): support S = r_i + r_j + SEP_MARGIN
V2 separationForce(V2 offset, uint32_t self, uint32_t other, float S) {
    float d2 = dot(offset, offset);
    if (d2 >= sepSupport2(S)) return V2{ 0.0f, 0.0f };
    float w = 1.0f - std::sqrt(d2) / S;
    float magnitude = K_SEP * w * w;
    float dxy2 = dot(offset, offset);
    if (dxy2 <= SEP_STACKED_XY2) return magnitude * pairDirection(self, other);
    return magnitude * offset * (1.0f / std::sqrt(dxy2));
}

// contactCorrection (cpu_crowd_sim.cpp:141-148): contact distance D = r_i + r_j
V2 contactCorrection(V2 offset, uint32_t self, uint32_t other, float D) {
    float dxy2 = dot(offset, offset);
    if (dxy2 >= contact2(D)) return V2{ 0.0f, 0.0f };
    float dxy = std::sqrt(dxy2);
    float shift = 0.5f * (D - dxy);
    if (dxy2 <= SEP_STACKED_XY2) return shift * pairDirection(self, other);
    return shift * (offset / dxy);
}

float rootStep(float br[4], float t, G3 g, V2 u) {   // cpu_crowd_sim.cpp:173-178; br = (a, f(a), b, f(b))
    if (g.w > 0.0f) { br[0] = t; br[1] = g.w; } else { br[2] = t; br[3] = g.w; }
    float du = g.gx * u.x + g.gy * u.y;
    float tn = du < 0.0f ? t - g.w / du : -1.0f;
    return tn > br[0] && tn < br[2] ? tn : (br[0] * br[3] - br[2] * br[1]) / (br[3] - br[1]);
}

float newtonStep(float t, G3 g, V2 u) {   // cpu_crowd_sim.cpp:180-183
    float du = g.gx * u.x + g.gy * u.y;
    return du < 0.0f ? t - g.w / du : t;
}

struct SlotState { float x, y, w, vx, vy, acc, z, vz; };   // position xy, active flag, velocity xy, projection displacement (velocity.w); the seated step's z

// ---- the seated 3D step's constants (sim_params.hpp; folded like the reference) ----
constexpr float BODY_RADIUS = 0.0064f;
constexpr float PROJ_MAX_SHIFT = 0x1.4f8b58p-12f;        // PROJ_MAX_SHIFT_RADII * BODY_RADIUS (0.00032)
constexpr float FEELER_REACH = 0x1.0624dep-7f;           // FEELER_REACH_RADII * BODY_RADIUS (0.008)
constexpr float PLANE_PROBE_REACH = 0x1.0624dep-6f;      // PLANE_PROBE_RADII * BODY_RADIUS (0.016)
constexpr float GRAVITY_DT = 0x1.ebb3b0p-1f;             // GRAVITY * DT (0.9603553)
constexpr float FLOOR_SETTLE_REACH = 0x1.3a92a4p-7f;     // FLOOR_SETTLE_RADII * BODY_RADIUS (0.0096)
constexpr float STEP_UP_FLOOR_REACH = 0x1.3a92a4p-5f;    // STEP_UP_FLOOR_REACH_RADII * BODY_RADIUS (0.0384)
constexpr float STEP_UP_TOL = 1e-6f;
constexpr float TORSO_HEIGHT = 0.05f;
constexpr float TAU_RELAX_REF = 0x1.d45d2cp-6f, RELAX_STEP_REF = 0x1.68de1cp-1f, OBSTACLE_GAIN_REF = 20.0f;
constexpr float FENCE_HALF_MID = 0.5f * (0.328395f + 0.335432f);
constexpr float FENCE_HALF_THK_COLL = 0.5f * (0.335432f - 0.328395f) + BODY_RADIUS;
constexpr int FENCE_CCD_BISECTIONS = 8, STEP_UP_SLIDE_TRIES = 3;
const float STEP_UP_SLIDE_DIRS[3][2] = { { 1.0f, 0.0f }, { 0.86602540f, 0.5f }, { 0.5f, 0.86602540f } };
enum : uint8_t { C_EXIT = 1, C_FENCE_CCD = 2, C_FENCE = 4, C_FLOOR = 8, C_WALL = 16, C_AGENT = 32, C_SWEEP = 64, C_STEP_UP = 128 };

struct V3 { float x, y, z; };
inline V3 v3(V2 a, float z) { return V3{ a.x, a.y, z }; }
inline V2 xy(V3 a) { return V2{ a.x, a.y }; }
inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }   // glm: (x x' + y y') + z z'
inline float length3(V3 a) { return std::sqrt(dot3(a, a)); }

float sdBox2(V2 p, float h) {   // sdBox2 (src/sdf.cpp:45-48), half size (h, h)
    const V2 q{ std::abs(p.x) - h, std::abs(p.y) - h };
    const V2 m{ std::max(q.x, 0.0f), std::max(q.y, 0.0f) };
    return length(m) + std::min(std::max(q.x, q.y), 0.0f);
}
V2 sdBox2Grad(V2 p, float h) {   // src/sdf.cpp:50-56
    const V2 q{ std::abs(p.x) - h, std::abs(p.y) - h };
    const V2 sg{ p.x < 0.0f ? -1.0f : 1.0f, p.y < 0.0f ? -1.0f : 1.0f };
    V2 g;
    if (std::max(q.x, q.y) > 0.0f) {
        const V2 m{ std::max(q.x, 0.0f), std::max(q.y, 0.0f) };
        g = m / length(m);
    } else {
        g = q.x > q.y ? V2{ 1.0f, 0.0f } : V2{ 0.0f, 1.0f };
    }
    return V2{ sg.x * g.x, sg.y * g.y };
}
// the pitch fence (src/sdf.cpp:64-82: the closed ring about the origin, no gates), inflated by BODY_RADIUS
float sdFence(V3 p) { return std::abs(sdBox2(V2{ p.x - 0.0f, p.y - 0.0f }, FENCE_HALF_MID)) - FENCE_HALF_THK_COLL; }
V2 sdFenceGradient(V3 p) {
    const V2 rel{ p.x - 0.0f, p.y - 0.0f };
    const float box = sdBox2(rel, FENCE_HALF_MID);
    return (box < 0.0f ? -1.0f : 1.0f) * sdBox2Grad(rel, FENCE_HALF_MID);
}
bool fenceCCD(V3 pOld, V3& pNew, V3& vel) {   // cpu_crowd_sim.cpp:94-112
    const float sOld = sdBox2(xy(pOld), FENCE_HALF_MID), sNew = sdBox2(xy(pNew), FENCE_HALF_MID);
    if (sOld * sNew >= 0.0f) return false;
    const float side = sOld > 0.0f ? 1.0f : -1.0f;
    float lo = 0.0f;
    if (side * sOld > FENCE_HALF_THK_COLL) {
        float hi = 1.0f;
        for (int i = 0; i < FENCE_CCD_BISECTIONS; ++i) {
            const float t = 0.5f * (lo + hi);
            if (side * sdBox2(mix2(xy(pOld), xy(pNew), t), FENCE_HALF_MID) > FENCE_HALF_THK_COLL) lo = t;
            else hi = t;
        }
    }
    const V2 q = mix2(xy(pOld), xy(pNew), lo);
    pNew.x = q.x; pNew.y = q.y;
    const V2 n = sdFenceGradient(pNew);
    const float vn = dot(xy(vel), n);
    if (vn < 0.0f) { const V2 u = xy(vel) - n * vn; vel.x = u.x; vel.y = u.y; }
    return true;
}

} // namespace

struct oracle_ctx {
    // scene (copied)
    oracle_scene sc;
    std::vector<float> sdf, invCapS, stair, exitRects, walls;
    std::vector<uint32_t> navInfo, navSlot;
    std::vector<uint16_t> navDist;
    oracle_params par;
    int n = 0, threads = 1;
    // agent state, ping-pong in agent order
    std::vector<float> x[2], y[2], vx[2], vy[2], spd[2], w[2];
    std::vector<float> z[2], vz[2];   // the seated step's heights and vertical velocities
    std::vector<uint8_t> contacts;    // the seated step's contact bits per agent
    std::vector<float> r, v0;
    int readIdx = 0, writeIdx = 1;
    bool bitonic = false;   // SortMode::GpuBitonic instead of the stable counting sort
    std::vector<float> driveX, driveY;
    bool drive = false;
    // grid
    std::vector<uint32_t> key, id, keyTmp, idTmp, bucketStart, cellStart, cellCount;
    // slot order
    std::vector<SlotState> slotA, slotB;
    std::vector<float> planes;   // 4 per slot
    std::vector<float> egX, egY;   // e_goal per agent (reference: goalDirection, evaluated on demand)
    // exit choice (nav::ChoiceState)
    std::vector<uint32_t> assign;
    uint32_t queue[128] = {}, switches = 0, forced = 0, evaluations = 0, eligible = 0;
    uint32_t choiceStep = 0;
    std::vector<uint8_t> pendEligible;
    std::vector<uint32_t> pendBest, pendBin, hist, threshold, ahead;
    // evacuation
    std::vector<int32_t> evacStep, evacExit;
    int steps = 0;
    double times[8] = {};

    // ---------------- sampling ----------------
    // mapExactGrad (cpu_crowd_sim.cpp:239-262) on the 2D grid; map() shares its arithmetic (exact weights)
    G3 mapExactGrad(V2 p) const {
        const float fx = (p.x - sc.sdfOX) * sc.sdfInvH, fy = (p.y - sc.sdfOY) * sc.sdfInvH;
        const bool outside = fx < -0.5f || fy < -0.5f || fx > float(sc.sdfNX) - 0.5f || fy > float(sc.sdfNY) - 0.5f;
        const int ix = std::clamp(int(std::floor(fx)), 0, sc.sdfNX - 2), iy = std::clamp(int(std::floor(fy)), 0, sc.sdfNY - 2);
        const float wx = fx - float(ix), wy = fy - float(iy);
        const float* row0 = &sdf[size_t(iy) * size_t(sc.sdfNX)];
        const float* row1 = row0 + sc.sdfNX;
        const float c00 = row0[ix], c10 = row0[ix + 1], c01 = row1[ix], c11 = row1[ix + 1];
        const float x0 = mixf(c00, c10, wx), x1 = mixf(c01, c11, wx);
        if (outside) return G3{ 0.0f, 0.0f, 1.0f };
        return G3{ mixf(c10 - c00, c11 - c01, wy) * sc.sdfInvH, (x1 - x0) * sc.sdfInvH, mixf(x0, x1, wy) };
    }
    float map(V2 p) const { return mapExactGrad(p).w; }

    // calcGradient / calcNormal (cpu_crowd_sim.cpp:297-305): the tetrahedral stencil's xy rows
    V2 calcGradient(V2 p) const {
        const float H = OBSTACLE_GRAD_H;
        const float m1 = map(v2(p.x + H, p.y + -H)), m2 = map(v2(p.x + -H, p.y + -H));
        const float m3 = map(v2(p.x + -H, p.y + H)), m4 = map(v2(p.x + H, p.y + H));
        const V2 g{ ((m1 + -m2) + -m3) + m4, ((-m1 + -m2) + m3) + m4 };
        return g * OBSTACLE_GRAD_SCALE;
    }

    // ---------------- navigation (exit_choice.cpp:512-578) ----------------
    bool fit(size_t i) const { return (navInfo[i] & NAV_FIT) != 0u; }
    bool has(size_t i) const { return (navInfo[i] & NAV_HAS) != 0u; }
    int choiceCell(V2 p) const {   // choiceCell (exit_choice.cpp:512-548), no surface rule
        const float gx = (p.x - sc.navOX) * sc.navInvCell, gy = (p.y - sc.navOY) * sc.navInvCell;
        const int cx = std::clamp(int(std::floor(gx)), 0, sc.navRX - 1), cy = std::clamp(int(std::floor(gy)), 0, sc.navRY - 1);
        int i = cx + cy * sc.navRX;
        if (!fit(size_t(i))) {
            float best = 1e9f;
            for (int r = 1; r <= 2 && best == 1e9f; r++)
                for (int dy = -r; dy <= r; dy++)
                    for (int dx = -r; dx <= r; dx++) {
                        if (std::max(std::abs(dx), std::abs(dy)) != r) continue;
                        const int mx = cx + dx, my = cy + dy;
                        if (mx < 0 || my < 0 || mx >= sc.navRX || my >= sc.navRY) continue;
                        const int j = mx + my * sc.navRX;
                        if (!fit(size_t(j)) || !has(size_t(j))) continue;
                        const float qx = gx - (float(mx) + 0.5f), qy = gy - (float(my) + 0.5f);
                        const float d2 = qx * qx + qy * qy;
                        if (d2 < best) { best = d2; i = j; }
                    }
        }
        if (sc.venue3d && !surfaceReaches(p, anchorCentre(uint32_t(i)))) {   // the surface rule (exit_choice.cpp:533-546)
            float best = 1e9f;
            int pick = i;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    const int mx = cx + dx, my = cy + dy;
                    if (mx < 0 || my < 0 || mx >= sc.navRX || my >= sc.navRY) continue;
                    const int j = mx + my * sc.navRX;
                    if (j == i || !fit(size_t(j)) || !has(size_t(j))) continue;
                    const float qx = gx - (float(mx) + 0.5f), qy = gy - (float(my) + 0.5f);
                    const float d2 = qx * qx + qy * qy;
                    if (d2 < best && surfaceReaches(p, anchorCentre(uint32_t(j)))) { best = d2; pick = j; }
                }
            i = pick;
        }
        return i;
    }
    // surfaceReaches (nav_field.cpp:381-385): the same venue cell, or a move the step-up scan allows
    bool surfaceReaches(V2 p, V2 c) const {
        float kc[2], kp[2], n[2];
        vo::cellKey(sc.venue, c.x, c.y, kc);
        vo::cellKey(sc.venue, p.x, p.y, kp);
        if (kc[0] == kp[0] && kc[1] == kp[1]) return true;
        return vo::stepUpScan(sc.venue, p.x, p.y, c.x, c.y, vo::floorHeight(sc.venue, p.x, p.y), n) >= 1.0f;
    }
    uint32_t slotExit(size_t cell, int k) const { return (navSlot[cell * size_t(sc.K) + size_t(k)] >> CHOICE_EXIT_SHIFT) & 63u; }
    uint32_t slotAnchor(size_t cell, int k) const { return navSlot[cell * size_t(sc.K) + size_t(k)] & CHOICE_ANCHOR_MASK; }
    uint32_t slotDistQ(size_t cell, int k) const { return navDist[cell * size_t(sc.K) + size_t(k)]; }
    float slotDist(size_t cell, int k) const {
        const uint32_t q = slotDistQ(cell, k);
        return q == CHOICE_DIST_EMPTY ? std::numeric_limits<float>::infinity() : float(q) * sc.distQuantum;
    }
    V2 anchorCentre(uint32_t a) const {   // exit_choice.cpp:550-555
        const uint32_t res = uint32_t(sc.navRX);
        const float cx = float(a % res) + 0.5f, cy = float(a / res) + 0.5f;
        const float px = cx * sc.navCell, py = cy * sc.navCell;
        return V2{ sc.navOX + px, sc.navOY + py };
    }
    V2 choiceDirection(V2 p, uint32_t exit) const {   // exit_choice.cpp:557-566
        const size_t c = size_t(choiceCell(p));
        for (int k = 0; k < sc.K; ++k) {
            if (slotExit(c, k) != exit) continue;
            const V2 d = anchorCentre(slotAnchor(c, k)) - p;
            const float l = length(d);
            return l > GOAL_DEADZONE ? d / l : V2{ 0.0f, 0.0f };
        }
        return V2{ 0.0f, 0.0f };
    }
    // seated::wall (venue.cpp:763-813) on a rectangle world: over every wall box, the nearest point within reach
    // (strict <, so the first of equal distances); (direction away from it, clearance), WALL_NONE when none. On the
    // box's boundary the direction is the normal of the face p lies on (venue.cpp:806-810); strictly inside a box, the
    // normal of its nearest face and the negative depth.
    G3 wallQuery(V2 p, float reach) const {
        float best = reach * reach;
        int bk = -1;
        V2 bq{ 0.0f, 0.0f };
        for (size_t k = 0; k < walls.size() / 4; ++k) {
            const float* r = &walls[4 * k];
            const V2 q{ std::min(std::max(p.x, r[0]), r[2]), std::min(std::max(p.y, r[1]), r[3]) };
            const V2 dd = p - q;
            const float d2 = dot(dd, dd);
            if (d2 < best) { best = d2; bq = q; bk = int(k); }
        }
        if (bk < 0) return G3{ 0.0f, 0.0f, WALL_NONE };
        const float d = std::sqrt(best);
        if (d > 0.0f) { const V2 u = (p - bq) / d; return G3{ u.x, u.y, d }; }
        const float* r = &walls[4 * size_t(bk)];
        const V2 h{ p.x <= r[0] ? -1.0f : (p.x >= r[2] ? 1.0f : 0.0f), p.y <= r[1] ? -1.0f : (p.y >= r[3] ? 1.0f : 0.0f) };
        if (h.x != 0.0f || h.y != 0.0f) { const V2 u = h / length(h); return G3{ u.x, u.y, 0.0f }; }
        float dep = p.x - r[0];
        V2 u{ -1.0f, 0.0f };
        if (r[2] - p.x < dep) { dep = r[2] - p.x; u = V2{ 1.0f, 0.0f }; }
        if (p.y - r[1] < dep) { dep = p.y - r[1]; u = V2{ 0.0f, -1.0f }; }
        if (r[3] - p.y < dep) { dep = r[3] - p.y; u = V2{ 0.0f, 1.0f }; }
        return G3{ u.x, u.y, -dep };
    }
    // choiceGoal (exit_choice.cpp:568-578): the direction's component into the nearest wall within FEELER_REACH
    // (FEELER_REACH_RADII r per agent) removed, the rest renormalized
    V2 choiceGoal(V2 p, float rad, uint32_t exit, bool* applied = nullptr) const {
        V2 e = choiceDirection(p, exit);
        const float reach = FEELER_REACH_RADII * rad;
        const G3 w = wallQuery(p, reach);
        const V2 n{ w.gx, w.gy };
        const float into = dot(e, n);
        if (w.w < reach && into < 0.0f) {
            e = e - n * into;
            const float l = length(e);
            e = l > SAFE_NORMALIZE_EPS ? e / l : V2{ 0.0f, 0.0f };
            if (applied) *applied = true;
        }
        return e;
    }
    V2 goalDirection(V2 p, uint32_t agent) const {   // cpu_crowd_sim.cpp:213-218
        if (drive) return V2{ driveX[agent], driveY[agent] };
        if (sc.venue3d) return seatedChoiceGoal(V3{ p.x, p.y, z[readIdx][agent] }, assign[agent]);
        return choiceGoal(p, r[agent], assign[agent]);
    }
    // choiceGoal on the venue (exit_choice.cpp:568-578): its exact wall at the centre within FEELER_REACH
    V2 seatedChoiceGoal(V3 p, uint32_t exit) const {
        V2 e = choiceDirection(xy(p), exit);
        float w[3];
        vo::wall(sc.venue, p.x, p.y, p.z, FEELER_REACH, w);
        const V2 n{ w[0], w[1] };
        const float into = dot(e, n);
        if (w[2] < FEELER_REACH && into < 0.0f) {
            e = e - n * into;
            const float l = length(e);
            e = l > SAFE_NORMALIZE_EPS ? e / l : V2{ 0.0f, 0.0f };
        }
        return e;
    }
    float estimatedTime(float dist, uint32_t exit, uint32_t queued) const {   // exit_choice.cpp:580-584
        const float walk = dist * sc.secondsPerUnit;
        const float wait = float(queued) * invCapS[exit];
        return walk + wait;
    }
    int slotOf(size_t cell, uint32_t exit) const {
        int k = -1;
        for (int i = 0; i < sc.K; ++i)
            if (slotExit(cell, i) == exit) k = i;
        return k;
    }
    uint32_t aheadOf(const uint32_t* aRead, uint32_t exit, uint32_t q) const {
        return aRead[size_t(exit) * CHOICE_AHEAD_BINS + std::min(q >> CHOICE_AHEAD_SHIFT, uint32_t(CHOICE_AHEAD_BINS - 1))];
    }
    int bestCandidate(size_t cell, const uint32_t* aRead, float& eBest) const {   // exit_choice.cpp:593-603
        int best = -1;
        eBest = 0.0f;
        for (int k = 0; k < sc.K; ++k) {
            const uint32_t ex = slotExit(cell, k);
            if (ex == CHOICE_EMPTY) break;
            const float e = estimatedTime(slotDist(cell, k), ex, aheadOf(aRead, ex, slotDistQ(cell, k)));
            if (best < 0 || e < eBest) { best = k; eBest = e; }
        }
        return best;
    }

    // exit_choice.cpp:683-791, sequential in agent order
    void chooseExits() {
        const uint32_t qRead = (choiceStep & 1u) * uint32_t(CHOICE_QUEUE_STRIDE), qWrite = uint32_t(CHOICE_QUEUE_STRIDE) - qRead;
        const bool evalStep = choiceStep > 0u && choiceStep % uint32_t(CHOICE_STEPS) == 0u;
        const size_t perParity = size_t(CHOICE_QUEUE_STRIDE) * CHOICE_AHEAD_BINS;
        uint32_t* aRead = ahead.data() + size_t(choiceStep & 1u) * perParity;
        uint32_t* aWrite = ahead.data() + size_t(1u - (choiceStep & 1u)) * perParity;
        for (int e = 0; e < sc.E; ++e) {
            uint32_t run = 0;
            for (int b = 0; b < CHOICE_AHEAD_BINS; ++b) {
                const size_t i = size_t(e) * CHOICE_AHEAD_BINS + size_t(b);
                const uint32_t h = aRead[i];
                aRead[i] = run;
                run += h;
                aWrite[i] = 0u;
            }
        }
        const float* px = x[readIdx].data();
        const float* py = y[readIdx].data();
        const float* pw = w[readIdx].data();
        if (evalStep) {
            for (int i = 0; i < n; ++i) {   // evaluate (exit_choice.cpp:605-623)
                pendEligible[size_t(i)] = 0;
                if (pw[i] == 0.0f) continue;
                const size_t c = size_t(choiceCell(V2{ px[i], py[i] }));
                if (slotExit(c, 0) == CHOICE_EMPTY) continue;
                const uint32_t current = assign[size_t(i)];
                const int curK = slotOf(c, current);
                if (curK < 0) continue;
                ++evaluations;
                float eBest;
                const int best = bestCandidate(c, aRead, eBest);
                const float eCur = estimatedTime(slotDist(c, curK), current, aheadOf(aRead, current, slotDistQ(c, curK)));
                const float gain = eCur - eBest;
                if (best != curK && gain > std::max(CHOICE_REL_MARGIN * eCur, CHOICE_ABS_MARGIN_S)) {
                    pendEligible[size_t(i)] = 1;
                    pendBest[size_t(i)] = slotExit(c, best);
                    pendBin[size_t(i)] = std::min(uint32_t(gain), uint32_t(CHOICE_GAIN_BINS - 1));
                    ++hist[size_t(current) * CHOICE_GAIN_BINS + pendBin[size_t(i)]];
                    ++eligible;
                }
            }
            for (int e = 0; e < sc.E; ++e) {   // switchThreshold (exit_choice.cpp:625-639)
                uint32_t* h = hist.data() + size_t(e) * CHOICE_GAIN_BINS;
                const uint32_t queued = queue[qRead + uint32_t(e)];
                const uint32_t cap = std::max(CHOICE_MIN_SWITCHES, uint32_t(float(queued) * CHOICE_SWITCH_SHARE));
                uint32_t cum = 0, thr = uint32_t(CHOICE_GAIN_BINS);
                for (int b = CHOICE_GAIN_BINS - 1; b >= 0; --b) {
                    const uint32_t cnt = h[b];
                    if (cnt == 0) continue;
                    if (cum + cnt > cap) {
                        if (cum == 0) thr = uint32_t(b);
                        break;
                    }
                    cum += cnt;
                    thr = uint32_t(b);
                }
                threshold[size_t(e)] = thr;
                std::fill(h, h + CHOICE_GAIN_BINS, 0u);
            }
        }
        std::fill(queue + qWrite, queue + qWrite + CHOICE_QUEUE_STRIDE, 0u);
        for (int i = 0; i < n; ++i) {   // apply (exit_choice.cpp:641-656) and the counts of the next step
            if (pw[i] == 0.0f) continue;
            const size_t c = size_t(choiceCell(V2{ px[i], py[i] }));
            const uint32_t current = assign[size_t(i)];
            uint32_t ex = current;
            const bool hasPath = slotExit(c, 0) != CHOICE_EMPTY;
            const int curK = hasPath ? slotOf(c, current) : -1;
            if (hasPath && curK < 0) {
                float eBest;
                ex = slotExit(c, bestCandidate(c, aRead, eBest));
                forced += current != CHOICE_NONE ? 1u : 0u;
            } else if (evalStep && curK >= 0 && pendEligible[size_t(i)] && pendBin[size_t(i)] >= threshold[current]) {
                ex = pendBest[size_t(i)];
                ++switches;
            }
            assign[size_t(i)] = ex;
            if (ex != CHOICE_NONE) ++queue[qWrite + ex];
            const int k = slotOf(c, ex);
            if (k >= 0) ++aWrite[size_t(ex) * CHOICE_AHEAD_BINS + std::min(slotDistQ(c, k) >> CHOICE_AHEAD_SHIFT, uint32_t(CHOICE_AHEAD_BINS - 1))];
        }
        ++choiceStep;
    }

    // ---------------- static contacts (cpu_crowd_sim.cpp:338-500, 2D) ----------------
    float wallContact(V2 np, float fc, V2 u, float fr, float reach, float recastTol, V2& away) const {   // :338-363
        float br[4] = { 0.0f, fc, reach, fr };
        float t = reach * fc / (fc - fr);
        G3 g{ 0.0f, 0.0f, 0.0f };
        for (int i = 0; i < WALL_ROOT_STEPS; ++i) {
            g = mapExactGrad(np + u * t);
            t = rootStep(br, t, g, u);
        }
        float e = t;   // a 2D gradient has n_z = 0: always a wall
        away = -u;
        V2 nn = safeNormalize(V2{ g.gx, g.gy });
        V2 u2 = -nn;
        float t2 = 0.0f, t3 = t * dot(u, u2);
        G3 g2{ 0.0f, 0.0f, 0.0f };
        for (int i = 0; i < WALL_RECAST_STEPS; ++i) {
            t2 = t3;
            g2 = mapExactGrad(np + u2 * t2);
            t3 = newtonStep(t2, g2, u2);
        }
        bool ok = (g2.gx * u2.x + g2.gy * u2.y) < 0.0f && std::abs(t3 - t2) < recastTol && t3 > 0.0f;
        if (ok && t3 < e) { e = t3; away = nn; }
        return e;
    }

    // clearWalls (:365-421); wall = (normal, clearance - r)
    bool clearWalls(V2& np, V2& nv, float rad, float wall[3]) const {
        bool moved = false;
        wall[0] = wall[1] = wall[2] = 0.0f;
        const float reach = FEELER_REACH_RADII * rad;
        const float recastTol = WALL_RECAST_TOL_RADII * rad;
        for (int pass = 0; pass < WALL_PASSES; ++pass) {
            float fc = map(np);
            float fr1 = 0.0f, fr2 = 0.0f;
            int i1 = -1, i2 = -1;
            for (int i = 0; i < 16; ++i) {
                const V2 f = np + V2{ FEELER_DIRS[i][0] * reach, FEELER_DIRS[i][1] * reach };
                float v = map(f);
                if (v < fr1) { fr2 = fr1; i2 = i1; fr1 = v; i1 = i; }
                else if (v < fr2) { fr2 = v; i2 = i; }
            }
            if (i1 < 0) break;
            V2 hNorm;
            float clearance;
            if (fc > 0.0f) {
                clearance = wallContact(np, fc, V2{ FEELER_DIRS[i1][0], FEELER_DIRS[i1][1] }, fr1, reach, recastTol, hNorm);
                if (i2 >= 0) {
                    V2 h2;
                    float c2 = wallContact(np, fc, V2{ FEELER_DIRS[i2][0], FEELER_DIRS[i2][1] }, fr2, reach, recastTol, h2);
                    if (c2 < clearance) { clearance = c2; hNorm = h2; }
                }
            } else {
                G3 g = mapExactGrad(np);
                hNorm = safeNormalize(V2{ g.gx, g.gy });
                clearance = g.w / std::max(length(V2{ g.gx, g.gy }), CENTRE_GRAD_MIN);
            }
            if (clearance >= rad) {
                if (clearance < 1e9f) { wall[0] = hNorm.x; wall[1] = hNorm.y; wall[2] = clearance - rad; }
                break;
            }
            np = np + hNorm * (rad - clearance);
            moved = true;
            wall[0] = hNorm.x; wall[1] = hNorm.y; wall[2] = 0.0f;
            float vDotN = dot(nv, hNorm);
            if (vDotN < 0.0f) nv = nv - hNorm * vDotN;
        }
        return moved;
    }

    void wallPlaneProbe(V2 np, float rad, float wall[3]) const {   // :423-442
        wall[0] = wall[1] = wall[2] = 0.0f;
        const float reach = PLANE_PROBE_RADII * rad;
        float fc = map(np);
        float fr = 0.0f;
        int i1 = -1;
        for (int i = 0; i < 16; ++i) {
            const V2 f = np + V2{ FEELER_DIRS[i][0] * reach, FEELER_DIRS[i][1] * reach };
            float v = map(f);
            if (v < fr) { fr = v; i1 = i; }
        }
        if (i1 < 0 || fc <= 0.0f) return;
        V2 away;
        float c = wallContact(np, fc, V2{ FEELER_DIRS[i1][0], FEELER_DIRS[i1][1] }, fr, reach, WALL_RECAST_TOL_RADII * rad, away);
        if (c < 1e9f) { wall[0] = away.x; wall[1] = away.y; wall[2] = std::max(c - rad, 0.0f); }
    }

    bool solid(V2 p) const { return mapExactGrad(p).w <= 0.0f; }   // bakedSolid (:444) without the fence

    bool sweepBack(V2 pOld, V2& np, V2& nv) const {   // :446-461 at one floor height
        if (!solid(np)) return false;
        float lo = 0.0f, hi = 1.0f;
        if (!solid(np)) lo = 1.0f;
        for (int i = 0; i < STATIC_SWEEP_BISECTIONS; ++i) {
            float t = 0.5f * (lo + hi);
            if (solid(mix2(pOld, np, t))) hi = t;
            else lo = t;
        }
        V2 d = safeNormalize(np - pOld);
        np = mix2(pOld, np, lo);
        float vd = dot(nv, d);
        if (vd > 0.0f) nv = nv - d * vd;
        return true;
    }

    // resolveStatic (:465-500): on a flat floor the centre is 'grounded' unless it lies in a wall
    bool resolveStatic(V2 pOld, V2& np, V2& nv, float rad, float wall[3]) const {
        bool grounded = !solid(np);
        if (!grounded && sweepBack(pOld, np, nv)) grounded = !solid(np);
        const bool walls = clearWalls(np, nv, rad, wall);
        if (walls && grounded) {
            float w2[3];
            clearWalls(np, nv, rad, w2);
            if (w2[0] != 0.0f || w2[1] != 0.0f) { wall[0] = w2[0]; wall[1] = w2[1]; wall[2] = w2[2]; }
        }
        return grounded;
    }

    // ---------------- the passes ----------------
    int exitOf(V2 p) const {
        for (int e = 0; e < sc.E; ++e) {
            const float* r = &exitRects[size_t(e) * 4];
            if (p.x >= r[0] && p.x <= r[2] && p.y >= r[1] && p.y <= r[3]) return e;
        }
        return -1;
    }
    // seated::inclineFreeSpeed / inclineHeadwayTime (src/venue.cpp:652-670): the smoothstep of |g| / INCLINE_REF_SLOPE
    // blends the level value into the stair value of the slope's sign
    static float inclineBlend(float g) {
        const float u = std::clamp(std::abs(g) * INCLINE_INV_REF_SLOPE, 0.0f, 1.0f);
        const float uu = u * u, w = 3.0f - 2.0f * u;
        return uu * w;
    }
    static float inclineFreeSpeed(float g) {
        const float t = inclineBlend(g);
        const float dv = (g > 0.0f ? V_STAIR_UP : V_STAIR_DOWN) - V_MAX;
        const float k = dv * t;
        return V_MAX + k;
    }
    static float inclineHeadwayTime(float g) {
        const float t = inclineBlend(g);
        const float dt = (g > 0.0f ? T_HEADWAY_STAIR_UP : T_HEADWAY_STAIR_DOWN) - T_HEADWAY;
        const float k = dt * t;
        return T_HEADWAY + k;
    }
    float stairFactor(V2 p) const {
        const float gx = (p.x - sc.navOX) * sc.navInvCell, gy = (p.y - sc.navOY) * sc.navInvCell;
        const int cx = std::clamp(int(std::floor(gx)), 0, sc.navRX - 1), cy = std::clamp(int(std::floor(gy)), 0, sc.navRY - 1);
        return stair[size_t(cx) + size_t(cy) * size_t(sc.navRX)];
    }

    // PHYSICS_PASS 1 (cpu_crowd_sim.cpp:655-777)
    void predict(uint32_t slot) {
        const uint32_t self = id[slot];
        const uint32_t cid = key[slot];
        if (cid == INACTIVE_CELL) return;
        const float* px = x[readIdx].data();
        const float* py = y[readIdx].data();
        const V2 p{ px[self], py[self] };
        const V2 v{ vx[readIdx][self], vy[readIdx][self] };
        const float ri = r[self];
        if (exitOf(p) >= 0 || p.x < sc.xmin || p.x > sc.xmax || p.y < sc.ymin || p.y > sc.ymax) {
            slotA[slot] = SlotState{ PARK_COORD, PARK_COORD, 0.0f, 0.0f, 0.0f, 0.0f };
            return;
        }
        const V2 eGoal{ egX[self], egY[self] };
        const int cx = int(cid % uint32_t(sc.hashNX));
        const int cy = int(cid / uint32_t(sc.hashNX));
        V2 sep{ 0.0f, 0.0f };
        float sAhead = 1e9f;   // the smallest gap (centre distance - contact distance) to an agent ahead
        const int R = sc.percRing;
        for (int yy = -R; yy <= R; yy++) {
            for (int xx = -R; xx <= R; xx++) {
                const int nx = cx + xx, ny = cy + yy;
                if (nx < 0 || nx >= sc.hashNX || ny < 0 || ny >= sc.hashNY) continue;
                const size_t c = size_t(nx) + size_t(ny) * size_t(sc.hashNX);
                const uint32_t st = cellStart[c], cnt = cellCount[c];
                for (uint32_t k = st; k < st + cnt; k++) {
                    if (k == slot) continue;
                    const uint32_t other = id[k];
                    const V2 op{ px[other], py[other] };
                    const V2 offset = p - op;
                    const float dxy2 = dot(offset, offset);
                    if (dxy2 < R_PERC2) {
                        const float D = ri + r[other];
                        sep = sep + separationForce(offset, self, other, D + SEP_MARGIN);
                        const float ahead = -dot(offset, eGoal);
                        const float lateral = std::abs(offset.x * eGoal.y - offset.y * eGoal.x);
                        const float dxy = std::sqrt(dxy2);
                        const float gap = dxy - D;
                        if (ahead > 0.0f && lateral < D && gap < sAhead) {
                            const V2 eOther{ egX[other], egY[other] };
                            const float aheadOther = dot(offset, eOther);
                            const bool mutual = aheadOther > 0.0f && std::abs(offset.x * eOther.y - offset.y * eOther.x) < D;
                            if (!mutual || ahead > aheadOther || (ahead == aheadOther && self > other)) sAhead = gap;
                        }
                    }
                }
            }
        }
        float vFree, vDes;
        if (sc.venue) {   // the default venue model's incline speed and stair spacing (cpu_crowd_sim.cpp:727-733)
            const float slope = oracle_venue_slope(sc.venue, p.x, p.y, eGoal.x, eGoal.y);
            vFree = inclineFreeSpeed(slope);
            vDes = (slope != 0.0f ? std::clamp(sAhead / inclineHeadwayTime(slope), 0.0f, vFree)
                                  : std::clamp(sAhead * HEADWAY_RATE, 0.0f, vFree)) * length(eGoal);
        } else {
            vFree = stairFactor(p) * v0[self];
            vDes = std::clamp(sAhead * HEADWAY_RATE, 0.0f, vFree) * length(eGoal);
        }
        const V2 eDir = safeNormalize(eGoal + sep);

        V2 aObs{ 0.0f, 0.0f };   // obstacle avoidance (:739-748)
        const float dTorso = map(p);
        if (dTorso < OBSTACLE_RANGE) {
            const V2 norm = safeNormalize(calcGradient(p));
            if (length(norm) > OBSTACLE_MIN_NORMAL) aObs = safeNormalize(norm) * (OBSTACLE_RANGE - dTorso) * W_OBS * par.obstacleGain;
        }
        const V2 vEq = eDir * vDes + aObs * par.tau;   // kinematics (:752-766)
        V2 nv = vEq + (v - vEq) * par.relaxStep;
        const float along = dot(nv, eGoal);
        if (along > vDes) nv = nv - eGoal * (along - vDes);
        if (length(nv) > v0[self]) nv = safeNormalize(nv) * v0[self];
        V2 np = p + nv * DT;

        float wall[3];
        const bool grounded = resolveStatic(p, np, nv, ri, wall);
        if (wall[0] == 0.0f && wall[1] == 0.0f) wallPlaneProbe(np, ri, wall);
        slotA[slot] = SlotState{ np.x, np.y, 1.0f, nv.x, nv.y, 0.0f };
        const V2 wn{ wall[0], wall[1] };
        float* pl = &planes[size_t(slot) * 4];
        pl[0] = wn.x; pl[1] = wn.y; pl[2] = dot(np, wn) - wall[2];
        pl[3] = (nonzero(wn) ? PLANE_WALL : 0.0f) + (grounded ? PLANE_GROUNDED : 0.0f);
    }

    // projectSlot (cpu_crowd_sim.cpp:781-822): one Jacobi iteration, reads 'in'
    SlotState projectSlot(const std::vector<SlotState>& in, uint32_t slot, uint32_t self, uint32_t cid) const {
        SlotState b = in[slot];
        if (b.w == 0.0f) return b;
        const int cx = int(cid % uint32_t(sc.hashNX));
        const int cy = int(cid / uint32_t(sc.hashNX));
        const float ri = r[self];
        V2 corr{ 0.0f, 0.0f };
        int nContacts = 0;
        const int R = sc.projRing;
        for (int yy = -R; yy <= R; yy++) {
            for (int xx = -R; xx <= R; xx++) {
                const int nx = cx + xx, ny = cy + yy;
                if (nx < 0 || nx >= sc.hashNX || ny < 0 || ny >= sc.hashNY) continue;
                const size_t c = size_t(nx) + size_t(ny) * size_t(sc.hashNX);
                const uint32_t st = cellStart[c], cnt = cellCount[c];
                for (uint32_t k = st; k < st + cnt; k++) {
                    if (k == slot) continue;
                    const SlotState& o = in[k];
                    const V2 offset = V2{ b.x, b.y } - V2{ o.x, o.y };
                    const uint32_t other = id[k];
                    const float D = ri + r[other];
                    if (o.w != 0.0f && dot(offset, offset) < contact2(D)) {
                        corr = corr + contactCorrection(offset, self, other, D);
                        ++nContacts;
                    }
                }
            }
        }
        corr = corr * (PROJ_RELAX / float(std::max(nContacts, 1)));
        const float maxShift = PROJ_MAX_SHIFT_RADII * ri;
        float len = length(corr);
        if (len > maxShift) corr = corr * (maxShift / len);
        V2 q = V2{ b.x, b.y } + corr;
        const float* pl = &planes[size_t(slot) * 4];
        if (std::fmod(pl[3], 2.0f) == PLANE_WALL) {
            float s = pl[2] - dot(q, V2{ pl[0], pl[1] });
            if (s > 0.0f) q = q + V2{ pl[0], pl[1] } * s;
        }
        const V2 d = q - V2{ b.x, b.y };
        b.x = q.x; b.y = q.y;
        b.acc += length(d);
        return b;
    }

    // PHYSICS_PASS 3 (cpu_crowd_sim.cpp:835-866)
    void finalPass(const std::vector<SlotState>& in, uint32_t slot, bool projectHere) {
        const uint32_t self = id[slot];
        const uint32_t cid = key[slot];
        const int wi = writeIdx, ri0 = readIdx;
        if (cid == INACTIVE_CELL) {
            x[wi][self] = x[ri0][self]; y[wi][self] = y[ri0][self]; w[wi][self] = w[ri0][self];
            vx[wi][self] = vx[ri0][self]; vy[wi][self] = vy[ri0][self]; spd[wi][self] = spd[ri0][self];
            return;
        }
        const SlotState b = projectHere ? projectSlot(in, slot, self, cid) : in[slot];
        if (b.w == 0.0f) {
            x[wi][self] = b.x; y[wi][self] = b.y; w[wi][self] = 0.0f;
            vx[wi][self] = b.vx; vy[wi][self] = b.vy; spd[wi][self] = b.acc;
            return;
        }
        V2 np{ b.x, b.y }, nv{ b.vx, b.vy };
        const V2 pOld{ x[ri0][self], y[ri0][self] };
        if (b.acc > 0.0f) {
            float wall[3];
            resolveStatic(pOld, np, nv, r[self], wall);
        }
        x[wi][self] = np.x; y[wi][self] = np.y; w[wi][self] = 1.0f;
        vx[wi][self] = nv.x; vy[wi][self] = nv.y; spd[wi][self] = length(nv);
    }

    // ================ the seated 3D step (cpu_crowd_sim.cpp, the VENUE_ANALYTIC branches) ================
    static void setXY(V3& v, V2 a) { v.x = a.x; v.y = a.y; }
    float seatedMap(V3 p) const { float g[4]; vo::sdfGrad(sc.venue, p.x, p.y, p.z, g); return g[3]; }   // map: seated::sdf
    // calcGradient (cpu_crowd_sim.cpp:297-303): the tetrahedral stencil k = (1, -1) in 3D
    V3 seatedGradient(V3 p) const {
        const float H = OBSTACLE_GRAD_H;
        const float m1 = seatedMap(V3{ p.x + H, p.y + -H, p.z + -H }), m2 = seatedMap(V3{ p.x + -H, p.y + -H, p.z + H });
        const float m3 = seatedMap(V3{ p.x + -H, p.y + H, p.z + -H }), m4 = seatedMap(V3{ p.x + H, p.y + H, p.z + H });
        // kxyy m1 + kyyx m2 + kyxy m3 + kxxx m4, then * OBSTACLE_GRAD_SCALE
        const V3 g{ ((1.0f * m1 + -1.0f * m2) + -1.0f * m3) + 1.0f * m4, ((-1.0f * m1 + -1.0f * m2) + 1.0f * m3) + 1.0f * m4,
                    ((-1.0f * m1 + 1.0f * m2) + -1.0f * m3) + 1.0f * m4 };
        return V3{ g.x * OBSTACLE_GRAD_SCALE, g.y * OBSTACLE_GRAD_SCALE, g.z * OBSTACLE_GRAD_SCALE };
    }
    // floorUnderReach's venue branch (cpu_crowd_sim.cpp:265-271)
    bool floorUnderReach(V3 c, float zFrom, float down, float& zFloor) const {
        const float h = vo::floorHeight(sc.venue, c.x, c.y);
        zFloor = h;
        if (c.z <= h) return h <= std::max(zFrom, c.z) + BODY_RADIUS;
        return h >= c.z - down;
    }
    bool restOnFloor(V3& np, V3& nv, float zFrom) const {   // cpu_crowd_sim.cpp:318-326
        float zF;
        if (floorUnderReach(np, zFrom, BODY_RADIUS, zF) && np.z < zF + BODY_RADIUS) { np.z = zF + BODY_RADIUS; nv.z = 0.0f; return true; }
        return false;
    }
    bool settleOnFloor(V3& np, V3& nv, float zFrom) const {   // cpu_crowd_sim.cpp:328-336
        float zF;
        if (floorUnderReach(np, zFrom, FLOOR_SETTLE_REACH, zF)) { np.z = zF + BODY_RADIUS; nv.z = 0.0f; return true; }
        return false;
    }
    bool bakedSolid(V3 p) const { return sdFence(p) >= 0.0f && seatedMap(p) <= 0.0f; }   // cpu_crowd_sim.cpp:445
    bool sweepBack(V3 pOld, V3& np, V3& nv) const {   // cpu_crowd_sim.cpp:447-463
        if (!bakedSolid(np)) return false;
        float lo = 0.0f, hi = 1.0f;
        if (!bakedSolid(V3{ np.x, np.y, pOld.z })) lo = 1.0f;
        for (int i = 0; i < STATIC_SWEEP_BISECTIONS; ++i) {
            const float t = 0.5f * (lo + hi);
            if (bakedSolid(v3(mix2(xy(pOld), xy(np), t), pOld.z))) hi = t;
            else lo = t;
        }
        const V2 d = safeNormalize(xy(np) - xy(pOld));
        np = v3(mix2(xy(pOld), xy(np), lo), pOld.z);
        const float vd = dot(xy(nv), d);
        if (vd > 0.0f) setXY(nv, xy(nv) - d * vd);
        nv.z = 0.0f;
        return true;
    }
    // clearWalls' venue branch (cpu_crowd_sim.cpp:367-381): the exact wall at the body's top
    bool clearWalls(V3& np, V3& nv, V3& wall) const {
        bool moved = false;
        wall = V3{ 0.0f, 0.0f, 0.0f };
        const float above = vo::wallAbove(sc.venue), clear = vo::wallClear(sc.venue);
        for (int pass = 0; pass < WALL_PASSES; ++pass) {
            float w[3];
            vo::wall(sc.venue, np.x, np.y, np.z + above, FEELER_REACH, w);
            const V2 n{ w[0], w[1] };
            if (w[2] >= clear) {
                if (w[2] < WALL_NONE) wall = v3(n, w[2] - clear);
                break;
            }
            setXY(np, xy(np) + n * (clear - w[2]));
            moved = true;
            wall = v3(n, 0.0f);
            const float vDotN = dot(xy(nv), n);
            if (vDotN < 0.0f) setXY(nv, xy(nv) - n * vDotN);
        }
        return moved;
    }
    V3 wallPlaneProbe(V3 np) const {   // cpu_crowd_sim.cpp:420-424
        float w[3];
        vo::wall(sc.venue, np.x, np.y, np.z + vo::wallAbove(sc.venue), PLANE_PROBE_REACH, w);
        return w[2] >= 0.0f && w[2] < WALL_NONE ? V3{ w[0], w[1], std::max(w[2] - vo::wallClear(sc.venue), 0.0f) } : V3{ 0.0f, 0.0f, 0.0f };
    }
    // resolveStatic (cpu_crowd_sim.cpp:467-500)
    bool resolveStatic3(V3 pOld, V3& np, V3& nv, bool settle, V3& wall, uint8_t& contact) const {
        if (fenceCCD(pOld, np, nv)) contact |= C_FENCE_CCD;
        const float dFence = sdFence(np);
        if (dFence < 0.0f) {
            const V2 n = sdFenceGradient(np);
            setXY(np, xy(np) - n * dFence);
            const float vDotN = dot(xy(nv), n);
            if (vDotN < 0.0f) setXY(nv, xy(nv) - n * vDotN);
            contact |= C_FENCE;
        }
        bool grounded = settle ? settleOnFloor(np, nv, pOld.z) : restOnFloor(np, nv, pOld.z);
        if (!grounded && sweepBack(pOld, np, nv)) {
            contact |= C_SWEEP;
            grounded = restOnFloor(np, nv, pOld.z);
        }
        if (grounded) contact |= C_FLOOR;
        const bool walls = clearWalls(np, nv, wall);
        if (walls) {
            contact |= C_WALL;
            if (grounded) {
                if (settleOnFloor(np, nv, pOld.z)) contact |= C_FLOOR;
                V3 w2;
                const bool again = clearWalls(np, nv, w2);
                if (w2.x != 0.0f || w2.y != 0.0f) wall = w2;
                if (again && restOnFloor(np, nv, pOld.z)) contact |= C_FLOOR;
            } else if (restOnFloor(np, nv, pOld.z)) {
                contact |= C_FLOOR;
            }
        }
        return grounded;
    }
    float scan(V2 a, V2 b, float f0, V2& n) const {
        float nn[2];
        const float t = vo::stepUpScan(sc.venue, a.x, a.y, b.x, b.y, f0, nn);
        n = V2{ nn[0], nn[1] };
        return t;
    }
    // venueBlockedWall (cpu_crowd_sim.cpp:591-605)
    void venueBlockedWall(float f0, V3 pOld, V3& np, V3& nv) const {
        float w[3];
        vo::wall(sc.venue, np.x, np.y, np.z + vo::wallAbove(sc.venue), FEELER_REACH, w);
        const float clear = vo::wallClear(sc.venue);
        if (w[2] >= clear) return;
        const V2 n{ w[0], w[1] };
        const V2 pushed = xy(np) + n * (clear - w[2]);
        V2 nn;
        if (scan(xy(np), pushed, f0, nn) < 1.0f) return;
        const float vDotN = dot(xy(nv), n);
        if (vDotN < 0.0f) setXY(nv, xy(nv) - n * vDotN);
        float zF;
        if (floorUnderReach(V3{ pushed.x, pushed.y, pOld.z }, pOld.z, FLOOR_SETTLE_REACH, zF)) {
            np = v3(pushed, zF + BODY_RADIUS);
            nv.z = 0.0f;
        } else {
            setXY(np, pushed);
        }
    }
    // venueEdgeMargin (cpu_crowd_sim.cpp:610-636)
    void venueEdgeMargin(V3 pOld, float f0, V3& np, V3& nv) const {
        const float margin = vo::faceMargin(sc.venue);
        float feet = np.z - BODY_RADIUS;
        const float fl = vo::floorHeight(sc.venue, np.x, np.y);
        float w[3];
        if (feet > fl + margin) {
            vo::wall(sc.venue, np.x, np.y, fl + STEP_UP_TOL, margin, w);
            if (w[2] < margin) {
                const V2 into = xy(np) - V2{ w[0], w[1] } * (w[2] + margin);
                const float top = vo::floorHeight(sc.venue, into.x, into.y);
                if (std::abs(top - feet) <= margin && top <= f0 + STEP_UP_TOL) { setXY(np, into); return; }
            }
            if (np.z <= pOld.z) {
                vo::wall(sc.venue, np.x, np.y, feet + STEP_UP_TOL, margin, w);
                if (w[2] < margin) setXY(np, xy(np) + V2{ w[0], w[1] } * (margin - w[2]));
                return;
            }
            float zF;
            if (!floorUnderReach(V3{ np.x, np.y, pOld.z }, pOld.z, FLOOR_SETTLE_REACH, zF)) { np.z = pOld.z; return; }
            np.z = zF + BODY_RADIUS;
            nv.z = 0.0f;
            feet = zF;
        }
        vo::wall(sc.venue, np.x, np.y, feet + STEP_UP_TOL, margin, w);
        if (w[2] >= margin) return;
        setXY(np, xy(np) + V2{ w[0], w[1] } * (margin - w[2]));
        np.z = vo::floorHeight(sc.venue, np.x, np.y) + BODY_RADIUS;
    }
    // limitStepUp with the venue's scan (cpu_crowd_sim.cpp:555-587)
    bool limitStepUp(V3 pOld, V3& np, V3& nv) const {
        float f0;
        if (!floorUnderReach(pOld, pOld.z, STEP_UP_FLOOR_REACH, f0)) f0 = pOld.z - STEP_UP_FLOOR_REACH;
        V2 n;
        const float t = scan(xy(pOld), xy(np), f0, n);
        if (t >= 1.0f) {
            venueEdgeMargin(pOld, f0, np, nv);
            return false;
        }
        V2 q = mix2(xy(pOld), xy(np), t);
        const V2 rest = xy(np) - q;
        const V2 tangent = rest - n * dot(rest, n);
        const float lt = length(tangent);
        if (lt > 0.0f) {
            for (int k = 0; k < STEP_UP_SLIDE_TRIES; ++k) {
                const V2 slide = (q + tangent * STEP_UP_SLIDE_DIRS[k][0]) + n * (lt * STEP_UP_SLIDE_DIRS[k][1]);
                V2 n2;
                if (sdFence(V3{ slide.x, slide.y, 0.0f }) >= 0.0f && scan(q, slide, f0, n2) >= 1.0f) { q = slide; break; }
            }
        }
        const float vn = dot(xy(nv), n);
        if (vn < 0.0f) setXY(nv, xy(nv) - n * vn);
        float zF;
        if (floorUnderReach(V3{ q.x, q.y, pOld.z }, pOld.z, FLOOR_SETTLE_REACH, zF)) {
            np = v3(q, zF + BODY_RADIUS);
            nv.z = 0.0f;
        } else {
            np = v3(q, std::min(np.z, pOld.z));
        }
        venueBlockedWall(f0, pOld, np, nv);
        venueEdgeMargin(pOld, f0, np, nv);
        return true;
    }
    // separationForce / contactCorrection over 3D offsets (cpu_crowd_sim.cpp:131-148): the reference's global radius
    static V2 separation3(V3 offset, uint32_t self, uint32_t other) {
        const float d2 = dot3(offset, offset);
        if (d2 >= R_SEP2) return V2{ 0.0f, 0.0f };
        const float w = 1.0f - std::sqrt(d2) / R_SEP;
        const float magnitude = K_SEP * w * w;
        const float dxy2 = dot(xy(offset), xy(offset));
        if (dxy2 <= SEP_STACKED_XY2) return magnitude * pairDirection(self, other);
        return magnitude * xy(offset) * (1.0f / std::sqrt(dxy2));
    }
    static V2 contact3(V3 offset, uint32_t self, uint32_t other) {
        const float dxy2 = dot(xy(offset), xy(offset));
        if (std::abs(offset.z) >= BODY_DIAM || dxy2 >= BODY_DIAM2) return V2{ 0.0f, 0.0f };
        const float dxy = std::sqrt(dxy2);
        const float shift = 0.5f * (BODY_DIAM - dxy);
        if (dxy2 <= SEP_STACKED_XY2) return shift * pairDirection(self, other);
        return shift * (xy(offset) / dxy);
    }
    // PHYSICS_PASS 1 on the venue (cpu_crowd_sim.cpp:655-777)
    void predict3(uint32_t slot) {
        const uint32_t self = id[slot];
        const uint32_t cid = key[slot];
        uint8_t& contact = contacts[self];
        contact = 0;
        if (cid == INACTIVE_CELL) return;
        const int ri = readIdx;
        const V3 p{ x[ri][self], y[ri][self], z[ri][self] };
        const V3 v{ vx[ri][self], vy[ri][self], vz[ri][self] };
        const float m = std::max(std::abs(p.x), std::abs(p.y));
        if ((m > sc.exitDepth && p.z < sc.exitMaxZ) || m > sc.domainMax) {
            slotA[slot] = SlotState{ PARK_COORD, PARK_COORD, 0.0f, 0.0f, 0.0f, 0.0f, PARK_COORD, 0.0f };
            contact = C_EXIT;
            return;
        }
        const V2 eGoal{ egX[self], egY[self] };
        const int cx = int(cid % uint32_t(sc.hashNX)), cy = int(cid / uint32_t(sc.hashNX));
        V2 sep{ 0.0f, 0.0f };
        float sAhead = 1e9f;
        for (int yy = -sc.percRing; yy <= sc.percRing; yy++)
            for (int xx = -sc.percRing; xx <= sc.percRing; xx++) {
                const int nx = cx + xx, ny = cy + yy;
                if (nx < 0 || nx >= sc.hashNX || ny < 0 || ny >= sc.hashNY) continue;
                const size_t c = size_t(nx) + size_t(ny) * size_t(sc.hashNX);
                for (uint32_t k = cellStart[c]; k < cellStart[c] + cellCount[c]; k++) {
                    if (k == slot) continue;
                    const uint32_t other = id[k];
                    const V3 op{ x[ri][other], y[ri][other], z[ri][other] };
                    const V3 offset{ p.x - op.x, p.y - op.y, p.z - op.z };
                    const float dxy2 = dot(xy(offset), xy(offset));
                    if (dxy2 < R_PERC2) {
                        sep = sep + separation3(offset, self, other);
                        const float ahead = -dot(xy(offset), eGoal);
                        const float lateral = std::abs(offset.x * eGoal.y - offset.y * eGoal.x);
                        const float dxy = std::sqrt(dxy2);
                        if (ahead > 0.0f && lateral < BODY_DIAM && dxy < sAhead) {
                            const V2 eOther{ egX[other], egY[other] };
                            const float aheadOther = dot(xy(offset), eOther);
                            const bool mutual = aheadOther > 0.0f && std::abs(offset.x * eOther.y - offset.y * eOther.x) < BODY_DIAM;
                            if (!mutual || ahead > aheadOther || (ahead == aheadOther && self > other)) sAhead = dxy;
                        }
                    }
                }
            }
        const float slope = vo::floorSlope(sc.venue, p.x, p.y, eGoal.x, eGoal.y);
        const float vFree = inclineFreeSpeed(slope);
        const float vDes = (slope != 0.0f ? std::clamp((sAhead - BODY_DIAM) / inclineHeadwayTime(slope), 0.0f, vFree)
                                          : std::clamp((sAhead - BODY_DIAM) * HEADWAY_RATE, 0.0f, vFree)) * length(eGoal);
        const V2 eDir = safeNormalize(eGoal + sep);
        // obstacle avoidance at torso height: the venue or the fence, the nearer (sceneDist / sceneNormal, :307-316)
        V2 aObs{ 0.0f, 0.0f };
        const V3 torso{ p.x, p.y, p.z + TORSO_HEIGHT };
        const float dStadium = seatedMap(torso);
        const float dFence = sdFence(torso) + BODY_RADIUS;
        const bool onFence = dFence < dStadium;
        const float dTorso = onFence ? dFence : dStadium;
        if (dTorso < OBSTACLE_RANGE) {
            V2 norm;
            if (onFence) norm = sdFenceGradient(torso);
            else {
                const V3 g = seatedGradient(torso);
                const float l = length3(g);
                norm = l > SAFE_NORMALIZE_EPS ? V2{ g.x / l, g.y / l } : V2{ 0.0f, 0.0f };
            }
            if (length(norm) > OBSTACLE_MIN_NORMAL) aObs = safeNormalize(norm) * (OBSTACLE_RANGE - dTorso) * W_OBS * OBSTACLE_GAIN_REF;
        }
        const V2 vEq = eDir * vDes + aObs * TAU_RELAX_REF;
        V3 nv = v;
        setXY(nv, vEq + (xy(nv) - vEq) * RELAX_STEP_REF);
        const float along = dot(xy(nv), eGoal);
        if (along > vDes) setXY(nv, xy(nv) - eGoal * (along - vDes));
        nv.z -= GRAVITY_DT;
        nv.z = std::min(nv.z, 0.0f);
        if (length(xy(nv)) > V_MAX) setXY(nv, safeNormalize(xy(nv)) * V_MAX);
        V3 np{ p.x + nv.x * DT, p.y + nv.y * DT, p.z + nv.z * DT };
        V3 wall;
        const bool grounded = resolveStatic3(p, np, nv, false, wall, contact);
        if (wall.x == 0.0f && wall.y == 0.0f) wall = wallPlaneProbe(np);
        slotA[slot] = SlotState{ np.x, np.y, 1.0f, nv.x, nv.y, 0.0f, np.z, nv.z };
        const V2 wn{ wall.x, wall.y };
        float* pl = &planes[size_t(slot) * 4];
        pl[0] = wn.x; pl[1] = wn.y; pl[2] = dot(xy(np), wn) - wall.z;
        pl[3] = (nonzero(wn) ? PLANE_WALL : 0.0f) + (grounded ? PLANE_GROUNDED : 0.0f);
    }
    // PHYSICS_PASS 2 on the venue (cpu_crowd_sim.cpp:780-820)
    SlotState projectSlot3(const std::vector<SlotState>& in, uint32_t slot, uint32_t self, uint32_t cid) {
        SlotState b = in[slot];
        if (b.w == 0.0f) return b;
        const int cx = int(cid % uint32_t(sc.hashNX)), cy = int(cid / uint32_t(sc.hashNX));
        V2 corr{ 0.0f, 0.0f };
        int nContacts = 0;
        for (int yy = -sc.projRing; yy <= sc.projRing; yy++)
            for (int xx = -sc.projRing; xx <= sc.projRing; xx++) {
                const int nx = cx + xx, ny = cy + yy;
                if (nx < 0 || nx >= sc.hashNX || ny < 0 || ny >= sc.hashNY) continue;
                const size_t c = size_t(nx) + size_t(ny) * size_t(sc.hashNX);
                for (uint32_t k = cellStart[c]; k < cellStart[c] + cellCount[c]; k++) {
                    if (k == slot) continue;
                    const SlotState& o = in[k];
                    const V3 offset{ b.x - o.x, b.y - o.y, b.z - o.z };
                    if (o.w != 0.0f && std::abs(offset.z) < BODY_DIAM && dot(xy(offset), xy(offset)) < BODY_DIAM2) {
                        corr = corr + contact3(offset, self, id[k]);
                        ++nContacts;
                    }
                }
            }
        corr = corr * (PROJ_RELAX / float(std::max(nContacts, 1)));
        const float len = length(corr);
        if (len > PROJ_MAX_SHIFT) corr = corr * (PROJ_MAX_SHIFT / len);
        V3 q{ b.x, b.y, b.z };
        setXY(q, xy(q) + corr);
        const float dFence = sdFence(q);
        if (dFence < 0.0f) setXY(q, xy(q) - sdFenceGradient(q) * dFence);
        const float* pl = &planes[size_t(slot) * 4];
        if (std::fmod(pl[3], 2.0f) == PLANE_WALL) {
            const float sgap = pl[2] - dot(xy(q), V2{ pl[0], pl[1] });
            if (sgap > 0.0f) setXY(q, xy(q) + V2{ pl[0], pl[1] } * sgap);
        }
        const V2 d{ q.x - b.x, q.y - b.y };
        if (nonzero(d)) contacts[self] |= C_AGENT;
        b.x = q.x; b.y = q.y;
        b.acc += length(d);
        return b;
    }
    // PHYSICS_PASS 3 on the venue (cpu_crowd_sim.cpp:835-866)
    void finalPass3(const std::vector<SlotState>& in, uint32_t slot, bool projectHere) {
        const uint32_t self = id[slot];
        const uint32_t cid = key[slot];
        const int wi = writeIdx, ri0 = readIdx;
        if (cid == INACTIVE_CELL) {
            x[wi][self] = x[ri0][self]; y[wi][self] = y[ri0][self]; z[wi][self] = z[ri0][self]; w[wi][self] = w[ri0][self];
            vx[wi][self] = vx[ri0][self]; vy[wi][self] = vy[ri0][self]; vz[wi][self] = vz[ri0][self]; spd[wi][self] = spd[ri0][self];
            return;
        }
        const SlotState b = projectHere ? projectSlot3(in, slot, self, cid) : in[slot];
        if (b.w == 0.0f) {
            x[wi][self] = b.x; y[wi][self] = b.y; z[wi][self] = b.z; w[wi][self] = 0.0f;
            vx[wi][self] = b.vx; vy[wi][self] = b.vy; vz[wi][self] = b.vz; spd[wi][self] = b.acc;
            return;
        }
        V3 np{ b.x, b.y, b.z }, nv{ b.vx, b.vy, b.vz };
        const V3 pOld{ x[ri0][self], y[ri0][self], z[ri0][self] };
        if (b.acc > 0.0f) {
            V3 wall;
            resolveStatic3(pOld, np, nv, planes[size_t(slot) * 4 + 3] >= PLANE_GROUNDED, wall, contacts[self]);
        }
        if (limitStepUp(pOld, np, nv)) contacts[self] |= C_STEP_UP;
        x[wi][self] = np.x; y[wi][self] = np.y; z[wi][self] = np.z; w[wi][self] = 1.0f;
        vx[wi][self] = nv.x; vy[wi][self] = nv.y; vz[wi][self] = nv.z; spd[wi][self] = length3(nv);
    }
};

namespace {
template <class Body>
void parallelFor(int threads, int n, bool dynamic, const Body& body) {
    if (dynamic) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
        for (int i = 0; i < n; ++i) body(i);
    } else {
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int i = 0; i < n; ++i) body(i);
    }
}
double msBetween(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}
} // namespace

extern "C" {

void oracle_goal(const oracle_ctx* c, int n, const float* x, const float* y, const float* r, const uint32_t* exit, int wall_rule,
                 float* ex, float* ey, uint8_t* applied) {
    for (int i = 0; i < n; ++i) {
        bool a = false;
        const V2 p{ x[i], y[i] };
        const V2 e = wall_rule ? c->choiceGoal(p, r[i], exit[i], &a) : c->choiceDirection(p, exit[i]);
        ex[i] = e.x;
        ey[i] = e.y;
        if (applied) applied[i] = a ? 1 : 0;
    }
}

void oracle_sincos(int n, const float* x, float* s, float* c) {
    for (int i = 0; i < n; ++i) sincosCephes(x[i], &s[i], &c[i]);
}

void oracle_choice_cell(const oracle_ctx* c, int n, const float* x, const float* y, int32_t* cell) {
    for (int i = 0; i < n; ++i) cell[i] = c->choiceCell(V2{ x[i], y[i] });
}

void oracle_exit_at(const oracle_ctx* c, int n, const float* x, const float* y, int32_t* exit) {
    for (int i = 0; i < n; ++i) exit[i] = c->exitOf(V2{ x[i], y[i] });
}

void oracle_wall(const oracle_ctx* c, int n, const float* x, const float* y, float reach, float* wx, float* wy, float* wd) {
    for (int i = 0; i < n; ++i) {
        const G3 w = c->wallQuery(V2{ x[i], y[i] }, reach);
        wx[i] = w.gx;
        wy[i] = w.gy;
        wd[i] = w.w;
    }
}

oracle_ctx* oracle_create(const oracle_scene* s, const oracle_params* p, int n) {
    if (!s || !p || n <= 0) return nullptr;
    oracle_ctx* c = new oracle_ctx();
    c->sc = *s;
    c->par = *p;
    c->n = n;
    c->threads = p->threads > 0 ? p->threads : omp_get_max_threads();
    c->sdf.assign(s->sdf, s->sdf + size_t(s->sdfNX) * s->sdfNY);
    const size_t nc = size_t(s->navRX) * s->navRY;
    c->navInfo.assign(s->navInfo, s->navInfo + nc);
    c->navSlot.assign(s->navSlot, s->navSlot + nc * size_t(s->K));
    c->navDist.assign(s->navDist, s->navDist + nc * size_t(s->K));
    c->stair.assign(s->stair, s->stair + nc);
    c->invCapS.assign(s->invCapS, s->invCapS + s->E);
    c->exitRects.assign(s->exitRects, s->exitRects + size_t(s->E) * 4);
    if (s->nWalls > 0 && s->wallRects) c->walls.assign(s->wallRects, s->wallRects + size_t(s->nWalls) * 4);
    c->sc.wallRects = nullptr;
    c->sc.sdf = nullptr; c->sc.navInfo = nullptr; c->sc.navSlot = nullptr; c->sc.navDist = nullptr;
    c->sc.stair = nullptr; c->sc.invCapS = nullptr; c->sc.exitRects = nullptr;
    for (int h = 0; h < 2; ++h) {
        c->x[h].assign(size_t(n), 0.0f); c->y[h].assign(size_t(n), 0.0f); c->vx[h].assign(size_t(n), 0.0f);
        c->vy[h].assign(size_t(n), 0.0f); c->spd[h].assign(size_t(n), 0.0f); c->w[h].assign(size_t(n), 0.0f);
        c->z[h].assign(size_t(n), 0.0f); c->vz[h].assign(size_t(n), 0.0f);
    }
    c->contacts.assign(size_t(n), 0);
    c->r.assign(size_t(n), 0.0f); c->v0.assign(size_t(n), 0.0f);
    const size_t cells = size_t(s->hashNX) * s->hashNY;
    c->key.assign(size_t(n), 0u); c->id.assign(size_t(n), 0u); c->keyTmp.assign(size_t(n), 0u); c->idTmp.assign(size_t(n), 0u);
    c->bucketStart.assign(cells + 2, 0u);
    c->cellStart.assign(cells, 0u); c->cellCount.assign(cells, 0u);
    c->slotA.assign(size_t(n), SlotState{}); c->slotB.assign(size_t(n), SlotState{});
    c->planes.assign(size_t(n) * 4, 0.0f);
    c->egX.assign(size_t(n), 0.0f); c->egY.assign(size_t(n), 0.0f);
    return c;
}

void oracle_destroy(oracle_ctx* c) { delete c; }

void oracle_set_state(oracle_ctx* c, const float* x, const float* y, const float* vx, const float* vy, const float* w,
                      const float* r, const float* v0) {
    const size_t n = size_t(c->n);
    for (int h = 0; h < 2; ++h) {
        c->x[h].assign(x, x + n); c->y[h].assign(y, y + n); c->vx[h].assign(vx, vx + n); c->vy[h].assign(vy, vy + n);
        c->w[h].assign(w, w + n);
        for (size_t i = 0; i < n; ++i) c->spd[h][i] = std::sqrt(vx[i] * vx[i] + vy[i] * vy[i]);
    }
    c->r.assign(r, r + n); c->v0.assign(v0, v0 + n);
    c->readIdx = 0; c->writeIdx = 1;
    // a fresh exit choice (nav::ChoiceState::reset, exit_choice.cpp:671-679)
    c->assign.assign(n, CHOICE_NONE);
    std::memset(c->queue, 0, sizeof(c->queue));
    c->switches = c->forced = c->evaluations = c->eligible = 0;
    c->choiceStep = 0;
    c->pendEligible.assign(n, 0); c->pendBest.assign(n, 0u); c->pendBin.assign(n, 0u);
    c->hist.assign(size_t(CHOICE_MAX_EXITS) * CHOICE_GAIN_BINS, 0u);
    c->threshold.assign(size_t(CHOICE_MAX_EXITS), uint32_t(CHOICE_GAIN_BINS));
    c->ahead.assign(size_t(2) * CHOICE_QUEUE_STRIDE * CHOICE_AHEAD_BINS, 0u);
    c->evacStep.assign(n, -1); c->evacExit.assign(n, -1);
    c->steps = 0;
}

void oracle_restore_state(oracle_ctx* c, const float* x, const float* y, const float* vx, const float* vy, const float* w) {
    const size_t n = size_t(c->n);
    for (int h = 0; h < 2; ++h) {
        c->x[h].assign(x, x + n); c->y[h].assign(y, y + n); c->vx[h].assign(vx, vx + n); c->vy[h].assign(vy, vy + n);
        c->w[h].assign(w, w + n);
        for (size_t i = 0; i < n; ++i) c->spd[h][i] = std::sqrt(vx[i] * vx[i] + vy[i] * vy[i]);
    }
    c->readIdx = 0; c->writeIdx = 1;
}

void oracle_choice_step(oracle_ctx* c) {
    if (c->sc.E > 0) c->chooseExits();
}

void oracle_set_choice_state(oracle_ctx* c, uint32_t step, const uint32_t* assign, const uint32_t* queue128, const uint32_t* ahead) {
    c->choiceStep = step;
    if (assign) c->assign.assign(assign, assign + c->n);
    if (queue128) std::memcpy(c->queue, queue128, sizeof(c->queue));
    if (ahead) c->ahead.assign(ahead, ahead + size_t(2) * CHOICE_QUEUE_STRIDE * CHOICE_AHEAD_BINS);
}

// bitonic_sort.comp's network (cpu_crowd_sim.cpp:1012-1029) on (key, id) pairs, n a power of two
static void bitonicNetwork(int n, uint32_t* key, uint32_t* id) {
    for (int k = 2; k <= n; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1)
            for (int i = 0; i < n; ++i) {
                const int p = i ^ j;
                if (p <= i) continue;
                const bool asc = (i & k) == 0;
                if ((asc && key[i] > key[p]) || (!asc && key[i] < key[p])) {
                    std::swap(key[i], key[p]);
                    std::swap(id[i], id[p]);
                }
            }
}

// the network alone on keys in agent order: the sorted keys and agents; -1 unless n is a power of two
int oracle_bitonic_sort(int n, const uint32_t* keys, uint32_t* sortedKeys, uint32_t* ids) {
    if (n < 1 || (n & (n - 1)) != 0) return -1;
    for (int i = 0; i < n; ++i) { sortedKeys[i] = keys[i]; ids[i] = uint32_t(i); }
    bitonicNetwork(n, sortedKeys, ids);
    return 0;
}

// 0: the stable counting sort (SortMode::Counting), 1: the GPU's bitonic network (SortMode::GpuBitonic, a
// power-of-two agent count); returns -1 for a mode the agent count does not allow
int oracle_set_sort_mode(oracle_ctx* c, int mode) {
    if (mode == 1 && (c->n < 2 || (c->n & (c->n - 1)) != 0)) return -1;
    c->bitonic = mode == 1;
    return 0;
}

void oracle_set_drive(oracle_ctx* c, const float* dx, const float* dy) {
    if (!dx || !dy) { c->drive = false; return; }
    c->driveX.assign(dx, dx + c->n);
    c->driveY.assign(dy, dy + c->n);
    c->drive = true;
}

void oracle_step(oracle_ctx* c, int nsteps) {
    using Clock = std::chrono::steady_clock;
    const int n = c->n;
    const int T = c->threads;
    for (int s = 0; s < nsteps; ++s) {
        const auto t0 = Clock::now();
        // hash (calc_grid.comp, cpu_crowd_sim.cpp:983-993): a NaN position is reset in the input
        {
            float* px = c->x[c->readIdx].data();
            float* py = c->y[c->readIdx].data();
            float* pw = c->w[c->readIdx].data();
            const oracle_scene& sc = c->sc;
            float* pz = c->z[c->readIdx].data();
            parallelFor(T, n, false, [&](int i) {
                if (std::isnan(px[i])) { px[i] = 0.0f; py[i] = 0.0f; pz[i] = 0.0f; pw[i] = 1.0f; }
                uint32_t k = INACTIVE_CELL;
                if (pw[i] != 0.0f) {
                    const int gx = std::clamp(int((px[i] - sc.xmin) / sc.hashSpanX * float(sc.hashNX)), 0, sc.hashNX - 1);
                    const int gy = std::clamp(int((py[i] - sc.ymin) / sc.hashSpanY * float(sc.hashNY)), 0, sc.hashNY - 1);
                    k = uint32_t(gx + gy * sc.hashNX);
                }
                c->keyTmp[size_t(i)] = k;
                c->idTmp[size_t(i)] = uint32_t(i);
            });
        }
        const auto t1 = Clock::now();
        if (c->bitonic) {
            // bitonic_sort.comp's network (cpu_crowd_sim.cpp:1012-1029): stages k = 2 .. n, passes j = k/2 .. 1; the
            // lower index i of each pair (i, i ^ j) compares keys only and swaps the pair out of its stage's
            // direction (ascending where (i & k) == 0)
            std::copy(c->keyTmp.begin(), c->keyTmp.begin() + n, c->key.begin());
            std::copy(c->idTmp.begin(), c->idTmp.begin() + n, c->id.begin());
            bitonicNetwork(n, c->key.data(), c->id.data());
        } else {   // stable counting sort (cpu_crowd_sim.cpp:997-1006): ascending (cell, agent), inactive at the tail
            const uint32_t cells = uint32_t(c->sc.hashNX) * uint32_t(c->sc.hashNY);
            auto bucketOf = [cells](uint32_t g) { return g == INACTIVE_CELL ? cells : g; };
            std::fill(c->bucketStart.begin(), c->bucketStart.end(), 0u);
            for (int i = 0; i < n; ++i) ++c->bucketStart[bucketOf(c->keyTmp[size_t(i)]) + 1];
            for (size_t b = 1; b < c->bucketStart.size(); ++b) c->bucketStart[b] += c->bucketStart[b - 1];
            for (int i = 0; i < n; ++i) {
                const uint32_t pos = c->bucketStart[bucketOf(c->keyTmp[size_t(i)])]++;
                c->key[pos] = c->keyTmp[size_t(i)];
                c->id[pos] = c->idTmp[size_t(i)];
            }
        }
        const auto t2 = Clock::now();
        // cell ranges (reorder.comp, cpu_crowd_sim.cpp:1034-1043)
        std::fill(c->cellStart.begin(), c->cellStart.end(), 0u);
        std::fill(c->cellCount.begin(), c->cellCount.end(), 0u);
        for (int i = 0; i < n; ++i) {
            const uint32_t g = c->key[size_t(i)];
            if (g == INACTIVE_CELL) continue;
            ++c->cellCount[g];
            if (i == 0 || g != c->key[size_t(i) - 1]) c->cellStart[g] = uint32_t(i);
        }
        const auto t3 = Clock::now();
        // exit choice (PHYSICS_PASS 0), then every agent's e_goal from the start-of-step state
        if (!c->drive && c->sc.E > 0) c->chooseExits();
        {
            const float* px = c->x[c->readIdx].data();
            const float* py = c->y[c->readIdx].data();
            const float* pw = c->w[c->readIdx].data();
            parallelFor(T, n, true, [&](int i) {
                if (pw[i] == 0.0f) return;
                const V2 e = c->goalDirection(V2{ px[i], py[i] }, uint32_t(i));
                c->egX[size_t(i)] = e.x;
                c->egY[size_t(i)] = e.y;
            });
        }
        const auto t4 = Clock::now();
        const bool seated = c->sc.venue3d != 0;
        parallelFor(T, n, true, [&](int slot) {
            if (seated) c->predict3(uint32_t(slot));
            else c->predict(uint32_t(slot));
        });
        const auto t5 = Clock::now();
        std::vector<SlotState>* in = &c->slotA;
        std::vector<SlotState>* out = &c->slotB;
        for (int it = 1; it < c->par.projIters; ++it) {
            parallelFor(T, n, true, [&](int slot) {
                const uint32_t cid = c->key[size_t(slot)];
                if (cid == INACTIVE_CELL) return;
                (*out)[size_t(slot)] = seated ? c->projectSlot3(*in, uint32_t(slot), c->id[size_t(slot)], cid)
                                              : c->projectSlot(*in, uint32_t(slot), c->id[size_t(slot)], cid);
            });
            std::swap(in, out);
        }
        const auto t6 = Clock::now();
        const bool projectInFinal = c->par.projIters > 0;
        parallelFor(T, n, true, [&](int slot) {
            if (seated) c->finalPass3(*in, uint32_t(slot), projectInFinal);
            else c->finalPass(*in, uint32_t(slot), projectInFinal);
        });
        // evacuation events: active at the step's input, parked at its output
        {
            const float* wIn = c->w[c->readIdx].data();
            const float* wOut = c->w[c->writeIdx].data();
            const float* px = c->x[c->readIdx].data();
            const float* py = c->y[c->readIdx].data();
            for (int i = 0; i < n; ++i)
                if (wIn[i] != 0.0f && wOut[i] == 0.0f) {
                    c->evacStep[size_t(i)] = c->steps;
                    // the exit: the rectangle it entered; on the venue its opening by the last active position
                    // (metrics_deposit.comp exitSlot, seated::exitByLat)
                    const int e = seated ? vo::exitByLat(c->sc.venue, px[i], py[i]) : c->exitOf(V2{ px[i], py[i] });
                    c->evacExit[size_t(i)] = e >= 0 ? e : c->sc.E;
                }
        }
        std::swap(c->readIdx, c->writeIdx);
        ++c->steps;
        const auto t7 = Clock::now();
        c->times[0] = msBetween(t0, t1); c->times[1] = msBetween(t1, t2); c->times[2] = msBetween(t2, t3);
        c->times[3] = msBetween(t3, t4); c->times[4] = msBetween(t4, t5); c->times[5] = msBetween(t5, t6);
        c->times[6] = msBetween(t6, t7); c->times[7] = msBetween(t0, t7);
    }
}

void oracle_set_state3(oracle_ctx* c, const float* x, const float* y, const float* z, const float* vx, const float* vy, const float* vz,
                       const float* w) {
    const size_t n = size_t(c->n);
    std::vector<float> r(n, BODY_RADIUS), v0(n, V_MAX);
    oracle_set_state(c, x, y, vx, vy, w, r.data(), v0.data());
    for (int h = 0; h < 2; ++h) {
        c->z[h].assign(z, z + n);
        c->vz[h].assign(vz, vz + n);
        for (size_t i = 0; i < n; ++i) c->spd[h][i] = std::sqrt((vx[i] * vx[i] + vy[i] * vy[i]) + vz[i] * vz[i]);
    }
}

void oracle_get_state3(const oracle_ctx* c, float* x, float* y, float* z, float* vx, float* vy, float* vz, float* speed, float* w) {
    const int h = c->readIdx;
    std::copy(c->x[h].begin(), c->x[h].end(), x); std::copy(c->y[h].begin(), c->y[h].end(), y);
    std::copy(c->z[h].begin(), c->z[h].end(), z); std::copy(c->vx[h].begin(), c->vx[h].end(), vx);
    std::copy(c->vy[h].begin(), c->vy[h].end(), vy); std::copy(c->vz[h].begin(), c->vz[h].end(), vz);
    std::copy(c->spd[h].begin(), c->spd[h].end(), speed); std::copy(c->w[h].begin(), c->w[h].end(), w);
}

uint64_t oracle_fnv64(const void* data, uint64_t bytes) {   // FNV-1a 64 (the reference's cache checksum)
    const unsigned char* b = static_cast<const unsigned char*>(data);
    uint64_t h = 0xcbf29ce484222325ull;
    for (uint64_t i = 0; i < bytes; ++i) { h ^= b[i]; h *= 0x100000001b3ull; }
    return h;
}

void oracle_get_contacts(const oracle_ctx* c, uint8_t* contacts) { std::copy(c->contacts.begin(), c->contacts.end(), contacts); }

void oracle_get_state(const oracle_ctx* c, float* x, float* y, float* vx, float* vy, float* speed, float* w) {
    const int h = c->readIdx;
    const size_t n = size_t(c->n);
    if (x) std::copy(c->x[h].begin(), c->x[h].end(), x);
    if (y) std::copy(c->y[h].begin(), c->y[h].end(), y);
    if (vx) std::copy(c->vx[h].begin(), c->vx[h].end(), vx);
    if (vy) std::copy(c->vy[h].begin(), c->vy[h].end(), vy);
    if (speed) std::copy(c->spd[h].begin(), c->spd[h].end(), speed);
    if (w) std::copy(c->w[h].begin(), c->w[h].end(), w);
    (void)n;
}
void oracle_get_grid(const oracle_ctx* c, uint32_t* key, uint32_t* id) {
    std::copy(c->key.begin(), c->key.end(), key);
    std::copy(c->id.begin(), c->id.end(), id);
}
void oracle_get_cells(const oracle_ctx* c, uint32_t* start, uint32_t* count) {
    std::copy(c->cellStart.begin(), c->cellStart.end(), start);
    std::copy(c->cellCount.begin(), c->cellCount.end(), count);
}
void oracle_get_assign(const oracle_ctx* c, uint32_t* a) { std::copy(c->assign.begin(), c->assign.end(), a); }
void oracle_get_choice_counters(const oracle_ctx* c, uint32_t* o) {
    std::memcpy(o, c->queue, sizeof(c->queue));
    o[128] = c->switches; o[129] = c->forced; o[130] = c->evaluations; o[131] = c->eligible;
}
void oracle_get_evac(const oracle_ctx* c, int32_t* step, int32_t* exitId) {
    std::copy(c->evacStep.begin(), c->evacStep.end(), step);
    std::copy(c->evacExit.begin(), c->evacExit.end(), exitId);
}
int oracle_num_steps(const oracle_ctx* c) { return c->steps; }
void oracle_get_times(const oracle_ctx* c, double* t8) { std::memcpy(t8, c->times, sizeof(c->times)); }

void oracle_incline(int n, const float* g, float* v_free, float* t_headway) {
    for (int i = 0; i < n; ++i) {
        v_free[i] = oracle_ctx::inclineFreeSpeed(g[i]);
        t_headway[i] = oracle_ctx::inclineHeadwayTime(g[i]);
    }
}

int oracle_constants(float* out, int cap) {
    const float v[] = { DT, K_SEP, R_SEP, BODY_DIAM, SEP_MARGIN, HEADWAY_RATE, R_PERC2, PROJ_MAX_SHIFT_RADII, W_OBS,
                        GOAL_DEADZONE, OBSTACLE_RANGE, OBSTACLE_MIN_NORMAL, OBSTACLE_GRAD_H, OBSTACLE_GRAD_SCALE,
                        SAFE_NORMALIZE_EPS, SEP_STACKED_XY2, FEELER_REACH_RADII, WALL_RECAST_TOL_RADII, CENTRE_GRAD_MIN,
                        PLANE_PROBE_RADII, PARK_COORD, HASH_TO_RADIANS, CHOICE_REL_MARGIN, CHOICE_ABS_MARGIN_S,
                        CHOICE_SWITCH_SHARE, float(CHOICE_STEPS), V_MAX, T_HEADWAY, V_STAIR_UP, V_STAIR_DOWN,
                        INCLINE_INV_REF_SLOPE, T_HEADWAY_STAIR_UP, T_HEADWAY_STAIR_DOWN };
    const int m = int(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < m && i < cap; ++i) out[i] = v[i];
    return m;
}

} // extern "C"

// the stacked-pair direction of two agents (pairDirection), for the adjoint's FP64 reference
extern "C" void oracle_pair_direction(uint32_t self, uint32_t other, float* dx, float* dy) {
    const V2 d = pairDirection(self, other);
    *dx = d.x;
    *dy = d.y;
}
