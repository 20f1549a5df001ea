// Model constants of the per-step crowd pipeline, in the reference's model units
// (1 length unit = 31.25 m, 1 time unit = 17.490671 s).
//
// Every value below is the bit pattern the reference's shaders and CPU port use
// (include/swarm/sim_params.hpp:56-535): constants that the reference folds from
// decimal literals in double and rounds once (glsl_fold, sim_params.hpp:28-54) are
// written here as their folded float, in hex where the decimal would not round-trip.
// A CPU test recomputes every folded value from its decimal factors.
//
// Generalisations for flat 2D floors with per-agent radius r_i and free speed v0_i
// (the reference has global BODY_RADIUS 0.0064 and V_MAX 0.75):
//   contact distance        D_ij = r_i + r_j                  (BODY_DIAM at r = BODY_RADIUS)
//   separation support      S_ij = D_ij + SEP_MARGIN           (R_SEP = 0.016 at r = BODY_RADIUS, exact)
//   per-agent reaches       FEELER_REACH_RADII * r_i, PLANE_PROBE_RADII * r_i,
//                           PROJ_MAX_SHIFT_RADII * r_i, WALL_RECAST_TOL_RADII * r_i
//   speed clamp             |v| <= v0_i                        (V_MAX at v0 = V_MAX)
//   free speed              vFree = stair factor(x) * v0_i
#pragma once
#include <cstdint>

#define SWARM_HD __host__ __device__ inline

namespace swarm {

// --- model scales (sim_params.hpp:77-78) ---
constexpr float WORLD_SCALE_M_PER_UNIT = 31.25f;
constexpr float TIME_SCALE_S_PER_UNIT = 17.490671f;
constexpr double M_PER_UNIT = 31.25;
constexpr double S_PER_UNIT = 17.490671;

// --- body model (sim_params.hpp:86-102) ---
constexpr float BODY_RADIUS = 0.0064f;       // 0.20 m: reference global radius
constexpr float PERSONAL_RADIUS = 0.008f;    // 0.25 m
constexpr float R_SEP = 0.016f;              // 2 * PERSONAL_RADIUS
constexpr float BODY_DIAM = 0.0128f;         // 2 * BODY_RADIUS
// R_SEP - BODY_DIAM; exact by Sterbenz (0.0128 <= 0.016 <= 0.0256), so BODY_DIAM + SEP_MARGIN == R_SEP exactly
constexpr float SEP_MARGIN = R_SEP - BODY_DIAM;
constexpr float K_SEP = 100.0f;
// R_SEP * R_SEP and BODY_DIAM * BODY_DIAM as the reference folds them (the squares of the decimals, one float step
// from the float squares, sim_params.hpp:502,512): a pair at the reference's distances is tested against these
constexpr float R_SEP2 = 0x1.0c6f7ap-12f;
constexpr float BODY_DIAM2 = 0x1.5798eep-13f;

// --- step and kinematics (sim_params.hpp:105-148) ---
constexpr float DT = 0.01f;
constexpr float V_MAX = 0.75f;               // 1.34 m/s, the reference's global free speed
constexpr float TAU_RELAX_S = 0.5f;
constexpr float T_HEADWAY_S = 0.3f;
constexpr float TAU_RELAX = 0x1.d45d2cp-6f;  // 0.5 / 17.490671 folded (0.028586667)
constexpr float T_HEADWAY = 0x1.1904b4p-6f;  // 0.3 / 17.490671 folded (0.017152)
constexpr float HEADWAY_RATE = 0x1.d26afcp+5f;   // 1.0 / T_HEADWAY folded (58.30224)
constexpr float RELAX_STEP = 0x1.68de1cp-1f;     // exp(-DT / TAU_RELAX) folded in double (0.70481956)
// --- the incline speed and stair spacing of the seated venues' default model (sim_params.hpp:429-481; ModelOptions
// {QueueAware, inclineSpeed, StairFlow::Empirical}, exit_choice.cpp:51-58) ---
constexpr float INCLINE_INV_REF_SLOPE = 0x1.f66da2p+0f;   // 1.0 / 0.50952545 (tan 27 deg) folded (1.9626104)
constexpr float R_STAIR_UP = 0x1.bb3950p-2f;              // 0.58 / 1.34 folded: V_STAIR_UP / V_MAX, per agent (0.43283582)
constexpr float R_STAIR_DOWN = 0x1.f4898ep-2f;            // 0.655 / 1.34 folded: V_STAIR_DOWN / V_MAX (0.48880598)
constexpr float T_HEADWAY_STAIR_UP = 0x1.24ba3cp-5f;      // 0.625 / 17.490671 folded, EMPIRICAL (0.035733335)
constexpr float T_HEADWAY_STAIR_DOWN = 0x1.306fc4p-5f;    // 0.65 / 17.490671 folded, EMPIRICAL (0.03716267)
constexpr float R_PERC = 0.045f;
constexpr float R_PERC2 = 0x1.096bbap-9f;        // R_PERC * R_PERC folded (0.002025)

// --- hard-body projection (sim_params.hpp:160-163) ---
constexpr int PROJ_ITERATIONS = 32;
constexpr float PROJ_RELAX = 1.0f;
constexpr float PROJ_MAX_SHIFT_RADII = 0.05f;

// --- steering (sim_params.hpp:171-185) ---
constexpr float W_OBS = 10.0f;
constexpr float GOAL_DEADZONE = 1e-4f;
constexpr float WALL_NONE = 1e9f;   // the wall query's clearance when no wall lies within reach (venue.hpp:251)
constexpr float OBSTACLE_RANGE = 0.025f;
constexpr float OBSTACLE_GAIN = 20.0f;          // the wall-avoidance stiffness (a differentiable global)
constexpr float OBSTACLE_MIN_NORMAL = 0.01f;
constexpr float OBSTACLE_GRAD_H = 0.0015f;
constexpr float OBSTACLE_GRAD_SCALE = 0x1.4d5556p+7f;   // 0.25 / OBSTACLE_GRAD_H folded (166.66667)
constexpr float SAFE_NORMALIZE_EPS = 1e-8f;
constexpr float SEP_STACKED_XY2 = 1e-12f;

// --- static wall contacts (sim_params.hpp:205-227) ---
constexpr float FEELER_REACH_RADII = 1.25f;
constexpr int FEELER_COUNT = 16;
constexpr int WALL_ROOT_STEPS = 3;
constexpr int WALL_RECAST_STEPS = 2;
constexpr float WALL_RECAST_TOL_RADII = 0.02f;
constexpr int WALL_PASSES = 3;
constexpr float CENTRE_GRAD_MIN = 1e-3f;
constexpr int STATIC_SWEEP_BISECTIONS = 8;
constexpr float PLANE_PROBE_RADII = 2.5f;
constexpr float NO_CONTACT = 1e9f;            // wallContact's "no wall" clearance

// --- the seated venues' own 3D step (sim_params.hpp:112-117,173-199,205-294; the reference's global radius) ---
constexpr float FEELER_REACH = 0x1.0624dep-7f;          // FEELER_REACH_RADII * BODY_RADIUS folded (0.008)
constexpr float PLANE_PROBE_REACH = 0x1.0624dep-6f;     // PLANE_PROBE_RADII * BODY_RADIUS folded (0.016)
constexpr float PROJ_MAX_SHIFT = 0x1.4f8b58p-12f;       // PROJ_MAX_SHIFT_RADII * BODY_RADIUS folded (0.00032)
constexpr float GRAVITY_DT = 0x1.ebb3b0p-1f;            // 9.81 m/s^2 in model units (96.03553) * DT, folded
constexpr float TORSO_HEIGHT = 0.05f;                   // the obstacle avoidance's sample, above the centre
constexpr float FLOOR_SETTLE_REACH = 0x1.3a92a4p-7f;    // 1.5 BODY_RADIUS folded (0.0096): a floor found this far down
constexpr float STEP_UP_FLOOR_REACH = 0x1.3a92a4p-5f;   // 6 BODY_RADIUS folded (0.0384): the step-up's start floor search
constexpr float MAX_STEP_UP = 0x1.3a92a4p-7f;           // 0.30 m / 31.25 folded (0.0096)
constexpr int STEP_UP_SLIDE_TRIES = 3;                  // slides along a blocking face: 0, 30 and 60 degrees off it

// --- exit / park (sim_params.hpp:351,393) ---
constexpr float PARK_COORD = 1000.0f;
constexpr uint32_t INACTIVE_CELL = 0xFFFFFFFFu;

// lowbias32 angle scale: 6.28318530718 / 2^32 folded (sim_params.hpp:507)
constexpr float HASH_TO_RADIANS = 0x1.921fb6p-30f;

// --- queue-aware exit choice (sim_params.hpp:412-427, exit_choice.hpp:104-137) ---
constexpr int CHOICE_STEPS = 17;              // int(T_CHOICE / DT + 0.5), T_CHOICE = 3 s
constexpr float CHOICE_REL_MARGIN = 0.10f;
constexpr float CHOICE_ABS_MARGIN_S = 5.0f;
constexpr float CHOICE_SWITCH_SHARE = 0.10f;
constexpr uint32_t CHOICE_MIN_SWITCHES = 1u;
constexpr int CHOICE_GAIN_BINS = 256;
constexpr float EXIT_FLOW_P_M_MIN = 83.6f;
constexpr float CHOICE_S_PER_UNIT = 0x1.752264p+4f;   // 31.25 / 1.34 folded (23.320896)
constexpr float CHOICE_DIST_QUANTUM = 1.0f / 4096.0f;   // 2^-12 model units
constexpr int CHOICE_K = 24;                 // candidate slots at most
// the exit-type rule of a two-type table (a seated venue's avenues, type 0, and doors, type 1): the 2 nearest avenues and the
// nearest door are always candidates (exit_choice.hpp:111-115)
constexpr int CHOICE_TYPE_MIN[2] = { 2, 1 };
constexpr int CHOICE_TYPE_MIN_MAX = 2;
constexpr int CHOICE_EXIT_SHIFT = 25;
constexpr uint32_t CHOICE_ANCHOR_MASK = (1u << CHOICE_EXIT_SHIFT) - 1u;
constexpr uint32_t CHOICE_EMPTY = 63u;
constexpr uint32_t CHOICE_DIST_EMPTY = 0xFFFFu;
constexpr int CHOICE_MAX_EXITS = 63;
constexpr int CHOICE_QUEUE_STRIDE = 64;
constexpr uint32_t CHOICE_NONE = 0xFFFFFFFFu;
constexpr int CHOICE_AHEAD_BINS = 1024;
constexpr int CHOICE_AHEAD_SHIFT = 6;
// per nav cell info word: the cell's class and whether it holds candidates
constexpr uint32_t NAV_FIT = 0x80000000u;
constexpr uint32_t NAV_HAS = 0x40000000u;

// --- wall planes (physics.comp slotPlanes flags) ---
constexpr uint32_t PLANE_WALL = 1u, PLANE_GROUNDED = 2u;

// 16 feeler directions every 22.5 degrees (sim_params.hpp:208-210), the same literals
struct F2 { float x, y; };
constexpr F2 FEELER_DIRS[FEELER_COUNT] = {
    { 1.0f, 0.0f }, { 0.92387953f, 0.38268343f }, { 0.70710678f, 0.70710678f }, { 0.38268343f, 0.92387953f },
    { 0.0f, 1.0f }, { -0.38268343f, 0.92387953f }, { -0.70710678f, 0.70710678f }, { -0.92387953f, 0.38268343f },
    { -1.0f, 0.0f }, { -0.92387953f, -0.38268343f }, { -0.70710678f, -0.70710678f }, { -0.38268343f, -0.92387953f },
    { 0.0f, -1.0f }, { 0.38268343f, -0.92387953f }, { 0.70710678f, -0.70710678f }, { 0.92387953f, -0.38268343f } };

// lowbias32 (physics.comp:338-343)
SWARM_HD uint32_t hash32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Float sine and cosine on [0, 2 pi) for the stacked-pair direction. One implementation, compiled
// identically for host and device (no FMA contraction): Cody-Waite reduction by pi/2 in three parts
// and the Cephes single-precision minimax polynomials (S. L. Moshier, Cephes Math Library, sinf.c / cosf.c).
SWARM_HD void sincos_cephes(float x, float* s, float* c) {
    const float FOPI = 1.27323954473516f;       // 4 / pi
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
    default: sn = -pc; cs = ps; break;   // j == 6
    }
    *s = sn; *c = cs;
}

} // namespace swarm
