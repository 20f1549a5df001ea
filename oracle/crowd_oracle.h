/* Oracle of the per-step crowd pipeline: a CPU restatement of the reference's CpuCrowdSim
 * (src/cpu_crowd_sim.cpp, include/swarm/cpu_crowd_sim.hpp) restricted to flat 2D floors.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the GPU path is compared against; only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it. The product path never
 * links, loads or calls it.
 *
 * Plain C ABI (ctypes): every array is caller-owned and copied on the call. */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

/* The scene tables (the same arrays the GPU path uploads). */
typedef struct {
    float xmin, ymin, xmax, ymax;          /* domain: agents outside are parked */
    float hashSpanX, hashSpanY;            /* spatial hash: cell = clamp(int((p - min) / span * n)) */
    int hashNX, hashNY;
    int percRing, projRing;                /* neighbour rings of the predict pass and the projection */
    int sdfNX, sdfNY;                      /* wall distance grid [sdfNY][sdfNX], texel (i, j) at (ox + i h, oy + j h) */
    float sdfOX, sdfOY, sdfInvH;
    const float* sdf;
    int navRX, navRY;                      /* navigation grid and candidate table (dense per cell) */
    float navOX, navOY, navCell, navInvCell;
    int K, E;                              /* candidate slots per cell, exits */
    const uint32_t* navInfo;               /* per cell: NAV_FIT 0x80000000 | NAV_HAS 0x40000000 */
    const uint32_t* navSlot;               /* [cell * K + k]: anchor | exit << 25 */
    const uint16_t* navDist;               /* [cell * K + k]: distance quanta */
    const float* invCapS;                  /* per exit: seconds per queued agent */
    float secondsPerUnit, distQuantum;
    const float* stair;                    /* per nav cell: free-speed factor */
    const float* exitRects;                /* E x (lox, loy, hix, hiy) */
    int nWalls;                            /* wall boxes (e_goal's wall rule queries them exactly) */
    const float* wallRects;                /* nWalls x (lox, loy, hix, hiy) */
    const void* venue;                     /* a seated venue's plan: its layout (oracle_venue_new), else NULL */
    int venue3d;                           /* 1: the venue's own 3D step (below); needs venue */
    float exitDepth, exitMaxZ, domainMax;  /* its exit test: max(|x|, |y|) > exitDepth below exitMaxZ, or > domainMax */
} oracle_scene;

typedef struct {
    float relaxStep;      /* exp(-DT / tau), folded on the host */
    float tau;            /* relaxation time (model time) */
    float obstacleGain;   /* wall-avoidance gain (OBSTACLE_GAIN) */
    int projIters;        /* Jacobi iterations (PROJ_ITERATIONS) */
    int threads;          /* OpenMP threads, <= 0: all */
} oracle_params;

/* The analytic seated venue (venue_oracle.cpp) for the step: on a venue plan the free speed and the headway rule's time gap
 * follow the floor's slope along e_goal (the reference's default venue model, cpu_crowd_sim.cpp:727-733), at the
 * reference's global free speed V_MAX (a venue plan does not read v0). The layout arrays are those of
 * oracle_venue_eval. */
void* oracle_venue_new(const float* lf, const int* li, const float* gwLo, const float* gwHi, const int* gwRow, const int* gwSeated,
                       const float* doorLo, const float* doorHi);
void oracle_venue_free(void* venue);
float oracle_venue_slope(const void* venue, float x, float y, float ex, float ey);
/* seated::inclineFreeSpeed and inclineHeadwayTime at EMPIRICAL's time gaps (src/venue.cpp:652-670) at n slopes */
void oracle_incline(int n, const float* g, float* v_free, float* t_headway);

/* The seated 3D step (venue3d): the reference's CpuCrowdSim on an analytic venue (src/cpu_crowd_sim.cpp, the seated
 * branches: the exit test, e_goal with the surface rule and the wall term, separation and the headway rule over 3D
 * offsets, the incline rules, obstacle avoidance at torso height against the venue and the pitch fence, gravity, the static
 * contacts (fence CCD, floor rest and settle, sweep-back, the exact walls), the wall plane, the projection with the fence
 * and the height test, the final contacts and the step-up limit with the venue's face margin). Global BODY_RADIUS and
 * V_MAX (r and v0 are not read). Positions and velocities carry z. */
void oracle_set_state3(oracle_ctx* c, const float* x, const float* y, const float* z, const float* vx, const float* vy, const float* vz,
                       const float* w);
void oracle_get_state3(const oracle_ctx* c, float* x, float* y, float* z, float* vx, float* vy, float* vz, float* speed, float* w);
/* FNV-1a 64 of a byte range (the checksum the reference's caches use) */
uint64_t oracle_fnv64(const void* data, uint64_t bytes);
/* the contact bits of the last step per agent (CpuCrowdSim::ContactBits) */
void oracle_get_contacts(const oracle_ctx* c, uint8_t* contacts);

oracle_ctx* oracle_create(const oracle_scene* scene, const oracle_params* params, int n);
void oracle_destroy(oracle_ctx* c);
/* A full state (both halves), resets the exit choice and the evacuation record (CpuCrowdSim::setState) */
void oracle_set_state(oracle_ctx* c, const float* x, const float* y, const float* vx, const float* vy, const float* w,
                      const float* r, const float* v0);
/* The agent state only (both halves), the exit choice runs on (CpuCrowdSim::restoreAgentState) */
void oracle_restore_state(oracle_ctx* c, const float* x, const float* y, const float* vx, const float* vy, const float* w);
/* The exit choice pass alone on the current state (CpuCrowdSim::stepExitChoice) */
void oracle_choice_step(oracle_ctx* c);
/* Sets the choice state: step counter, assigned exits (n), queue counters (128), queue-ahead histograms
 * (2 x 64 x 1024); NULL arrays are left as they are */
void oracle_set_choice_state(oracle_ctx* c, uint32_t step, const uint32_t* assign, const uint32_t* queue128, const uint32_t* ahead);
/* Forced desired direction per agent (CpuCrowdSim::setDriveTable); NULL: the navigation table */
void oracle_set_drive(oracle_ctx* c, const float* dx, const float* dy);
void oracle_step(oracle_ctx* c, int nsteps);
void oracle_get_state(const oracle_ctx* c, float* x, float* y, float* vx, float* vy, float* speed, float* w);
/* sorted (cell, agent) pairs and the cell table of the last step's input */
void oracle_get_grid(const oracle_ctx* c, uint32_t* key, uint32_t* id);
void oracle_get_cells(const oracle_ctx* c, uint32_t* start, uint32_t* count);
void oracle_get_assign(const oracle_ctx* c, uint32_t* assign);
/* queue[128], switches, forced, evaluations, eligible */
void oracle_get_choice_counters(const oracle_ctx* c, uint32_t* out132);
/* evacuation: per agent the step it left at (-1: not yet) and its exit (E: out of the domain) */
void oracle_get_evac(const oracle_ctx* c, int32_t* step, int32_t* exitId);
int oracle_num_steps(const oracle_ctx* c);
/* wall clock of the last step per stage, ms: hash, sort, cells, choice, predict, project, final, total */
void oracle_get_times(const oracle_ctx* c, double* t8);
/* e_goal of agents at (x, y) with radius r assigned to exit[i]: the navigation table's direction (choiceDirection,
 * exit_choice.cpp:557-566), with the wall rule when wall_rule != 0 (choiceGoal, :568-578), n points; applied[i]: the
 * wall rule changed it. The choice state is not touched. */
void oracle_goal(const oracle_ctx* c, int n, const float* x, const float* y, const float* r, const uint32_t* exit, int wall_rule,
                 float* ex, float* ey, uint8_t* applied);
/* The stacked-pair direction's sine and cosine (pairDirection's cos/sin, cpu_crowd_sim.cpp:123-129) at n angles */
void oracle_sincos(int n, const float* x, float* s, float* c);
/* choiceCell (exit_choice.cpp:512-548) at n points */
void oracle_choice_cell(const oracle_ctx* c, int n, const float* x, const float* y, int32_t* cell);
/* the exit rectangle holding each point (RectGeometry::exitAt, nav_field.cpp:155-159; -1: none), n points */
void oracle_exit_at(const oracle_ctx* c, int n, const float* x, const float* y, int32_t* exit);
/* The exact wall query (seated::wall, venue.cpp:763-813, over the scene's wall boxes): direction away from the
 * nearest wall within reach and the clearance (WALL_NONE 1e9 when none), n points */
void oracle_wall(const oracle_ctx* c, int n, const float* x, const float* y, float reach, float* wx, float* wy, float* wd);
/* exported model constants, for the constant-parity test */
int oracle_constants(float* out, int cap);

#ifdef __cplusplus
}
#endif
