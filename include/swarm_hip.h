/* swarm_hip.h - C ABI of the MI355X (gfx950) per-step crowd pipeline.
 *
 * The drop-in boundary for the reference's GL dispatch sequence: the reference's C++ simulator
 * class CrowdSim (include/swarm/crowd_sim.hpp:53-194) drives calc_grid.comp, bitonic_sort*.comp,
 * reorder.comp and physics.comp through OpenGL; here the same calls drive hand-written HIP kernels.
 * Plain pointers and sizes only. Every host array is caller-owned and copied on the call.
 * Return codes: SWARM_OK (0) or a negative SWARM_E*; swarm_last_error() holds the message.
 * One thread per context; independent contexts may run on different threads. Readbacks
 * synchronise the context's own stream only.
 *
 * Units: the reference's model units (1 length unit = 31.25 m, 1 time unit = 17.490671 s).
 * Agent state is SoA FP32: x, y, vx, vy, speed (velocity.w), w (1 active, 0 evacuated).
 */
#ifndef SWARM_HIP_H
#define SWARM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWARM_OK 0
#define SWARM_EINVAL (-1)   /* bad argument (agent count, iterations, layout, null pointer) */
#define SWARM_ESIZE (-2)    /* size mismatch */
#define SWARM_ENOMEM (-3)   /* device allocation failed */
#define SWARM_ESCENE (-4)   /* invalid scene */
#define SWARM_EHIP (-5)     /* HIP runtime error or no device */
#define SWARM_ENONFINITE (-6)  /* swarm_adjoint_backward: a gradient value is inf or NaN (the outputs are still written) */

typedef struct swarm_scene swarm_scene;
typedef struct swarm_ctx swarm_ctx;

/* ---------------- scene preparation (input tables; not part of the step) ---------------- */

/* Benchmark floors */
#define SWARM_LAYOUT_ROOM 0          /* 40 x 30 m room, two 2 m exits */
#define SWARM_LAYOUT_BOWL 1          /* 200 x 150 m seated bowl, 240 rows, 16 gangways, 8 exits */
#define SWARM_LAYOUT_CORRIDOR 2      /* 60 x 20 m corridor, one 1.5 m exit */
#define SWARM_LAYOUT_BOWL_EIGHTH 3   /* the bowl cropped to 1/8 */
#define SWARM_LAYOUT_PILLAR 4        /* 30 x 30 m room with a 4 x 4 m pillar */
#define SWARM_LAYOUT_BOWL_QUARTER 5  /* the bowl cropped to 1/4 */

/* A rectangle world (replaces RectGeometry, include/swarm/nav_field.hpp:135-160, plus the SDF bake
 * of src/sdf.cpp and the stair bands). Rects are (lox, loy, hix, hiy), model units. */
typedef struct {
    float xmin, ymin, xmax, ymax;
    float hash_cell, sdf_spacing, sdf_clamp;
    int nav_res_x, nav_res_y;
    float nav_origin_x, nav_origin_y, nav_cell, nav_inv_cell;
    float nav_body;
    double metres_per_unit;
    int n_walls, n_exits, n_stairs;
    const float* walls;          /* n_walls x 4 */
    const float* exits;          /* n_exits x 4 */
    const float* exit_width_m;   /* n_exits (0: the rect's longer side) */
    const float* stairs;         /* n_stairs x 4 */
    const float* stair_factor;   /* n_stairs */
} swarm_scene_desc;

/* Sizes and scalars of a built scene */
typedef struct {
    float xmin, ymin, xmax, ymax, hash_span_x, hash_span_y;
    int hash_nx, hash_ny, perc_ring, proj_ring;
    int sdf_nx, sdf_ny;
    float sdf_ox, sdf_oy, sdf_h, sdf_inv_h;
    int sdf_min_win;
    int nav_res_x, nav_res_y;
    float nav_origin_x, nav_origin_y, nav_cell, nav_inv_cell;
    int K, E;
    float seconds_per_unit, dist_quantum;
    double build_seconds;
    int n_walls, wall_bin_entries;   /* wall boxes; entries of their index on the hash grid */
    int from_cache;                  /* 1: the tables were loaded from the scene cache */
} swarm_scene_info;

/* nav::build + buildChoiceTable (src/nav_field.cpp:344-366, src/exit_choice.cpp:143-317) and the SDF bake
 * (src/sdf.cpp) for a rectangle world. max_radius / max_v0 size the neighbour rings. threads <= 0: all. */
int swarm_scene_build(const swarm_scene_desc* desc, float max_radius, float max_v0, int threads, swarm_scene** out);
/* The same for a benchmark floor (SWARM_LAYOUT_*) */
int swarm_scene_build_layout(int layout, float max_radius, float max_v0, int threads, swarm_scene** out);
/* The same through the checksummed scene cache in cache_dir (the reference's navigation / choice caches,
 * src/nav_field.cpp:468-688, src/exit_choice.cpp:335-445): the file scene_<key>.bin, key = FNV-1a 64 of everything the
 * tables depend on, holds magic + payload + the payload's FNV-1a checksum; a valid file is loaded (info.from_cache 1),
 * an absent, foreign or damaged one is replaced by a fresh build (written to a temporary file, then renamed). NULL or
 * "": no cache. The tables are the same bit for bit either way. */
int swarm_scene_build_cached(const swarm_scene_desc* desc, float max_radius, float max_v0, int threads, const char* cache_dir,
                             swarm_scene** out);
int swarm_scene_build_layout_cached(int layout, float max_radius, float max_v0, int threads, const char* cache_dir, swarm_scene** out);
int swarm_scene_info_get(const swarm_scene* s, swarm_scene_info* info);
/* Copies of the tables; any pointer may be NULL. Sizes: sdf, sdf_min sdf_nx*sdf_ny; nav_info, stair
 * nav_res_x*nav_res_y; nav_slot, nav_dist cells*K; inv_cap_s E; exit_rects 4E. */
int swarm_scene_read(const swarm_scene* s, float* sdf, float* sdf_min, uint32_t* nav_info, uint32_t* nav_slot,
                     uint16_t* nav_dist, float* inv_cap_s, float* stair, float* exit_rects);
/* The wall boxes (n_walls x (lox, loy, hix, hiy)) and their index on the hash grid (per hash cell, ascending, the
 * walls e_goal's wall rule can reach from it: bin_start hash_nx * hash_ny + 1, bin_list wall_bin_entries). Any
 * pointer may be NULL. The reference queries its analytic venue's walls directly (seated::wall, src/venue.cpp:763-813). */
int swarm_scene_read_walls(const swarm_scene* s, float* rects, uint32_t* bin_start, uint32_t* bin_list);
/* The floor's safety-metric grid (MetricVenue, metrics.hpp:145-162; its walkable area per cell as
 * src/metrics.cpp:143-164 samples it, there over the terrace's centred 64 x 64 m square): 1 m cells covering the
 * floor from whole metres, g = xy * cells_per_unit + off per axis, and each cell's walkable area in m^2 from
 * 40 x 40 samples against the floor's walls. Host only; any
 * pointer may be NULL; walkable_m2 holds res_x * res_y values, row-major in y. */
int swarm_scene_metric_grid(const swarm_scene* s, int* res_x, int* res_y, float* cells_per_unit, float* off_x, float* off_y,
                            float* walkable_m2);
void swarm_scene_destroy(swarm_scene* s);

/* Seeded agents on a benchmark floor (replaces CrowdSim's spawners, src/crowd_sim.cpp:485-808): 'active' of its
 * slots, padded to 'total' with parked agents; r ~ U[0.20, 0.25] m, v0 ~ N(mean, sd) m/s clipped to [0.5, 2.0]. */
int swarm_place_agents(int layout, size_t active, size_t total, uint64_t seed, double speed_mean, double speed_sd,
                       float* x, float* y, float* r, float* v0, float* w);
/* The reference's seat spawn (seatedSpawnSeats, src/crowd_sim.cpp:652-674): round(load * seat_count) of seat_count seats
 * chosen by a partial Fisher-Yates on mt19937's raw output from seed, in seat order, padded with -1 to n (chosen[n]);
 * SWARM_EINVAL when more seats are chosen than n. The same seats as the reference for the same arguments. */
int swarm_spawn_seats(int seat_count, int n, double load, uint32_t seed, int32_t* chosen);
/* An open floor (SWARM_LAYOUT_ROOM, _CORRIDOR, _PILLAR): 'active' agents placed uniformly with rejection (mt19937's raw
 * output), then relaxed by Gauss-Seidel sweeps over overlapping pairs (relaxSpawn, src/crowd_sim.cpp:544-597, with
 * per-agent radii), padded to 'total' with parked agents; r ~ U[0.20, 0.25] m, v0 ~ N(mean, sd) m/s clipped to
 * [0.5, 2.0]. sweeps / max_overlap (may be NULL): the relaxation's sweeps and its last largest overlap (model units). */
int swarm_spawn_uniform(int layout, size_t active, size_t total, uint32_t seed, double speed_mean, double speed_sd, float* x, float* y,
                        float* r, float* v0, float* w, int* sweeps, float* max_overlap);
/* Slots of a benchmark floor (capacity query; x, y may be NULL) */
int swarm_layout_slots(int layout, float* x, float* y, size_t cap, size_t* count);

/* tau (model time) and exp(-DT / tau) for a relaxation time in seconds, folded from decimals (sim_params.hpp:128-129,510) */
int swarm_model_params(float tau_s, float* tau, float* relax_step);

/* ---------------- the analytic seated venues S, M and L (include/swarm/venue.hpp, src/venue.cpp) ---------------- */

#define SWARM_VENUE_MAX_GANGWAYS 8
/* SeatedLayout (venue.hpp:103-145) and the Venue fields around it (venue.hpp:149-166), model units; the exit sizing
 * rule's figures in metres. Fields up to corner_door_lo are the layout the heightfield queries read. */
typedef struct {
    float stand_front, row0, row_depth, clearway;
    int rows;
    float conc_front, conc_back, outer_face, landing_end;
    float avenue_half, lat0, gangway_lo, gangway_hi, exit_lo, exit_hi, corner_inset;
    float seat_inset, face_margin, wall_above, wall_clear;
    float first_row, row_rise, seat_top, barrier_top, concourse_z, wall_top, rake;
    int max_band_steps;
    float seat_width, domain_half;
    int gangways;
    float gw_lo[SWARM_VENUE_MAX_GANGWAYS], gw_hi[SWARM_VENUE_MAX_GANGWAYS];
    int gw_row[SWARM_VENUE_MAX_GANGWAYS], gw_seated[SWARM_VENUE_MAX_GANGWAYS];
    int doors;
    float door_lo[SWARM_VENUE_MAX_GANGWAYS], door_hi[SWARM_VENUE_MAX_GANGWAYS];
    float corner_door_lo;   /* >= 1e8: no corner door */
    float domain_min, domain_span, exit_depth, exit_max_z;
    int grid_res, metric_grid_res, exit_count, gangway_count, seat_total;
    double door_width_m, corner_door_width_m, total_exit_width_m, required_exit_width_m;
} swarm_venue_layout;

/* venueByName + buildSeatedLayout (src/venue.cpp:121-327): name "S", "M" or "L" (SWARM_EINVAL otherwise) */
int swarm_venue_layout_get(const char* name, swarm_venue_layout* out);
/* seated::seats (src/venue.cpp:1054-1083): every seat in the order quadrant, wing, row, segment, index; x, y, z the spawn
 * position (clearway centre, rest height), where 5 per seat (quadrant, wing, row, segment, index). Any array may be NULL
 * (count only); otherwise it holds cap entries and SWARM_ESIZE is returned when the venue has more seats. */
int swarm_venue_seats(const char* name, float* x, float* y, float* z, int32_t* where, size_t cap, size_t* count);
/* The heightfield queries of src/venue.cpp:481-935 (the GLSL venue_seated.glsl inside physics.comp), evaluated on the GPU
 * at n points p (n x 3) with plan directions e (n x 2): sdf_grad n x 4 (seated::sdfGrad: gradient, signed distance),
 * floor_h n (floorHeight), cell n x 3 (cellAt: kind, row, band), wall n x 9 (seated::wall at reach[0], [1], [2]: unit
 * direction away from the wall, clearance; 1e9: none within reach), slope n (floorSlope), exit_id n (exitByLat),
 * gangway n (gangwayAt). Output pointers may be NULL. Synchronous on the current device. */
int swarm_venue_probe(const char* name, size_t n, const float* p, const float* e, const float* reach3, float* sdf_grad, float* floor_h,
                      int32_t* cell, float* wall, float* slope, int32_t* exit_id, int32_t* gangway);
/* The venue's plan as a rectangle world for the step (the 2D cut of the analytic venue): walls the footprints of the seat
 * blocks, barrier, side and outer walls and the fenced pitch; the domain ends at the outer wall's outer face (the reference's
 * exit test, max(|x|, |y|) > exitDepth); exit rectangles beyond it, one per avenue and door, the L-shaped corner door as
 * two (its second halves after the reference's exits). The plan carries the layout: the step applies the reference's
 * default venue model (defaultModelOptions, src/exit_choice.cpp:51-58) on the gangway inclines, the free speed and the
 * headway rule's time gap by the floor's slope along e_goal (cpu_crowd_sim.cpp:727-733, the agent's v0 in place of
 * V_MAX). Built as swarm_scene_build_cached. */
int swarm_scene_build_venue(const char* name, float max_radius, float max_v0, int threads, const char* cache_dir, swarm_scene** out);
/* A scene's venue: has_venue 1 for a venue plan (swarm_scene_build_venue), 2 for a venue's own 3D step
 * (swarm_scene_build_seated), 0 otherwise (layout zeroed). layout (may be NULL): the fields up to corner_door_lo, and
 * for a 3D scene also domain_min, domain_span, exit_depth, exit_max_z, grid_res and exit_count; the rest zero. */
int swarm_scene_read_venue(const swarm_scene* s, int* has_venue, swarm_venue_layout* layout);
/* The reference's own 3D step on its seated venue S, M or L (CrowdSim on a seated Venue with defaultModelOptions): the
 * venue's neighbour grid (gridRes cells over its domain), its navigation field on the heightfield (nav::forVenue: 0.05 m
 * cells, SeatedGeometry, src/nav_field.cpp:41-101) and candidate table with the exit-type rule (nav::choiceForVenue,
 * src/exit_choice.cpp:143-333), bit for bit the reference's; the heightfield is evaluated exactly in the step. Agents
 * carry z (swarm_upload3 / swarm_state3) and the reference's global radius and free speed. Through the scene cache in
 * cache_dir (NULL or "": none). */
int swarm_scene_build_seated(const char* name, int threads, const char* cache_dir, swarm_scene** out);
/* The reference's exit id of each plan exit (ref_exit[cap]; count: the plan's exits) */
int swarm_venue_exit_map(const char* name, int32_t* ref_exit, int cap, int* count);

/* ---------------- the plain step (CrowdSim) ---------------- */

typedef struct {
    float tau_s;           /* relaxation time, seconds (0.5) */
    float obstacle_gain;   /* wall-avoidance gain (20) */
    int proj_iters;        /* Jacobi iterations (32) */
    int use_graph;         /* capture steps in a hipGraph (1) */
} swarm_config;

/* CrowdSim(shaderDir, numBoids, venue, options) (crowd_sim.hpp:62-73); n: any positive agent count */
int swarm_create(const swarm_scene* scene, int n, const swarm_config* cfg, swarm_ctx** out);
/* CrowdSim::upload / CpuCrowdSim::setState (crowd_sim.hpp:90, cpu_crowd_sim.hpp:91): both halves; resets the exit
 * choice, the evacuation record and the step counter */
int swarm_upload(swarm_ctx* c, const float* x, const float* y, const float* vx, const float* vy, const float* w,
                 const float* r, const float* v0);
/* CpuCrowdSim::restoreAgentState (cpu_crowd_sim.hpp:94): the agent state only */
int swarm_restore_state(swarm_ctx* c, const float* x, const float* y, const float* vx, const float* vy, const float* w);
/* CpuCrowdSim::setDriveTable / physics.comp TEST_DRIVE_TABLE (cpu_crowd_sim.hpp:121): forced e_goal; NULL clears */
int swarm_set_drive(swarm_ctx* c, const float* dx, const float* dy);
/* CrowdSim::step (crowd_sim.hpp:91, crowd_sim.cpp:316-412), nsteps times; asynchronous */
int swarm_step(swarm_ctx* c, int nsteps);
/* CrowdSim::stepExitChoice / CpuCrowdSim::stepExitChoice (crowd_sim.hpp:116, cpu_crowd_sim.hpp:104): the exit choice pass alone */
int swarm_step_exit_choice(swarm_ctx* c);
/* The choice state, as the reference's choice harness sets it (tools/choice_check.hpp:88-100: setAssign,
 * setCounters, setAhead): step counter, assigned exits (n), queue counters (128), queue-ahead histograms
 * (2 x 64 x 1024); NULL arrays are left as they are */
int swarm_set_choice_state(swarm_ctx* c, uint32_t step, const uint32_t* assign, const uint32_t* queue128, const uint32_t* ahead);
int swarm_sync(swarm_ctx* c);
/* CrowdSim::setProjectionIterations (crowd_sim.hpp:125) */
int swarm_set_proj_iters(swarm_ctx* c, int iterations);
/* currentStateBuffer readback (crowd_sim.hpp:132; CpuCrowdSim::state, cpu_crowd_sim.hpp:129) */
int swarm_read_state(swarm_ctx* c, float* x, float* y, float* vx, float* vy, float* speed, float* w);
/* A seated venue's own 3D step (a context on a swarm_scene_build_seated scene): CrowdSim::upload / currentStateBuffer
 * with GpuBoid's full position and velocity (gpu_types.hpp:17-20): x, y, z, w and vx, vy, vz, speed; radii and free
 * speeds are the reference's global BODY_RADIUS and V_MAX. upload3 resets the exit choice and the evacuation record;
 * either refuses (SWARM_EINVAL) a context on any other scene. Output pointers may be NULL. */
int swarm_upload3(swarm_ctx* c, const float* x, const float* y, const float* z, const float* vx, const float* vy, const float* vz,
                  const float* w);
int swarm_read_state3(swarm_ctx* c, float* x, float* y, float* z, float* vx, float* vy, float* vz, float* speed, float* w);
/* gridBuffer readback (crowd_sim.hpp:134; CpuCrowdSim::grid, cpu_crowd_sim.hpp:131): sorted (cell, agent) pairs of the last step's input */
int swarm_read_grid(swarm_ctx* c, uint32_t* key, uint32_t* id);
/* cellBuffer readback (crowd_sim.hpp:135; CpuCrowdSim::cells, cpu_crowd_sim.hpp:132): start and count per cell */
int swarm_read_cells(swarm_ctx* c, uint32_t* start, uint32_t* count);
/* assignBuffer readback (crowd_sim.hpp:109; CpuCrowdSim::assignedExits, cpu_crowd_sim.hpp:100) */
int swarm_read_assign(swarm_ctx* c, uint32_t* assign);
/* readChoiceCounters (crowd_sim.hpp:115): queue[128], switches, forced, evaluations, eligible */
int swarm_read_choice_counters(swarm_ctx* c, uint32_t* out132);
/* Evacuation record (metrics_deposit.comp:96-104's events): per agent its step and exit (-1: none; E: domain), per exit counts (E + 1) */
int swarm_read_evac(swarm_ctx* c, int32_t* step, int32_t* exit_id, uint32_t* per_exit);
/* previousStateBuffer readback (crowd_sim.hpp:133): the state one step earlier (before the first step: the upload) */
int swarm_read_prev_state(swarm_ctx* c, float* x, float* y, float* vx, float* vy, float* speed, float* w);
/* CrowdSim::sortGrid (crowd_sim.hpp:99): the sort stage alone on keys of the caller's own. keys: n cell indices
 * (< swarm_num_cells) or 0xFFFFFFFF (inactive); out: the sorted keys and the permutation (ascending (key, index),
 * the stable counting sort's order). Leaves the context's own grid readback holding this sort. */
int swarm_sort_grid(swarm_ctx* c, const uint32_t* keys, uint32_t* sorted_keys, uint32_t* ids);
/* The step's navigation lookups at m caller points, on the GPU (the reference's probe dispatch of physics.comp's
 * navChoiceCell / navChoiceDirection / navChoiceGoal, tools/test_gpu.cpp:4155-4191): per point its nav cell, the table's
 * direction to exit_id[i] (dir: 2 floats) and e_goal with the wall rule at radius r[i] (goal: 2 floats). Scenario 0's
 * scene; synchronises. */
int swarm_nav_probe(swarm_ctx* c, int m, const float* x, const float* y, const float* r, const uint32_t* exit_id, int32_t* cell,
                    float* dir, float* goal);
/* step counter, active agents at the last step's input, running sum of active agents over steps */
int swarm_read_control(swarm_ctx* c, uint32_t* step, uint32_t* n_active, uint64_t* active_sum);
/* CrowdSim::setTimingEnabled / readTimings (crowd_sim.hpp:142-143): per-pass HIP events (steps then run
 * without graph capture). ms8: mean ms per step of grid passes (hash, sort, cells, gather), exit choice,
 * predict, projection iterations, final, total; the mean ms of one projection iteration after the first (on = 2
 * only: a pair of events recorded by each of those dispatches, which slows the pass totals); the steps timed.
 * Switching the mode starts a new measurement. */
int swarm_set_timing(swarm_ctx* c, int on);
int swarm_read_timings(swarm_ctx* c, double* ms8);
/* nsteps steps between two HIP events on the context's stream (synchronises); *ms: their elapsed time */
int swarm_time_steps(swarm_ctx* c, int nsteps, double* ms);
/* A plain device-to-device copy kernel of 'bytes' (reps launches, mean ms per launch): the measured copy
 * bandwidth and the calibration of HBM traffic counters. No counterpart in the reference (measurement aid). */
int swarm_copy_probe(size_t bytes, int reps, double* ms);
/* ---------------- the batched step ----------------
 * No counterpart in the reference (its CrowdSim steps one venue; the batch is the ensemble form of
 * CrowdSim::step, crowd_sim.cpp:316-412). Scenario s runs scene layouts[scen_layout[s]] with scen_agents[s]
 * agents; every pass is one launch for all scenarios. The state arrays of swarm_upload / swarm_read_state /
 * swarm_read_assign / swarm_read_evac are the scenarios' arrays concatenated in scenario order; grid keys and
 * cell tables are over global cells (scenario s's cells follow scenario s-1's) and grid ids are global agent
 * indices. Each scenario's results equal its own plain context's, bit for bit. swarm_set_drive is refused. */
int swarm_batch_create(const swarm_scene* const* layouts, int n_layouts, const int* scen_layout, const int* scen_agents, int scenarios,
                       const swarm_config* cfg, swarm_ctx** out);
int swarm_num_scenarios(swarm_ctx* c);
/* readChoiceCounters per scenario: 132 words each (queue[128], switches, forced, evaluations, eligible) */
int swarm_read_choice_counters_all(swarm_ctx* c, uint32_t* out);
/* the evacuation counters per scenario: 64 words each (exit e's count at e, out-of-domain at E) */
int swarm_read_evac_counters(swarm_ctx* c, uint32_t* per_exit64);
/* Launch shape of the physics passes (predict, projection, final): block 0 (default 128), 64, 128, 192 or 256
 * threads; grid 0 (one slot per thread) or a persistent grid of that many blocks. Results do not depend on it.
 * The reference fixes its workgroup size in the shaders (physics.comp:22, local_size_x = 128). */
int swarm_set_physics_shape(swarm_ctx* c, int block, int grid);
int swarm_get_physics_shape(swarm_ctx* c, int* block, int* grid);
/* Candidate pruning in the projection: every 'every' iterations (default 8; 0: never) each slot drops the contact
 * candidates that cannot touch in the iterations left. Results are the same bit for bit either way. No counterpart
 * in the reference (its projection scans the 3 x 3 cell ring every iteration, physics.comp:1367-1418). */
int swarm_set_projection_prune(swarm_ctx* c, int every);
/* The projection's candidate lists' displacement budget, in maximal moves (PROJ_MAX_SHIFT_RADII r each) over a step.
 * The first iteration scans the cell ring and keeps, in ring order, every candidate within D + moves (shift_i +
 * shift_j) (a seated venue's own step: every body with a height overlap within D + 2 min(moves, P - 1) PROJ_MAX_SHIFT);
 * the later iterations walk that list and check every body's path length against moves x its shift; after an
 * iteration where one exceeds it, that scenario's iterations scan the ring again. moves >= 0 (infinity allowed),
 * default 12 (at the 63,424-agent bowl and at venue L almost no body moves that far in one step's projection). On the
 * plain and batched steps moves >= P needs no check (every iteration moves a body by at most one shift); a seated
 * venue's own step uses at most P - 1 moves and always checks (its fence and wall plane can push further). 0: no lists,
 * the ring every iteration, as physics.comp:1367-1418 does. Results are the same bit for bit for every budget. No
 * counterpart in the reference. */
int swarm_set_candidate_budget(swarm_ctx* c, float moves);
/* Lanes per slot in the projection's iterations over the candidate lists (1 default, 2, 4 or 8): the lanes load
 * and test a slot's candidates together and add the contact terms in candidate order, so results are the same bit
 * for bit. More lanes pay where slots are few and candidates many (4 lanes: 4,096 agents in the 40 x 30 m room,
 * 308 -> 289 us per step; 1 lane at the 63,424-agent bowl); swarm_autotune picks it. No counterpart in the
 * reference. */
int swarm_set_projection_lanes(swarm_ctx* c, int lanes);
int swarm_get_projection_lanes(swarm_ctx* c, int* lanes);
/* Lanes per slot (1, 2, 4 or 8; default 4) in the projection's first iteration, the ring scan that also collects
 * each slot's contact candidates: a slot's lanes take interleaved slots of each ring row, add the contact terms in
 * ring order and write the candidates at their rank in ring order, so results are the same bit for bit. No
 * counterpart in the reference. */
int swarm_set_projection_collect_lanes(swarm_ctx* c, int lanes);
/* The projection's contact candidates after the last step, in slot order: count[n] (NBR_OVERFLOW 0xFFFFFFFF: more
 * than 48, the ring scan is used instead), index and contact distance [48 x n], candidate-major (entry j of slot s at
 * j * n + s). Any pointer may be NULL. No counterpart in the reference (its projection rescans the ring). */
int swarm_read_candidates(swarm_ctx* c, uint32_t* count, uint32_t* index, float* dist);
/* The key sort (CpuCrowdSim::SortMode, cpu_crowd_sim.hpp:32-45): 0 (default) the stable counting sort, ascending
 * (cell, agent) (SortMode::Counting); 1 the reference GPU's bitonic network (bitonic_sort.comp,
 * bitonic_sort_shared.comp; SortMode::GpuBitonic), whose unstable within-cell order is the GPU reference's own and
 * changes the order of every neighbour sum. Mode 1 needs a power-of-two agent count (crowd_sim.cpp:99-100). */
int swarm_set_sort_mode(swarm_ctx* c, int mode);
/* Lanes per slot (1, 2, 4 or 8; default 4) in predict (PHYSICS_PASS 1, cpu_crowd_sim.cpp:655-777): a slot's lanes
 * take interleaved slots of each perception-ring row and add the separation terms in ring order (the headway gap is
 * a minimum, the same in any order); then every lane carries the same state and the wall contacts split their 16
 * feelers and two contact measurements over the lanes. Results are the same bit for bit. No counterpart in the
 * reference. */
int swarm_set_predict_lanes(swarm_ctx* c, int lanes);
/* The lanes per slot in force: the projection's list iterations, its first (collecting) iteration, predict. */
int swarm_get_lanes(swarm_ctx* c, int* list_lanes, int* collect_lanes, int* predict_lanes);
/* Times each candidate physics launch shape over 'steps' steps from the current state, then at the fastest shape
 * the lanes per slot of the projection's list iterations (1, 2, 4, 8), of predict and of the collecting iteration
 * (1, 2, 4 each), one setting at a time: every option twice, in interleaved rounds, each keeping its faster run;
 * keeps the fastest and restores the state; *best_ms: its mean ms per step. No counterpart in the reference. */
int swarm_autotune(swarm_ctx* c, int steps, double* best_ms);

/* ---- the safety metrics (class CrowdMetrics, include/swarm/metrics.hpp:181-227; plain contexts) ----
 * Write-only side outputs of each step: evacuations per exit (the floor's exit rectangles, then 'none'), each
 * agent's evacuation step, cloud-in-cell density and speed on a floor grid of 1 m cells in fixed point (one person =
 * 65536), and per cell the last, peak, occupied steps, steps at or above the LoS E / LoS F / crush threshold masses,
 * the 64-bit sums of mass and speed mass and the crush exposure. Counters: METRIC_COUNTER_WORDS = 68 words (step,
 * active, evacuated, activeAccum, exits[64]); cells: 16 words each (GpuMetricCell, gpu_types.hpp:127-137). */
/* On: every swarm_step is followed by the deposit and finalize passes (metrics_deposit.comp, metrics_finalize.comp;
 * CrowdMetrics::afterStep, metrics.hpp:197-199). The first enable makes the scene's grid (walkable areas sampled as
 * computeMetricVenue does) and clears it. */
int swarm_metrics_enable(swarm_ctx* c, int on);
int swarm_metrics_reset(swarm_ctx* c);                                       /* CrowdMetrics::reset (metrics.hpp:196) */
/* another grid: res_x x res_y cells, g = xy * cells_per_unit + off, 4 threshold masses per cell; then reset */
int swarm_metrics_configure(swarm_ctx* c, int res_x, int res_y, float cells_per_unit, float off_x, float off_y, const uint32_t* thresholds4);
int swarm_metrics_grid(swarm_ctx* c, int* res_x, int* res_y, float* cells_per_unit, float* off_x, float* off_y);
/* the two passes on the caller's states, n each, agent order (CrowdMetrics::record, metrics.hpp:201) */
int swarm_metrics_record(swarm_ctx* c, const float* px, const float* py, const float* pw, const float* cx, const float* cy, const float* cw,
                         const float* cvx, const float* cvy);
int swarm_read_metric_counters(swarm_ctx* c, uint32_t* out68);      /* readCounters (metrics.hpp:212) */
int swarm_read_metric_cells(swarm_ctx* c, uint32_t* out16);         /* readCells (metrics.hpp:214) */
int swarm_read_metric_evac_steps(swarm_ctx* c, uint32_t* out_n);    /* readEvacSteps (metrics.hpp:213) */
int swarm_read_metric_cell_info(swarm_ctx* c, float* walkable_m2, uint32_t* thresholds4);   /* MetricVenue (metrics.hpp:145-153) */

/* ---- the sort stage without allocation ----
 * No counterpart in the reference: a variant of swarm_sort_grid (CrowdSim::sortGrid) for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Every pointer is
 * DEVICE memory of the caller: keys[n] (a cell < cells, or 0xFFFFFFFF inactive; a value neither is not counted in
 * any cell), outputs sorted_keys[n] and ids[n] in ascending (key, index) order (swarm_sort_grid's), cell_start[cells]
 * and cell_count[cells] (swarm_read_cells'); any output may be NULL. stream: a hipStream_t, or NULL for the default
 * stream; the call is asynchronous on it, launches kernels only and allocates nothing: its working memory is one
 * static device array of SWARM_NOALLOC_SCRATCH_BYTES per GPU, held from the first call on that GPU to the end of the
 * process. Limits: n <= SWARM_NOALLOC_MAX_KEYS, cells <= SWARM_NOALLOC_MAX_CELLS (SWARM_EINVAL beyond). Calls that
 * overlap on one GPU (other streams, other threads) share that array and race: serialise them. */
#define SWARM_NOALLOC_MAX_KEYS 65536u
#define SWARM_NOALLOC_MAX_CELLS 262144u
#define SWARM_NOALLOC_SCRATCH_BYTES 2886912u
int swarm_sort_keys_noalloc(const uint32_t* keys, uint32_t n, uint32_t cells, uint32_t* sorted_keys, uint32_t* ids,
                            uint32_t* cell_start, uint32_t* cell_count, void* stream);
/* the message of this thread's last failed swarm_sort_keys_noalloc */
const char* swarm_noalloc_last_error(void);

/* the context's HIP stream (hipStream_t), for events around swarm_step */
int swarm_get_stream(swarm_ctx* c, void** stream);
int swarm_num_cells(swarm_ctx* c);
void swarm_destroy(swarm_ctx* c);

/* The message of the last failed call on this thread */
const char* swarm_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
