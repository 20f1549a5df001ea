// Scene preparation for the per-step pipeline: the input tables the step reads, built once on the
// host. Rectangle worlds (flat floor, axis-aligned wall boxes, exit rectangles, stair bands) are
// turned into
//   - a 2D FP32 signed-distance grid of the walls (the step's wall contacts and obstacle avoidance),
//   - the navigation grid and one any-angle distance field per exit (Lazy Theta* over legal moves,
//     following src/nav_field.cpp:184-366 of the reference), merged into a per-cell candidate table
//     (following src/exit_choice.cpp:143-317),
//   - a stair speed-factor field on the navigation grid,
// plus seeded agent placements for the benchmark floors.
#pragma once
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "seated_venue.h"

namespace swarm {

// Closed axis-aligned rectangle (RectGeometry's Rect, include/swarm/nav_field.hpp:135-138)
struct Rect {
    float lox, loy, hix, hiy;
    bool contains(float x, float y) const { return x >= lox && x <= hix && y >= loy && y <= hiy; }
};

// A navigation grid: res cells per axis from origin, cell edge 'cell', invCell = 1 / cell rounded once
// (nav::Grid, nav_field.hpp:74-84, generalised to a rectangle)
struct NavGrid {
    int resX = 0, resY = 0;
    float originX = 0.0f, originY = 0.0f;
    float cell = 0.0f, invCell = 0.0f;
    size_t cells() const { return size_t(resX) * size_t(resY); }
    int index(int ix, int iy) const { return ix + iy * resX; }
    void centre(int ix, int iy, float& x, float& y) const;   // origin + (i + 0.5) cell, in double, rounded once
};

struct SceneDesc {
    // domain: agents outside it are parked
    float xmin = 0, ymin = 0, xmax = 0, ymax = 0;
    float hashCell = 0.016f;     // spatial-hash cell (model units)
    float sdfSpacing = 0.0032f;  // SDF texel pitch; texel centres at xmin + (i + 0.5) pitch
    float sdfClamp = 0.05f;      // SDF values are clamped to this (well beyond every reach the step uses)
    NavGrid nav;                 // navigation grid (cell, origin, res)
    float navBody = 0.008f;      // a FIT cell's centre clears every wall by this
    double metresPerUnit = 31.25;
    std::vector<Rect> walls, exits, stairs;
    std::vector<float> stairFactor;   // per stair rect
    std::vector<float> exitWidthM;    // per exit, metres (0: the exit rect's longer side)
    // a seated venue's plan: its layout, for the incline rules of the step (the free speed and the headway time gap by the
    // floor's slope along e_goal on the gangway inclines)
    bool hasVenue = false;
    seated::Layout venue{};
};

// The tables the step reads (uploaded to the GPU as they are)
struct Scene {
    // domain and spatial hash (calc_grid.comp: cell = clamp(int((p - min) / span * res)))
    float xmin, ymin, xmax, ymax;
    float hashSpanX, hashSpanY;
    int hashNX, hashNY;
    int percRing, projRing;
    // signed-distance grid, row-major [sdfNY][sdfNX]; texel (i, j) at (sdfOX + i h, sdfOY + j h)
    int sdfNX, sdfNY;
    float sdfOX, sdfOY, sdfH, sdfInvH;
    std::vector<float> sdf;
    std::vector<float> sdfMin;   // per texel: min over [i - win, i + win + 1]^2 (a conservative 'no wall near' test)
    int sdfMinWin;
    // navigation / exit-choice table, dense per nav cell
    NavGrid nav;
    int K = 0, E = 0;                      // candidate slots per cell, exits
    std::vector<uint32_t> navInfo;         // NAV_FIT | NAV_HAS
    std::vector<uint32_t> navSlot;         // [cell * K + k]: anchor | exit << 25 (exit 63: empty)
    std::vector<uint16_t> navDist;         // [cell * K + k]: distance quanta (0xFFFF: empty)
    std::vector<float> invCapS;            // per exit: 60 / (w_e J)
    float secondsPerUnit = 0.0f, distQuantum = 0.0f;
    std::vector<float> stair;              // per nav cell: free-speed factor
    std::vector<Rect> exits;
    // the wall boxes (the exact wall query of e_goal's wall rule) and their index on the hash grid: per hash cell,
    // ascending, every wall within FEELER_REACH_RADII r_max + half a cell of the cell's rectangle
    std::vector<Rect> walls;
    std::vector<uint32_t> wallBinStart;    // hashNX * hashNY + 1
    std::vector<uint32_t> wallBinList;
    float maxRadius = 0.0f, maxV0 = 0.0f;  // used for the rings and sdfMin window
    // the safety metrics' floor grid (metrics.hpp:49-67, generalised to a rectangle off the origin): 1 m cells,
    // g = xy * metricCellsPerUnit + metricOff per axis is the position in cells from the centre of cell 0;
    // per cell its walkable floor area (m^2), sampled on 40 x 40 points as computeMetricVenue samples
    // (src/metrics.cpp:150-165): inside the domain and outside every wall
    int metricResX = 0, metricResY = 0;
    float metricCellsPerUnit = 0.0f, metricOffX = 0.0f, metricOffY = 0.0f;
    std::vector<float> metricWalkableM2;
    double buildSeconds = 0.0;
    int fromCache = 0;                     // the tables came from the scene cache
    int hasVenue = 0;                      // a seated venue's plan: the layout of its incline rules
    seated::Layout venue{};
    // the venue's own 3D step (buildSeatedScene): the heightfield is the scene (no SDF grid, walls or exit rectangles);
    // agents leave when max(|x|, |y|) passes exitDepth below exitMaxZ, or domainMax
    int venue3d = 0;
    float exitDepth = 0.0f, exitMaxZ = 0.0f, domainMax = 0.0f;
};

// Build every table of a scene. maxRadius / maxV0: the largest agent radius and free speed the scene
// will carry (they size the projection ring and the SDF-min window).
Scene buildScene(const SceneDesc& d, float maxRadius, float maxV0, int threads = 0);

// --- the checksummed scene cache (the reference's navigation and choice caches, src/nav_field.cpp:468-688,
// src/exit_choice.cpp:335-445) ---
uint64_t sceneCacheKey(const SceneDesc& d, float maxRadius, float maxV0);
std::string sceneCacheFile(const std::string& dir, uint64_t key);
bool saveScene(const Scene& s, uint64_t key, const std::string& file);
bool loadScene(Scene& s, uint64_t key, const std::string& file);   // false: absent, foreign or damaged
// buildScene through the cache in 'dir' (empty: no cache): a valid file is loaded, else the scene is built and saved
Scene buildSceneCached(const SceneDesc& d, float maxRadius, float maxV0, int threads, const std::string& dir, int* fromCache);

// --- the navigation builder, exposed for known-answer tests ---
enum NavClass : uint8_t { NAV_CLASS_FIT = 0, NAV_CLASS_NOFIT = 1 };
struct NavField {
    NavGrid grid;
    std::vector<uint8_t> cls, out;
    std::vector<int8_t> exitRegion;
    int exitCount = 0;
};
NavField buildNavGrid(const SceneDesc& d, int threads = 0);
// nav::buildGrid on a seated venue's heightfield (SeatedGeometry, src/nav_field.cpp:55-101): per cell its floor, its
// class (the fence band, the exact walls at the body's top), its exit region (past the outer face: the avenue, or a door
// up to its landing's end) and its legal moves (no fence crossing, no drop or rise beyond MAX_STEP_UP, the step-up scan,
// rows entered only from the same row). floorZ (may be null): the floors
NavField buildSeatedNavGrid(const seated::Layout& L, const NavGrid& g, int threads = 0, std::vector<float>* floorZ = nullptr);
// The candidate table of a navigation field (buildChoiceTable, src/exit_choice.cpp:143-317): widthM the exits' clear
// widths (metres), exitType 0 / 1 per exit (two types: the type rule)
void buildCandidateTable(const NavField& f, const std::vector<double>& widthM, const std::vector<int>& exitType, double metresPerUnit,
                         int threads, Scene& s);
// One exit's field: dist (grid units; +inf: no path) and anchor (-1: none); exitFilter -1: every exit
void lazyThetaStar(const NavField& f, int exitFilter, std::vector<double>& g, std::vector<int32_t>& par,
                   std::vector<uint8_t>& closed);

// --- agent placements of the benchmark floors (model units) ---
struct Agents {
    std::vector<float> x, y, r, v0, w;   // w: 1 active, 0 parked (x, y = PARK_COORD)
    size_t size() const { return x.size(); }
};

// splitmix64: a portable, seeded generator (doubles in [0, 1) from the top 53 bits)
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next();
    double uniform();                 // [0, 1)
    double normal();                  // Box-Muller, cached pair not kept (one draw each call)
};

// Layout ids of the benchmark floors
enum Layout { LAYOUT_ROOM = 0, LAYOUT_BOWL = 1, LAYOUT_CORRIDOR = 2, LAYOUT_BOWL_EIGHTH = 3, LAYOUT_PILLAR = 4, LAYOUT_BOWL_QUARTER = 5 };
SceneDesc layoutDesc(int layout);
// Seat / lattice slots of a layout (agent centres in model units) in their fixed order
void layoutSlots(int layout, std::vector<float>& sx, std::vector<float>& sy);
// Agents on a layout: 'active' of its slots chosen (room-like layouts: seeded random choice of lattice
// sites with jitter; bowls: the first 'active' seats in slot order, jittered), padded with parked agents to
// 'total'. Radii U[0.20, 0.25] m, free speed N(1.34, 0.26) m/s clipped to [0.5, 2.0].
Agents placeAgents(int layout, size_t active, size_t total, uint64_t seed, double speedMean = 1.34, double speedSd = 0.26);

// --- spawners (spawn.cpp) ---
// The reference's seat spawn (seatedSpawnSeats, src/crowd_sim.cpp:652-674): the chosen seat per agent (-1: parked)
std::vector<int> spawnSeats(int seatCount, int numAgents, double load, uint32_t seed);
// Gauss-Seidel relaxation of overlapping discs in agent order (relaxSpawn, src/crowd_sim.cpp:544-597, with radii r_i):
// clampTo(i, x, y) keeps agent i on its band; returns the sweeps run, maxOverlap the last sweep's largest overlap
int relaxSpawnDiscs(std::vector<float>& x, std::vector<float>& y, const std::vector<float>& r,
                    const std::function<void(size_t, float&, float&)>& clampTo, float& maxOverlap);
// An open floor (room, corridor, pillar room): uniform placement with rejection, then the relaxation
Agents spawnUniform(int layout, size_t active, size_t total, uint32_t seed, double speedMean, double speedSd, int* sweeps,
                    float* maxOverlap);

} // namespace swarm
