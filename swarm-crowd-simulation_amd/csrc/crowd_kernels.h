// Buffer views and launchers of the per-step kernels (crowd_kernels.hip). The kernels serve the plain step
// (one scene) and the batched step (many scenes of uneven agent counts in one launch per pass): the batched
// instances read each slot's scene from a per-scenario table.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/swarm_adjoint.h"
#include "crowd_device.h"

namespace swarm {

constexpr int HASH_BLOCK = 256;    // agents per block of the grid passes (also the inactive bucket's span)
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
constexpr int PHYS_BLOCK = 128;    // slots per block of the physics passes (default)
constexpr int PHYS_MAX_BLOCK = 256;
constexpr int XCDS = 8;            // accelerator complex dies of an MI355X (one L2 each)
constexpr int NBR_CAP = 48;        // contact candidates per slot
constexpr int PROJ_CHUNK = 8;      // candidates loaded together by one projection thread
constexpr int PROJ_CHUNK_ENSEMBLE = 4;   // the same in the batched step's list iterations (more waves per CU)
constexpr int PROJ_PRUNE_EVERY = 8; // projection iterations between candidate prunings (0: never)
// The plain and batched steps' candidate budget, in maximal moves per step (the lists' reach assumes no body moves
// farther over the projection; checked every iteration). At the 63,424-agent bowl 99.99 % of the bodies' projection
// paths stay under 2.75 moves, and a path past 12 comes about once in 200 steps (the oracle, 400 steps).
constexpr float PROJ_LIST_MOVES = 12.0f;
constexpr int PROJ_LANES = 1;       // lanes per slot in the projection iterations that do not prune (autotuned)
constexpr int PROJ_COLLECT_LANES = 4;   // lanes per slot in the first projection iteration (ring scan + candidates)
constexpr int PRED_LANES = 4;           // lanes per slot in predict (its perception ring and wall contacts)
constexpr int ADJ_GRAPH_MAX_STEPS = 32; // adjoint segments up to this long replay and run backwards as captured graphs
constexpr int BITONIC_TILE = 2048;      // pairs a block sorts in LDS in the bitonic sort mode (16 KB)
constexpr int BITONIC_BLOCK = 1024;
constexpr int ADJ_CHUNK = 4;            // contacts whose loads the adjoint projection issues together
constexpr size_t ADJ_TAPE_BUDGET = size_t(4) << 30;   // bytes of tape a run may keep instead of recomputing segments
constexpr size_t ADJ_DUAL_LIMIT = size_t(16) << 30;   // two segment tapes up to this many bytes (the pipelined backward)
// ring slots loaded together by one predict thread (same-box against 4: the 256-scenario ensemble 1.296 -> 1.279 ms
// per step, the plain bowl unchanged; 8 slower in both)
constexpr int PRED_CHUNK = 2;
constexpr uint32_t NBR_OVERFLOW = 0xFFFFFFFFu;

// control words on the device
constexpr int CTRL_STEP = 0;            // steps since the last upload (the exit choice's step counter)
constexpr int CTRL_NACTIVE = 1;         // active agents at the current step's input
constexpr int CTRL_ACTIVE_SUM_LO = 2;   // running sum of active agents over steps (uint64, low word first)
constexpr int CTRL_DONE_BLOCKS = 4;     // blocks of the final pass that have finished (reset by the last one)
constexpr int CTRL_LIST_BROKEN = 5;     // a seated step's agent moved past its candidate lists' budget (cleared by predict)
constexpr int CTRL_WORDS = 8;
// The counting sort ranks an agent within its cell by scanning the cell's slots: at most this many (a step's cells
// hold a few hard discs; swarm_sort_grid's caller keys beyond it take the stable network instead)
constexpr uint32_t FIX_RANK_MAX = 64;
// choice counters
constexpr int CNT_SWITCHES = 0, CNT_FORCED = 1, CNT_EVALUATIONS = 2, CNT_ELIGIBLE = 3;
// per-scenario strides of the exit-choice and evacuation state
constexpr int QUEUE_STRIDE = 128, COUNTER_STRIDE = 4, THRESHOLD_STRIDE = 64, PER_EXIT_STRIDE = 64;

// agent-order state, one half of the ping-pong
struct AgentHalf {
    float* x; float* y; float* vx; float* vy; float* spd; float* w;
    float* z; float* vz;   // heights and vertical velocities (a seated venue's own 3D step; unused elsewhere)
};

struct GridBuffers {
    uint32_t* key;      // agent order: global cell or INACTIVE_CELL
    uint32_t* loc;      // agent order: offset within its bucket (atomic order)
    uint32_t* bcnt;     // per bucket: count (cleared by the scan)
    uint32_t* bstart;   // per bucket: exclusive start (+ sentinel n); cells first: a cell's slots are [bstart[c], bstart[c + 1])
    uint32_t* tileSum;
    uint32_t* tmp;      // slot order, atomic order within buckets
    uint64_t* pairs;    // the bitonic sort mode's (key << 32 | agent) pairs, n
};

// the batched step's per-scenario tables (unused by the plain step)
struct BatchView {
    const SceneView* views;      // per layout
    const uint32_t* scenLayout;  // per scenario
    const uint32_t* cellBase;    // per scenario: its first global cell
    const uint32_t* agentBase;   // per scenario: its first global agent
    const uint32_t* agentScen;   // per agent
    uint32_t* slotScen;          // per slot (written by the gather)
};

// slot-order step buffers
struct StepBuffers {
    uint32_t* sid; uint32_t* skey;
    const uint32_t* bstart;                 // the bucket starts (cell ranges)
    float* sx; float* sy; float* sr;        // start-of-step positions and radii
    float* egx; float* egy;                 // e_goal
    float* pw; float* pvx; float* pvy;      // predicted: active flag, velocity
    float4* planes;                         // wall plane (normal, offset, flags)
    const float* avx; const float* avy;     // agent order, start of step
    const float* v0;                        // agent order
    // the projection's contact candidates, slot-interleaved: nbr[j * n + slot] (a pair's contact distance is the radii's
    // sum, formed where it is used); nbrD[j * n + slot]: on a seated venue's own step, the candidate's height offset;
    // nbrCount[slot] (NBR_OVERFLOW: more than NBR_CAP, the slot scans its ring instead)
    uint32_t* nbr; float* nbrD; uint32_t* nbrCount;
    // the step's control words; ctrl[CTRL_NACTIVE] is the active slot count: the inactive key sorts last, so
    // slots [ctrl[CTRL_NACTIVE], n) are the inactive ones
    const uint32_t* ctrl;
    // The candidate lists' displacement budget, in maximal moves (PROJ_MAX_SHIFT_RADII r each): > 0, the lists hold the
    // candidates within D + listMoves (shift_i + shift_j) and every iteration checks its slots' path lengths against
    // listMoves shift_i; a slot past it marks its scenario's listBroken word, and that scenario's later iterations scan
    // the ring. < 0: the lists hold every candidate within reach over all P iterations (no checks needed); 0: no lists,
    // the ring every iteration. listBroken (per scenario, nScen words) is set by predict: 0, or 1 for listMoves == 0.
    float listMoves;
    uint32_t* listBroken;
    int nScen;
};

// the projection's slot-order ping-pong
struct ProjBuffers {
    float2* p; float* acc;
};

struct ChoiceBuffers {
    uint32_t* assign;     // agent order
    uint32_t* pending;    // agent order: eligible | best << 16 | bin
    uint32_t* queue;      // per scenario QUEUE_STRIDE: [parity * 64 + exit]
    uint32_t* counters;   // per scenario: switches, forced, evaluations, eligible
    uint32_t* hist;       // per scenario: exitStride x 256
    uint32_t* threshold;  // per scenario THRESHOLD_STRIDE
    uint32_t* ahead;      // per scenario: 2 x exitStride x 1024
    int exitStride;       // exits per parity of 'ahead' and per scenario of 'hist'
};

struct EvacBuffers {
    int32_t* step;        // agent order: the step the agent left at (-1)
    int32_t* exit;        // agent order: its exit (E: out of the domain)
    uint32_t* perExit;    // per scenario PER_EXIT_STRIDE counters
    float2* inside;       // agent order: its position at that step's input (inside the exit)
    float2* before;       // agent order: its position one step earlier
};

// One step's tape of the adjoint's forward (swarm_adjoint.h's records), slot order
struct TapeStep {
    swarm_predict_rec* pred;
    swarm_final_rec* fin;
    uint64_t* mask;            // P x n: the contacts of iteration it among the slot's candidates, mask[(it - 1) * n + slot]
    uint64_t* capBits;         // bit it - 1: iteration it's correction was capped
    uint64_t* planeBits;       // bit it - 1: iteration it's wall plane applied
    uint64_t* overlap;         // the candidates overlapping at the step's end
    const uint32_t* assign;    // agent order: this step's exits
    int efixed;                // e_goal is a drive direction (or zero): no goal anchor
    uint32_t* err;             // bit 0: candidates overflowed; bit 1: more wall pushes than recorded
};

// records the wall pushes and the sweep of one resolveStatic into a predict or final record
struct WallTape {
    static constexpr bool on = true;
    uint32_t* flags;
    uint32_t* nev;
    swarm_wall_event* ev;
    float* sw;                 // sweep lo, dx, dy
    uint32_t* err;
    __device__ void push(V2 a, const ContactRec& c, bool vel) {
        const uint32_t k = *nev;
        if (k >= uint32_t(SWARM_TAPE_MAX_WALL)) { atomicOr(err, 2u); return; }
        ev[k] = swarm_wall_event{ a.x, a.y, c.u.x, c.u.y, c.g.x, c.g.y, c.x0.x, c.x0.y, c.t0, vel ? 1u : 0u };
        *nev = k + 1u;
    }
    __device__ void sweep(float lo, V2 d, bool vel) {
        sw[0] = lo; sw[1] = d.x; sw[2] = d.y;
        *flags |= SWARM_REC_SWEEP | (vel ? SWARM_REC_SWEEP_VEL : 0u);
    }
};

// XCD-aware block numbering for the slot-order kernels. The GPU hands consecutive workgroups to its XCDs in
// turn and each XCD has its own L2; renumbered, XCD x takes a contiguous 1/XCDS of the slots, so neighbours
// (slots are sorted by cell) are mostly read through one L2. A bijection on [0, gridDim.x).
__device__ inline int xcdBlock() {
    const int g = int(gridDim.x), b = int(blockIdx.x), q = g / XCDS;
    return b < q * XCDS ? (b % XCDS) * q + b / XCDS : b;
}

// the scene of an agent / slot: plain instances use the kernel's scene and no offsets
struct ScnRef {
    uint32_t sc, cellBase, agentBase;
};
template <bool B> __device__ inline const SceneView& sceneOf(const SceneView& s0, const BatchView& bt, uint32_t sc) {
    if constexpr (B) return bt.views[bt.scenLayout[sc]];
    else return s0;
}
template <bool B> __device__ inline ScnRef scnOfAgent(const BatchView& bt, uint32_t agent) {
    if constexpr (B) { const uint32_t sc = bt.agentScen[agent]; return ScnRef{ sc, bt.cellBase[sc], bt.agentBase[sc] }; }
    else return ScnRef{ 0u, 0u, 0u };
}
template <bool B> __device__ inline ScnRef scnOfSlot(const BatchView& bt, uint32_t slot) {
    if constexpr (B) { const uint32_t sc = bt.slotScen[slot]; return ScnRef{ sc, bt.cellBase[sc], bt.agentBase[sc] }; }
    else return ScnRef{ 0u, 0u, 0u };
}

// the plain step: 'bt' unused; the batched step: s unused, scenarios = the number of scenarios
struct StepLaunch {
    int n = 0;
    int scenarios = 1;
    bool batched = false;
    SceneView s{};
    BatchView bt{};
    ModelParams mp{};
    int maxExits = 0;     // the largest exit count over the scenes
    uint32_t cells = 0;   // global cells
    int physBlock = 0;    // threads per block of the physics passes (0: PHYS_BLOCK)
    int physGrid = 0;     // blocks of the physics passes (0: one slot per thread; else a persistent grid)
    int pruneEvery = 0;   // > 0: the projection drops candidates out of reach every that many iterations
    int projLanes = 1;    // lanes per slot in the projection iterations that do not prune (1, 2, 4 or 8)
    int collectLanes = 1; // lanes per slot in the first projection iteration (1, 2, 4 or 8)
    int predictLanes = 1; // lanes per slot in predict (1, 2, 4 or 8)
    bool bitonic = false; // the sort: the reference GPU's bitonic network instead of the stable counting sort
    float listMoves = 0.0f;   // a seated venue's own step: its candidate lists' budget in maximal moves (0: no lists)
};

void launchGridPasses(hipStream_t st, const StepLaunch& L, AgentHalf rd, const GridBuffers& g, StepBuffers& b, const float* rad,
                      uint32_t* ctrl);
// the sort stage alone on the keys already in g.key (CrowdSim::sortGrid)
void launchSortKeys(hipStream_t st, const StepLaunch& L, AgentHalf rd, const GridBuffers& g, StepBuffers& b, const float* rad,
                    uint32_t* ctrl);
void launchChoice(hipStream_t st, const StepLaunch& L, AgentHalf rd, const ChoiceBuffers& c, StepBuffers& b, const float* driveX,
                  const float* driveY, uint32_t* ctrl);
// stamps (optional, 2 events): recorded after predict and after the projection iterations before the final pass
// stamps: events after predict and after the projection; iterStamps: a pair around each projection iteration after
// the first (2 (P - 2) events)
void launchPhysics(hipStream_t st, const StepLaunch& L, StepBuffers& b, ProjBuffers pa, ProjBuffers pb, AgentHalf rd, AgentHalf wr,
                   EvacBuffers ev, uint32_t* ctrl, hipEvent_t* stamps = nullptr, hipEvent_t* iterStamps = nullptr);
// The physics passes of the adjoint's taped forward: positions of every iteration into q ((P + 1) x n: q[0]
// predict's output, q[it] iteration it's), the accumulator through acc0 / acc1, the records into tp
void launchPhysicsRecord(hipStream_t st, const StepLaunch& L, StepBuffers& b, float2* q, float* acc0, float* acc1, AgentHalf rd,
                         AgentHalf wr, EvacBuffers ev, uint32_t* ctrl, const TapeStep& tp);
// The adjoint's loss terms: the overlap term of the step just run (weight * (D - d)^2 per overlapping candidate
// pair, tp: also the overlap masks) and the exit term of the state s; both added to *loss
void launchOverlap(hipStream_t st, const StepLaunch& L, StepBuffers& b, AgentHalf wr, float weight, double* loss, const TapeStep* tp);
void launchExitLoss(hipStream_t st, const StepLaunch& L, AgentHalf s, const uint32_t* assign, swarm_exit_rec* out, double* loss);
void launchStepIncrement(hipStream_t st, uint32_t* ctrl);
// count arrays of n 32-bit words, src[a] -> dst[a], in one launch
constexpr int TAPE_COPY_MAX = 8;
struct TapeCopy {
    const float* src[TAPE_COPY_MAX];
    float* dst[TAPE_COPY_MAX];
    int count;
};
void launchTapeCopy(hipStream_t st, int n, const TapeCopy& c);
void launchCopyProbe(hipStream_t st, const void* src, void* dst, size_t bytes);
// The stable sort of n caller keys (cells < C, or INACTIVE_CELL) into ascending (key, index) order on the bitonic
// network with whole 64-bit compares (a total order: no rank loops, O(n log^2 n) whatever the keys), and the cell
// table from the counting pass. Scratch: loc n, bcnt keySortBuckets, bstart keySortBuckets (+ 1), tileSum
// keySortTiles, pairs keySortPow2(n), ctrl CTRL_WORDS. Any output may be nullptr.
struct KeySortScratch {
    uint32_t* loc; uint32_t* bcnt; uint32_t* bstart; uint32_t* tileSum; uint64_t* pairs; uint32_t* ctrl;
};
uint32_t keySortPow2(uint32_t n);
uint32_t keySortBuckets(uint32_t n, uint32_t C);
uint32_t keySortTiles(uint32_t n, uint32_t C);
void launchStableKeySort(hipStream_t st, uint32_t n, uint32_t C, const uint32_t* key, const KeySortScratch& s, uint32_t* sortedKeys,
                         uint32_t* ids, uint32_t* cellStart, uint32_t* cellCount);
// slot order from sorted (key << 32 | agent) pairs: ids, keys, and the state gathered (as the step's gather)
void launchSortedGather(hipStream_t st, const StepLaunch& L, AgentHalf rd, const uint64_t* pairs, StepBuffers& b, const float* rad);
// The navigation lookups of the step at caller points (the reference's probe dispatch of physics.comp's nav section,
// tools/test_gpu.cpp:4155-4191): choiceCell, choiceDirection and e_goal with the wall rule for exit ex[i], radius r[i]
void launchNavProbe(hipStream_t st, const SceneView& s, int n, const float* x, const float* y, const float* r, const uint32_t* ex,
                    int32_t* cell, float* dir, float* goal);

} // namespace swarm
