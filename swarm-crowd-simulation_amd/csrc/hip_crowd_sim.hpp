// HipCrowdSim: the host side of the GPU step, with the surface of the reference's CrowdSim
// (include/swarm/crowd_sim.hpp:53-194) and CpuCrowdSim (include/swarm/cpu_crowd_sim.hpp:80-134):
// construct from a scene, upload a state, step, read the state, grid and cell table back. It owns every
// device buffer and one HIP stream; readbacks synchronise that stream only.
//
// A batch: many independent scenarios (each one of a few scene layouts, with its own agents) stepped together,
// one launch per pass for all of them. Agents are concatenated scenario by scenario; cell keys are global cell
// indices (scenario-major), so the counting sort and the cell ranges are segmented by scenario; every scenario
// keeps its own exit-choice state and evacuation counters; results are each scenario's plain results bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "adjoint.h"
#include "crowd_kernels.h"
#include "metrics.h"
#include "scene.hpp"
#include "seated_kernels.h"

namespace swarm {

struct SimOptions {
    float tauS = TAU_RELAX_S;       // relaxation time, seconds
    float obstacleGain = OBSTACLE_GAIN;
    int projIters = PROJ_ITERATIONS;
    bool useGraph = true;           // capture the step (two steps: both ping-pong parities) in a hipGraph
};

// tau (model time) and exp(-DT / tau), folded from the decimal values in double and rounded once
void modelParamsFor(float tauS, float& tau, float& relaxStep);

struct DeviceScene {
    SceneView view{};
    std::vector<void*> owned;
    void upload(const Scene& s);
    void release();
};

class HipCrowdSim {
public:
    // the plain step of one scene
    HipCrowdSim(const Scene& scene, int numAgents, const SimOptions& opt);
    // a batch: scenario s runs scene layouts[scenLayout[s]] with scenAgents[s] agents
    HipCrowdSim(const std::vector<const Scene*>& layouts, const std::vector<int>& scenLayout, const std::vector<int>& scenAgents,
                const SimOptions& opt);
    ~HipCrowdSim();
    HipCrowdSim(const HipCrowdSim&) = delete;
    HipCrowdSim& operator=(const HipCrowdSim&) = delete;

    // A full state (both halves); resets the exit choice, the evacuation record and the step counter (CrowdSim::upload)
    void upload(const float* x, const float* y, const float* vx, const float* vy, const float* w, const float* r, const float* v0);
    // The agent state only (both halves), as the reference's benchmark snapshot restore: the exit choice runs on
    void restoreAgentState(const float* x, const float* y, const float* vx, const float* vy, const float* w);
    void setDrive(const float* dx, const float* dy);   // nullptr: the navigation table (CpuCrowdSim::setDriveTable); plain only
    void step(int nsteps);
    // step(nsteps) between two HIP events on the context's stream; returns their elapsed ms (synchronises)
    double timeSteps(int nsteps);
    // The exit choice pass alone on the current state (CrowdSim::stepExitChoice); advances the step counter
    void stepExitChoice();
    // Sets the exit choice's state of scenario 0 (tools and tests: the reference's choice harness setters,
    // tools/choice_check.hpp:88-100): step counter, assigned exits, queue counters (128), queue-ahead histograms
    // (2 x 64 x 1024); nullptr arrays are left as they are
    void setChoiceState(uint32_t step, const uint32_t* assign, const uint32_t* queue128, const uint32_t* ahead);
    void sync();
    void setProjectionIterations(int iterations);
    // Launch shape of the physics passes: threads per block, blocks (0: one slot per thread; else a persistent grid)
    void setPhysicsShape(int block, int grid);
    int physicsBlock() const { return physBlock_; }
    int physicsGrid() const { return physGrid_; }
    // Candidate pruning in the projection every 'every' iterations (0: off); results are the same bit for bit
    void setPruneEvery(int every);
    // The displacement budget of the projection's candidate lists, in maximal moves over the step (default 12; a seated
    // venue's own step at most P - 1); 0: no lists, the ring every iteration; results are the same bit for bit
    void setCandidateBudget(float moves);
    // the plain and batched steps' budget for this step (StepBuffers::listMoves)
    float stepListMoves() const;
    // Lanes per slot in the projection's non-pruning iterations (1, 2, 4, 8); results are the same bit for bit
    void setProjectionLanes(int lanes);
    int projectionLanes() const { return projLanes_; }
    void setCollectLanes(int lanes);
    void setPredictLanes(int lanes);
    void setSortMode(int mode);   // 0: stable counting sort (SortMode::Counting), 1: the GPU's bitonic network
    int predictLanes() const { return predictLanes_; }
    int collectLanes() const { return collectLanes_; }
    // Times every candidate launch shape of the physics passes over 'steps' steps from the current state, then the
    // projection's lanes per slot at the fastest shape; keeps the fastest and puts the state back as it was (a
    // slot's result depends on neither). Returns the fastest setting's mean ms per step.
    double autotune(int steps);

    // Per-pass GPU timing (CrowdSim::setTimingEnabled / readTimings, crowd_sim.hpp:141-149): while enabled every
    // step runs eagerly with HIP events between its passes on the context's stream. readTimings: the mean ms per
    // step of the grid passes (hash, sort, cell ranges, gather), exit choice, predict, projection (all
    // PROJ_ITERATIONS - 1 launches), final, total, and the mean ms of one projection launch; returns the steps timed
    void setTimingEnabled(bool on, bool launches = false);   // launches: also time each later projection iteration
    int readTimings(double* ms8);

    void readState(float* x, float* y, float* vx, float* vy, float* speed, float* w);
    // A seated venue's own 3D step (a scene of buildSeatedScene): the state with heights and vertical velocities, at the
    // reference's global radius and free speed (upload3 resets the exit choice and the evacuation record as upload does)
    void upload3(const float* x, const float* y, const float* z, const float* vx, const float* vy, const float* vz, const float* w);
    void readState3(float* x, float* y, float* z, float* vx, float* vy, float* vz, float* speed, float* w);
    bool seated3d() const { return venue3d_; }
    void readPrevState(float* x, float* y, float* vx, float* vy, float* speed, float* w);   // previousStateBuffer
    // The sort stage alone on the caller's keys (CrowdSim::sortGrid): n keys, each a cell (< numCells) or
    // INACTIVE_CELL; returns the sorted keys and the permutation (the grid readback then holds this sort)
    void sortGrid(const uint32_t* keys, uint32_t* sortedKeys, uint32_t* ids);
    void readGrid(uint32_t* key, uint32_t* id);
    void readCandidates(uint32_t* count, uint32_t* index, float* dist);   // slot order; NBR_CAP x n, candidate-major
    void readCells(uint32_t* start, uint32_t* count);       // the reference's cell table over the global cells
    void readAssign(uint32_t* assign);
    std::vector<uint32_t> queueCounts();                     // every scenario's queue counts (QUEUE_STRIDE words each)
    void readChoiceCounters(uint32_t* out132);               // scenario 0: queue[128], switches, forced, evaluations, eligible
    void readChoiceCountersAll(uint32_t* out);               // every scenario: 132 words each
    void readEvac(int32_t* step, int32_t* exitId, uint32_t* perExit);   // perExit: scenario 0's E + 1 counters
    void readEvacCounters(uint32_t* perExit64);              // every scenario: PER_EXIT_STRIDE words each
    void readControl(uint32_t* step, uint32_t* nActive, uint64_t* activeSum);
    // The step's navigation lookups at m caller points (scenario 0's scene): choiceCell, choiceDirection for exit ex[i]
    // (2 floats each) and e_goal with the wall rule at radius r[i] (2 floats each)
    void navProbe(int m, const float* x, const float* y, const float* r, const uint32_t* ex, int32_t* cell, float* dir, float* goal);

    // ---- the adjoint of K steps (swarm_adjoint.h) ----
    // Runs K steps from the current state with the loss terms, keeping the state every 'every' steps; returns the
    // loss. The radii of the run are radiusScale x the uploaded radii.
    double adjointForward(int K, int every, float radiusScale, float overlapWeight);
    double adjointForwardRun(int K, int every, float radiusScale, float overlapWeight);
    // dL/dv0 (n, agent order) and dL/d(tau_s, radius scale, wall gain) of the last adjointForward; per scenario in
    // a batch (globals: 3 per scenario). The state afterwards is the forward's final state, as the forward left it.
    // Throws NonFiniteGradient (after writing the outputs) when any gradient value is inf or NaN.
    void adjointBackward(float* gradV0, double* gradGlobals);
    void adjointBackwardRun(float* gradV0, double* gradGlobals);
    // one step's tape of the last adjointForward (recomputed from its checkpoint)
    void adjointTape(int step, const swarm_tape_view& out, swarm_exit_rec* exit);
    void adjointTapeRun(int step, const swarm_tape_view& out, swarm_exit_rec* exit);
    // the taped forward's candidate budget (StepBuffers::listMoves)
    float recListMoves() const;
    // the relaxation time (seconds) and wall gain of the next steps
    void setModel(float tauS, float obstacleGain);
    // Loss terms of the next adjointBackward beyond the built-in ones: position adjoints at evacuation crossings
    // (agent a: gin at the input of step step[a], where it is inside its exit; gout at the input of step
    // step[a] - 1; step < 1: none), whether the exit term counts, a truncation window (> 0: the state adjoint
    // is cleared every 'window' steps going backwards, so every path is at most that long; 0: the exact adjoint),
    // and 'local': each agent's own chain only (no transfer across agents through contacts, separation, headway or
    // overlaps), cut at every step where the agent touched another (its position is then held by the crowd): a
    // stable, biased estimator for horizons where the exact one explodes
    void setAdjointTerms(const int32_t* step, const float* gin, const float* gout, bool exitTerm, int window, bool local = false);
    // per agent: the step it left at, its exit, its position at that step's input and one step earlier (2 floats each)
    void readEvacPoints(int32_t* step, int32_t* exitId, float* inside, float* before);

    // ---- the safety metrics (CrowdMetrics, swarm/metrics.hpp:168-214; plain contexts) ----
    // On: every step is followed by the deposit and finalize passes (in the captured graph). The first enable
    // allocates the floor grid of the scene (1 m cells over its domain, thresholds from its walkable areas) and
    // clears it.
    void setMetricsEnabled(bool on);
    bool metricsEnabled() const { return metricsOn_; }
    // Zeroes every counter and accumulator and the metric step (CrowdMetrics::reset)
    void resetMetrics();
    // Another grid: resX x resY cells, g = xy * cellsPerUnit + off, the caller's 4 threshold masses per cell (then reset)
    void configureMetrics(int resX, int resY, float cellsPerUnit, float offX, float offY, const uint32_t* thresholds4);
    void metricGrid(int* resX, int* resY, float* cellsPerUnit, float* offX, float* offY) const;
    // The two passes on the caller's states (CrowdMetrics::record; agent order, n values each)
    void recordMetrics(const float* px, const float* py, const float* pw, const float* cx, const float* cy, const float* cw,
                       const float* cvx, const float* cvy);
    void readMetricCounters(uint32_t* out);    // METRIC_COUNTER_WORDS
    void readMetricCells(uint32_t* out);       // METRIC_CELL_WORDS per cell
    void readMetricEvacSteps(uint32_t* out);   // n
    void readMetricCellInfo(float* walkableM2, uint32_t* thresholds4);

    int numAgents() const { return n_; }
    int numCells() const { return C_; }
    int numScenarios() const { return S_; }
    int numExits() const { return scenes_[0].view.E; }
    hipStream_t stream() const { return stream_; }
    const SceneView& view() const { return scenes_[0].view; }

private:
    std::vector<uint32_t> aheadTotals(uint32_t parity);
    void init(const std::vector<const Scene*>& layouts, const std::vector<int>& scenLayout, const std::vector<int>& scenAgents, bool batched);
    void enqueueStep(int parity);
    void buildGraph();
    void dropGraph();
    void dropAdjointGraphs();
    void backwardSegment(int seg, int len, int buf);
    StepLaunch launchDesc() const;
    template <class T> T* alloc(size_t count);
    void collectTimings();
    // every buffer a step writes that outlives the step (agent halves, exit choice, evacuation, control)
    std::vector<std::pair<void*, size_t>> persistentState() const;
    // adjoint helpers
    void snapshotTo(char* dst);
    void restoreFrom(const char* src, hipStream_t st);
    void allocAdjoint(int every, bool dual);
    void tapedStep(int buf, int k, hipStream_t st);
    void tapeSegment(int seg);
    void recordSegment(int seg, int buf, hipStream_t st);
    int tapeBuf(int seg) const { return dual_ ? (seg & 1) : 0; }
    void checkTapeErrors();
public:
    size_t adjointBytes() const { return adjBytes_; }   // device memory the adjoint holds (tape, buffers, checkpoints)
    // a run whose whole tape fits this many bytes is taped in its forward and never recomputed
    void setTapeBudget(size_t bytes) { tapeBudget_ = bytes; }
    size_t tapeBytesPerStep() const;
    int adjointSegments() const { return adjSegs_; }
private:
    AdjStep adjStepView(int buf, int k) const;

    std::vector<DeviceScene> scenes_;
    SimOptions opt_;
    ModelParams mp_{};
    bool batched_ = false;
    bool venue3d_ = false;          // the scene is a seated venue's own 3D step
    SeatedSlots seatedSlots_{};
    int n_ = 0, C_ = 0, nb_ = 0, S_ = 1, maxExits_ = 0, exitStride_ = CHOICE_QUEUE_STRIDE;
    int physBlock_ = 0, physGrid_ = 0;
    int pruneEvery_ = PROJ_PRUNE_EVERY;
    float listMoves_ = PROJ_LIST_MOVES;
    bool recBudgetOff_ = false;     // a taped step broke the budget once: the adjoint tapes whole lists from then on
    int projLanes_ = PROJ_LANES;
    int collectLanes_ = PROJ_COLLECT_LANES;
    int predictLanes_ = PRED_LANES;
    bool bitonic_ = false;
    hipStream_t stream_ = nullptr;
    std::vector<void*> bufs_;
    AgentHalf half_[2]{};
    float* rad_ = nullptr; float* v0_ = nullptr;
    float* driveX_ = nullptr; float* driveY_ = nullptr;
    bool drive_ = false;
    GridBuffers grid_{};
    StepBuffers sb_{};
    ProjBuffers pa_{}, pb_{};
    ChoiceBuffers ch_{};
    EvacBuffers ev_{};
    BatchView bt_{};
    uint32_t* ctrl_ = nullptr;
    int parity_ = 0;   // the half holding the current state
    // the step's graphs (two steps, both parities), without and with the adjoint's loss terms
    hipGraph_t graph_[2]{};
    hipGraphExec_t graphExec_[2]{};
    float lossGraphOvw_ = 0.0f;          // the overlap weight captured in graph_[1]
    // the adjoint's graphs: a full segment's taped replay (by starting parity) and backward (by which half of the
    // state adjoint's ping-pong holds the output), for the current segment length and loss terms
    // (per tape buffer, then parity / half)
    hipGraph_t replayGraph_[2][2]{}, backGraph_[2][2]{};
    hipGraphExec_t replayExec_[2][2]{}, backExec_[2][2]{};
    int* stateBase_ = nullptr;           // device: the first state of the segment a captured backward runs
    float* adjG0_ = nullptr;             // adj_.gx as allocated (the ping-pong's reference half)
    // timing: events per stamp (start, sorted, choice, predict, projection, final) per recorded step
    bool timing_ = false;
    bool timeLaunches_ = false;          // timing mode 2: an event pair in each later projection iteration's dispatch
    std::vector<std::vector<hipEvent_t>> pendingStamps_;
    double timeSum_[8] = {};
    int timedSteps_ = 0;

    // the adjoint: loss terms in the step, checkpoints, a segment's tape, the backward's buffers
    std::vector<float> r0_;              // the uploaded radii
    bool lossOn_ = false;
    float overlapWeight_ = 0.0f;
    double* lossAcc_ = nullptr;          // device
    swarm_exit_rec* exitRec_ = nullptr;  // device, n
    float* finX_ = nullptr;              // device, n: the forward's final positions
    float* finY_ = nullptr;
    int32_t* seedStep_ = nullptr;        // device, n: evacuation-crossing seeds (setAdjointTerms)
    float2* seedIn_ = nullptr;
    float2* seedOut_ = nullptr;
    bool seedsOn_ = false, exitTermOn_ = true;
    int window_ = 0;
    bool local_ = false;
    int adjK_ = 0, adjEvery_ = 0, adjSegs_ = 0, adjCap_ = 0;
    int tapedSeg_[2] = { -1, -1 };      // the segment each tape buffer holds (-1: none)
    // two segment tapes: the backward of a segment runs on the context's stream while the replay of the segment before
    // it, from its checkpoint into the other tape, runs on stream2_ (the replays keep to stream2_, one after another)
    bool dual_ = false;
    hipStream_t stream2_ = nullptr;
    hipEvent_t evReplay_[2]{}, evBack_[2]{}, evMain_ = nullptr;
    int tapeP_ = 0;                      // the projection iterations the tape's mask and position arrays are sized for
    char* finalSnap_ = nullptr;          // device: the forward's final state, put back after a recompute
    int finalParity_ = 0;
    void restoreFinal();
    float adjScale_ = 1.0f;
    size_t snapBytes_ = 0;
    std::vector<char*> ckpt_;            // device snapshots, one per segment
    std::vector<int> ckptParity_;
    struct Tape {
        float *inx, *iny, *invx, *invy, *inw, *outx, *outy;
        uint32_t *sid, *skey, *bstart;
        float *sx, *sy, *sr, *egx, *egy, *pw, *pvx, *pvy;
        float4* planes;
        uint32_t* nbr;
        uint32_t* nbrCount;
        swarm_predict_rec* pred;
        swarm_final_rec* fin;
        uint64_t *mask, *capBits, *planeBits, *overlap;
        float2* q;
        float *acc0, *acc1;
        uint32_t* err;
    } tape_[2]{};
    AdjBufs adj_{};
    std::vector<void*> adjBufs_;         // the adjoint's own allocations
    size_t adjBytes_ = 0;                // their bytes, checkpoints included
    std::vector<void*> tapeBufs_;        // the tape's (freed when a longer segment needs a larger one)
    size_t tapeBytes_ = 0;
    size_t tapeBudget_ = ADJ_TAPE_BUDGET;

    // the safety metrics
    void allocMetrics();
    bool metricsOn_ = false;
    MetricGrid mgrid_{};
    MetricBuffers mbuf_{};
    std::vector<float> mWalkable_;       // the scene's walkable area per cell (host)
    std::vector<uint32_t> mThr_;         // 4 threshold masses per cell (host)
    float* mScratch_ = nullptr;          // recordMetrics: 8 arrays of n
    std::vector<void*> mBufs_;           // the metrics' own allocations
};

// An adjoint whose gradient overflowed FP32 (the outputs are written; the C ABI returns SWARM_ENONFINITE)
struct NonFiniteGradient : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// A taped step whose candidate lists broke their displacement budget (caught inside: the call runs again with whole
// lists)
struct TapeBudgetBroken : std::runtime_error {
    TapeBudgetBroken() : std::runtime_error("adjoint: a taped step's candidate lists broke their budget") {}
};

#define SWARM_HIP_CHECK(expr)                                                                              \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + \
                                                       " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

} // namespace swarm
