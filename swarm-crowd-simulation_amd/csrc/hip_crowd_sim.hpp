// HipCrowdSim: the host side of the GPU step, with the surface of the reference's CrowdSim
// (include/swarm/crowd_sim.hpp:53-194) and CpuCrowdSim (include/swarm/cpu_crowd_sim.hpp:80-134):
// construct from a scene, upload a state, step, read the state, grid and cell table back. It owns every
// device buffer and one HIP stream; readbacks synchronise that stream only.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "crowd_kernels.h"
#include "scene.hpp"

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
    HipCrowdSim(const Scene& scene, int numAgents, const SimOptions& opt);
    ~HipCrowdSim();
    HipCrowdSim(const HipCrowdSim&) = delete;
    HipCrowdSim& operator=(const HipCrowdSim&) = delete;

    // A full state (both halves); resets the exit choice, the evacuation record and the step counter (CrowdSim::upload)
    void upload(const float* x, const float* y, const float* vx, const float* vy, const float* w, const float* r, const float* v0);
    // The agent state only (both halves), as the reference's benchmark snapshot restore: the exit choice runs on
    void restoreAgentState(const float* x, const float* y, const float* vx, const float* vy, const float* w);
    void setDrive(const float* dx, const float* dy);   // nullptr: the navigation table (CpuCrowdSim::setDriveTable)
    void step(int nsteps);
    // step(nsteps) between two HIP events on the context's stream; returns their elapsed ms (synchronises)
    double timeSteps(int nsteps);
    // The exit choice pass alone on the current state (CrowdSim::stepExitChoice); advances the step counter
    void stepExitChoice();
    // Sets the exit choice's state (tools and tests: the reference's choice harness setters,
    // tools/choice_check.hpp:88-100): step counter, assigned exits, queue counters (128), queue-ahead histograms
    // (2 x 64 x 1024); nullptr arrays are left as they are
    void setChoiceState(uint32_t step, const uint32_t* assign, const uint32_t* queue128, const uint32_t* ahead);
    void sync();
    void setProjectionIterations(int iterations);

    void readState(float* x, float* y, float* vx, float* vy, float* speed, float* w);
    void readGrid(uint32_t* key, uint32_t* id);
    void readCells(uint32_t* start, uint32_t* count);
    void readAssign(uint32_t* assign);
    void readChoiceCounters(uint32_t* out132);
    void readEvac(int32_t* step, int32_t* exitId, uint32_t* perExit);
    void readControl(uint32_t* step, uint32_t* nActive, uint64_t* activeSum);

    // Per-pass GPU timing (CrowdSim::setTimingEnabled / readTimings, crowd_sim.hpp:141-149): while enabled every
    // step runs eagerly with HIP events between its passes on the context's stream. readTimings: the mean ms per
    // step of hash, sort (scan + scatter + in-bucket order + gather), exit choice, predict, projection (all
    // PROJ_ITERATIONS - 1 launches), final, total, and the mean ms of one projection launch; returns the steps timed
    void setTimingEnabled(bool on);
    int readTimings(double* ms8);

    int numAgents() const { return n_; }
    int numCells() const { return C_; }
    int numExits() const { return scene_.view.E; }
    hipStream_t stream() const { return stream_; }
    const SceneView& view() const { return scene_.view; }

private:
    void enqueueStep(int parity);
    void buildGraph();
    void dropGraph();
    template <class T> T* alloc(size_t count);

    DeviceScene scene_;
    SimOptions opt_;
    ModelParams mp_{};
    int n_ = 0, C_ = 0, nb_ = 0;
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
    uint32_t* ctrl_ = nullptr;
    int parity_ = 0;   // the half holding the current state
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graphExec_ = nullptr;
    // timing: events per stamp (hash, sort, choice, predict, project, final, end) per recorded step
    bool timing_ = false;
    std::vector<hipEvent_t> events_;
    std::vector<std::vector<hipEvent_t>> pendingStamps_;
    double timeSum_[8] = {};
    int timedSteps_ = 0;
    void collectTimings();
};

#define SWARM_HIP_CHECK(expr)                                                                              \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + \
                                                       " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

} // namespace swarm
