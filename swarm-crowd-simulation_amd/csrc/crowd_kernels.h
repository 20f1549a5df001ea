// Buffer views and launchers of the per-step kernels (crowd_kernels.hip)
#pragma once
#include <hip/hip_runtime.h>

#include "crowd_device.h"

namespace swarm {

constexpr int HASH_BLOCK = 256;    // agents per block of the grid passes (also the inactive bucket's span)
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
constexpr int PHYS_BLOCK = 128;    // slots per block of the physics passes

// control words on the device
constexpr int CTRL_STEP = 0;            // steps since the last upload (the exit choice's step counter)
constexpr int CTRL_NACTIVE = 1;         // active agents at the current step's input
constexpr int CTRL_ACTIVE_SUM_LO = 2;   // running sum of active agents over steps (uint32 pair, low word)
constexpr int CTRL_WORDS = 8;
// choice counters
constexpr int CNT_SWITCHES = 0, CNT_FORCED = 1, CNT_EVALUATIONS = 2, CNT_ELIGIBLE = 3;

// agent-order state, one half of the ping-pong
struct AgentHalf {
    float* x; float* y; float* vx; float* vy; float* spd; float* w;
};

struct GridBuffers {
    uint32_t* key;      // agent order: cell or INACTIVE_CELL
    uint32_t* loc;      // agent order: offset within its bucket (atomic order)
    uint32_t* bcnt;     // per bucket: count (cleared by the scan)
    uint32_t* bstart;   // per bucket: exclusive start (+ sentinel n)
    uint32_t* tileSum;
    uint32_t* tmp;      // slot order, atomic order within buckets
};

// slot-order step buffers
struct StepBuffers {
    uint32_t* sid; uint32_t* skey;
    uint32_t* cellStart; uint32_t* cellCount;
    float* sx; float* sy; float* sr;        // start-of-step positions and radii
    float* egx; float* egy;                 // e_goal
    float* pw; float* pvx; float* pvy;      // predicted: active flag, velocity
    float4* planes;                         // wall plane (normal, offset, flags)
    const float* avx; const float* avy;     // agent order, start of step
    const float* v0;                        // agent order
    // the projection's contact candidates per slot, slot-interleaved: nbr[j * n + slot] (NBR_CAP entries);
    // nbrCount[slot] (NBR_OVERFLOW: more than NBR_CAP, the slot scans the ring instead)
    uint32_t* nbr; uint32_t* nbrCount;
};

constexpr int NBR_CAP = 48;
constexpr uint32_t NBR_OVERFLOW = 0xFFFFFFFFu;

// the projection's slot-order ping-pong
struct ProjBuffers {
    float* x; float* y; float* acc;
};

struct ChoiceBuffers {
    uint32_t* assign;     // agent order
    uint32_t* pending;    // agent order: eligible | best << 16 | bin
    uint32_t* queue;      // 128
    uint32_t* counters;   // switches, forced, evaluations, eligible
    uint32_t* hist;       // 64 x 256
    uint32_t* threshold;  // 64
    uint32_t* ahead;      // 2 x 64 x 1024
};

struct EvacBuffers {
    int32_t* step;        // agent order: the step the agent left at (-1)
    int32_t* exit;        // agent order: its exit (E: out of the domain)
    uint32_t* perExit;    // E + 1 counters
};

void launchGridPasses(hipStream_t st, int n, const SceneView& s, AgentHalf rd, const GridBuffers& g, StepBuffers& b,
                      const float* rad, uint32_t* ctrl);
void launchChoice(hipStream_t st, int n, const SceneView& s, AgentHalf rd, const ChoiceBuffers& c, StepBuffers& b,
                  const float* driveX, const float* driveY, uint32_t* ctrl);
// stamps (optional, 2 events): recorded after predict and after the projection iterations
void launchPhysics(hipStream_t st, int n, const SceneView& s, const ModelParams& mp, StepBuffers& b, ProjBuffers pa,
                   ProjBuffers pb, AgentHalf rd, AgentHalf wr, EvacBuffers ev, uint32_t* ctrl, hipEvent_t* stamps = nullptr);
void launchStepIncrement(hipStream_t st, uint32_t* ctrl);
void launchCopyProbe(hipStream_t st, const void* src, void* dst, size_t bytes);

} // namespace swarm
