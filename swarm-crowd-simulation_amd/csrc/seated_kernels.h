// Launcher of a seated venue's own 3D step (seated_kernels.hip): the physics passes after the plain step's grid passes
// and exit choice.
#pragma once
#include <hip/hip_runtime.h>

#include "crowd_kernels.h"

namespace swarm {

// slot order: the predicted height and vertical velocity (the projection moves only x and y)
struct SeatedSlots {
    float* pz;
    float* pvz;
};

// predict, PROJ_ITERATIONS - 1 Jacobi iterations, the final pass; stamps (may be null): after predict, after the
// iterations
void launchSeatedPhysics(hipStream_t st, const StepLaunch& L, StepBuffers& b, SeatedSlots q, ProjBuffers pa, ProjBuffers pb,
                         AgentHalf rd, AgentHalf wr, EvacBuffers ev, uint32_t* ctrl, hipEvent_t* stamps);

}  // namespace swarm
