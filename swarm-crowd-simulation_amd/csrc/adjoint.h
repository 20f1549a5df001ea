// The adjoint of the crowd step: the backward kernels' views and launchers (adjoint.hip).
//
// The backward of one step runs in slot order of that step's tape, from the adjoint of its output state
// (agent order: positions, velocities) to the adjoint of its input state, accumulating the parameter adjoints:
//   final      the overlap term's gradient at the step's end, the final pass's wall pushes reversed, and phase a
//              of the last projection iteration
//   project    per iteration it = P .. 1: phase b of it (each slot gathers the contact Jacobians' transposes over
//              its candidates: the contact sets are symmetric, so every term is a gather), then phase a of it - 1
//              for the same slot (its own cap, wall plane and mean)
//   predict a  predict's own chain per slot (walls, clamps, relaxation, obstacle avoidance, goal direction)
//   predict b  the neighbour terms (separation and headway) gathered over the perception ring
#pragma once
#include "crowd_kernels.h"

namespace swarm {

struct AdjStep {
    int n, P;
    const uint32_t* sid;
    const uint32_t* skey;
    const uint32_t* bstart;
    const float *sx, *sy, *sr;
    const float *invx, *invy;           // the step's input velocity (agent order)
    const float *outx, *outy;           // the step's output positions (agent order)
    const swarm_predict_rec* pred;
    const swarm_final_rec* fin;
    const uint32_t* nbr;
    const uint32_t* cnt;
    const uint64_t *mask, *capBits, *planeBits, *overlap;
    const float2* q;                    // (P + 1) x n
    const float4* planes;
    const float* v0;                    // agent order
    int local;                          // 1: each agent's own chain only (no contact, separation, headway or
                                        // overlap terms across agents), ended at steps where it touched another
};

struct AdjBufs {
    float *gx, *gy, *gvx, *gvy;         // agent order: the adjoint of the step's output
    float *hx, *hy, *hvx, *hvy;         // agent order: the adjoint of the step's input
    float2 *gq, *gq2;                   // slot: the adjoint of an iteration's positions (ping-pong)
    float2 *w[2], *gself[2];            // slot: phase a of an iteration (ping-pong by iteration parity)
    float2 *gpv, *gsep, *gp, *gpo;      // slot
    float *gpc, *gr, *gs;               // slot
    float *gv0, *gR;                    // agent order, summed over steps
    float *gtau, *ggain;                // per scenario
    uint32_t *flags, *lead;             // slot: the step's predict flags and headway leader (NBR_OVERFLOW: none),
                                        // copied out of the predict records by the step's first backward kernel
};

// the adjoint of the exit term of the final state s
void launchAdjointSeed(hipStream_t st, int n, AgentHalf s, const swarm_exit_rec* rec, AdjBufs g);
// evacuation-crossing seeds at state 'state': gin for agents that left at that step, gout for those that left at
// the next one (added to g.gx, g.gy)
void launchAdjointEventSeed(hipStream_t st, int n, int state, const int32_t* seedStep, const float2* gin, const float2* gout, AdjBufs g);
// the same with state = *base + rel (a captured segment reads its segment's first state from the device)
void launchAdjointEventSeedRel(hipStream_t st, int n, const int* base, int rel, const int32_t* seedStep, const float2* gin,
                               const float2* gout, AdjBufs g);
// *base = value, in stream order
void launchAdjointSetBase(hipStream_t st, int* base, int value);
// zeroes the state adjoint g.gx .. g.gvy
void launchAdjointZero(hipStream_t st, int n, AdjBufs g);
// one step backwards (g.gx.. in, g.hx.. out)
void launchAdjointStep(hipStream_t st, const StepLaunch& L, const AdjStep& t, AdjBufs g, float overlapWeight);

} // namespace swarm
