/* swarm_adjoint.h - the adjoint of K steps of the crowd pipeline: its C ABI and its tape format.
 *
 * No counterpart in the reference (its CrowdSim has no derivative; the reference's finite-difference tools,
 * tools/fd, perturb whole runs). The adjoint differentiates the plain step (CrowdSim::step,
 * src/crowd_sim.cpp:316-412) with every discrete decision of a step frozen at the forward's values: the
 * neighbour and contact sets, the headway leader and clamp branch, the caps and clamps, the exit choices and
 * goal anchors, the wall contacts (each push is a move along a fixed direction to the clearance measured by a
 * fixed local linear model of the distance field), the wall planes, the park events.
 *
 * Loss (model units): sum over agents active at step K of D_exit(x_K)^2, with D_exit(x) = g(a) + |x - a| (a the
 * goal anchor of the agent's cell for its exit, g(a) the anchor's geodesic distance to that exit), plus
 * overlap_weight * sum over steps and over contact-candidate pairs overlapping at the step's end of
 * (r_i + r_j - |x_i - x_j|)^2.
 * Gradients: per-agent desired speed v0_i, the relaxation time tau (seconds), the radius scale (r_i = scale *
 * the uploaded radius) and the wall-avoidance gain.
 */
#ifndef SWARM_ADJOINT_H
#define SWARM_ADJOINT_H

#include <stddef.h>
#include <stdint.h>

#include "swarm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SWARM_TAPE_MAX_WALL 6    /* wall pushes per resolveStatic (two clearWalls calls of three passes) */
#define SWARM_TAPE_NBR 48        /* contact candidates per slot */
#define SWARM_TAPE_MAX_ITERS 32  /* projection iterations the adjoint supports */

/* predict record flags */
#define SWARM_REC_ACTIVE 0x1u      /* the slot holds an active agent (its cell is a real cell) */
#define SWARM_REC_PARKED 0x2u      /* parked this step (its outputs are constants) */
#define SWARM_REC_EFIXED 0x4u      /* e_goal is a constant (a forced drive direction, no exits, or the dead zone) */
#define SWARM_REC_LEADER 0x8u      /* has a headway leader (slot 'leader') */
#define SWARM_REC_VC_ZERO 0x10u    /* vc = 0 */
#define SWARM_REC_VC_RATE 0x20u    /* vc = sAhead * HEADWAY_RATE */
#define SWARM_REC_VC_FREE 0x40u    /* vc = stair * v0 */
#define SWARM_REC_EDIR_ZERO 0x80u  /* e_goal + sep below SAFE_NORMALIZE_EPS */
#define SWARM_REC_OBST 0x100u      /* obstacle avoidance active */
#define SWARM_REC_ALONG 0x200u     /* the along-goal cap applied */
#define SWARM_REC_CLAMP 0x400u     /* the speed clamp |v| <= v0 applied */
#define SWARM_REC_SWEEP 0x800u     /* sweepBack moved the position */
#define SWARM_REC_SWEEP_VEL 0x1000u
#define SWARM_REC_MOVED 0x2000u    /* final: the projection moved the agent, resolveStatic ran again */
#define SWARM_REC_GOAL_WALL 0x4000u /* e_goal's wall rule applied: e = unit(e0 - n (e0 . n)), n = goal_wall */

/* A wall push: x' = x + a (r - t(x)), t(x) = t0 - g.(x - x0) / (g.u); vel: v' = v - a (a.v) */
typedef struct {
    float ax, ay, ux, uy, gx, gy, x0, y0, t0;
    uint32_t vel;
} swarm_wall_event;

typedef struct {
    uint32_t flags, leader;
    float anchor_x, anchor_y;   /* e_goal = unit(anchor - p) unless SWARM_REC_EFIXED */
    float egoal_x, egoal_y;     /* e_goal as used */
    float stair;                /* free-speed factor under the start position */
    float sep_x, sep_y;         /* the separation sum (FP32) */
    float sweep_lo, sweep_dx, sweep_dy;   /* x' = x0 (1 - lo) + x lo; v' = v - d (d.v) when SWARM_REC_SWEEP_VEL */
    uint32_t n_events;
    uint32_t plane_kind;        /* 0 none; 1 offset n.x; 2 offset n.x - (t(x) - r) */
    float plane_nx, plane_ny;
    float pc_ux, pc_uy, pc_gx, pc_gy, pc_x0, pc_y0, pc_t0;   /* the plane's clearance model (kind 2) */
    float goal_wall_x, goal_wall_y;   /* the wall direction of e_goal's wall rule (SWARM_REC_GOAL_WALL) */
    swarm_wall_event ev[SWARM_TAPE_MAX_WALL];
} swarm_predict_rec;

typedef struct {
    uint32_t flags, n_events;
    float sweep_lo, sweep_dx, sweep_dy;
    swarm_wall_event ev[SWARM_TAPE_MAX_WALL];
} swarm_final_rec;

/* The exit term of an agent at step K: D = base + |x - anchor| (flags 1: counted) */
typedef struct {
    float anchor_x, anchor_y, base;
    uint32_t flags;
} swarm_exit_rec;

/* One step's tape, slot order, as swarm_adjoint_tape copies it (host arrays the caller allocates; n the
 * context's agent count, P its projection iterations): the step's input state (agent order), the slot order
 * (sid), the records, the start positions and radii (FP32), the contact candidates, the per-iteration contact
 * masks over the candidates (bit k: candidate k is a contact; mask[(it - 1) * n + slot], it = 1..P), the
 * iterations whose cap / wall plane applied (bit it - 1), the candidates overlapping at the step's end. */
typedef struct {
    float *x, *y, *vx, *vy, *w;                 /* n, agent order */
    uint32_t* sid;                              /* n */
    swarm_predict_rec* pred;                    /* n */
    swarm_final_rec* fin;                       /* n */
    float *sx, *sy, *sr;                        /* n */
    uint32_t* nbr;                              /* SWARM_TAPE_NBR * n, nbr[k * n + slot] */
    uint32_t* nbr_count;                        /* n */
    uint64_t* mask;                             /* P * n */
    uint64_t *cap_bits, *plane_bits, *overlap;  /* n */
} swarm_tape_view;

/* Runs K steps from the uploaded state (the plain step, hipGraph replays), keeps the state every
 * checkpoint_every steps, and returns the loss. radius_scale multiplies the uploaded radii for this run. The last
 * segment runs taped, so the backward starts without recomputing it; a run whose whole tape fits the tape budget
 * (swarm_adjoint_set_tape_budget), or checkpoint_every >= K, is taped whole in the forward and never recomputed. */
int swarm_adjoint_forward(swarm_ctx* c, int K, int checkpoint_every, float radius_scale, float overlap_weight, double* loss);
/* After swarm_adjoint_forward: recomputes each earlier segment from its checkpoint with the tape and runs the
 * adjoint kernels backwards. grad_v0: n (agent order); grad_globals: dL/dtau_s, dL/dradius_scale, dL/dwall_gain.
 * Afterwards the context holds the forward's final state again (a recompute puts it back). Returns
 * SWARM_ENONFINITE, with both outputs written, when any gradient value is inf or NaN (the pathwise derivative
 * overflowed FP32: a horizon beyond what FP32 holds in a chaotic crowd). swarm_upload, swarm_set_proj_iters and a
 * swarm_set_model that changes the model end the forward's run: a backward then returns SWARM_EINVAL. */
int swarm_adjoint_backward(swarm_ctx* c, float* grad_v0, double* grad_globals);
/* Device bytes the context's adjoint holds (tape, adjoint buffers, checkpoints). */
int swarm_adjoint_memory(swarm_ctx* c, size_t* bytes);
/* A forward whose whole tape (all K steps) fits 'bytes' is taped whole and never recomputed, whatever the
 * checkpoint interval (default 4 GiB: K = 200 at 4,096 agents takes 1.3 GB); 0: always checkpoint and recompute. */
int swarm_adjoint_set_tape_budget(swarm_ctx* c, size_t bytes);
/* The number of segments of the last swarm_adjoint_forward (1: taped whole). */
int swarm_adjoint_segments(swarm_ctx* c, int* segments);
/* The tape of step 'step' (0-based) of the last forward (recomputed from its checkpoint), and for step K - 1 also
 * the exit records (exit may be NULL). Any pointer of the view may be NULL. */
int swarm_adjoint_tape(swarm_ctx* c, int step, const swarm_tape_view* out, swarm_exit_rec* exit);
/* Loss terms of the next swarm_adjoint_backward beyond the built-in ones (for a loss on evacuation times, e.g. exit-flow
 * curves): per agent, the position adjoint gin (2 floats) at the input of step step[a] (the agent is inside its exit
 * there) and gout (2 floats) at the input of step step[a] - 1 (step < 1: none; NULL arrays: no such terms).
 * exit_term: whether the built-in exit-distance term counts. window > 0 truncates the adjoint: its state is cleared
 * every 'window' steps going backwards, so each path of the gradient spans at most that many steps (0: exact).
 * local != 0: each agent's own chain only (no transfer across agents through contacts, separation, headway or
 * overlaps), cut at every step where the agent touched another agent in the projection (its position is then held
 * by the crowd): a stable, biased estimator for horizons where the exact gradient explodes. */
int swarm_adjoint_set_terms(swarm_ctx* c, const int32_t* step, const float* gin, const float* gout, int exit_term, int window,
                            int local);
/* Per agent: the step it left at (-1: none), its exit, its position at that step's input (inside the exit) and one
 * step earlier (x, y pairs): the crossing of its exit's boundary lies between the two. */
int swarm_read_evac_points(swarm_ctx* c, int32_t* step, int32_t* exit_id, float* inside, float* before);
/* The model parameters of the next steps: relaxation time (seconds) and wall-avoidance gain */
int swarm_set_model(swarm_ctx* c, float tau_s, float obstacle_gain);

#ifdef __cplusplus
}
#endif

#endif
