// The seated venues' own 3D step on the GPU (gfx950): predict, the projection's Jacobi iterations and the final pass
// of the reference's step on an analytic venue (CpuCrowdSim / physics.comp with VENUE_ANALYTIC), one slot per thread.
// The grid passes and the exit choice are the plain step's (crowd_kernels.hip, with the venue's surface rule and wall
// term); agents carry a height and a vertical velocity (AgentHalf z, vz; the slots' predicted ones in SeatedSlots).
// Bit-identical to the reference's CPU step: every expression is its FP32 operation order (seated_step.h).
#include "crowd_kernels.h"
#include "seated_kernels.h"
#include "seated_step.h"

namespace swarm {

using namespace seatedstep;

namespace {

constexpr int SEATED_BLOCK = 128;

inline int blocksOf(int n, int b) { return (n + b - 1) / b; }

constexpr int SEATED_LIST_CHUNK = 4;   // list candidates whose loads one thread issues together
constexpr int SEATED_RING_CHUNK = 4;   // ring slots whose loads one thread issues together

// The contact candidates (the plain step's list buffers, slot-interleaved: index nbr[j * n + slot], height offset
// nbrD[j * n + slot]). The first iteration's ring scan keeps, in ring order, every slot with a height overlap (heights
// stay fixed through the projection) whose centre lies within D + 2 B of the slot's (predict's positions), B the
// displacement budget (min(listMoves, P - 1) maximal moves: an iteration's contacts are tested at positions at most
// P - 1 moves from predict's, a move is at most PROJ_MAX_SHIFT unless the fence or the wall plane pushes further), and
// every iteration checks its slots' path lengths (acc, the sum of the moves) against B. A slot past it sets CTRL_LIST_BROKEN and the iterations after
// it scan the ring again; so a list holds, in the ring's order, every slot that can touch, and the sums are the ring
// scan's bit for bit.
struct ListBudget {
    int on;         // 0: every iteration scans the ring (no lists)
    float budget;   // B, model units
    float reach2;   // the collection's squared reach, ((D + 2 B) 1.001 + 1e-6)^2: FP32 rounding of the tests absorbed
};

// The contacts of the 3 x 3 ring's slots whose heights overlap (cpu_crowd_sim.cpp:780-820's neighbour loop), in ring
// order. COLLECT: also the slot's candidate list
template <bool COLLECT>
__device__ inline void ringContactsSeated(const SceneView& s, const StepBuffers& b, const SeatedSlots& q, ProjBuffers in, int n, int slot,
                                          uint32_t key, uint32_t self, V2 p, V2& corr, int& nContacts, float reach2 = 0.0f) {
    const float z = q.pz[slot];
    const int cx = int(key % uint32_t(s.hashNX)), cy = int(key / uint32_t(s.hashNX));
    uint32_t cnt = 0;
    // a ring row's slots are one range in the ring's order (consecutive keys); loaded SEATED_RING_CHUNK at a time
    const int x0 = cx - s.projRing < 0 ? 0 : cx - s.projRing;
    const int x1 = cx + s.projRing >= s.hashNX ? s.hashNX - 1 : cx + s.projRing;
    for (int yy = -s.projRing; yy <= s.projRing; yy++) {
        const int ny = cy + yy;
        if (ny < 0 || ny >= s.hashNY) continue;
        const uint32_t row = uint32_t(ny * s.hashNX);
        const uint32_t k0 = b.bstart[row + uint32_t(x0)], k1 = b.bstart[row + uint32_t(x1) + 1u];
        for (uint32_t kc = k0; kc < k1; kc += SEATED_RING_CHUNK) {
            float2 o[SEATED_RING_CHUNK];
            float oz[SEATED_RING_CHUNK], ow[SEATED_RING_CHUNK];
#pragma unroll
            for (int u = 0; u < SEATED_RING_CHUNK; ++u) {
                const uint32_t k = kc + uint32_t(u) < k1 ? kc + uint32_t(u) : kc;
                o[u] = in.p[k];
                oz[u] = q.pz[k];
                ow[u] = b.pw[k];
            }
#pragma unroll
            for (int u = 0; u < SEATED_RING_CHUNK; ++u) {
                const uint32_t k = kc + uint32_t(u);
                if (k >= k1) break;
                if (int(k) == slot) continue;
                const V3 offset{ p.x - o[u].x, p.y - o[u].y, z - oz[u] };
                if (ow[u] == 0.0f || !(fabsf(offset.z) < BODY_DIAM)) continue;
                const float d2 = dot2(xy(offset), xy(offset));
                if (d2 < BODY_DIAM2) {
                    corr = corr + contact3(offset, self, b.sid[k]);
                    ++nContacts;
                }
                if constexpr (COLLECT) {
                    if (d2 < reach2) {
                        if (cnt < uint32_t(NBR_CAP)) {
                            b.nbr[size_t(cnt) * size_t(n) + size_t(slot)] = k;
                            b.nbrD[size_t(cnt) * size_t(n) + size_t(slot)] = offset.z;
                        }
                        ++cnt;
                    }
                }
            }
        }
    }
    if constexpr (COLLECT) b.nbrCount[slot] = cnt <= uint32_t(NBR_CAP) ? cnt : NBR_OVERFLOW;
}

// the same contacts from the slot's candidate list (cnt entries)
__device__ inline void listContactsSeated(const StepBuffers& b, ProjBuffers in, int n, int slot, uint32_t self, uint32_t cnt, V2 p,
                                          V2& corr, int& nContacts) {
    for (uint32_t j0 = 0; j0 < cnt; j0 += SEATED_LIST_CHUNK) {
        uint32_t kk[SEATED_LIST_CHUNK];
        float dz[SEATED_LIST_CHUNK];
        float2 o[SEATED_LIST_CHUNK];
#pragma unroll
        for (int u = 0; u < SEATED_LIST_CHUNK; ++u) {
            kk[u] = uint32_t(slot);
            dz[u] = 0.0f;
            if (j0 + u < cnt) {
                kk[u] = b.nbr[size_t(j0 + u) * size_t(n) + size_t(slot)];
                dz[u] = b.nbrD[size_t(j0 + u) * size_t(n) + size_t(slot)];
            }
        }
#pragma unroll
        for (int u = 0; u < SEATED_LIST_CHUNK; ++u) o[u] = in.p[kk[u]];
#pragma unroll
        for (int u = 0; u < SEATED_LIST_CHUNK; ++u) {
            if (j0 + u >= cnt) break;
            const V3 offset{ p.x - o[u].x, p.y - o[u].y, dz[u] };
            if (dot2(xy(offset), xy(offset)) < BODY_DIAM2) {
                corr = corr + contact3(offset, self, b.sid[kk[u]]);
                ++nContacts;
            }
        }
    }
}

// The rest of one Jacobi iteration for one slot (cpu_crowd_sim.cpp:780-820): the mean correction capped at
// PROJ_MAX_SHIFT, the fence, the slot's wall plane; p and acc are updated
__device__ inline void finishSeated(float4 pl, V2 corr, int nContacts, V2& p, float& acc) {
    corr = corr * (PROJ_RELAX / float(nContacts > 1 ? nContacts : 1));
    const float len = len2(corr);
    if (len > PROJ_MAX_SHIFT) corr = corr * (PROJ_MAX_SHIFT / len);
    V2 r = p + corr;
    const float dFence = seated::sdFence(seated::F2{ r.x, r.y });
    if (dFence < 0.0f) {
        const seated::F2 g = seated::sdFenceGradient(seated::F2{ r.x, r.y });
        r = r - V2{ g.x, g.y } * dFence;
    }
    if (fmodf(pl.w, 2.0f) == 1.0f) {   // PLANE_WALL
        const float sg = pl.z - dot2(r, V2{ pl.x, pl.y });
        if (sg > 0.0f) r = r + V2{ pl.x, pl.y } * sg;
    }
    const V2 d = r - p;
    p = r;
    acc += len2(d);
}

// One Jacobi iteration for one slot: COLLECT scans the ring and writes the slot's candidate list (the first
// iteration), LISTS walks the list (the ring again where the lists broke or the slot's overflowed), RING scans the ring
enum class ProjMode { Ring, Collect, Lists };
// The slot's own words (self: its agent; listed: its candidate count; pl: its wall plane) are the caller's.
__device__ inline void projectSeated(ProjMode mode, bool broken, const SceneView& s, const StepBuffers& b, const SeatedSlots& q,
                                     ProjBuffers in, int n, int slot, uint32_t self, uint32_t listed, float4 pl, float reach2, V2& p,
                                     float& acc) {
    V2 corr{ 0.0f, 0.0f };
    int nContacts = 0;
    if (mode == ProjMode::Collect) {
        ringContactsSeated<true>(s, b, q, in, n, slot, b.skey[slot], self, p, corr, nContacts, reach2);
    } else {
        const uint32_t cnt = mode == ProjMode::Lists && !broken ? listed : NBR_OVERFLOW;
        if (cnt == NBR_OVERFLOW) ringContactsSeated<false>(s, b, q, in, n, slot, b.skey[slot], self, p, corr, nContacts);
        else listContactsSeated(b, in, n, slot, self, cnt, p, corr, nContacts);
    }
    finishSeated(pl, corr, nContacts, p, acc);
}

// PHYSICS_PASS 1 on the venue (cpu_crowd_sim.cpp:655-777): the exit test, separation and the headway rule over the
// 3 x 3 ring (3D offsets), the incline rules, obstacle avoidance at torso height against the venue and the fence, the
// kinematics with gravity, the static contacts and the wall plane
// LDS: the venue's cell table read from the block's LDS (venueCellsInLds)
template <bool LDS>
__global__ void __launch_bounds__(SEATED_BLOCK) k_seated_predict(int n, SceneView s, ModelParams mp, StepBuffers b, AgentHalf rd,
                                                                SeatedSlots q, ProjBuffers out, uint32_t* __restrict__ ctrl) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ctrl[CTRL_LIST_BROKEN] = 0u;   // the step's candidate lists start whole
    seated::CellTable cells = s.venueCells;
    if constexpr (LDS) cells = venueCellsInLds(s);
    const int slot = xcdBlock() * SEATED_BLOCK + threadIdx.x;
    if (slot >= n || slot >= int(b.ctrl[CTRL_NACTIVE])) return;
    const uint32_t key = b.skey[slot];
    if (key == INACTIVE_CELL) return;
    const Venue3 V{ *s.venue, cells };
    const uint32_t self = b.sid[slot];
    const V3 p{ b.sx[slot], b.sy[slot], rd.z[self] };
    const V3 v{ rd.vx[self], rd.vy[self], rd.vz[self] };
    const float m = mxf(fabsf(p.x), fabsf(p.y));
    if ((m > s.exitDepth && p.z < s.exitMaxZ) || m > s.domainMax) {   // parked: it left the venue
        out.p[slot] = make_float2(PARK_COORD, PARK_COORD);
        out.acc[slot] = 0.0f;
        b.pw[slot] = 0.0f;
        b.pvx[slot] = 0.0f;
        b.pvy[slot] = 0.0f;
        q.pz[slot] = PARK_COORD;
        q.pvz[slot] = 0.0f;
        b.planes[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const V2 eGoal{ b.egx[slot], b.egy[slot] };
    const int cx = int(key % uint32_t(s.hashNX)), cy = int(key / uint32_t(s.hashNX));
    V2 sep{ 0.0f, 0.0f };
    float sAhead = 1e9f;   // the horizontal centre distance to the nearest agent ahead
    // a ring row's cells are consecutive keys, so their slots are one range, in the ring's order; its slots' words are
    // loaded SEATED_RING_CHUNK at a time, then used in order
    const int x0 = cx - s.percRing < 0 ? 0 : cx - s.percRing;
    const int x1 = cx + s.percRing >= s.hashNX ? s.hashNX - 1 : cx + s.percRing;
    for (int yy = -s.percRing; yy <= s.percRing; yy++) {
        const int ny = cy + yy;
        if (ny < 0 || ny >= s.hashNY) continue;
        const uint32_t row = uint32_t(ny * s.hashNX);
        const uint32_t k0 = b.bstart[row + uint32_t(x0)], k1 = b.bstart[row + uint32_t(x1) + 1u];
        for (uint32_t kc = k0; kc < k1; kc += SEATED_RING_CHUNK) {
            uint32_t other[SEATED_RING_CHUNK];
            float ox[SEATED_RING_CHUNK], oy[SEATED_RING_CHUNK], oz[SEATED_RING_CHUNK];
#pragma unroll
            for (int u = 0; u < SEATED_RING_CHUNK; ++u) {
                const uint32_t k = kc + uint32_t(u) < k1 ? kc + uint32_t(u) : kc;
                other[u] = b.sid[k];
                ox[u] = b.sx[k];
                oy[u] = b.sy[k];
            }
#pragma unroll
            for (int u = 0; u < SEATED_RING_CHUNK; ++u) oz[u] = rd.z[other[u]];
#pragma unroll
            for (int u = 0; u < SEATED_RING_CHUNK; ++u) {
                const uint32_t k = kc + uint32_t(u);
                if (k >= k1) break;
                if (int(k) == slot) continue;
                const V3 offset{ p.x - ox[u], p.y - oy[u], p.z - oz[u] };
                const float dxy2 = dot2(xy(offset), xy(offset));
                if (dxy2 < R_PERC2) {
                    sep = sep + separation3(offset, self, other[u]);
                    const float ahead = -dot2(xy(offset), eGoal);
                    const float lateral = fabsf(offset.x * eGoal.y - offset.y * eGoal.x);
                    const float dxy = sqrtf(dxy2);
                    if (ahead > 0.0f && lateral < BODY_DIAM && dxy < sAhead) {
                        const V2 eOther{ b.egx[k], b.egy[k] };
                        const float aheadOther = dot2(xy(offset), eOther);
                        const bool mutual = aheadOther > 0.0f && fabsf(offset.x * eOther.y - offset.y * eOther.x) < BODY_DIAM;
                        if (!mutual || ahead > aheadOther || (ahead == aheadOther && self > other[u])) sAhead = dxy;
                    }
                }
            }
        }
    }
    // the incline speed and the stair spacing by the floor's slope along e_goal (the reference's default venue model)
    const float slope = seated::floorSlopeC(V.L, V.cs, seated::F2{ p.x, p.y }, seated::F2{ eGoal.x, eGoal.y });
    const float vFree = inclineFreeSpeed(V_MAX, slope);
    const float vc = slope != 0.0f ? (sAhead - BODY_DIAM) / inclineHeadwayTime(slope) : (sAhead - BODY_DIAM) * HEADWAY_RATE;
    const float vDes = seated::cl(vc, 0.0f, vFree) * len2(eGoal);
    const V2 eDir = safeNormalize(eGoal + sep);
    // obstacle avoidance TORSO_HEIGHT above the centre: the venue's distance or the fence's, the nearer
    V2 aObs{ 0.0f, 0.0f };
    const V3 torso{ p.x, p.y, p.z + TORSO_HEIGHT };
    const float dStadium = venueMap(V, torso);
    const float dFence = fence(torso) + BODY_RADIUS;
    const bool onFence = dFence < dStadium;
    const float dTorso = onFence ? dFence : dStadium;
    if (dTorso < OBSTACLE_RANGE) {
        const V2 norm = onFence ? fenceGrad(torso) : venueNormalXY(V, torso);
        if (len2(norm) > OBSTACLE_MIN_NORMAL) aObs = safeNormalize(norm) * (OBSTACLE_RANGE - dTorso) * W_OBS * mp.obstacleGain;
    }
    const V2 vEq = eDir * vDes + aObs * mp.tau;
    V3 nv = v;
    setXY(nv, vEq + (xy(nv) - vEq) * mp.relaxStep);
    const float along = dot2(xy(nv), eGoal);
    if (along > vDes) setXY(nv, xy(nv) - eGoal * (along - vDes));
    nv.z -= GRAVITY_DT;
    nv.z = mnf(nv.z, 0.0f);   // no upward self-propulsion
    if (len2(xy(nv)) > V_MAX) setXY(nv, safeNormalize(xy(nv)) * V_MAX);
    V3 np{ p.x + nv.x * DT, p.y + nv.y * DT, p.z + nv.z * DT };
    V3 wall;
    const bool grounded = resolveStatic(V, p, np, nv, false, wall);
    if (wall.x == 0.0f && wall.y == 0.0f) wall = wallPlaneProbe(V, np);
    out.p[slot] = make_float2(np.x, np.y);
    out.acc[slot] = 0.0f;
    b.pw[slot] = 1.0f;
    b.pvx[slot] = nv.x;
    b.pvy[slot] = nv.y;
    q.pz[slot] = np.z;
    q.pvz[slot] = nv.z;
    const V2 wn{ wall.x, wall.y };
    b.planes[slot] = make_float4(wn.x, wn.y, dot2(xy(np), wn) - wall.z, (nonzero(wn) ? 1.0f : 0.0f) + (grounded ? 2.0f : 0.0f));
}

// PHYSICS_PASS 2 on the venue (cpu_crowd_sim.cpp:822-829): Jacobi iteration it of P, in -> out. The first collects the
// candidate lists; every iteration checks its slots' path lengths against the lists' budget. Slots past the step's
// active count are inactive (the inactive key sorts last) and left alone.
__global__ void __launch_bounds__(SEATED_BLOCK) k_seated_project(int n, SceneView s, StepBuffers b, SeatedSlots q, ProjBuffers in,
                                                                ProjBuffers out, int it, ListBudget lb, uint32_t* __restrict__ ctrl) {
    const int slot = xcdBlock() * SEATED_BLOCK + threadIdx.x;
    if (slot >= n || slot >= int(b.ctrl[CTRL_NACTIVE])) return;
    const bool broken = it > 1 && ctrl[CTRL_LIST_BROKEN] != 0u;
    // every word of the slot in one load round, before the branches on them
    const float2 p0 = in.p[slot];
    const float acc0 = in.acc[slot];
    const float pw = b.pw[slot];
    const uint32_t self = b.sid[slot];
    const uint32_t listed = b.nbrCount[slot];
    const float4 pl = b.planes[slot];
    asm volatile("" ::"v"(p0.x), "v"(p0.y), "v"(acc0), "v"(pw), "v"(self), "v"(listed), "v"(pl.x), "v"(pl.y), "v"(pl.z), "v"(pl.w));
    V2 p{ p0.x, p0.y };
    float acc = acc0;
    if (pw != 0.0f) {
        if (!lb.on) {
            projectSeated(ProjMode::Ring, false, s, b, q, in, n, slot, self, listed, pl, 0.0f, p, acc);
        } else {
            projectSeated(it == 1 ? ProjMode::Collect : ProjMode::Lists, broken, s, b, q, in, n, slot, self, listed, pl, lb.reach2, p, acc);
            if (!broken && acc > lb.budget) atomicOr(&ctrl[CTRL_LIST_BROKEN], 1u);
        }
    }
    out.p[slot] = make_float2(p.x, p.y);
    out.acc[slot] = acc;
}

// PHYSICS_PASS 3 on the venue (cpu_crowd_sim.cpp:835-866): the last iteration, the final static contacts of every agent
// the projection moved, the step-up limit of every agent; the scatter to agent order with the evacuation events
template <bool LDS>
__global__ void __launch_bounds__(SEATED_BLOCK) k_seated_final(int n, SceneView s, StepBuffers b, SeatedSlots q, ProjBuffers in,
                                                              int projectHere, AgentHalf rd, AgentHalf wr, EvacBuffers ev,
                                                              uint32_t* __restrict__ ctrl) {
    seated::CellTable cells = s.venueCells;
    if constexpr (LDS) cells = venueCellsInLds(s);
    const uint32_t step = ctrl[CTRL_STEP];
    const int active = int(ctrl[CTRL_NACTIVE]);
    const int slot = xcdBlock() * SEATED_BLOCK + threadIdx.x;
    if (slot < n) {
        const uint32_t agent = b.sid[slot];
        if (slot >= active) {   // an inactive slot (the inactive key sorts last): its agent is carried over
            wr.x[agent] = rd.x[agent]; wr.y[agent] = rd.y[agent]; wr.z[agent] = rd.z[agent]; wr.w[agent] = rd.w[agent];
            wr.vx[agent] = rd.vx[agent]; wr.vy[agent] = rd.vy[agent]; wr.vz[agent] = rd.vz[agent]; wr.spd[agent] = rd.spd[agent];
        } else {
            const Venue3 V{ *s.venue, cells };
            // every word of the slot in one load round, before the branches on them
            const float2 p0 = in.p[slot];
            const float acc0 = in.acc[slot];
            const float pw = b.pw[slot];
            const uint32_t listed = b.nbrCount[slot];
            const float4 pl = b.planes[slot];
            const float sx = b.sx[slot], sy = b.sy[slot];
            asm volatile("" ::"v"(p0.x), "v"(p0.y), "v"(acc0), "v"(pw), "v"(listed), "v"(pl.x), "v"(pl.y), "v"(pl.z), "v"(pl.w), "v"(sx),
                         "v"(sy));
            V2 p{ p0.x, p0.y };
            float acc = acc0;
            if (pw != 0.0f && projectHere) {   // iteration P: the lists when earlier iterations collected them
                const ProjMode mode = projectHere > 1 ? ProjMode::Lists : ProjMode::Ring;
                projectSeated(mode, mode == ProjMode::Lists && ctrl[CTRL_LIST_BROKEN] != 0u, s, b, q, in, n, slot, agent, listed, pl, 0.0f,
                              p, acc);
            }
            if (pw == 0.0f) {   // parked this step: an evacuation event, its exit by the last active position
                ev.inside[agent] = make_float2(sx, sy);
                ev.before[agent] = make_float2(wr.x[agent], wr.y[agent]);
                wr.x[agent] = p.x; wr.y[agent] = p.y; wr.z[agent] = q.pz[slot]; wr.w[agent] = 0.0f;
                wr.vx[agent] = 0.0f; wr.vy[agent] = 0.0f; wr.vz[agent] = 0.0f; wr.spd[agent] = acc;
                const int e = seated::exitByLat(V.L, seated::F2{ sx, sy });
                const int ex = e >= 0 ? e : s.E;
                ev.step[agent] = int(step);
                ev.exit[agent] = ex;
                atomicAdd(&ev.perExit[uint32_t(ex)], 1u);
            } else {
                V3 np{ p.x, p.y, q.pz[slot] }, nv{ b.pvx[slot], b.pvy[slot], q.pvz[slot] };
                const V3 pOld{ sx, sy, rd.z[agent] };
                if (acc > 0.0f) {
                    V3 wall;
                    resolveStatic(V, pOld, np, nv, pl.w >= 2.0f, wall);   // PLANE_GROUNDED
                }
                limitStepUp(V, pOld, np, nv);
                wr.x[agent] = np.x; wr.y[agent] = np.y; wr.z[agent] = np.z; wr.w[agent] = 1.0f;
                wr.vx[agent] = nv.x; wr.vy[agent] = nv.y; wr.vz[agent] = nv.z; wr.spd[agent] = len3(nv);
            }
        }
    }
    // the last block to finish advances the step counter: every block has read it by then
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&ctrl[CTRL_DONE_BLOCKS], 1u) == gridDim.x - 1u) {
            ctrl[CTRL_DONE_BLOCKS] = 0u;
            ctrl[CTRL_STEP] = step + 1u;
            *reinterpret_cast<unsigned long long*>(ctrl + CTRL_ACTIVE_SUM_LO) += ctrl[CTRL_NACTIVE];
        }
    }
}

}  // namespace

void launchSeatedPhysics(hipStream_t st, const StepLaunch& L, StepBuffers& b, SeatedSlots q, ProjBuffers pa, ProjBuffers pb,
                         AgentHalf rd, AgentHalf wr, EvacBuffers ev, uint32_t* ctrl, hipEvent_t* stamps) {
    const int n = L.n;
    const int g = blocksOf(n, SEATED_BLOCK);
    const int P = L.mp.projIters;
    ListBudget lb{ L.listMoves > 0.0f && P >= 2 ? 1 : 0, 0.0f, 0.0f };
    if (lb.on) {
        lb.budget = (L.listMoves < float(P - 1) ? L.listMoves : float(P - 1)) * PROJ_MAX_SHIFT;
        const float r = (BODY_DIAM + 2.0f * lb.budget) * 1.001f + 1e-6f;
        lb.reach2 = r * r;
    }
    const int lds = venueLdsBytes(L.s);   // the cell table in each block's LDS where it fits
    if (lds > 0) k_seated_predict<true><<<g, SEATED_BLOCK, lds, st>>>(n, L.s, L.mp, b, rd, q, pa, ctrl);
    else k_seated_predict<false><<<g, SEATED_BLOCK, 0, st>>>(n, L.s, L.mp, b, rd, q, pa, ctrl);
    if (stamps) (void)hipEventRecord(stamps[0], st);
    ProjBuffers in = pa, out = pb;
    for (int it = 1; it < P; ++it) {
        k_seated_project<<<g, SEATED_BLOCK, 0, st>>>(n, L.s, b, q, in, out, it, lb, ctrl);
        const ProjBuffers t = in;
        in = out;
        out = t;
    }
    if (stamps) (void)hipEventRecord(stamps[1], st);
    // the final pass runs iteration P: on the lists when iteration 1 collected them, else on the ring
    const int here = lb.on ? 2 : (P > 0 ? 1 : 0);
    if (lds > 0) k_seated_final<true><<<g, SEATED_BLOCK, lds, st>>>(n, L.s, b, q, in, here, rd, wr, ev, ctrl);
    else k_seated_final<false><<<g, SEATED_BLOCK, 0, st>>>(n, L.s, b, q, in, here, rd, wr, ev, ctrl);
}

}  // namespace swarm
