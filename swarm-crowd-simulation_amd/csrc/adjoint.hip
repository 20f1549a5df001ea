// The adjoint of the crowd step (gfx950): reverse-mode kernels over one step's tape. Every discrete decision
// of the forward is read from the tape (swarm_adjoint.h); the continuous values a derivative needs are recomputed
// in FP32 from the tape's inputs with the forward's own expressions. Adjoints are FP32 and accumulate in FP32.
//
// Notation: for a slot i, g* is the adjoint (dL/d*) of a forward quantity; a wall push is
// x' = x + a (r - t(x)) with t(x) = t0 - g.(x - x0) / (g.u), so dx'/dx = I + a g^T / (g.u) and dx'/dr = a.
#include "adjoint.h"

namespace swarm {

namespace {

// Sums v over the wave into acc[sc] when the wave's lanes share one scenario, else one atomic per lane
__device__ inline void scenarioAdd(float* acc, uint32_t sc, float v, bool valid) {
    const uint32_t s0 = __shfl(sc, 0, 64);
    const bool uniform = __ballot(valid && sc != s0) == 0ull;
    if (uniform) {
        float t = valid ? v : 0.0f;
        for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
        if (__lane_id() == 0 && t != 0.0f) atomicAdd(&acc[s0], t);
    } else if (valid && v != 0.0f) {
        atomicAdd(&acc[sc], v);
    }
}

// a wall push reversed: g (position adjoint), gv (velocity adjoint), gr (radius adjoint)
__device__ inline void reversePush(const swarm_wall_event& e, V2& g, V2& gv, float& gr) {
    const V2 a{ e.ax, e.ay }, gg{ e.gx, e.gy }, u{ e.ux, e.uy };
    const float ag = dot2(a, g);
    gr += ag;
    const float gu = dot2(gg, u);
    if (gu != 0.0f) g = g + gg * (ag / gu);
    if (e.vel) gv = gv - a * dot2(a, gv);
}

__device__ inline V2 sdfGrad(const SceneView& s, V2 p) {
    const G3 g = mapExactGrad(s, p);
    return V2{ g.gx, g.gy };
}

// phase a of projection iteration 'it' for an active, not parked slot: from the adjoint gqi of the iteration's
// output position to the adjoint through the wall plane (gself, the identity path of q' = q + c) and the
// adjoint of the contact sum (w); the plane offset's and the radius's adjoints accumulate
template <bool B>
__device__ inline void projectPhaseA(const AdjStep& t, const BatchView& bt, int i, int it, V2 gqi, float& gpc, float& gr, V2& gself, V2& w) {
    const size_t n = size_t(t.n);
    const uint64_t mask = t.mask[size_t(it - 1) * n + size_t(i)];
    const float2 qi = t.q[size_t(it - 1) * n + size_t(i)];
    const float ri = t.sr[i];
    const ScnRef sr = scnOfAgent<B>(bt, t.sid[i]);   // the tape's own slot order (the live slot map is the last step's)
    const uint32_t self = t.sid[i] - sr.agentBase;
    V2 u{ 0.0f, 0.0f };
    int nc = 0;
    for (uint64_t m = mask; m; m &= m - 1ull) {
        const int k = __ffsll((unsigned long long)m) - 1;
        const uint32_t j = t.nbr[size_t(k) * n + size_t(i)];
        const float2 qj = t.q[size_t(it - 1) * n + size_t(j)];
        const V2 off{ qi.x - qj.x, qi.y - qj.y };
        const float d2 = dot2(off, off);
        const float D = ri + t.sr[j];
        const float d = sqrtf(d2);
        const float shift = 0.5f * (D - d);
        u = u + (d2 <= SEP_STACKED_XY2 ? shift * pairDirection(self, t.sid[j] - sr.agentBase) : shift * (off / d));
        ++nc;
    }
    const float inv = PROJ_RELAX / float(nc > 1 ? nc : 1);
    const V2 c = u * inv;
    V2 g = gqi;
    if ((t.planeBits[i] >> (it - 1)) & 1ull) {
        const float4 pl = t.planes[i];
        const V2 nn{ pl.x, pl.y };
        const float ng = dot2(nn, g);
        gpc += ng;
        g = g - nn * ng;
    }
    gself = g;
    if ((t.capBits[i] >> (it - 1)) & 1ull) {
        const float L = len2(c);
        const V2 ch = c / L;
        const float cg = dot2(ch, g);
        gr += PROJ_MAX_SHIFT_RADII * cg;
        g = (g - ch * cg) * (PROJ_MAX_SHIFT_RADII * ri / L);
    }
    w = g * inv;
}

// the overlap term's gradient at the step's end, the final pass reversed, phase a of iteration P
template <bool B>
__global__ void __launch_bounds__(256) k_adj_final(AdjStep t, BatchView bt, AdjBufs g, float ovw) {
    const int i = xcdBlock() * int(blockDim.x) + int(threadIdx.x);
    if (i >= t.n) return;
    const uint32_t a = t.sid[i];
    const swarm_predict_rec& pr = t.pred[i];
    const V2 z{ 0.0f, 0.0f };
    g.flags[i] = pr.flags;
    g.lead[i] = (pr.flags & SWARM_REC_LEADER) ? pr.leader : NBR_OVERFLOW;
    g.gpc[i] = 0.0f; g.gr[i] = 0.0f; g.gs[i] = 0.0f;
    g.gsep[i] = make_float2(0.0f, 0.0f); g.gp[i] = make_float2(0.0f, 0.0f); g.gpo[i] = make_float2(0.0f, 0.0f);
    g.gpv[i] = make_float2(0.0f, 0.0f);
    const int P = t.P;
    float2* wP = g.w[P & 1];
    float2* sP = g.gself[P & 1];
    wP[i] = make_float2(0.0f, 0.0f);
    if (!(pr.flags & SWARM_REC_ACTIVE)) {   // inactive: carried
        g.hx[a] = g.gx[a]; g.hy[a] = g.gy[a]; g.hvx[a] = g.gvx[a]; g.hvy[a] = g.gvy[a];
        return;
    }
    if (pr.flags & SWARM_REC_PARKED) {   // parked this step: constant outputs
        sP[i] = make_float2(0.0f, 0.0f);
        g.hvx[a] = 0.0f; g.hvy[a] = 0.0f;
        return;
    }
    V2 gp{ g.gx[a], g.gy[a] }, gv{ g.gvx[a], g.gvy[a] };
    if (t.local) {   // local: an agent that touched another this step is held by the crowd; its own chain ends here
        uint64_t touched = 0ull;
        for (int it = 1; it <= P; ++it) touched |= t.mask[size_t(it - 1) * size_t(t.n) + size_t(i)];
        if (touched) { gp = z; gv = z; }
    }
    float gr = 0.0f;
    const V2 pi{ t.outx[a], t.outy[a] };
    for (uint64_t m = t.local ? 0ull : t.overlap[i]; m; m &= m - 1ull) {
        const int k = __ffsll((unsigned long long)m) - 1;
        const uint32_t j = t.nbr[size_t(k) * size_t(t.n) + size_t(i)];
        const uint32_t aj = t.sid[j];
        const V2 dl = pi - V2{ t.outx[aj], t.outy[aj] };
        const float d = len2(dl);
        const float ov = t.sr[i] + t.sr[j] - d;
        if (d > 0.0f) gp = gp - dl * (2.0f * ovw * ov / d);
        gr += 2.0f * ovw * ov;
    }
    const swarm_final_rec& fr = t.fin[i];
    V2 gpo = z;
    if (fr.flags & SWARM_REC_MOVED) {
        for (int e = int(fr.n_events) - 1; e >= 0; --e) reversePush(fr.ev[e], gp, gv, gr);
        if (fr.flags & SWARM_REC_SWEEP) {
            const float lo = fr.sweep_lo;
            gpo = gp * (1.0f - lo);
            gp = gp * lo;
            if (fr.flags & SWARM_REC_SWEEP_VEL) {
                const V2 d{ fr.sweep_dx, fr.sweep_dy };
                gv = gv - d * dot2(d, gv);
            }
        }
    }
    g.gpv[i] = make_float2(gv.x, gv.y);
    g.gpo[i] = make_float2(gpo.x, gpo.y);
    float gpc = 0.0f;
    V2 gself, w;
    projectPhaseA<B>(t, bt, i, P, gp, gpc, gr, gself, w);
    sP[i] = make_float2(gself.x, gself.y);
    wP[i] = make_float2(w.x, w.y);
    g.gpc[i] = gpc;
    g.gr[i] = gr;
}

// phase b of iteration it (the adjoint of its input positions), then phase a of iteration it - 1
template <bool B>
__global__ void __launch_bounds__(256) k_adj_project(AdjStep t, BatchView bt, AdjBufs g, int it) {
    const int i = xcdBlock() * int(blockDim.x) + int(threadIdx.x);
    if (i >= t.n) return;
    const uint32_t pf = g.flags[i];
    if (!(pf & SWARM_REC_ACTIVE)) return;
    const size_t n = size_t(t.n);
    const float2* wv = g.w[it & 1];
    const float2 s = g.gself[it & 1][i];
    V2 gq{ s.x, s.y };
    float2* wNext = g.w[(it - 1) & 1];
    float2* sNext = g.gself[(it - 1) & 1];
    if (pf & SWARM_REC_PARKED) {   // carried: no contacts, nobody's contact
        if (it > 1) { sNext[i] = make_float2(gq.x, gq.y); wNext[i] = make_float2(0.0f, 0.0f); }
        else g.gq[i] = make_float2(gq.x, gq.y);
        return;
    }
    const ScnRef sr = scnOfAgent<B>(bt, t.sid[i]);   // the tape's own slot order (the live slot map is the last step's)
    const uint32_t self = t.sid[i] - sr.agentBase;
    const float2 qi = t.q[size_t(it - 1) * n + size_t(i)];
    const float ri = t.sr[i];
    const V2 wi{ wv[i].x, wv[i].y };
    float gr = g.gr[i];
    // the contacts in chunks of ADJ_CHUNK: a chunk's index loads, then its position, radius and adjoint loads, are
    // all in flight together; the sums still run in contact order
    uint64_t m = t.local ? 0ull : t.mask[size_t(it - 1) * n + size_t(i)];
    while (m) {
        uint32_t jj[ADJ_CHUNK];
        int cnt = 0;
#pragma unroll
        for (int u = 0; u < ADJ_CHUNK; ++u) {
            if (m) {
                const int k = __ffsll((unsigned long long)m) - 1;
                m &= m - 1ull;
                jj[u] = t.nbr[size_t(k) * n + size_t(i)];
                ++cnt;
            } else {
                jj[u] = uint32_t(i);
            }
        }
        float2 qq[ADJ_CHUNK], ww[ADJ_CHUNK];
        float rr[ADJ_CHUNK];
#pragma unroll
        for (int u = 0; u < ADJ_CHUNK; ++u) {
            qq[u] = t.q[size_t(it - 1) * n + size_t(jj[u])];
            rr[u] = t.sr[jj[u]];
            ww[u] = wv[jj[u]];
        }
#pragma unroll
        for (int u = 0; u < ADJ_CHUNK; ++u) {
            if (u >= cnt) break;
            const uint32_t j = jj[u];
            const float2 qj = qq[u];
            const V2 off{ qi.x - qj.x, qi.y - qj.y };
            const float d2 = dot2(off, off);
            const float D = ri + rr[u];
            const V2 dw = wi - V2{ ww[u].x, ww[u].y };
            if (d2 <= SEP_STACKED_XY2) {
                const V2 dir = pairDirection(self, t.sid[j] - sr.agentBase);
                const float dd = dot2(dir, dw);
                gr += 0.5f * dd;
                if (d2 > 0.0f) gq = gq - off * (0.5f * dd / sqrtf(d2));
            } else {
                const float d = sqrtf(d2);
                const V2 nh = off / d;
                const float nd = dot2(nh, dw);
                gq = gq + (dw - nh * nd) * (0.5f * (D / d - 1.0f)) - nh * (0.5f * nd);
                gr += 0.5f * nd;
            }
        }
    }
    if (it > 1) {
        float gpc = g.gpc[i];
        V2 gself, w;
        projectPhaseA<B>(t, bt, i, it - 1, gq, gpc, gr, gself, w);
        sNext[i] = make_float2(gself.x, gself.y);
        wNext[i] = make_float2(w.x, w.y);
        g.gpc[i] = gpc;
    } else {
        g.gq[i] = make_float2(gq.x, gq.y);
    }
    g.gr[i] = gr;
}

// predict's own chain reversed, per active slot not parked
template <bool B>
__global__ void __launch_bounds__(256) k_adj_predict_a(AdjStep t, SceneView s0, BatchView bt, ModelParams mp, AdjBufs g) {
    const int i = xcdBlock() * int(blockDim.x) + int(threadIdx.x);
    const bool live = i < t.n && (g.flags[i] & SWARM_REC_ACTIVE) && !(g.flags[i] & SWARM_REC_PARKED);
    float gtau = 0.0f, ggain = 0.0f;
    uint32_t scn = 0u;
    if (live) {
        const swarm_predict_rec& pr = t.pred[i];
        const uint32_t f = pr.flags;
        const ScnRef sr = scnOfAgent<B>(bt, t.sid[i]);   // the tape's own slot order (the live slot map is the last step's)
        scn = sr.sc;
        const SceneView& s = sceneOf<B>(s0, bt, sr.sc);
        const uint32_t a = t.sid[i];
        const V2 p{ t.sx[i], t.sy[i] };
        const float ri = t.sr[i];
        const V2 v{ t.invx[a], t.invy[a] };
        const V2 e{ pr.egoal_x, pr.egoal_y };
        const V2 sep{ pr.sep_x, pr.sep_y };
        const float v0a = t.v0[a];
        // the forward's values (FP32, the forward's expressions)
        const float vFree = pr.stair * v0a;
        float sAhead = 0.0f;
        V2 nl{ 0.0f, 0.0f };
        if (f & SWARM_REC_LEADER) {
            const uint32_t j = pr.leader;
            const V2 dl = p - V2{ t.sx[j], t.sy[j] };
            const float d = len2(dl);
            sAhead = d - (ri + t.sr[j]);
            nl = dl / d;
        }
        const float vc = (f & SWARM_REC_VC_ZERO) ? 0.0f : ((f & SWARM_REC_VC_FREE) ? vFree : sAhead * HEADWAY_RATE);
        const float le = len2(e);
        const float vDes = vc * le;
        const V2 uv = e + sep;
        const float lu = len2(uv);
        const V2 eDir = (f & SWARM_REC_EDIR_ZERO) ? V2{ 0.0f, 0.0f } : uv / lu;
        V2 aObs{ 0.0f, 0.0f }, n1{ 0.0f, 0.0f }, gradG{ 0.0f, 0.0f };
        float dT = 0.0f, lg = 1.0f;
        if (f & SWARM_REC_OBST) {
            dT = mapSdf(s, p);
            gradG = calcGradient(s, p);
            lg = len2(gradG);
            n1 = gradG / lg;
            aObs = n1 * ((OBSTACLE_RANGE - dT) * W_OBS * mp.obstacleGain);
        }
        const V2 vEq = eDir * vDes + aObs * mp.tau;
        const float relax = mp.relaxStep;
        const V2 nv1 = vEq + (v - vEq) * relax;
        const float along = dot2(nv1, e);
        const V2 nv2 = (f & SWARM_REC_ALONG) ? nv1 - e * (along - vDes) : nv1;
        const float l2 = len2(nv2);
        // reverse
        const float2 gq0 = g.gq[i];
        V2 gx{ gq0.x, gq0.y };
        const float2 gpv = g.gpv[i];
        V2 gvv{ gpv.x, gpv.y };
        float gr = g.gr[i];
        const float gpc = g.gpc[i];
        if (pr.plane_kind != 0u) {   // plane offset: n.x (kind 1), n.x - (t(x) - r) (kind 2)
            gx = gx + V2{ pr.plane_nx, pr.plane_ny } * gpc;
            if (pr.plane_kind == 2u) {
                const V2 pg{ pr.pc_gx, pr.pc_gy };
                const float gu = dot2(pg, V2{ pr.pc_ux, pr.pc_uy });
                if (gu != 0.0f) gx = gx + pg * (gpc / gu);
                gr += gpc;
            }
        }
        for (int k = int(pr.n_events) - 1; k >= 0; --k) reversePush(pr.ev[k], gx, gvv, gr);
        V2 gp{ 0.0f, 0.0f };
        if (f & SWARM_REC_SWEEP) {
            const float lo = pr.sweep_lo;
            gp = gx * (1.0f - lo);
            gx = gx * lo;
            if (f & SWARM_REC_SWEEP_VEL) {
                const V2 d{ pr.sweep_dx, pr.sweep_dy };
                gvv = gvv - d * dot2(d, gvv);
            }
        }
        gp = gp + gx;                       // np = p + nv DT
        const V2 gnv3 = gvv + gx * DT;
        float gv0 = 0.0f;
        V2 gnv2 = gnv3;
        if (f & SWARM_REC_CLAMP) {          // nv3 = v0 nv2 / |nv2|
            const V2 nh = nv2 / l2;
            const float ng = dot2(nh, gnv3);
            gv0 += ng;
            gnv2 = (gnv3 - nh * ng) * (v0a / l2);
        }
        V2 gnv1 = gnv2, ge{ 0.0f, 0.0f };
        float gvDes = 0.0f;
        if (f & SWARM_REC_ALONG) {          // nv2 = nv1 - e (e.nv1 - vDes)
            const float eg = dot2(e, gnv2);
            gnv1 = gnv2 - e * eg;
            ge = ge + gnv2 * (vDes - along) - nv1 * eg;
            gvDes += eg;
        }
        const V2 gv = gnv1 * relax;         // nv1 = vEq + (v - vEq) relax
        const V2 gvEq = gnv1 * (1.0f - relax);
        gtau += dot2(v - vEq, gnv1) * relax * DT / (mp.tau * mp.tau);
        const V2 geDir = gvEq * vDes;       // vEq = eDir vDes + aObs tau
        gvDes += dot2(eDir, gvEq);
        const V2 gaObs = gvEq * mp.tau;
        gtau += dot2(aObs, gvEq);
        if (f & SWARM_REC_OBST) {           // aObs = n1 (RANGE - d) W gain, n1 = G / |G|, G the 4-sample gradient
            const float kk = (OBSTACLE_RANGE - dT) * W_OBS;
            const float na = dot2(n1, gaObs);
            ggain += kk * na;
            const float gdT = -W_OBS * mp.obstacleGain * na;
            const V2 gn1 = gaObs * (kk * mp.obstacleGain);
            const V2 gG = (gn1 - n1 * dot2(n1, gn1)) / lg;
            const float S = OBSTACLE_GRAD_SCALE, H = OBSTACLE_GRAD_H;
            gp = gp + sdfGrad(s, V2{ p.x + H, p.y - H }) * (S * (gG.x - gG.y)) + sdfGrad(s, V2{ p.x - H, p.y - H }) * (S * (-gG.x - gG.y)) +
                 sdfGrad(s, V2{ p.x - H, p.y + H }) * (S * (-gG.x + gG.y)) + sdfGrad(s, V2{ p.x + H, p.y + H }) * (S * (gG.x + gG.y)) +
                 sdfGrad(s, p) * gdT;
        }
        V2 gsep{ 0.0f, 0.0f };
        if (!(f & SWARM_REC_EDIR_ZERO)) {   // eDir = (e + sep) / |e + sep|
            const V2 gu = (geDir - eDir * dot2(eDir, geDir)) / lu;
            ge = ge + gu;
            gsep = gu;
        }
        const float gvc = gvDes * le;       // vDes = vc |e|
        float gs = 0.0f;
        if (f & SWARM_REC_VC_FREE) gv0 += pr.stair * gvc;
        if (f & SWARM_REC_VC_RATE) {        // vc = (|p - p_lead| - D) rate
            gs = gvc * HEADWAY_RATE;
            gp = gp + nl * gs;
            gr -= gs;
        }
        if (!(f & SWARM_REC_EFIXED)) {      // e0 = (anchor - p) / |anchor - p|
            const V2 d = V2{ pr.anchor_x, pr.anchor_y } - p;
            const float ld = len2(d);
            const V2 ee = d / ld;
            if (f & SWARM_REC_GOAL_WALL) {  // e1 = e0 - n (e0 . n), e = e1 / |e1| (n frozen)
                const V2 nw{ pr.goal_wall_x, pr.goal_wall_y };
                const V2 e1 = ee - nw * dot2(ee, nw);
                const float l1 = len2(e1);
                const V2 eu = e1 / l1;
                const V2 g1 = (ge - eu * dot2(eu, ge)) / l1;
                ge = g1 - nw * dot2(nw, g1);
            }
            gp = gp - (ge - ee * dot2(ee, ge)) / ld;
        }
        const float2 po = g.gpo[i];
        g.gp[i] = make_float2(gp.x + po.x, gp.y + po.y);
        g.gsep[i] = t.local ? make_float2(0.0f, 0.0f) : make_float2(gsep.x, gsep.y);
        g.gs[i] = gs;
        g.gr[i] = gr;
        g.hvx[a] = gv.x; g.hvy[a] = gv.y;
        g.gv0[a] += gv0;
    }
    scenarioAdd(g.gtau, scn, gtau, live);
    scenarioAdd(g.ggain, scn, ggain, live);
}

// the neighbour terms of predict (separation, headway), gathered over the perception ring; the step's input adjoint
template <bool B>
__global__ void __launch_bounds__(256) k_adj_predict_b(AdjStep t, SceneView s0, BatchView bt, AdjBufs g) {
    const int i = xcdBlock() * int(blockDim.x) + int(threadIdx.x);
    if (i >= t.n) return;
    if (!(g.flags[i] & SWARM_REC_ACTIVE)) return;
    const uint32_t a = t.sid[i];
    const ScnRef sr = scnOfAgent<B>(bt, t.sid[i]);   // the tape's own slot order (the live slot map is the last step's)
    const SceneView& s = sceneOf<B>(s0, bt, sr.sc);
    const uint32_t self = a - sr.agentBase;
    const float2 gp0 = g.gp[i];
    V2 gp{ gp0.x, gp0.y };
    float gr = g.gr[i];
    const V2 p{ t.sx[i], t.sy[i] };
    const float ri = t.sr[i];
    const float2 gsi2 = g.gsep[i];
    const V2 gsi{ gsi2.x, gsi2.y };
    const uint32_t cid = t.skey[i] - sr.cellBase;
    const int cx = int(cid % uint32_t(s.hashNX)), cy = int(cid / uint32_t(s.hashNX));
    const int R = t.local ? -1 : s.percRing;   // local: no neighbour terms
    const int x0 = cx - R < 0 ? 0 : cx - R;
    const int x1 = cx + R > s.hashNX - 1 ? s.hashNX - 1 : cx + R;
    for (int yy = -R; yy <= R; yy++) {
        const int ny = cy + yy;
        if (ny < 0 || ny >= s.hashNY) continue;
        const uint32_t row = sr.cellBase + uint32_t(ny * s.hashNX);
        const uint32_t st = t.bstart[row + uint32_t(x0)], en = t.bstart[row + uint32_t(x1) + 1u];
        for (uint32_t k = st; k < en; ++k) {
            if (k == uint32_t(i)) continue;
            const V2 off = p - V2{ t.sx[k], t.sy[k] };
            const float d2 = dot2(off, off);
            if (!(d2 < R_PERC2)) continue;
            const float S = ri + t.sr[k] + SEP_MARGIN;
            if (d2 < sepSupport2(S)) {   // separation: F = K (1 - d/S)^2 n (a stacked pair: the hash direction)
                const float d = sqrtf(d2);
                const float wgt = 1.0f - d / S;
                const float mag = K_SEP * wgt * wgt;
                const float dmd = -2.0f * K_SEP * wgt / S;
                const float dmS = 2.0f * K_SEP * wgt * d / (S * S);
                const float2 gsk = g.gsep[k];
                const V2 dg = gsi - V2{ gsk.x, gsk.y };
                if (d2 <= SEP_STACKED_XY2) {
                    const V2 dir = pairDirection(self, t.sid[k] - sr.agentBase);
                    const float dd = dot2(dir, dg);
                    gr += dmS * dd;
                    if (d > 0.0f) gp = gp + off * (dmd * dd / d);
                } else {
                    const V2 nh = off / d;
                    const float nd = dot2(nh, dg);
                    gp = gp + (dg - nh * nd) * (mag / d) + nh * (dmd * nd);
                    gr += dmS * nd;
                }
            }
            if (g.lead[k] == uint32_t(i)) {   // slot k's headway leader is this slot
                const float gsk = g.gs[k];
                if (gsk != 0.0f) {
                    gp = gp + off * (gsk / sqrtf(d2));
                    gr -= gsk;
                }
            }
        }
    }
    g.hx[a] = gp.x; g.hy[a] = gp.y;
    g.gR[a] += gr;
}

// the exit term's adjoint: dL/dx = 2 D (x - anchor) / |x - anchor|
__global__ void __launch_bounds__(256) k_adj_seed(int n, AgentHalf s, const swarm_exit_rec* __restrict__ rec, AdjBufs g) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    float gx = 0.0f, gy = 0.0f;
    const swarm_exit_rec r = rec[i];
    if (r.flags) {
        const V2 d = V2{ s.x[i], s.y[i] } - V2{ r.anchor_x, r.anchor_y };
        const float l = len2(d);
        const float D = r.base + l;
        if (l > 0.0f) { gx = 2.0f * D * d.x / l; gy = 2.0f * D * d.y / l; }
    }
    g.gx[i] = gx; g.gy[i] = gy; g.gvx[i] = 0.0f; g.gvy[i] = 0.0f;
}

// position adjoints of evacuation crossings at state 'state' (the input of that step): the inside point of agents
// that left at step 'state', the point before of agents that left at step 'state + 1'
__global__ void __launch_bounds__(256) k_adj_event_seed(int n, int state, const int32_t* __restrict__ seedStep,
                                                        const float2* __restrict__ gin, const float2* __restrict__ gout, AdjBufs g) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int32_t s = seedStep[i];
    if (s == state) { g.gx[i] += gin[i].x; g.gy[i] += gin[i].y; }
    else if (s == state + 1) { g.gx[i] += gout[i].x; g.gy[i] += gout[i].y; }
}

// the same with the state taken from a device word (base + rel): one captured graph serves every segment
__global__ void __launch_bounds__(256) k_adj_event_seed_rel(int n, const int* __restrict__ base, int rel,
                                                            const int32_t* __restrict__ seedStep, const float2* __restrict__ gin,
                                                            const float2* __restrict__ gout, AdjBufs g) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int state = *base + rel;
    const int32_t s = seedStep[i];
    if (s == state) { g.gx[i] += gin[i].x; g.gy[i] += gin[i].y; }
    else if (s == state + 1) { g.gx[i] += gout[i].x; g.gy[i] += gout[i].y; }
}

__global__ void k_adj_set_int(int* __restrict__ p, int v) { *p = v; }

__global__ void __launch_bounds__(256) k_adj_zero(int n, AdjBufs g) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    g.gx[i] = 0.0f; g.gy[i] = 0.0f; g.gvx[i] = 0.0f; g.gvy[i] = 0.0f;
}

inline int blocks256(int n) { return (n + 255) / 256; }

template <bool B>
void adjointStep(hipStream_t st, const StepLaunch& L, const AdjStep& t, AdjBufs g, float ovw) {
    const int gr = blocks256(t.n);
    k_adj_final<B><<<gr, 256, 0, st>>>(t, L.bt, g, ovw);
    for (int it = t.P; it >= 1; --it) k_adj_project<B><<<gr, 256, 0, st>>>(t, L.bt, g, it);
    k_adj_predict_a<B><<<gr, 256, 0, st>>>(t, L.s, L.bt, L.mp, g);
    k_adj_predict_b<B><<<gr, 256, 0, st>>>(t, L.s, L.bt, g);
}

} // namespace

void launchAdjointSeed(hipStream_t st, int n, AgentHalf s, const swarm_exit_rec* rec, AdjBufs g) {
    k_adj_seed<<<blocks256(n), 256, 0, st>>>(n, s, rec, g);
}

void launchAdjointEventSeed(hipStream_t st, int n, int state, const int32_t* seedStep, const float2* gin, const float2* gout, AdjBufs g) {
    k_adj_event_seed<<<blocks256(n), 256, 0, st>>>(n, state, seedStep, gin, gout, g);
}

void launchAdjointEventSeedRel(hipStream_t st, int n, const int* base, int rel, const int32_t* seedStep, const float2* gin,
                               const float2* gout, AdjBufs g) {
    k_adj_event_seed_rel<<<blocks256(n), 256, 0, st>>>(n, base, rel, seedStep, gin, gout, g);
}

void launchAdjointSetBase(hipStream_t st, int* base, int value) { k_adj_set_int<<<1, 1, 0, st>>>(base, value); }

void launchAdjointZero(hipStream_t st, int n, AdjBufs g) { k_adj_zero<<<blocks256(n), 256, 0, st>>>(n, g); }

void launchAdjointStep(hipStream_t st, const StepLaunch& L, const AdjStep& t, AdjBufs g, float overlapWeight) {
    if (L.batched) adjointStep<true>(st, L, t, g, overlapWeight);
    else adjointStep<false>(st, L, t, g, overlapWeight);
}

} // namespace swarm
