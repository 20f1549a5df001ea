// FP64 reference of the adjoint (TEST INFRASTRUCTURE ONLY: the checker of the GPU adjoint; only the tests load it).
//
// The model the adjoint differentiates: K crowd steps with every discrete decision frozen at the FP32 forward's
// values, read from the tape the GPU path records (include/swarm_adjoint.h): neighbour and contact sets, the
// headway leader and clamp branch, the caps and clamps, the goal anchors, the wall pushes (each a move along a
// fixed direction to the clearance of a fixed local linear model of the distance field), the wall planes, the
// park events. Along those decisions this file evaluates the continuous maps in double precision, written out
// here from the forward's definitions (cpu_crowd_sim.cpp: predict :655-777, projectSlot :781-831, finalPass
// :835-866; the wall model of resolveStatic :465-500 linearised per push), with forward-mode dual numbers for
// directional derivatives and plain doubles for central finite differences.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../include/swarm_adjoint.h"
#include "crowd_oracle.h"

extern "C" void oracle_pair_direction(uint32_t self, uint32_t other, float* dx, float* dy);

namespace {

// the model's constants (include/swarm/sim_params.hpp as the oracle folds them), as doubles of the float values
constexpr float R_PERC2_F = 0x1.096bbap-9f, SEP_STACKED_XY2_F = 1e-12f;
constexpr float R_SEP_F = 0.016f, R_SEP2_F = 0x1.0c6f7ap-12f, BODY_DIAM_F = 0.0128f;
constexpr float SEP_MARGIN_F = R_SEP_F - BODY_DIAM_F;
const double DT_STEP = double(0.01f);              // the position update's DT (float constant)
const double DT_RELAX = 0.01;                      // exp(-DT / tau): the decimal definition
const double T_UNIT = 17.490671;                   // seconds per model time unit
const double K_SEP = 100.0, SEP_MARGIN = double(SEP_MARGIN_F);
const double HEADWAY_RATE = double(0x1.d26afcp+5f);
const double OBSTACLE_RANGE = double(0.025f), W_OBS = 10.0;
const double GRAD_H = double(0.0015f), GRAD_SCALE = double(0x1.4d5556p+7f);
const double PROJ_MAX_SHIFT = double(0.05f), PARK = 1000.0;
const double STACKED_XY2 = double(SEP_STACKED_XY2_F);

// a dual number with N tangents (N = 0: a plain double)
template <int N> struct Dual {
    double v;
    double d[N > 0 ? N : 1];
};
template <int N> inline Dual<N> cst(double v) { Dual<N> r; r.v = v; for (int i = 0; i < N; ++i) r.d[i] = 0.0; return r; }
template <int N> inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v + b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v - b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> inline Dual<N> operator-(const Dual<N>& a) { Dual<N> r; r.v = -a.v; for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
template <int N> inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v * b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> inline Dual<N> operator*(const Dual<N>& a, double s) { Dual<N> r; r.v = a.v * s; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * s; return r; }
template <int N> inline Dual<N> operator*(double s, const Dual<N>& a) { return a * s; }
template <int N> inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
template <int N> inline Dual<N> operator/(const Dual<N>& a, double s) { return a * (1.0 / s); }
template <int N> inline Dual<N> operator+(const Dual<N>& a, double s) { Dual<N> r = a; r.v += s; return r; }
template <int N> inline Dual<N> operator-(const Dual<N>& a, double s) { Dual<N> r = a; r.v -= s; return r; }
template <int N> inline Dual<N> operator-(double s, const Dual<N>& a) { Dual<N> r = -a; r.v += s; return r; }
template <int N> inline Dual<N> dsqrt(const Dual<N>& a) {
    Dual<N> r; r.v = std::sqrt(a.v);
    const double k = r.v > 0.0 ? 0.5 / r.v : 0.0;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * k;
    return r;
}
template <int N> inline Dual<N> dexp(const Dual<N>& a) { Dual<N> r; r.v = std::exp(a.v); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * r.v; return r; }

template <int N> struct W2 { Dual<N> x, y; };
template <int N> inline W2<N> operator+(const W2<N>& a, const W2<N>& b) { return { a.x + b.x, a.y + b.y }; }
template <int N> inline W2<N> operator-(const W2<N>& a, const W2<N>& b) { return { a.x - b.x, a.y - b.y }; }
template <int N> inline W2<N> operator*(const W2<N>& a, const Dual<N>& s) { return { a.x * s, a.y * s }; }
template <int N> inline W2<N> operator*(const W2<N>& a, double s) { return { a.x * s, a.y * s }; }
template <int N> inline W2<N> operator/(const W2<N>& a, const Dual<N>& s) { return { a.x / s, a.y / s }; }
template <int N> inline Dual<N> dot(const W2<N>& a, const W2<N>& b) { return a.x * b.x + a.y * b.y; }
template <int N> inline Dual<N> dot(const W2<N>& a, double bx, double by) { return a.x * bx + a.y * by; }
template <int N> inline Dual<N> len(const W2<N>& a) { return dsqrt(dot(a, a)); }
template <int N> inline W2<N> cst2(double x, double y) { return { cst<N>(x), cst<N>(y) }; }
template <int N> inline W2<N> mul(const W2<N>& a, double x, double y) { return { a.x * x, a.y * y }; }   // componentwise by constants
template <int N> inline W2<N> along(double x, double y, const Dual<N>& s) { return { s * x, s * y }; }

struct SepLink { uint32_t j; int stacked; double dx, dy; };

struct StepRec {
    std::vector<uint32_t> sid;
    std::vector<swarm_predict_rec> pred;
    std::vector<swarm_final_rec> fin;
    std::vector<float> x, y, vx, vy;    // the FP32 input state (agent order)
    std::vector<uint32_t> nbr, cnt;
    std::vector<uint64_t> mask, cap, plane, ov;
    std::vector<std::vector<SepLink>> sep;   // per slot: the separation links of predict
};

} // namespace

struct adjref {
    int n = 0, P = 0, K = 0;
    double ovw = 0.0;
    // the SDF grid
    int sdfNX = 0, sdfNY = 0;
    double sdfOX = 0, sdfOY = 0, sdfInvH = 0;
    std::vector<float> sdf;
    std::vector<float> x0, y0, vx0, vy0, w0, r0, v0;
    double tauS = 0.5, scale = 1.0, gain = 20.0;
    std::vector<StepRec> steps;
    std::vector<swarm_exit_rec> exitRec;
    std::vector<double> stepErr;   // the last plain run: per step, the largest |x64 - x32| of its input state
    std::vector<int> stepArg;      // and its agent
};

namespace {

// bilinear distance field (the grid's value and exact weights; mapExactGrad's 2D form) in the dual type
template <int N> Dual<N> sdfAt(const adjref& A, const W2<N>& p) {
    const double fx = (p.x.v - A.sdfOX) * A.sdfInvH, fy = (p.y.v - A.sdfOY) * A.sdfInvH;
    if (fx < -0.5 || fy < -0.5 || fx > double(A.sdfNX) - 0.5 || fy > double(A.sdfNY) - 0.5) return cst<N>(1.0);
    int ix = int(std::floor(fx)), iy = int(std::floor(fy));
    ix = ix < 0 ? 0 : (ix > A.sdfNX - 2 ? A.sdfNX - 2 : ix);
    iy = iy < 0 ? 0 : (iy > A.sdfNY - 2 ? A.sdfNY - 2 : iy);
    const Dual<N> wx = (p.x - A.sdfOX) * A.sdfInvH - double(ix), wy = (p.y - A.sdfOY) * A.sdfInvH - double(iy);
    const float* r0 = A.sdf.data() + size_t(iy) * size_t(A.sdfNX);
    const float* r1 = r0 + A.sdfNX;
    const double c00 = r0[ix], c10 = r0[ix + 1], c01 = r1[ix], c11 = r1[ix + 1];
    const Dual<N> a = (1.0 - wx) * c00 + wx * c10, b = (1.0 - wx) * c01 + wx * c11;
    return (1.0 - wy) * a + wy * b;
}

// one step of the frozen model: X, Y, VX, VY (agent order) in and out; the overlap term into loss
template <int N>
void stepModel(const adjref& A, const StepRec& t, std::vector<Dual<N>>& X, std::vector<Dual<N>>& Y, std::vector<Dual<N>>& VX,
               std::vector<Dual<N>>& VY, const std::vector<Dual<N>>& V0, const std::vector<Dual<N>>& R, const Dual<N>& tau,
               const Dual<N>& relax, const Dual<N>& gain, Dual<N>& loss) {
    const int n = A.n, P = A.P;
    const size_t nn = static_cast<size_t>(n);
    std::vector<W2<N>> q0(nn), pv(nn);
    std::vector<Dual<N>> pc(nn);
    // predict
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        const swarm_predict_rec& pr = t.pred[size_t(i)];
        const uint32_t f = pr.flags;
        if (!(f & SWARM_REC_ACTIVE)) continue;
        if (f & SWARM_REC_PARKED) { q0[size_t(i)] = cst2<N>(PARK, PARK); pv[size_t(i)] = cst2<N>(0.0, 0.0); continue; }
        const uint32_t a = t.sid[size_t(i)];
        const W2<N> p{ X[a], Y[a] }, v{ VX[a], VY[a] };
        const Dual<N>& ri = R[a];
        W2<N> e;
        if (f & SWARM_REC_EFIXED) e = cst2<N>(pr.egoal_x, pr.egoal_y);
        else {
            const W2<N> d = cst2<N>(pr.anchor_x, pr.anchor_y) - p;
            e = d / len(d);
            if (f & SWARM_REC_GOAL_WALL) {   // e_goal's wall rule with the recorded wall direction
                const W2<N> nw = cst2<N>(pr.goal_wall_x, pr.goal_wall_y);
                const W2<N> e1 = e - nw * dot(e, nw);
                e = e1 / len(e1);
            }
        }
        W2<N> sep = cst2<N>(0.0, 0.0);
        for (const SepLink& L : t.sep[size_t(i)]) {
            const uint32_t b = t.sid[L.j];
            const W2<N> off = p - W2<N>{ X[b], Y[b] };
            const Dual<N> d = len(off);
            const Dual<N> S = ri + R[b] + SEP_MARGIN;
            const Dual<N> wgt = 1.0 - d / S;
            const Dual<N> mag = wgt * wgt * K_SEP;
            sep = sep + (L.stacked ? along<N>(L.dx, L.dy, mag) : off * (mag / d));
        }
        Dual<N> vc = cst<N>(0.0);
        if (f & SWARM_REC_VC_FREE) vc = V0[a] * double(pr.stair);
        else if (f & SWARM_REC_VC_RATE) {
            const uint32_t j = pr.leader, b = t.sid[j];
            const W2<N> off = p - W2<N>{ X[b], Y[b] };
            vc = (len(off) - (ri + R[b])) * HEADWAY_RATE;
        }
        const Dual<N> vDes = vc * len(e);
        W2<N> eDir = cst2<N>(0.0, 0.0);
        if (!(f & SWARM_REC_EDIR_ZERO)) {
            const W2<N> u = e + sep;
            eDir = u / len(u);
        }
        W2<N> aObs = cst2<N>(0.0, 0.0);
        if (f & SWARM_REC_OBST) {
            const Dual<N> dT = sdfAt(A, p);
            const Dual<N> m1 = sdfAt(A, W2<N>{ p.x + GRAD_H, p.y - GRAD_H }), m2 = sdfAt(A, W2<N>{ p.x - GRAD_H, p.y - GRAD_H });
            const Dual<N> m3 = sdfAt(A, W2<N>{ p.x - GRAD_H, p.y + GRAD_H }), m4 = sdfAt(A, W2<N>{ p.x + GRAD_H, p.y + GRAD_H });
            const W2<N> G{ (m1 - m2 - m3 + m4) * GRAD_SCALE, (m4 + m3 - m1 - m2) * GRAD_SCALE };
            const W2<N> n1 = G / len(G);
            aObs = n1 * ((OBSTACLE_RANGE - dT) * W_OBS * gain);
        }
        const W2<N> vEq = eDir * vDes + aObs * tau;
        const W2<N> nv1 = vEq + (v - vEq) * relax;
        W2<N> nv = nv1;
        if (f & SWARM_REC_ALONG) nv = nv1 - e * (dot(nv1, e) - vDes);
        if (f & SWARM_REC_CLAMP) nv = nv * (V0[a] / len(nv));
        W2<N> np = p + nv * DT_STEP;
        if (f & SWARM_REC_SWEEP) {
            np = p * (1.0 - double(pr.sweep_lo)) + np * double(pr.sweep_lo);
            if (f & SWARM_REC_SWEEP_VEL) nv = nv - along<N>(pr.sweep_dx, pr.sweep_dy, dot(nv, double(pr.sweep_dx), double(pr.sweep_dy)));
        }
        for (uint32_t k = 0; k < pr.n_events; ++k) {
            const swarm_wall_event& ev = pr.ev[k];
            const double gu = double(ev.gx) * ev.ux + double(ev.gy) * ev.uy;
            Dual<N> tt = cst<N>(ev.t0);
            if (gu != 0.0) tt = ev.t0 - dot(np - cst2<N>(ev.x0, ev.y0), double(ev.gx), double(ev.gy)) / gu;
            np = np + along<N>(ev.ax, ev.ay, ri - tt);
            if (ev.vel) nv = nv - along<N>(ev.ax, ev.ay, dot(nv, double(ev.ax), double(ev.ay)));
        }
        Dual<N> c = cst<N>(0.0);
        if (pr.plane_kind != 0u) {
            c = dot(np, double(pr.plane_nx), double(pr.plane_ny));
            if (pr.plane_kind == 2u) {
                const double gu = double(pr.pc_gx) * pr.pc_ux + double(pr.pc_gy) * pr.pc_uy;
                Dual<N> tt = cst<N>(pr.pc_t0);
                if (gu != 0.0) tt = pr.pc_t0 - dot(np - cst2<N>(pr.pc_x0, pr.pc_y0), double(pr.pc_gx), double(pr.pc_gy)) / gu;
                c = c - (tt - ri);
            }
        }
        q0[size_t(i)] = np;
        pv[size_t(i)] = nv;
        pc[size_t(i)] = c;
    }
    // the Jacobi projection
    std::vector<W2<N>> qa = q0, qb = q0;
    for (int it = 1; it <= P; ++it) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < n; ++i) {
            const swarm_predict_rec& pr = t.pred[size_t(i)];
            if (!(pr.flags & SWARM_REC_ACTIVE) || (pr.flags & SWARM_REC_PARKED)) { qb[size_t(i)] = qa[size_t(i)]; continue; }
            const uint32_t a = t.sid[size_t(i)];
            const W2<N>& qi = qa[size_t(i)];
            W2<N> u = cst2<N>(0.0, 0.0);
            int nc = 0;
            for (uint64_t m = t.mask[size_t(it - 1) * size_t(n) + size_t(i)]; m; m &= m - 1ull) {
                const int k = __builtin_ctzll(m);
                const uint32_t j = t.nbr[size_t(k) * size_t(n) + size_t(i)];
                const W2<N> off = qi - qa[j];
                const Dual<N> d2 = dot(off, off);
                const Dual<N> D = R[a] + R[t.sid[j]];
                if (d2.v <= STACKED_XY2) {
                    float dx, dy;
                    oracle_pair_direction(a, t.sid[j], &dx, &dy);
                    u = u + along<N>(dx, dy, (D - dsqrt(d2)) * 0.5);
                } else {
                    const Dual<N> d = dsqrt(d2);
                    u = u + off * ((D - d) * 0.5 / d);
                }
                ++nc;
            }
            W2<N> cc = u * (1.0 / double(nc > 1 ? nc : 1));
            if ((t.cap[size_t(i)] >> (it - 1)) & 1ull) cc = cc * (R[a] * PROJ_MAX_SHIFT / len(cc));
            W2<N> q = qi + cc;
            if ((t.plane[size_t(i)] >> (it - 1)) & 1ull) {
                const double nx = pr.plane_nx, ny = pr.plane_ny;
                q = q + along<N>(nx, ny, pc[size_t(i)] - dot(q, nx, ny));
            }
            qb[size_t(i)] = q;
        }
        std::swap(qa, qb);
    }
    // final: the last iteration's position, the wall pushes again, the output state
    std::vector<Dual<N>> Xn = X, Yn = Y, VXn = VX, VYn = VY;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        const swarm_predict_rec& pr = t.pred[size_t(i)];
        if (!(pr.flags & SWARM_REC_ACTIVE)) continue;
        const uint32_t a = t.sid[size_t(i)];
        if (pr.flags & SWARM_REC_PARKED) {
            Xn[a] = cst<N>(PARK); Yn[a] = cst<N>(PARK); VXn[a] = cst<N>(0.0); VYn[a] = cst<N>(0.0);
            continue;
        }
        W2<N> np = qa[size_t(i)], nv = pv[size_t(i)];
        const swarm_final_rec& fr = t.fin[size_t(i)];
        if (fr.flags & SWARM_REC_MOVED) {
            const W2<N> p{ X[a], Y[a] };
            if (fr.flags & SWARM_REC_SWEEP) {
                np = p * (1.0 - double(fr.sweep_lo)) + np * double(fr.sweep_lo);
                if (fr.flags & SWARM_REC_SWEEP_VEL) nv = nv - along<N>(fr.sweep_dx, fr.sweep_dy, dot(nv, double(fr.sweep_dx), double(fr.sweep_dy)));
            }
            for (uint32_t k = 0; k < fr.n_events; ++k) {
                const swarm_wall_event& ev = fr.ev[k];
                const double gu = double(ev.gx) * ev.ux + double(ev.gy) * ev.uy;
                Dual<N> tt = cst<N>(ev.t0);
                if (gu != 0.0) tt = ev.t0 - dot(np - cst2<N>(ev.x0, ev.y0), double(ev.gx), double(ev.gy)) / gu;
                np = np + along<N>(ev.ax, ev.ay, R[a] - tt);
                if (ev.vel) nv = nv - along<N>(ev.ax, ev.ay, dot(nv, double(ev.ax), double(ev.ay)));
            }
        }
        Xn[a] = np.x; Yn[a] = np.y; VXn[a] = nv.x; VYn[a] = nv.y;
    }
    // the overlap term at the step's end
    Dual<N> ovl = cst<N>(0.0);
    for (int i = 0; i < n; ++i) {
        for (uint64_t m = t.ov[size_t(i)]; m; m &= m - 1ull) {
            const int k = __builtin_ctzll(m);
            const uint32_t j = t.nbr[size_t(k) * size_t(n) + size_t(i)];
            if (uint32_t(i) >= j) continue;
            const uint32_t a = t.sid[size_t(i)], b = t.sid[j];
            const W2<N> off = W2<N>{ Xn[a], Yn[a] } - W2<N>{ Xn[b], Yn[b] };
            const Dual<N> o = R[a] + R[b] - len(off);
            ovl = ovl + o * o;
        }
    }
    loss = loss + ovl * A.ovw;
    X.swap(Xn); Y.swap(Yn); VX.swap(VXn); VY.swap(VYn);
}

// the whole run at parameters + the tangent directions (N > 0) or + eps * dir (N = 0)
// anchored: every step starts from the tape's FP32 input state (the tangents carry over), so the derivative is
// the product of the step Jacobians along the recorded trajectory, not along the FP64 model's own trajectory
template <int N>
Dual<N> runModel(const adjref& A, const double* dirs, int ndir, double eps, double* trajErr, bool anchored = false) {
    const int n = A.n;
    auto param = [&](double base, int index) {
        Dual<N> r = cst<N>(base);
        if (N == 0) { if (dirs) r.v += eps * dirs[index]; }
        else for (int k = 0; k < ndir && k < N; ++k) r.d[k] = dirs[size_t(k) * size_t(n + 3) + size_t(index)];
        return r;
    };
    const size_t nn = static_cast<size_t>(n);
    std::vector<Dual<N>> V0(nn), R(nn), X(nn), Y(nn), VX(nn), VY(nn);
    const Dual<N> tauS = param(A.tauS, n), scale = param(A.scale, n + 1), gain = param(A.gain, n + 2);
    for (int a = 0; a < n; ++a) {
        V0[size_t(a)] = param(A.v0[size_t(a)], a);
        R[size_t(a)] = scale * double(A.r0[size_t(a)]);
        X[size_t(a)] = cst<N>(A.x0[size_t(a)]); Y[size_t(a)] = cst<N>(A.y0[size_t(a)]);
        VX[size_t(a)] = cst<N>(A.vx0[size_t(a)]); VY[size_t(a)] = cst<N>(A.vy0[size_t(a)]);
    }
    const Dual<N> tau = tauS / T_UNIT;
    const Dual<N> relax = dexp(-(cst<N>(DT_RELAX) / tau));
    Dual<N> loss = cst<N>(0.0);
    double err = 0.0;
    std::vector<double> se(size_t(A.K), 0.0);
    std::vector<int> sa(size_t(A.K), -1);
    for (int s = 0; s < A.K; ++s) {
        const StepRec& t = A.steps[size_t(s)];
        for (int a = 0; a < n; ++a) {
            if (!(t.x[size_t(a)] < 999.0f)) continue;
            const double ex = std::fabs(X[size_t(a)].v - double(t.x[size_t(a)])), ey = std::fabs(Y[size_t(a)].v - double(t.y[size_t(a)]));
            const double e = std::max(ex, ey);
            if (e > se[size_t(s)]) { se[size_t(s)] = e; sa[size_t(s)] = a; }
        }
        err = std::max(err, se[size_t(s)]);
        if (anchored)
            for (int a = 0; a < n; ++a) {
                X[size_t(a)].v = t.x[size_t(a)]; Y[size_t(a)].v = t.y[size_t(a)];
                VX[size_t(a)].v = t.vx[size_t(a)]; VY[size_t(a)].v = t.vy[size_t(a)];
            }
        stepModel<N>(A, t, X, Y, VX, VY, V0, R, tau, relax, gain, loss);
    }
    if (N == 0 && !dirs) {
        const_cast<adjref&>(A).stepErr = se;
        const_cast<adjref&>(A).stepArg = sa;
    }
    for (int a = 0; a < n; ++a) {
        const swarm_exit_rec& e = A.exitRec[size_t(a)];
        if (!e.flags) continue;
        const W2<N> d = W2<N>{ X[size_t(a)], Y[size_t(a)] } - cst2<N>(e.anchor_x, e.anchor_y);
        const Dual<N> D = len(d) + double(e.base);
        loss = loss + D * D;
    }
    if (trajErr) *trajErr = err;
    return loss;
}

} // namespace

extern "C" {

adjref* adjref_create(const oracle_scene* sc, int n, int P, int K, float overlapWeight, const float* x, const float* y, const float* vx,
                      const float* vy, const float* w, const float* r0, const float* v0, float tauS, float radiusScale, float gain) {
    adjref* A = new adjref;
    A->n = n; A->P = P; A->K = K; A->ovw = overlapWeight;
    A->sdfNX = sc->sdfNX; A->sdfNY = sc->sdfNY; A->sdfOX = sc->sdfOX; A->sdfOY = sc->sdfOY; A->sdfInvH = sc->sdfInvH;
    A->sdf.assign(sc->sdf, sc->sdf + size_t(sc->sdfNX) * size_t(sc->sdfNY));
    A->x0.assign(x, x + n); A->y0.assign(y, y + n); A->vx0.assign(vx, vx + n); A->vy0.assign(vy, vy + n); A->w0.assign(w, w + n);
    A->r0.assign(r0, r0 + n); A->v0.assign(v0, v0 + n);
    A->tauS = tauS; A->scale = radiusScale; A->gain = gain;
    A->steps.resize(size_t(K));
    return A;
}

void adjref_destroy(adjref* A) { delete A; }

// one step's tape (host arrays of swarm_adjoint_tape); the separation links are decided here in FP32 from the
// tape's start positions and radii, as the forward decides them
int adjref_add_step(adjref* A, int step, const swarm_tape_view* v) {
    if (!A || step < 0 || step >= A->K) return -1;
    const int n = A->n;
    StepRec& t = A->steps[size_t(step)];
    t.sid.assign(v->sid, v->sid + n);
    t.pred.assign(v->pred, v->pred + n);
    t.fin.assign(v->fin, v->fin + n);
    t.x.assign(v->x, v->x + n); t.y.assign(v->y, v->y + n);
    t.vx.assign(v->vx, v->vx + n); t.vy.assign(v->vy, v->vy + n);
    t.nbr.assign(v->nbr, v->nbr + size_t(SWARM_TAPE_NBR) * size_t(n));
    t.cnt.assign(v->nbr_count, v->nbr_count + n);
    t.mask.assign(v->mask, v->mask + size_t(A->P) * size_t(n));
    t.cap.assign(v->cap_bits, v->cap_bits + n);
    t.plane.assign(v->plane_bits, v->plane_bits + n);
    t.ov.assign(v->overlap, v->overlap + n);
    // separation links: a uniform grid of the perception radius over the active slots
    t.sep.assign(size_t(n), {});
    const float cell = std::sqrt(R_PERC2_F);
    float minx = 1e30f, miny = 1e30f, maxx = -1e30f, maxy = -1e30f;
    std::vector<int> act;
    for (int i = 0; i < n; ++i)
        if ((t.pred[size_t(i)].flags & SWARM_REC_ACTIVE)) {
            act.push_back(i);
            minx = std::min(minx, v->sx[i]); miny = std::min(miny, v->sy[i]);
            maxx = std::max(maxx, v->sx[i]); maxy = std::max(maxy, v->sy[i]);
        }
    if (act.empty()) return 0;
    const int gx = std::max(1, int((maxx - minx) / cell) + 1), gy = std::max(1, int((maxy - miny) / cell) + 1);
    std::vector<std::vector<int>> grid(size_t(gx) * size_t(gy));
    auto cx = [&](float x) { int c = int((x - minx) / cell); return c < 0 ? 0 : (c >= gx ? gx - 1 : c); };
    auto cy = [&](float y) { int c = int((y - miny) / cell); return c < 0 ? 0 : (c >= gy ? gy - 1 : c); };
    for (int i : act) grid[size_t(cy(v->sy[i])) * size_t(gx) + size_t(cx(v->sx[i]))].push_back(i);
#pragma omp parallel for schedule(dynamic, 64)
    for (int ii = 0; ii < int(act.size()); ++ii) {
        const int i = act[size_t(ii)];
        if (t.pred[size_t(i)].flags & SWARM_REC_PARKED) continue;
        const int x0 = cx(v->sx[i]), y0 = cy(v->sy[i]);
        for (int yy = y0 - 1; yy <= y0 + 1; ++yy)
            for (int xx = x0 - 1; xx <= x0 + 1; ++xx) {
                if (xx < 0 || yy < 0 || xx >= gx || yy >= gy) continue;
                for (int j : grid[size_t(yy) * size_t(gx) + size_t(xx)]) {
                    if (j == i) continue;
                    const float ox = v->sx[i] - v->sx[j], oy = v->sy[i] - v->sy[j];
                    const float d2 = ox * ox + oy * oy;
                    if (!(d2 < R_PERC2_F)) continue;
                    const float S = (v->sr[i] + v->sr[j]) + SEP_MARGIN_F;
                    const float s2 = S == R_SEP_F ? R_SEP2_F : S * S;
                    if (d2 >= s2) continue;
                    SepLink L{ uint32_t(j), d2 <= SEP_STACKED_XY2_F ? 1 : 0, 0.0, 0.0 };
                    if (L.stacked) {
                        float dx, dy;
                        oracle_pair_direction(t.sid[size_t(i)], t.sid[size_t(j)], &dx, &dy);
                        L.dx = dx; L.dy = dy;
                    }
                    t.sep[size_t(i)].push_back(L);
                }
            }
    }
    return 0;
}

void adjref_set_exit(adjref* A, const swarm_exit_rec* e) { A->exitRec.assign(e, e + A->n); }

// the loss at parameters + eps * dir (dir: n + 3 values: v0, tau_s, radius scale, wall gain; NULL: the base
// point); trajErr: the largest |x64 - x32| of the run's input states against the tape's FP32 states
double adjref_loss(adjref* A, const double* dir, double eps, double* trajErr) {
    return runModel<0>(*A, dir, dir ? 1 : 0, eps, trajErr).v;
}

// the last base-point run's per-step trajectory error and its agent (K values each)
void adjref_step_errors(const adjref* A, double* err, int* agent) {
    for (size_t s = 0; s < A->stepErr.size(); ++s) { err[s] = A->stepErr[s]; agent[s] = A->stepArg[s]; }
}

// directional derivatives of the loss along ndir directions (ndir x (n + 3)), forward mode
// (anchored: along the tape's FP32 trajectory)
void adjref_directional(adjref* A, int ndir, const double* dirs, double* out, double* loss, int anchored) {
    for (int k0 = 0; k0 < ndir; k0 += 16) {
        const int m = std::min(16, ndir - k0);
        const Dual<16> L = runModel<16>(*A, dirs + size_t(k0) * size_t(A->n + 3), m, 0.0, nullptr, anchored != 0);
        for (int k = 0; k < m; ++k) out[k0 + k] = L.d[k];
        if (loss) *loss = L.v;
    }
}

} // extern "C"
