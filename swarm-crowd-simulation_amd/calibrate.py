"""Calibration of the desired-speed distribution and the relaxation time from exit-flow curves.

The workload: many scenarios of the seated-bowl quarter (60 rows, 4 gangways, 2 exits of 2.4 m), each with its
own seeded placement and its own fixed standard-normal draws z_i, run as one batch. An agent's desired speed is
v0_i = clip(mean + sd * z_i, 0.5, 2.0) m/s. Target exit-flow curves come from a forward run at hidden parameters;
gradient descent on (mean, sd, tau) fits them, with gradients from the batched adjoint of K steps.

Exit flow: an agent's exit time tau_i is the step where its crossing into its exit's rectangle happened, refined
by where the last step's segment enters the rectangle (the last slab entry along the segment); the smoothed flow
of exit e at time c is the sum over its agents of a Gaussian of c - tau_i (5 s full width at half maximum), sampled
every 2.5 s. The loss is the sum over exits and samples of the squared difference between the scenario-averaged
flow and the target's: each scenario's own curve carries the crowd's chaotic scatter (a 1e-4 change of one parameter
already moves it by about as much as the whole calibration signal), which the ensemble average suppresses.

Gradients: dL/dtau_i from the flow, then the position adjoints of each crossing's two endpoints (the adjoint's
evacuation-crossing terms), the batched adjoint backwards to dL/dv0_i and dL/dtau, and the chain rule to the three
parameters. The crowd's dynamics are unstable over tens of steps (contacts), so the exact pathwise gradient over K
steps is dominated by exponentially growing terms; a gradient path must stay within a window of about 25 steps.

The default estimator ('headway') keeps that window and still sees the exits' capacity. Per scenario and exit, sort
the crossings by time and write tau_k = tau_first + sum_{j <= k} h_j with headways h_j = tau_j - tau_{j-1}; then
dL/dtheta = sum_j c_j dh_j/dtheta (+ c_1 dtau_first/dtheta) with c_j = sum_{k >= j} dL/dtau_k, the loss weight of
every crossing from j on. A headway is set by the few seconds before it, so its derivative is estimated by the
exact adjoint truncated to the window, over the same window for both crossings of the pair (pairs whose four
crossing points fall in one window block: the adjoint is cleared at every multiple of the window). In seeds: +c_j
at crossing j and -c_j at crossing j - 1 for every such pair. Plain per-crossing seeds dL/dtau_k under the same
truncation see only shifts of crossing times within the window, which leave a saturated exit's flow unchanged, and
miss how each parameter changes the exit's discharge rate (measured: they point away from the finite differences of
the loss). Alternatives: 'tau' (plain seeds, truncated exact adjoint) and 'local' (the adjoint's local mode: each
agent's own chain, cut at every step where it touched another agent).
"""
import time

import numpy as np

import swarm_hip as sh

FWHM_S = 5.0
SIGMA_S = FWHM_S / (2.0 * np.sqrt(2.0 * np.log(2.0)))
SAMPLE_S = 2.5
STEP_S = 0.01 * sh.S_PER_UNIT          # seconds per step
SPEED_TO_MODEL = sh.S_PER_UNIT / sh.M_PER_UNIT


def crossing_times(step, exit_id, inside, before, rects, n_exits):
    """Per agent: the exit time (s), whether it counts, and its partial derivatives with respect to the crossing's
    inside and before positions (n x 2 each)."""
    n = step.size
    tau = np.zeros(n)
    d_in = np.zeros((n, 2))
    d_bf = np.zeros((n, 2))
    ok = (step >= 1) & (exit_id >= 0) & (exit_id < n_exits)
    idx = np.nonzero(ok)[0]
    if idx.size:
        r = rects[exit_id[idx]]                      # lox, loy, hix, hiy
        b = before[idx].astype(np.float64)
        i = inside[idx].astype(np.float64)
        dl = i - b
        alpha = np.zeros((idx.size, 2))
        face = np.zeros((idx.size, 2))
        lo, hi = r[:, 0:2], r[:, 2:4]
        below, above = b < lo, b > hi
        face = np.where(below, lo, np.where(above, hi, b))
        with np.errstate(divide="ignore", invalid="ignore"):
            alpha = np.where(below | above, (face - b) / dl, 0.0)
        alpha = np.nan_to_num(alpha)
        k = np.argmax(alpha, axis=1)                 # the last slab the segment enters
        rows = np.arange(idx.size)
        a = np.clip(alpha[rows, k], 0.0, 1.0)
        tau[idx] = (step[idx] - 1 + a) * STEP_S
        crossing = (alpha[rows, k] > 0.0) & (alpha[rows, k] < 1.0) & (dl[rows, k] != 0.0)
        den = np.where(crossing, dl[rows, k], 1.0)
        d_in[idx, k] = np.where(crossing, -a / den * STEP_S, 0.0)
        d_bf[idx, k] = np.where(crossing, (a - 1.0) / den * STEP_S, 0.0)
    return tau, ok, d_in, d_bf


def flows(tau, ok, exit_id, scen, n_scen, n_exits, samples):
    """The smoothed flow (persons / s) per scenario, exit and sample time, and each agent's Gaussian weights."""
    F = np.zeros((n_scen, n_exits, samples.size))
    idx = np.nonzero(ok)[0]
    g = np.exp(-0.5 * ((samples[None, :] - tau[idx, None]) / SIGMA_S) ** 2) / (SIGMA_S * np.sqrt(2.0 * np.pi))
    np.add.at(F, (scen[idx], exit_id[idx]), g)
    return F, idx, g


class Calibration:
    """The batched scenarios and their fixed draws; run(theta) gives the loss and its gradient."""

    def __init__(self, scenarios=64, agents=15832, K=400, every=25, estimator="headway", window=25, seed=1337):
        if estimator not in ("headway", "tau", "local"):
            raise ValueError("estimator: 'headway', 'tau' or 'local'")
        if estimator != "local" and window <= 0:
            raise ValueError("the exact adjoint over %d steps of this crowd needs a window > 0" % K)
        self.layout = sh.LAYOUT_BOWL_QUARTER
        self.scene = sh.Scene.layout(self.layout)
        self.S, self.N, self.K, self.every = scenarios, agents, K, every
        self.estimator = estimator
        self.window = 0 if estimator == "local" else window
        self.local = estimator == "local"
        parts = [sh.place_agents(self.layout, agents, agents, seed + s) for s in range(scenarios)]
        self.base = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
        self.z = np.random.default_rng(seed).standard_normal(scenarios * agents)
        self.scen = np.repeat(np.arange(scenarios), agents)
        self.sim = sh.BatchSim([self.scene], np.zeros(scenarios, np.int32), np.full(scenarios, agents, np.int32))
        self.E = self.scene.info.E
        self.rects = self.scene.tables()["exit_rects"].reshape(-1, 4)[: self.E].astype(np.float64)
        self.samples = np.arange(SAMPLE_S, K * STEP_S + 1e-9, SAMPLE_S)
        self.target = None

    def speeds(self, mean, sd):
        raw = mean + sd * self.z
        v = np.clip(raw, 0.5, 2.0)
        free = (raw > 0.5) & (raw < 2.0)
        return (v * SPEED_TO_MODEL).astype(np.float32), free

    def forward(self, theta):
        mean, sd, tau_s = (float(t) for t in theta)
        v0, free = self.speeds(mean, sd)
        a = dict(self.base)
        a["v0"] = v0
        self.sim.set_model(tau_s, 20.0)
        self.sim.upload(a)
        self.sim.adjoint_forward(self.K, self.every, 1.0, 0.0)
        step, ex, ins, bef = self.sim.evac_points()
        tau, ok, d_in, d_bf = crossing_times(step, ex, ins, bef, self.rects, self.E)
        F, idx, g = flows(tau, ok, ex, self.scen, self.S, self.E, self.samples)
        return dict(F=F, tau=tau, ok=ok, d_in=d_in, d_bf=d_bf, idx=idx, g=g, step=step, ex=ex, free=free)

    def set_target(self, theta):
        self.target = self.forward(theta)["F"]

    def loss(self, theta):
        """The loss alone (one forward)."""
        R = self.forward(theta)["F"].mean(axis=0) - self.target.mean(axis=0)
        return float(np.sum(R ** 2))

    def headway_seeds(self, dtau, f):
        """Per-crossing seeds of the headway estimator: +c_j at crossing j and -c_j at j - 1 for consecutive
        crossings of one exit whose crossing points share a window block, c_j the suffix sum of dL/dtau from j on;
        +c_1 at each exit's first crossing."""
        idx = f["idx"]
        seed = np.zeros_like(dtau)
        if idx.size == 0:
            return seed
        key = self.scen[idx] * self.E + f["ex"][idx]
        order = np.lexsort((f["tau"][idx], key))
        ids, keys = idx[order], key[order]
        d = dtau[ids]
        rc = np.cumsum(d[::-1])[::-1]                       # suffix sums over the whole order
        ends = np.searchsorted(keys, keys, side="right")    # one past each group's last crossing
        c = rc - np.where(ends < d.size, rc[np.minimum(ends, d.size - 1)], 0.0)
        st = f["step"][ids]
        # the adjoint is cleared at state b * window; a seed at state s lives in block (s - 1) // window
        W = self.window
        pair = (keys[1:] == keys[:-1]) & ((st[:-1] - 2) // W == (st[1:] - 1) // W)
        np.add.at(seed, ids[1:][pair], c[1:][pair])
        np.add.at(seed, ids[:-1][pair], -c[1:][pair])
        first = np.ones(ids.size, bool)
        first[1:] = keys[1:] != keys[:-1]
        np.add.at(seed, ids[first], c[first])
        return seed

    def run(self, theta):
        """Loss and gradient with respect to (mean m/s, sd m/s, tau s)."""
        f = self.forward(theta)
        R = f["F"].mean(axis=0) - self.target.mean(axis=0)     # exits x samples
        loss = float(np.sum(R ** 2))
        idx, g, ex = f["idx"], f["g"], f["ex"]
        # dL/dtau_i = sum_samples 2 R dG(c - tau)/dtau / S
        dG = g * (self.samples[None, :] - f["tau"][idx, None]) / SIGMA_S ** 2
        dtau = np.zeros(f["tau"].size)
        dtau[idx] = np.sum(2.0 * R[ex[idx], :] * dG, axis=1) / self.S
        seed = self.headway_seeds(dtau, f) if self.estimator == "headway" else dtau
        gin = (seed[:, None] * f["d_in"]).astype(np.float32)
        gout = (seed[:, None] * f["d_bf"]).astype(np.float32)
        st = np.where(f["ok"], f["step"], -1).astype(np.int32)
        self.sim.set_adjoint_terms(st, gin, gout, exit_term=False, window=self.window, local=self.local)
        gv0, gglob = self.sim.adjoint_backward()
        self.sim.set_adjoint_terms(None, None, None, exit_term=True, window=0)
        gv = (gv0.astype(np.float64) * SPEED_TO_MODEL * f["free"]).reshape(self.S, self.N)
        z = self.z.reshape(self.S, self.N)
        # each scenario's share of the gradient (its agents' crossings and its own relaxation-time adjoint)
        self.per_scenario = np.stack([gv.sum(axis=1), (gv * z).sum(axis=1), gglob.reshape(-1, 3)[:, 0]], axis=1)
        return loss, self.per_scenario.sum(axis=0)


def calibrate(hidden=(1.20, 0.30, 0.5), start=(1.34, 0.26, 0.7), iterations=80, lr=0.02, average=0.25, cal=None, log=None,
              **kw):
    """Fits (mean, sd, tau) to the hidden parameters' exit-flow curves by Adam on parameters scaled by their
    magnitudes, the step size decaying as a cosine over the iterations; the estimate is the mean of the last
    'average' fraction of the iterates (Polyak averaging: the loss and its gradient carry the crowd's chaotic
    scatter, and mean speed and relaxation time trade off along a narrow valley). Returns the history and timings."""
    cal = cal if cal is not None else Calibration(**kw)
    cal.set_target(hidden)
    scale = np.array([1.0, 0.3, 0.5])
    x = np.array(start, np.float64) / scale
    m = np.zeros(3)
    v = np.zeros(3)
    hist = []
    t0 = time.perf_counter()
    for it in range(1, iterations + 1):
        theta = x * scale
        loss, grad = cal.run(theta)
        gx = grad * scale
        m = 0.9 * m + 0.1 * gx
        v = 0.999 * v + 0.001 * gx * gx
        rate = lr * 0.5 * (1.0 + np.cos(np.pi * (it - 1) / iterations))
        step = rate * (m / (1 - 0.9 ** it)) / (np.sqrt(v / (1 - 0.999 ** it)) + 1e-12)
        x = x - step
        x[1] = max(x[1], 0.01 / scale[1])
        x[2] = max(x[2], 0.05 / scale[2])
        hist.append(dict(iteration=it, loss=loss, theta=theta.tolist(), grad=grad.tolist()))
        if log is not None:
            log(hist[-1])
    elapsed = time.perf_counter() - t0
    tail = max(1, int(round(average * iterations)))
    theta = np.mean([h["theta"] for h in hist[-tail:]], axis=0)
    final_loss, _ = cal.run(theta)
    return dict(theta=theta.tolist(), hidden=list(hidden), error=(theta - np.array(hidden)).tolist(),
                start=list(start), final_loss=final_loss, history=hist, seconds=elapsed, iterations=iterations,
                agents=cal.S * cal.N, K=cal.K)


def throughput(cal, theta, looped=8, repeats=2):
    """Agent-steps/s of one gradient evaluation (forward with checkpoints + adjoint) batched over all scenarios,
    of the forward alone, and of the same work looped one scenario at a time through an unbatched context (the
    first 'looped' scenarios, each uploaded, run forward and backward with its own crossing terms)."""
    mean, sd, tau_s = theta
    v0, _ = cal.speeds(mean, sd)
    fwd, full = [], []
    for _ in range(repeats):
        t0 = time.perf_counter()
        f = cal.forward(theta)
        t1 = time.perf_counter()
        cal.run(theta)
        t2 = time.perf_counter()
        fwd.append(t1 - t0)
        full.append(t2 - t1)
    plain = sh.Sim(cal.scene, cal.N)
    plain.set_model(tau_s, 20.0)
    one = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        for s in range(looped):
            sl = slice(s * cal.N, (s + 1) * cal.N)
            a = {k: v[sl] for k, v in cal.base.items()}
            a["v0"] = v0[sl]
            plain.upload(a)
            plain.adjoint_forward(cal.K, cal.every, 1.0, 0.0)
            step, ex, ins, bef = plain.evac_points()
            tau, ok, d_in, d_bf = crossing_times(step, ex, ins, bef, cal.rects, cal.E)
            st = np.where(ok, step, -1).astype(np.int32)
            w = np.full(cal.N, 1e-3)
            plain.set_adjoint_terms(st, (w[:, None] * d_in).astype(np.float32), (w[:, None] * d_bf).astype(np.float32),
                                    exit_term=False, window=cal.window, local=cal.local)
            plain.adjoint_backward()
        one.append((time.perf_counter() - t0) / looped)
    work = cal.S * cal.N * cal.K
    tf, tg, t1s = min(fwd), min(full), min(one)
    return dict(forward_s=tf, gradient_s=tg, adjoint_vs_forward=(tg - tf) / tf,
                batched_agent_steps_per_s=work / tg, looped_agent_steps_per_s=cal.N * cal.K / t1s,
                looped_s_per_scenario=t1s, batched_vs_looped=(work / tg) / (cal.N * cal.K / t1s),
                looped_sample=looped)


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="exit-flow calibration of (speed mean, speed sd, relaxation time)")
    ap.add_argument("--scenarios", type=int, default=64)
    ap.add_argument("--agents", type=int, default=15832)
    ap.add_argument("--K", type=int, default=400)
    ap.add_argument("--every", type=int, default=25)
    ap.add_argument("--iterations", type=int, default=80)
    ap.add_argument("--lr", type=float, default=0.02)
    ap.add_argument("--window", type=int, default=25)
    ap.add_argument("--estimator", default="headway", choices=("headway", "tau", "local"))
    ap.add_argument("--looped", type=int, default=8)
    args = ap.parse_args()
    cal = Calibration(scenarios=args.scenarios, agents=args.agents, K=args.K, every=args.every,
                      estimator=args.estimator, window=args.window)
    res = calibrate(iterations=args.iterations, lr=args.lr, cal=cal,
                    log=lambda h: print(json.dumps({k: (round(v, 5) if isinstance(v, float) else
                                                        [round(x, 5) for x in v] if isinstance(v, list) else v)
                                                    for k, v in h.items()}), flush=True))
    # the loss's scatter floor: the hidden parameters moved by 1e-4 (the crowd's chaos alone)
    res["noise_floor"] = [cal.loss(np.array(res["hidden"]) + d) for d in np.eye(3) * 1e-4]
    res["throughput"] = throughput(cal, res["hidden"], looped=args.looped)
    res["estimator"] = args.estimator if args.estimator == "local" else "%s, window %d" % (args.estimator, args.window)
    print(json.dumps({k: v for k, v in res.items() if k != "history"}), flush=True)
