"""Calibration of the desired-speed distribution and the relaxation time from exit-flow curves.

The workload: many scenarios of the seated-bowl quarter (60 rows, 4 gangways, 2 exits of 2.4 m), each with its
own seeded placement and its own fixed standard-normal draws z_i, run as one batch. An agent's desired speed is
v0_i = clip(mean + sd * z_i, 0.5, 2.0) m/s. Target exit-flow curves come from a forward run at hidden parameters;
gradient descent on (mean, sd, tau) fits them, with gradients from the batched adjoint of K steps.

Exit flow: an agent's exit time tau_i is the step where its crossing into its exit's rectangle happened, refined
by where the last step's segment enters the rectangle (the last slab entry along the segment); the smoothed flow
of exit e at time c is the sum over its agents of a Gaussian of c - tau_i (5 s full width at half maximum), sampled
every 2.5 s. The loss is the sum over scenarios, exits and samples of the squared difference to the target.

Gradients: dL/dtau_i from the flow, then the position adjoints of each crossing's two endpoints (the adjoint's
evacuation-crossing terms), the batched adjoint backwards to dL/dv0_i and dL/dtau, and the chain rule to the three
parameters. The crowd's dynamics are unstable over tens of steps (contacts), so the exact pathwise gradient over K
steps is dominated by exponentially growing terms; the adjoint is truncated to a window of steps (each gradient path
spans at most 'window' steps), the usual estimator for long horizons of unstable dynamics.
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

    def __init__(self, scenarios=64, agents=15832, K=400, every=25, window=25, seed=1337):
        self.layout = sh.LAYOUT_BOWL_QUARTER
        self.scene = sh.Scene.layout(self.layout)
        self.S, self.N, self.K, self.every, self.window = scenarios, agents, K, every, window
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

    def run(self, theta):
        """Loss and gradient with respect to (mean m/s, sd m/s, tau s)."""
        f = self.forward(theta)
        R = f["F"] - self.target
        loss = float(np.sum(R ** 2))
        idx, g, ex = f["idx"], f["g"], f["ex"]
        # dL/dtau_i = sum_samples 2 R * dG(c - tau)/dtau
        dG = g * (self.samples[None, :] - f["tau"][idx, None]) / SIGMA_S ** 2
        dtau = np.zeros(f["tau"].size)
        dtau[idx] = np.sum(2.0 * R[self.scen[idx], ex[idx], :] * dG, axis=1)
        gin = (dtau[:, None] * f["d_in"]).astype(np.float32)
        gout = (dtau[:, None] * f["d_bf"]).astype(np.float32)
        st = np.where(f["ok"], f["step"], -1).astype(np.int32)
        self.sim.set_adjoint_terms(st, gin, gout, exit_term=False, window=self.window)
        gv0, gglob = self.sim.adjoint_backward()
        self.sim.set_adjoint_terms(None, None, None, exit_term=True, window=0)
        gv = gv0.astype(np.float64) * SPEED_TO_MODEL * f["free"]
        grad = np.array([gv.sum(), (gv * self.z).sum(), gglob.reshape(-1, 3)[:, 0].sum()])
        return loss, grad


def calibrate(hidden=(1.20, 0.30, 0.5), start=(1.34, 0.26, 0.7), iterations=30, lr=0.03, **kw):
    """Fits (mean, sd, tau) to the hidden parameters' exit-flow curves by Adam on parameters scaled by their
    magnitudes. Returns the history and timings."""
    cal = Calibration(**kw)
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
        step = lr * (m / (1 - 0.9 ** it)) / (np.sqrt(v / (1 - 0.999 ** it)) + 1e-12)
        x = x - step
        x[1] = max(x[1], 0.01 / scale[1])
        x[2] = max(x[2], 0.05 / scale[2])
        hist.append(dict(iteration=it, loss=loss, theta=theta.tolist(), grad=grad.tolist()))
    elapsed = time.perf_counter() - t0
    theta = x * scale
    return dict(theta=theta.tolist(), hidden=list(hidden), error=(theta - np.array(hidden)).tolist(), history=hist,
                seconds=elapsed, iterations=iterations, agents=cal.S * cal.N, K=cal.K)
