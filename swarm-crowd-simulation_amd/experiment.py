"""Evacuation experiments with the safety metrics, and the reference's on-disk formats.

The reference's headless experiment (tools/experiment/run_experiment.cpp) steps the pipeline with its metric
passes after every step, reads back every K steps and writes timeseries.csv, heatmap.csv and a summary; its
benchmarks keep agent states as raw GpuBoid arrays (tools/benchmark/bench_common.hpp:109-147). This module does
the same on a floor of this build: the step and the metric passes run on the GPU (swarm_hip); here are only the
readbacks and their conversion to real units.

- timeseries.csv (run_experiment.cpp:929-941): step, t_model, t_real_s, active, the cumulative evacuated per exit
  slot (the floor's exits, then 'other'), and the largest density over the floor's included cells of at least
  0.5 m^2 at that step (run_experiment.cpp:545-551; a rectangle floor is one zone, 'floor').
- heatmap.csv (run_experiment.cpp:942-961, cells :706-716): per included cell its centre in metres, walkable area,
  peak density, time at or above each threshold, occupied time, mean speed while occupied, exposure and
  person-seconds.
- summary.json: t50 / t90 / t100 as the reference takes them (run_experiment.cpp:660-669: the step by which
  ceil(p N0) agents have left, from the per-agent evacuation steps), per exit slot counts, the floor's peak
  density and exposure, and the hotspot cells in the reference's ranking (exposure, then time above LoS F, LoS E,
  peak density; run_experiment.cpp:744-751).
- Snapshots: n GpuBoid records of 32 bytes (position x, y, z, w; velocity x, y, z, w = speed), little-endian
  float32, model units; this build's floors are flat, so z = 0 (gpu_types.hpp:17-20).
"""
import json
import math
import os

import numpy as np

import swarm_hip as sh

DT_MODEL = float(np.float32(0.01))                               # sim_params.hpp:105, widened
DT_REAL_S = DT_MODEL * float(np.float32(sh.S_PER_UNIT))          # sim_params.hpp:389 (units::seconds)
MASS_ONE = 65536                                                 # metrics.hpp:78
THRESHOLD_NAMES = ("los_e", "los_f", "crush")                    # metrics.hpp:94
MIN_WALKABLE_M2 = 0.10                                           # metrics.hpp:138: a cell is included
PEAK_MIN_WALKABLE_M2 = 0.5                                       # metrics.hpp:143: peaks over cells this large
SNAPSHOT_RECORD = np.dtype([("px", "<f4"), ("py", "<f4"), ("pz", "<f4"), ("pw", "<f4"),
                            ("vx", "<f4"), ("vy", "<f4"), ("vz", "<f4"), ("vw", "<f4")])


def sum64(cells, k):
    """The 64-bit accumulator at words k, k + 1 of each cell record (metrics.hpp:174)."""
    return cells[:, k].astype(np.uint64) | (cells[:, k + 1].astype(np.uint64) << np.uint64(32))


def floor_max_density(last_mass, walkable):
    """persons/m^2: the densest included cell of at least PEAK_MIN_WALKABLE_M2 at the last recorded step."""
    big = (walkable >= np.float32(MIN_WALKABLE_M2)) & (walkable >= PEAK_MIN_WALKABLE_M2)
    if not big.any():
        return 0.0
    return float(np.max(last_mass[big].astype(np.float64) / MASS_ONE / walkable[big].astype(np.float64)))


def cell_rows(cells, walkable, grid, dt=DT_REAL_S):
    """The heatmap's per-cell figures (run_experiment.cpp:706-716) for every included cell, in cell order."""
    rx, ry, cpu, ox, oy = grid
    inc = np.nonzero(walkable >= np.float32(MIN_WALKABLE_M2))[0]
    area = walkable[inc].astype(np.float64)
    g = cells[inc]
    mass = sum64(g, 8)
    speed = sum64(g, 10)
    expo = sum64(g, 12)
    with np.errstate(invalid="ignore", divide="ignore"):
        mean_speed = np.where(mass > 0, speed.astype(np.float64) / mass.astype(np.float64), np.nan)
    ix, iy = inc % rx, inc // rx
    # cell ix is centred at g = ix, g = xy * cpu + off (model units): x = (ix - off) / cpu
    return dict(ix=ix, iy=iy, x_m=(ix - float(ox)) / float(cpu) * sh.M_PER_UNIT, y_m=(iy - float(oy)) / float(cpu) * sh.M_PER_UNIT,
                walkable_m2=area, peak_density_p_m2=g[:, 2].astype(np.float64) / MASS_ONE / area,
                time_above=np.stack([g[:, 4 + k] * dt for k in range(3)], 1), occupied_s=g[:, 3] * dt,
                mean_speed_occupied_m_s=mean_speed, exposure_person_s=expo.astype(np.float64) / MASS_ONE * dt,
                person_s=mass.astype(np.float64) / MASS_ONE * dt, cell=inc)


def evacuation_times(evac_steps, n0, dt=DT_REAL_S):
    """(t50, t90, t100) in steps and seconds: the step by which ceil(p N0) agents have left (-1 / NaN if never)."""
    left = np.sort(evac_steps[evac_steps != 0])
    out = {}
    for p, name in ((0.5, "t50"), (0.9, "t90"), (1.0, "t100")):
        k = math.ceil(p * n0)
        s = int(left[k - 1]) if 1 <= k <= left.size else -1
        out[name + "_step"] = s
        out[name + "_s"] = s * dt if s >= 0 else float("nan")
    return out


def hotspots(rows, top=10):
    """The reference's ranking: exposure, then time above LoS F, then LoS E, then peak density (all descending)."""
    order = np.lexsort((-rows["peak_density_p_m2"], -rows["time_above"][:, 0], -rows["time_above"][:, 1], -rows["exposure_person_s"]))
    return order[:top]


def save_snapshot(path, state):
    """state: dict of x, y, vx, vy, speed, w (agent order, model units) -> n GpuBoid records."""
    n = np.asarray(state["x"]).size
    rec = np.zeros(n, SNAPSHOT_RECORD)
    rec["px"], rec["py"], rec["pw"] = state["x"], state["y"], state["w"]
    rec["vx"], rec["vy"], rec["vw"] = state["vx"], state["vy"], state.get("speed", np.zeros(n, np.float32))
    rec.tofile(path)


def load_snapshot(path, n):
    """n GpuBoid records -> dict of x, y, vx, vy, speed, w; a file of another size is an error (bench_common.hpp:127-135)."""
    if os.path.getsize(path) != n * SNAPSHOT_RECORD.itemsize:
        raise ValueError(f"missing or wrong-size snapshot {path}: {os.path.getsize(path)} bytes for {n} agents")
    rec = np.fromfile(path, SNAPSHOT_RECORD, count=n)
    f = lambda k: np.ascontiguousarray(rec[k], np.float32)
    return dict(x=f("px"), y=f("py"), w=f("pw"), vx=f("vx"), vy=f("vy"), speed=f("vw"))


def snapshot_path(directory, state, n):
    return os.path.join(directory, f"{state}_{n}.bin")   # bench_common.hpp:112-114


class Experiment:
    """One evacuation run on a floor: the GPU step with the metric passes, read back every 'every' steps."""

    def __init__(self, layout, active, total, seed=1337, every=10, use_graph=True):
        self.scene = sh.Scene.layout(layout)
        self.E = self.scene.info.E
        self.sim = sh.Sim(self.scene, total, use_graph=use_graph)
        self.agents = sh.place_agents(layout, active, total, seed)
        self.sim.upload(self.agents)
        self.sim.metrics_enable(True)
        self.grid = self.sim.metrics_grid()
        self.walkable, _ = self.sim.metric_cell_info()
        self.n0 = int((self.agents["w"] != 0).sum())
        self.every = every
        self.samples = []

    def run(self, max_steps):
        done = 0
        while done < max_steps:
            k = min(self.every, max_steps - done)
            self.sim.step(k)
            done += k
            c = self.sim.metric_counters()
            cells = self.sim.metric_cells()
            self.samples.append(dict(step=int(c[0]), active=int(c[1]), exits=c[4:4 + self.E + 1].copy(),
                                     max_density=floor_max_density(cells[:, 0], self.walkable)))
            if c[1] == 0:
                break
        return self

    def write(self, out_dir):
        os.makedirs(out_dir, exist_ok=True)
        names = [f"exit{e}" for e in range(self.E)] + ["other"]
        with open(os.path.join(out_dir, "timeseries.csv"), "w") as f:
            f.write("step,t_model,t_real_s,active," + ",".join(f"cum_evacuated_{m}" for m in names) + ",max_density_floor_p_m2\n")
            for s in self.samples:
                f.write(f"{s['step']},{s['step'] * DT_MODEL:.6g},{s['step'] * DT_REAL_S:.6g},{s['active']},"
                        + ",".join(str(int(v)) for v in s["exits"]) + f",{s['max_density']:.6g}\n")
        cells = self.sim.metric_cells()
        rows = cell_rows(cells, self.walkable, self.grid)
        with open(os.path.join(out_dir, "heatmap.csv"), "w") as f:
            f.write("ix,iy,x_m,y_m,zone,walkable_m2,peak_density_p_m2," + ",".join(f"time_above_{t}_s" for t in THRESHOLD_NAMES)
                    + ",occupied_s,mean_speed_occupied_m_s,exposure_person_s,person_s\n")
            for j in range(rows["cell"].size):
                ms = rows["mean_speed_occupied_m_s"][j]
                f.write(f"{rows['ix'][j]},{rows['iy'][j]},{rows['x_m'][j]:.6g},{rows['y_m'][j]:.6g},floor,{rows['walkable_m2'][j]:.6g},"
                        f"{rows['peak_density_p_m2'][j]:.6g}," + ",".join(f"{t:.6g}" for t in rows["time_above"][j])
                        + f",{rows['occupied_s'][j]:.6g}," + ("" if not np.isfinite(ms) else f"{ms:.6g}")
                        + f",{rows['exposure_person_s'][j]:.6g},{rows['person_s'][j]:.6g}\n")
        counters = self.sim.metric_counters()
        times = evacuation_times(self.sim.metric_evac_steps(), self.n0)
        hot = hotspots(rows)
        summary = dict(n=int(self.sim.n), n0=self.n0, steps=int(counters[0]), evacuated=int(counters[2]), active=int(counters[1]),
                       exits={m: int(v) for m, v in zip(names, counters[4:4 + self.E + 1])}, dt_real_s=DT_REAL_S, **times,
                       peak_density_p_m2=float(rows["peak_density_p_m2"][rows["walkable_m2"] >= PEAK_MIN_WALKABLE_M2].max(initial=0.0)),
                       exposure_person_s=float(rows["exposure_person_s"].sum()), person_s=float(rows["person_s"].sum()),
                       hotspots=[dict(ix=int(rows["ix"][h]), iy=int(rows["iy"][h]), exposure_person_s=float(rows["exposure_person_s"][h]),
                                      peak_density_p_m2=float(rows["peak_density_p_m2"][h])) for h in hot])
        summary = {k: (None if isinstance(v, float) and not math.isfinite(v) else v) for k, v in summary.items()}
        with open(os.path.join(out_dir, "summary.json"), "w") as f:
            json.dump(summary, f, indent=1)
        return summary


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="an evacuation run with the safety metrics on one of this build's floors")
    ap.add_argument("--layout", type=int, default=sh.LAYOUT_ROOM)
    ap.add_argument("--active", type=int, default=4096)
    ap.add_argument("--total", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=1337)
    ap.add_argument("--max-steps", type=int, default=3000)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--out", default="experiment_out")
    a = ap.parse_args()
    x = Experiment(a.layout, a.active, a.total, a.seed, a.every).run(a.max_steps)
    print(json.dumps({k: v for k, v in x.write(a.out).items() if k != "hotspots"}))
