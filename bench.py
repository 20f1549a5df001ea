"""Benchmark of the per-step crowd pipeline on one MI355X.

Workload: the large plain configuration, 63,424 active agents (sort size 65,536: 2,112 parked) seated on the
200 x 150 m bowl (240 rows of 0.1 m boxes, 16 gangways, 8 exits of 2.4 m, stair bands at 0.6x), per-agent
radius U[0.20, 0.25] m and free speed N(1.34, 0.26) m/s clipped to [0.5, 2.0], seed 1337. A step is the whole
pipeline: hash, counting sort, cell table + gather, the queue-aware exit choice, predict, 31 Jacobi projection
iterations, final pass. Inputs are resident in HBM before the timed region; the timed region is K consecutive
steps (hipGraph replays) bracketed by HIP events on the pipeline's stream.

Prints one JSON line: agent-steps/s (the sum over the timed steps of the agents active at each step's input,
per second), the roofline of the dominant kernel (the projection iteration, HBM-bound), and the CPU baseline
(the oracle, a CPU restatement of the reference's CpuCrowdSim, on the same workload, host cores).
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "swarm-crowd-simulation_amd"))

import numpy as np  # noqa: E402

import swarm_hip as sh  # noqa: E402

HBM_PEAK_GBS = 8000.0
ACTIVE, TOTAL, SEED = 63424, 65536, 1337


def step_bytes(n_total, cells, active, iters=32):
    """Algorithmic HBM bytes of one plain step (compulsory traffic, neighbour reads cache-resident; the per-kernel
    terms of the per-step accounting): hash 20N, sort 16N, ranges 4N + 8C, gather 64A, predict 100A,
    projection 36A per iteration, final 64A, evacuation 8A."""
    return 40 * n_total + 8 * cells + (236 + 36 * (iters - 1)) * active


# one projection iteration's algorithmic bytes per active slot (the per-step accounting's 36 A per iteration: position
# and displacement accumulator read and written, radius, the wall plane's offset)
PROJ_BYTES_PER_SLOT = 36
# the same launch's bytes counted word by word as this build's kernel touches them once (key 4, x/y/acc 12, flag 4,
# radius 4, plane 16 read; x/y/acc 12 written): a secondary figure
PROJ_WORDS_PER_SLOT = 52
TRAFFIC_SUMMARY = os.path.join(REPO, "profiles", "traffic_bowl_63424.md")


def measured_traffic():
    """HBM bytes per projection launch from the FETCH_SIZE / WRITE_SIZE counter runs summarised in profiles/
    (each counter scaled by what it reads for a copy kernel of known size); None when absent."""
    try:
        for line in open(TRAFFIC_SUMMARY):
            if line.startswith("traffic_bytes_per_launch k_project:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_model():
    """The host CPU as lscpu names it: model name, family and model (a guest may report only the vendor's line)."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("model name", "cpu family", "model") and k not in info:
                info[k] = v.strip()
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    if "cpu family" in info:
        name += f" (family {info['cpu family']}, model {info.get('model', '?')})"
    return name


CPU_STAGES = ("hash", "sort", "cells", "choice", "predict", "project", "final", "total")


def cpu_snapshot_rate(scene, a, n, threads, warmup, measure, rounds=3):
    """The oracle (a CPU restatement of the reference's CpuCrowdSim: counting sort, OpenMP dynamic,64) in the
    reference benchmark's snapshot mode (tools/benchmark/benchmark_crowd.cpp:16-24): every step, warm-up and timed,
    starts from the same agent state, restored outside the timed region (the exit choice runs on); the median of the
    per-round medians over `rounds` rounds, with the median per-stage wall clock of the timed steps."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py as op

    tau, relax = sh.model_params(0.5)
    lay = scene.venue_layout()
    o = op.Oracle(scene.info, scene.tables(), n, relax, tau, threads=threads,
                  venue=None if lay is None else op.venue_layout_arrays_of(lay))
    o.set_state(a)
    active = int(np.sum(a["w"] > 0))
    medians, stages = [], []
    for _ in range(rounds):
        for _ in range(warmup):
            o.restore(a)
            o.step(1)
        times = []
        for _ in range(measure):
            o.restore(a)
            t0 = time.perf_counter()
            o.step(1)
            times.append(time.perf_counter() - t0)
            stages.append(o.times())
        medians.append(float(np.median(times)))
    ms = 1000.0 * float(np.median(medians))
    st = np.median(np.array(stages), axis=0)
    return {"agent_steps_per_s": active / (ms * 1e-3), "ms_per_step": ms, "threads": threads, "agents_active": active,
            "rounds": rounds, "warmup": warmup, "measured_per_round": measure,
            "stages_ms": {k: float(v) for k, v in zip(CPU_STAGES, st)}}


def gpu_plain_rate(scene, a, n, steps=100, warmup=10):
    """The plain step on the GPU for the same state (graph replays, HIP events), as agent-steps/s."""
    sim = sh.Sim(scene, n)
    sim.upload(a)
    sim.autotune(5)
    sim.step(warmup)
    sim.sync()
    _, _, s0 = sim.control()
    ms = sim.time_steps(steps)
    _, _, s1 = sim.control()
    sim.close()
    return {"agent_steps_per_s": (s1 - s0) / (ms * 1e-3), "ms_per_step": ms / steps}


def cpu_baseline(bowl_scene, bowl_agents, full=True):
    """The CPU baseline: the oracle on the host cores of this box, all logical cores and one thread, on the large
    floor (63,424 agents), the small room (4,096 agents) and one scenario of the ensemble (the one of median agent
    count), beside the GPU's rate for the same state."""
    cores = os.cpu_count() or 1
    out = {"cpu_model": cpu_model(), "logical_cores": cores, "protocol": "snapshot mode (benchmark_crowd.cpp:16-24): "
           "every step from the same restored agent state; median of per-round medians over 3 rounds", "configs": {}}
    bowl = {"all_cores": cpu_snapshot_rate(bowl_scene, bowl_agents, TOTAL, cores, 2, 5)}
    if full:
        bowl["one_thread"] = cpu_snapshot_rate(bowl_scene, bowl_agents, TOTAL, 1, 1, 3)
    out["configs"]["bowl_63424"] = bowl
    if full:
        room_scene = sh.Scene.layout(sh.LAYOUT_ROOM)
        room = sh.spawn_uniform(sh.LAYOUT_ROOM, 4096, 4096, SEED)[0]   # uniform with rejection, relaxed
        out["configs"]["room_4096"] = {"all_cores": cpu_snapshot_rate(room_scene, room, 4096, cores, 5, 20),
                                       "one_thread": cpu_snapshot_rate(room_scene, room, 4096, 1, 2, 8),
                                       "gpu": gpu_plain_rate(room_scene, room, 4096)}
        scen_layout, totals, actives, agents = sh.batch_workload()
        s = int(np.argsort(totals)[len(totals) // 2])
        lo = int(np.sum(totals[:s]))
        lay = sh.BATCH_LAYOUTS[scen_layout[s]]
        sc = sh.Scene.layout(lay)
        part = {k: v[lo:lo + int(totals[s])] for k, v in agents.items()}
        out["configs"]["ensemble_scenario"] = {"scenario": s, "layout": int(lay), "agents": int(totals[s]),
                                               "all_cores": cpu_snapshot_rate(sc, part, int(totals[s]), cores, 5, 20),
                                               "one_thread": cpu_snapshot_rate(sc, part, int(totals[s]), 1, 2, 8),
                                               "gpu": gpu_plain_rate(sc, part, int(totals[s]))}
    return out


# The reference's published GPU step on its venue L (AMD Radeon RX 7800 XT, OpenGL compute; its 3D model): ms per step and
# per pass, at spawn and congested (the state at t50 of its full-load run) (results/benchmarks/thesis/perf_tables.md:17-18,45-46)
REFERENCE_VENUE_L = {
    "hardware": "AMD Radeon RX 7800 XT (OpenGL 4.3 compute), the reference's 3D seated model",
    "source": "results/benchmarks/thesis/perf_tables.md:17-18,45-46",
    "spawn": {"active": 63424, "ms_per_step": 1.861, "passes_ms": {"hash": 0.0036, "sort": 0.1064, "cells_gather": 0.0087,
                                                                    "exit_choice": 0.2039, "predict": 0.7722, "projection": 0.6953,
                                                                    "final": 0.0663}},
    "congested": {"active": 31703, "ms_per_step": 3.928, "passes_ms": {"hash": 0.0050, "sort": 0.0919, "cells_gather": 0.0088,
                                                                        "exit_choice": 0.4564, "predict": 2.289, "projection": 0.6879,
                                                                        "final": 0.3720}},
}


def venue_bench(name="L", n=65536, steps=100, warmup=5, use_graph=True, cache_dir=None):
    """The plan of the reference's venue (its own seat spawn at full load, its global radius and free speed) at spawn and
    congested: the state when half of the seated crowd has left (t50), reached by stepping from spawn. Each state: warm-up,
    then 'steps' timed steps (graph replays, HIP events), and the per-pass split of an eager run."""
    t0 = time.perf_counter()
    scene = sh.Scene.venue(name, cache_dir=cache_dir)
    build_s = time.perf_counter() - t0
    a = sh.venue_agents(name, n, 1.0, SEED)
    seated = int(a["w"].sum())
    sim = sh.Sim(scene, n, use_graph=use_graph)
    sim.upload(a)
    sim.autotune(5)

    def timed(state):
        sim.step(warmup)
        sim.sync()
        _, act0, s0 = sim.control()
        ms = sim.time_steps(steps)
        _, _, s1 = sim.control()
        sim.set_timing(True)
        sim.step(5)
        t = sim.timings()
        sim.set_timing(False)
        out = {"active_at_start": int(act0), "ms_per_step": ms / steps, "agent_steps_per_s": (s1 - s0) / (ms * 1e-3),
               "passes_ms": dict(grid=t[0], choice=t[1], predict=t[2], projection=t[3], final=t[4], total_eager=t[5])}
        ref = REFERENCE_VENUE_L.get(state) if name == "L" else None
        if ref:
            out["speedup_vs_reference_gpu"] = ref["ms_per_step"] / out["ms_per_step"]
        return out

    res = {"venue": name, "agents_total": n, "seated": seated, "scene_build_s": build_s, "exits": scene.info.E,
           "nav_cells": scene.info.nav_res_x * scene.info.nav_res_y, "hash_cells": scene.info.hash_nx * scene.info.hash_ny}
    res["spawn"] = timed("spawn")
    sim.upload(a)
    chunk, stepped = 50, 0
    while stepped < 20000:   # to t50: half of the seated crowd gone
        sim.step(chunk)
        stepped += chunk
        if sim.control()[1] <= seated // 2:
            break
    res["congested_after_steps"] = stepped
    res["congested"] = timed("congested")
    if name == "L":
        res["reference"] = REFERENCE_VENUE_L
    sim.close()
    return res


def seated_agents(name, n, load=1.0, seed=SEED):
    """The reference's seated crowd (spawnSeatedCrowd, src/crowd_sim.cpp:677-697) for the venue's own 3D step: GpuBoid
    positions (x, y, z, w) at the chosen seats, at rest; the other slots parked at (1000, 1000, 1000)."""
    x, y, z, _ = sh.venue_seats(name)
    chosen = sh.spawn_seats(len(x), n, load, seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = 1000.0
    live = chosen >= 0
    pos[live, 0], pos[live, 1], pos[live, 2], pos[live, 3] = x[chosen[live]], y[chosen[live]], z[chosen[live]], 1.0
    return pos, np.zeros((n, 4), np.float32)


def seated_bench(scene, name="L", n=65536, steps=100, warmup=5, use_graph=True, build_s=None):
    """The reference's own 3D step on its venue (the default venue model: floors, gravity, fence, exact walls, step-up
    limit, queue-aware choice with the surface rule, incline rules), its seat spawn at full load, at spawn and congested
    (t50, reached by stepping from spawn): like for like with the reference's published venue timings."""
    pos, vel = seated_agents(name, n)
    seated = int(pos[:, 3].sum())
    sim = sh.Sim(scene, n, use_graph=use_graph)
    sim.upload3(pos, vel)

    def timed(state):
        sim.step(warmup)
        sim.sync()
        _, act0, s0 = sim.control()
        ms = sim.time_steps(steps)
        _, _, s1 = sim.control()
        sim.set_timing(True)
        sim.step(5)
        t = sim.timings()
        sim.set_timing(False)
        out = {"active_at_start": int(act0), "ms_per_step": ms / steps, "agent_steps_per_s": (s1 - s0) / (ms * 1e-3),
               "passes_ms": dict(grid=t[0], choice=t[1], predict=t[2], projection=t[3], final=t[4], total_eager=t[5])}
        ref = REFERENCE_VENUE_L.get(state) if name == "L" else None
        if ref:
            out["speedup_vs_reference_gpu"] = ref["ms_per_step"] / out["ms_per_step"]
        return out

    res = {"venue": name, "model": "the reference's own 3D step, default venue model", "agents_total": n, "seated": seated,
           "scene_build_s": build_s, "exits": scene.info.E, "nav_cells": scene.info.nav_res_x * scene.info.nav_res_y,
           "hash_cells": scene.info.hash_nx * scene.info.hash_ny}
    res["spawn"] = timed("spawn")
    sim.upload3(pos, vel)
    chunk, stepped = 50, 0
    while stepped < 20000:   # to t50: half of the seated crowd gone
        sim.step(chunk)
        stepped += chunk
        if sim.control()[1] <= seated // 2:
            break
    res["congested_after_steps"] = stepped
    res["congested"] = timed("congested")
    sim.close()
    return res


def batch_bench(steps, warmup, loop_steps, use_graph=True):
    """The ensemble (256 scenarios of the room, corridor, bowl-eighth and pillar floors, about 1.4 M agents) as one
    batched step, against the same scenarios stepped by a loop over plain contexts (each scenario's steps, then a
    synchronisation, one scenario after another)."""
    scen_layout, totals, actives, agents = sh.batch_workload()
    scenes = [sh.Scene.layout(l) for l in sh.BATCH_LAYOUTS]
    b = sh.BatchSim(scenes, scen_layout, totals, use_graph=use_graph)
    b.upload(agents)
    tune_ms = b.autotune(5)
    shape = b.physics_shape()
    lane_cfg = dict(zip(("list", "collect", "predict"), b.lanes()))
    b.step(warmup)
    b.sync()
    _, _, s0 = b.control()
    ms = b.time_steps(steps)
    _, _, s1 = b.control()
    batched = (s1 - s0) / (ms * 1e-3)
    cells = int(b.cells)
    b_step = step_bytes(int(totals.sum()), cells, (s1 - s0) / steps)
    b.close()
    plains = []
    for s in range(len(totals)):
        lo = int(np.sum(totals[:s]))
        p = sh.Sim(scenes[scen_layout[s]], int(totals[s]), use_graph=use_graph)
        p.upload({k: v[lo:lo + int(totals[s])] for k, v in agents.items()})
        p.step(2)
        plains.append(p)
    for p in plains:
        p.sync()
    sum0 = sum(p.control()[2] for p in plains)
    t0 = time.perf_counter()
    for p in plains:
        p.step(loop_steps)
        p.sync()
    dt = time.perf_counter() - t0
    sum1 = sum(p.control()[2] for p in plains)
    looped = (sum1 - sum0) / dt
    for p in plains:
        p.close()
    return {"value": batched, "unit": "agent-steps/s", "ms_per_step": ms / steps, "steps": steps,
            "scenarios": int(len(totals)), "agents_total": int(totals.sum()), "agents_active": int(actives.sum()),
            "physics_shape": list(shape), "lanes_per_slot": lane_cfg, "autotune_ms_per_step": tune_ms,
            "step_algorithmic_bytes": b_step, "step_frac": b_step / (ms / steps * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "looped_plain": {"value": looped, "unit": "agent-steps/s", "steps": loop_steps,
                             "ms_per_round": 1000.0 * dt / loop_steps},
            "speedup_vs_looped_plain": batched / looped}


def adjoint_bench(K=200, every=20, n=4096, use_graph=True):
    """The differentiable configuration: 4,096 agents placed uniformly in the 40 x 30 m room, K steps, checkpoints every 20. The
    forward is K plain steps (graph replays); the adjoint is the adjoint's forward (the same steps with the loss
    terms) plus its backward. By default the whole run's tape fits the tape budget, so the forward tapes it and
    nothing is recomputed; 'recompute' is the same run with a zero budget (each segment but the last replayed
    from its checkpoint with the tape, then the adjoint kernels)."""
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    a = sh.spawn_uniform(sh.LAYOUT_ROOM, n, n, SEED)[0]   # the small floor's placement: uniform with rejection, relaxed
    sim = sh.Sim(scene, n, use_graph=use_graph)
    sim.upload(a)
    tune_ms = sim.autotune(5)   # launch shape and lanes per slot for this crowd (the state is restored)
    sim.step(2)
    sim.upload(a)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(K)
    sim.sync()
    fwd = time.perf_counter() - t0
    _, _, steps_sum = sim.control()

    def backward():
        try:
            return sim.adjoint_backward()
        except sh.NonFiniteGradient as e:   # SWARM_ENONFINITE: the outputs are written all the same
            return e.grad_v0, e.grad_globals

    def timed_adjoint():
        sim.upload(a)
        sim.adjoint_forward(K, every, 1.0, 0.1)   # warm-up (allocations, graphs)
        backward()
        sim.upload(a)
        t0 = time.perf_counter()
        loss = sim.adjoint_forward(K, every, 1.0, 0.1)
        t1 = time.perf_counter()
        gv0, g = backward()
        t2 = time.perf_counter()
        return loss, gv0, g, t1 - t0, t2 - t1, sim.adjoint_segments(), sim.tape_bytes()

    loss, gv0, g, tf, tb, segs, mem = timed_adjoint()
    total = tf + tb
    finite = bool(np.isfinite(gv0).all() and np.isfinite(g).all())
    # the same K steps with the state adjoint cleared every 25 steps (the truncated estimator long horizons use)
    sim.set_adjoint_terms(window=25)
    sim.upload(a)
    t3 = time.perf_counter()
    sim.adjoint_forward(K, every, 1.0, 0.1)
    wv0, wg = backward()
    t4 = time.perf_counter()
    sim.set_adjoint_terms(window=0)
    windowed = {"window": 25, "ms": 1000.0 * (t4 - t3), "gradient_finite": bool(np.isfinite(wv0).all() and np.isfinite(wg).all()),
                "grad_v0_norm": float(np.linalg.norm(wv0.astype(np.float64)))}
    # checkpoints every 20 steps and recompute, whatever fits
    sim.set_tape_budget(0)
    loss_r, rv0, _, rf, rb, rsegs, _ = timed_adjoint()
    same = bool(np.array_equal(rv0.view(np.uint32), gv0.view(np.uint32)) and loss_r == loss)
    recompute = {"segments": rsegs, "adjoint_forward_ms": 1000.0 * rf, "adjoint_backward_ms": 1000.0 * rb,
                 "cost_vs_forward": (rf + rb) / fwd, "same_loss_and_grad_v0_bits": same}
    return {"K": K, "checkpoint_every": every, "segments": segs, "adjoint_memory_mb": mem / 1e6, "agents": n,
            "physics_shape": list(sim.physics_shape()), "lanes_per_slot": dict(zip(("list", "collect", "predict"), sim.lanes())),
            "autotune_ms_per_step": tune_ms, "forward_ms": 1000.0 * fwd,
            "adjoint_forward_ms": 1000.0 * tf, "adjoint_backward_ms": 1000.0 * tb,
            "cost_vs_forward": total / fwd, "agent_steps_per_s": steps_sum / total, "loss": loss,
            "gradient_finite": finite,
            "gradient_note": None if finite else "the exact pathwise gradient of a 200-step loss overflows FP32 in this crowd "
                                                 "(its sensitivity grows about exp(0.46 K)); the timing is the measured quantity",
            "windowed": windowed, "recompute": recompute}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-quick", action="store_true", help="the CPU baseline on the large floor at all cores only")
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--copy-probe", action="store_true", default=True)
    ap.add_argument("--no-copy-probe", dest="copy_probe", action="store_false")
    ap.add_argument("--batch-steps", type=int, default=40)
    ap.add_argument("--loop-steps", type=int, default=10)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-venue", action="store_true", help="skip the reference's venue L legs")
    ap.add_argument("--no-venue-plan", action="store_true", help="skip venue L's plan leg (the seated 3D leg stays)")
    ap.add_argument("--no-graph", action="store_true",
                    help="eager launches in every leg (for profilers that do not follow graph launches)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the step does not shard across GPUs here",
              file=sys.stderr)
        sys.exit(2)

    # venue L's own 3D scene (its navigation field on the heightfield, 16 M cells x 48 exits) is built on the host while
    # the GPU legs run; ctypes releases the interpreter during the build
    seated_build = {}
    if not args.no_venue:
        import tempfile
        import threading

        def build_seated():
            t = time.perf_counter()
            seated_build["scene"] = sh.Scene.seated("L", threads=12, cache_dir=os.path.join(tempfile.gettempdir(), "swarm_scene_cache"))
            seated_build["seconds"] = time.perf_counter() - t

        seated_thread = threading.Thread(target=build_seated, daemon=True)
        seated_thread.start()

    t0 = time.perf_counter()
    scene = sh.Scene.layout(sh.LAYOUT_BOWL)
    a = sh.place_agents(sh.LAYOUT_BOWL, ACTIVE, TOTAL, SEED)
    build_s = time.perf_counter() - t0
    sim = sh.Sim(scene, TOTAL, use_graph=not args.no_graph)
    sim.upload(a)
    print(f"# scene built in {build_s:.1f} s; hash grid {scene.info.hash_nx}x{scene.info.hash_ny}, "
          f"SDF {scene.info.sdf_nx}x{scene.info.sdf_ny}, nav {scene.info.nav_res_x}x{scene.info.nav_res_y} x {scene.info.E} exits",
          file=sys.stderr, flush=True)

    tune_ms = sim.autotune(10)   # launch shape and lanes per slot (the state is restored)
    shape, lanes = list(sim.physics_shape()), sim.projection_lanes()
    lane_cfg = dict(zip(("list", "collect", "predict"), sim.lanes()))
    sim.step(args.warmup)
    sim.sync()
    _, _, sum0 = sim.control()
    ms = sim.time_steps(args.steps)
    _, n_active, sum1 = sim.control()
    agent_steps = sum1 - sum0
    value = agent_steps / (ms / 1000.0)
    ms_per_step = ms / args.steps
    print(f"# timed {args.steps} steps: {ms_per_step:.4f} ms/step, {value / 1e6:.1f} M agent-steps/s", file=sys.stderr, flush=True)

    # a caller that hands the state over through host memory every step: read it back after each step and write it
    # again before the next (the PCIe-inclusive rate; the timed steps above keep the state in HBM)
    st = sim.state()
    t0 = time.perf_counter()
    rt_steps = 20
    for _ in range(rt_steps):
        sim.restore(st)
        sim.step(1)
        st = sim.state()
    host_rt_ms = 1000.0 * (time.perf_counter() - t0) / rt_steps
    print(f"# with the state through host memory every step: {host_rt_ms:.3f} ms/step", file=sys.stderr, flush=True)

    # the safety-metric passes (density deposit + per-cell accumulators) inside the same steps: the same upload and
    # warm-up replay the same trajectory, so the difference is the passes' cost
    sim.upload(a)
    sim.metrics_enable(True)
    sim.step(args.warmup)   # the same warm-up, now with the passes (their graph is built and warm before timing)
    sim.sync()
    ms_m = sim.time_steps(args.steps)
    mc = sim.metric_counters()
    rx, ry = sim.metrics_grid()[:2]
    sim.metrics_enable(False)
    metrics = {"ms_per_step": ms_m / args.steps, "overhead_vs_plain": ms_m / ms - 1.0, "grid_cells": [rx, ry],
               "recorded_steps": int(mc[0]), "active_last": int(mc[1])}
    print(f"# with the safety metrics: {metrics['ms_per_step']:.4f} ms/step ({100 * metrics['overhead_vs_plain']:+.1f}%)",
          file=sys.stderr, flush=True)

    # per-pass timing (eager launches with events between passes), then the dominant kernel's launch duration (a pair
    # of events recorded by each later projection iteration's dispatch, a separate run: they slow the pass totals)
    sim.set_timing(True)
    sim.step(max(1, args.profile_steps))
    t = sim.timings()
    sim.set_timing(True, launches=True)
    sim.step(max(1, args.profile_steps))
    proj_launch_ms = sim.timings()[6]
    sim.set_timing(False)
    passes = dict(grid=t[0], choice=t[1], predict=t[2], projection=t[3], final=t[4], total_eager=t[5])
    mean_active = agent_steps / args.steps
    proj_bytes = PROJ_BYTES_PER_SLOT * mean_active
    achieved = proj_bytes / (proj_launch_ms * 1e-3) / 1e9
    achieved_words = PROJ_WORDS_PER_SLOT * mean_active / (proj_launch_ms * 1e-3) / 1e9
    cells = scene.info.hash_nx * scene.info.hash_ny
    b_step = step_bytes(TOTAL, cells, mean_active)
    step_gbs = b_step / (ms_per_step * 1e-3) / 1e9
    print(f"# passes ms/step: " + ", ".join(f"{k} {v:.4f}" for k, v in passes.items()) +
          f"; one projection iteration after the first {proj_launch_ms * 1000:.2f} us (event pair per launch)", file=sys.stderr, flush=True)

    copy_gbs = None
    if args.copy_probe:   # a plain copy of 1 GiB: the measured HBM copy bandwidth (read + write bytes / time)
        nb = 1 << 30
        cms = sh.copy_probe(nb, 10)
        copy_gbs = 2.0 * nb / (cms * 1e-3) / 1e9
        print(f"# copy probe: 1 GiB in {cms:.3f} ms = {copy_gbs:.0f} GB/s (read + write)", file=sys.stderr, flush=True)

    batch = None
    if not args.no_batch:
        sim.close()
        batch = batch_bench(args.batch_steps, 5, args.loop_steps, use_graph=not args.no_graph)
        print(f"# batched ensemble: {batch['ms_per_step']:.3f} ms/step, {batch['value'] / 1e6:.1f} M agent-steps/s; "
              f"looped plain {batch['looped_plain']['value'] / 1e6:.1f} M; x{batch['speedup_vs_looped_plain']:.2f}",
              file=sys.stderr, flush=True)

    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_bench(use_graph=not args.no_graph)
        print(f"# adjoint K={adjoint['K']}: forward {adjoint['forward_ms']:.1f} ms, adjoint {adjoint['adjoint_forward_ms']:.1f} + "
              f"{adjoint['adjoint_backward_ms']:.1f} ms ({adjoint['cost_vs_forward']:.2f}x, {adjoint['segments']} segment(s)); "
              f"with recompute {adjoint['recompute']['cost_vs_forward']:.2f}x", file=sys.stderr, flush=True)

    venue = None
    seated = None
    if not args.no_venue:
        seated_thread.join()
        seated = seated_bench(seated_build["scene"], "L", use_graph=not args.no_graph, build_s=seated_build["seconds"])
        for st in ("spawn", "congested"):
            v = seated[st]
            print(f"# venue L (the reference's own 3D step), {st}: {v['active_at_start']} active, {v['ms_per_step']:.4f} ms/step, "
                  f"{v['agent_steps_per_s'] / 1e6:.1f} M agent-steps/s; the reference's GPU {REFERENCE_VENUE_L[st]['ms_per_step']} ms "
                  f"(x{v['speedup_vs_reference_gpu']:.2f})", file=sys.stderr, flush=True)
    if not args.no_venue and not args.no_venue_plan:
        import tempfile
        venue = venue_bench("L", use_graph=not args.no_graph, cache_dir=os.path.join(tempfile.gettempdir(), "swarm_scene_cache"))
        for st in ("spawn", "congested"):
            v = venue[st]
            print(f"# venue L plan, {st}: {v['active_at_start']} active, {v['ms_per_step']:.4f} ms/step, "
                  f"{v['agent_steps_per_s'] / 1e6:.1f} M agent-steps/s; the reference's GPU {REFERENCE_VENUE_L[st]['ms_per_step']} ms "
                  f"(x{v['speedup_vs_reference_gpu']:.2f})", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scene, a, full=not args.cpu_quick)
        for name, c in cpu["configs"].items():
            print(f"# cpu baseline {name}: " + ", ".join(f"{k} {v['ms_per_step']:.2f} ms/step ({v['agent_steps_per_s'] / 1e6:.2f} M agent-steps/s)"
                                                         for k, v in c.items() if isinstance(v, dict) and "ms_per_step" in v),
                  file=sys.stderr, flush=True)

    line = {
        "metric": "agent-steps/s @63,424 agents + %HBM roofline; batched & adjoint vs plain",
        "value": value,
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded seat placement on the generated bowl floor",
        "config": {"workload": "plain step, 63,424 agents on the 200 x 150 m seated bowl, sort size 65,536",
                   "agents_active": ACTIVE, "agents_total": TOTAL, "hash_cells": cells, "proj_iterations": 32,
                   "exits": scene.info.E, "parallelism": "single GPU", "physics_shape": shape, "projection_lanes": lanes,
                   "lanes_per_slot": lane_cfg,
                   "autotune_ms_per_step": tune_ms},
        "roofline": {"bound": "hbm", "kernel": "k_project (one Jacobi iteration after the first, over the candidate lists)", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "traffic_source": "FETCH_SIZE + WRITE_SIZE counter runs, scaled by a plain copy's counts (profiles/traffic_bowl_63424.md)",
                     "bytes_per_launch": proj_bytes, "bytes_per_slot": PROJ_BYTES_PER_SLOT, "launch_us": proj_launch_ms * 1000.0,
                     "launch_timing": "HIP events recorded by each later iteration's own dispatch, mean over the run",
                     "achieved_word_bytes": achieved_words, "word_bytes_per_slot": PROJ_WORDS_PER_SLOT,
                     "step_algorithmic_bytes": b_step, "step_achieved_gbs": step_gbs, "step_frac": step_gbs / HBM_PEAK_GBS,
                     "measured_copy_gbs": copy_gbs},
        "passes_ms": passes,
        "host_roundtrip": {"ms_per_step": host_rt_ms, "steps": rt_steps,
                           "what": "restore the state from host arrays, one step, read the state back (wall clock)"},
        "metrics": metrics,
        "batched": batch,
        "adjoint": adjoint,
        "venue_l_seated": seated,
        "venue_l": venue,
        "cpu_baseline": None if cpu is None else dict(
            value=cpu["configs"]["bowl_63424"]["all_cores"]["agent_steps_per_s"], unit="agent-steps/s", cores=cpu["logical_cores"],
            kind="port", sample="the 63,424-agent bowl: 3 rounds of 2 warm-up + 5 timed steps, each from the same restored state "
                                "(snapshot mode), median of the per-round medians; oracle = CPU restatement of CpuCrowdSim",
            cpu_model=cpu["cpu_model"], protocol=cpu["protocol"], configs=cpu["configs"]),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
