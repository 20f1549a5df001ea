"""GPU safety-metric passes (csrc/metrics.hip) through the C ABI.

- swarm_metrics_record on the reference's own recorded steps (tests/golden/metrics_kat.npz, written by the
  reference's MetricsReference, src/metrics.cpp:429-480): counters, per-cell masses each step, the evacuation steps
  and every per-cell accumulator at the end, bit for bit, on the reference's 64 x 64 grid;
- swarm_metrics_record on random states (positions far off the grid, parked agents, an offset rectangular grid)
  against the oracle (oracle/metrics_oracle.cpp);
- the passes inside the step (swarm_metrics_enable) against the oracle fed with the step's own input and output
  states, on the room through its evacuation and on the 63,424-agent bowl; their evacuation steps and exits against
  the pipeline's own evacuation counters; a reset in mid-run; captured (two steps a graph) against eager launches.
Integer atomics are order-free, so every comparison is exact. Exits are compared as their total where the
reference's slots are its terrace's avenue gaps (its floor has none of this build's exit rectangles).
"""
import numpy as np
import pytest

import oracle_py as op
import swarm_hip as sh
from common import golden, terrace_scene

pytestmark = pytest.mark.gpu

RES = 64
CPU = np.float32(31.25)
OFF = np.float32(0.5 * RES - 0.5)


def _split(states, k):
    prev = dict(x=states[k, :, 0], y=states[k, :, 1], w=states[k, :, 2])
    cur = dict(x=states[k + 1, :, 0], y=states[k + 1, :, 1], w=states[k + 1, :, 2], vx=states[k + 1, :, 3], vy=states[k + 1, :, 4])
    return prev, cur


def test_record_matches_reference_golden():
    g = golden("metrics_kat.npz")
    states = g["metrics__run__states"]
    steps, n = states.shape[0] - 1, states.shape[1]
    print(f"metrics record: reference run, n={n}, {steps} steps", flush=True)
    sim = sh.Sim(terrace_scene(), n)
    sim.metrics_configure(RES, RES, CPU, OFF, OFF, g["metrics__info__threshold_mass"])
    assert sim.metrics_grid() == (RES, RES, CPU, OFF, OFF)
    for k in range(steps):
        sim.metrics_record(*_split(states, k))
        c = sim.metric_counters()
        ref = g["metrics__run__counters"][k]
        assert list(c[:4]) == list(ref[:4]), f"step {k + 1} counters"
        assert int(c[4]) == int(ref[4]) and c[5:].sum() == 0   # no exit rectangles: every evacuation in slot E = 0
        cells = sim.metric_cells()
        assert np.array_equal(cells[:, 0], g["metrics__run__last_mass"][k]), f"step {k + 1} mass"
        assert np.array_equal(cells[:, 1], g["metrics__run__last_speed_mass"][k]), f"step {k + 1} speed mass"
    assert np.array_equal(sim.metric_evac_steps(), g["metrics__run__evac_step"])
    assert np.array_equal(sim.metric_cells(), g["metrics__run__cells"])
    # a reset starts the record again
    sim.metrics_reset()
    assert not sim.metric_counters().any() and not sim.metric_cells().any() and not sim.metric_evac_steps().any()
    sim.metrics_record(*_split(states, 0))
    assert sim.metric_counters()[0] == 1


def test_record_random_states_match_oracle():
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    rects = scene.tables()["exit_rects"][: 4 * scene.info.E].reshape(-1, 4)
    n, rx, ry = 5000, 37, 23
    cpu, ox, oy = np.float32(20.0), np.float32(3.25), np.float32(-1.5)
    rng = np.random.default_rng(17)
    thr = rng.integers(1, 6 * 65536, (rx * ry, 4)).astype(np.uint32)
    thr[rng.random(rx * ry) < 0.2] = 0xFFFFFFFF
    print(f"metrics record: random states, n={n}, grid {rx}x{ry}", flush=True)
    sim = sh.Sim(scene, n)
    sim.metrics_configure(rx, ry, cpu, ox, oy, thr)
    o = op.MetricsOracle(n, rx, ry, cpu, ox, oy, thr, rects)
    lo = np.array([(-ox - 10) / cpu, (-oy - 10) / cpu], np.float32)
    hi = np.array([(rx - ox + 10) / cpu, (ry - oy + 10) / cpu], np.float32)
    pos = lambda: rng.uniform(lo, hi, (n, 2)).astype(np.float32)
    p = pos()
    # some agents start inside the exits, so that their evacuations land in the exits' slots
    for e, r in enumerate(rects):
        sl = slice(100 * e, 100 * e + 100)
        p[sl, 0] = rng.uniform(r[0], r[2], 100)
        p[sl, 1] = rng.uniform(r[1], r[3], 100)
    w = (rng.random(n) < 0.9).astype(np.float32)
    active_sum = 0
    for k in range(6):
        q = pos()
        wq = w * (rng.random(n) < 0.85)
        v = rng.normal(0.0, 0.03, (n, 2)).astype(np.float32)
        prev = dict(x=p[:, 0], y=p[:, 1], w=w)
        cur = dict(x=q[:, 0], y=q[:, 1], w=wq.astype(np.float32), vx=v[:, 0], vy=v[:, 1])
        sim.metrics_record(prev, cur)
        o.record(prev, cur)
        active_sum += int((cur["w"] != 0).sum())
        assert np.array_equal(sim.metric_counters(), o.counters), f"record {k + 1} counters"
        assert np.array_equal(sim.metric_cells(), o.cells), f"record {k + 1} cells"
        p, w = q, cur["w"]
    assert np.array_equal(sim.metric_evac_steps(), o.evac_step)
    c = o.counters
    assert all(c[4 + e] > 0 for e in range(len(rects))) and c[4 + len(rects)] > 0
    assert o.cells[:, 4:8].all(axis=0).any() or o.cells[:, 4:8].any(axis=0).all()   # every threshold crossed somewhere
    # the 64-bit mass sums over all records hold exactly one person per active agent and record
    total = (o.cells[:, 8].astype(np.uint64) + (o.cells[:, 9].astype(np.uint64) << 32)).sum()
    assert int(total) == active_sum * 65536


def _in_step_run(layout, active, total, seed, steps, compare_every, reset_at=None):
    """One step at a time (eager launches), the oracle fed with each step's input and output."""
    scene = sh.Scene.layout(layout)
    E = scene.info.E
    rects = scene.tables()["exit_rects"][: 4 * E].reshape(-1, 4)
    a = sh.place_agents(layout, active, total, seed)
    sim = sh.Sim(scene, total)
    sim.upload(a)
    sim.metrics_enable(True)
    rx, ry, cpu, ox, oy = sim.metrics_grid()
    walk, thr = sim.metric_cell_info()
    assert (rx, ry, cpu, ox, oy) == scene.metric_grid()[:5]
    assert np.array_equal(walk, scene.metric_grid()[5].reshape(-1))
    assert np.array_equal(thr, op.metrics_thresholds(walk, walk >= np.float32(0.10)))   # MetricVenue::cellInfo's rule
    o = op.MetricsOracle(total, rx, ry, cpu, ox, oy, thr, rects)
    prev = sim.state()
    for k in range(1, steps + 1):
        if reset_at is not None and k == reset_at:
            sim.metrics_reset()
            o = op.MetricsOracle(total, rx, ry, cpu, ox, oy, thr, rects)
        sim.step(1)
        cur = sim.state()
        o.record(prev, cur)
        prev = cur
        if k % compare_every == 0 or k == steps:
            assert np.array_equal(sim.metric_counters(), o.counters), f"step {k} counters"
            assert np.array_equal(sim.metric_cells(), o.cells), f"step {k} cells"
    assert np.array_equal(sim.metric_evac_steps(), o.evac_step)
    return sim, o, E


def test_in_step_metrics_room_evacuation():
    print("metrics in step: room 4096, 400 steps", flush=True)
    sim, o, E = _in_step_run(sh.LAYOUT_ROOM, 4096, 4096, 99, 400, 25)
    st, ex, per = sim.evac()
    left = st >= 0
    assert left.sum() > 100 and o.counters[2] == left.sum()
    # the metrics' evacuation record agrees with the pipeline's own: the metric step counts from 1
    ms = sim.metric_evac_steps()
    assert np.array_equal(ms[left], st[left].astype(np.uint32) + 1) and not ms[~left].any()
    assert np.array_equal(o.counters[4:4 + E + 1], per)
    cells = o.cells
    mass_steps = cells[:, 8].astype(np.uint64) + (cells[:, 9].astype(np.uint64) << 32)
    assert cells[:, 3].max() > 0 and cells[:, 4].max() > 0 and mass_steps.sum() > 0


def test_in_step_metrics_bowl_and_reset():
    print("metrics in step: bowl 63424 of 65536, 12 steps, reset at 7", flush=True)
    sim, o, _ = _in_step_run(sh.LAYOUT_BOWL, 63424, 65536, 1337, 12, 3, reset_at=7)
    assert sim.metric_counters()[0] == 6 and sim.metric_counters()[1] == 63424


def test_metrics_graph_and_eager_identical():
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    a = sh.place_agents(sh.LAYOUT_ROOM, 4096, 4096, 5)
    out = []
    print("metrics: graph vs eager, room 4096, 151 steps", flush=True)
    for graph in (True, False):
        s = sh.Sim(scene, 4096, use_graph=graph)
        s.upload(a)
        s.metrics_enable(True)
        s.step(151)
        out.append((s.metric_counters(), s.metric_cells(), s.metric_evac_steps(), s.state()["x"]))
    for u, v in zip(out[0], out[1]):
        assert np.array_equal(u, v)
    # switching the passes off leaves the record where it was, on again continues it
    s.metrics_enable(False)
    s.step(10)
    assert np.array_equal(s.metric_counters(), out[1][0])
    s.metrics_enable(True)
    s.step(1)
    assert s.metric_counters()[0] == 152


def test_experiment_reports_room(tmp_path):
    """An evacuation run through experiment.py: the report's cumulative evacuations are the pipeline's own per-exit
    counts, its person-seconds the active agents of every step, and its evacuation times the reference's rule."""
    import experiment as ex

    print("metrics: experiment report, room 4096, 300 steps every 25", flush=True)
    x = ex.Experiment(sh.LAYOUT_ROOM, 4096, 4096, seed=99, every=25).run(300)
    summ = x.write(str(tmp_path))
    st, _, per = x.sim.evac()
    assert [s["step"] for s in x.samples] == list(range(25, 301, 25))
    assert np.array_equal(x.samples[-1]["exits"], per)
    assert summ["evacuated"] == int((st >= 0).sum()) > 100
    head = open(tmp_path / "timeseries.csv").readline().strip().split(",")
    assert head[:4] == ["step", "t_model", "t_real_s", "active"] and len(head) == 4 + x.E + 1 + 1
    lines = open(tmp_path / "heatmap.csv").read().strip().split("\n")
    walk = x.walkable
    assert len(lines) == 1 + int((walk >= np.float32(0.10)).sum())
    # person-seconds: every step's active agents, one person of mass each (cells below 0.10 m^2 hold a little)
    active_sum = 4096 * 300 - sum(300 - int(s) + 1 for s in x.sim.metric_evac_steps() if s)
    cells = x.sim.metric_cells()
    assert int(ex.sum64(cells, 8).sum()) == active_sum * 65536
    assert summ["person_s"] <= active_sum * ex.DT_REAL_S * (1 + 1e-9)
    t = ex.evacuation_times(x.sim.metric_evac_steps(), 4096)
    left = np.sort(st[st >= 0]) + 1
    if left.size >= 2048:
        assert t["t50_step"] == left[2047]
