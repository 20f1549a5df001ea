"""GPU parity of the batched step: every scenario of a batch equals its own plain run bit for bit.

The batched step has no counterpart in the reference; its oracle is the plain step (itself pinned to the
reference and the CPU restatement), run once per scenario on the same inputs. Compared per scenario: the agent
state, the sorted (cell, agent) pairs (global keys and ids less the scenario's bases), the cell table, the exit
choices and choice counters, and the evacuation record.
"""
import numpy as np
import pytest

import swarm_hip as sh
from common import assert_bits

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "vx", "vy", "speed", "w")
_SCENES = {}


def scene(layout):
    if layout not in _SCENES:
        _SCENES[layout] = sh.Scene.layout(layout)
    return _SCENES[layout]


def make_batch(layouts, scen_layout, totals, actives, seed=1337, **kw):
    scenes = [scene(l) for l in layouts]
    parts = [sh.place_agents(layouts[scen_layout[s]], int(actives[s]), int(totals[s]), seed + s) for s in range(len(totals))]
    agents = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    b = sh.BatchSim(scenes, scen_layout, totals, **kw)
    b.upload(agents)
    plains = []
    for s, p in enumerate(parts):
        sim = sh.Sim(scenes[scen_layout[s]], int(totals[s]), **kw)
        sim.upload(p)
        plains.append(sim)
    return b, plains


def compare(b, plains, what):
    st = b.state()
    key, idx = b.grid()
    cstart, ccount = b.cell_table()
    assign = b.assign()
    ev_step, ev_exit, _ = b.evac()
    ev_cnt = b.evac_counters()
    counters = b.choice_counters_all()
    inactive = key == sh.INACTIVE_CELL
    for s, p in enumerate(plains):
        lo, hi = int(b.agent_base[s]), int(b.agent_base[s]) + p.n
        ps = p.state()
        for f in FIELDS:
            assert_bits(st[f][lo:hi], ps[f], f"{what} scenario {s} {f}")
        # sorted pairs: the scenario's active slots, then its inactive agents in id order
        cb, nc = int(b.cell_base[s]), int(b.scen_cells[s])
        mine = (~inactive) & (key >= cb) & (key < cb + nc)
        pk, pi = p.grid()
        na = int(np.count_nonzero(pk != sh.INACTIVE_CELL))
        assert np.array_equal(key[mine] - cb, pk[:na]), f"{what} scenario {s} keys"
        assert np.array_equal(idx[mine] - lo, pi[:na]), f"{what} scenario {s} permutation"
        own_inactive = inactive & (idx >= lo) & (idx < hi)
        assert np.array_equal(idx[own_inactive] - lo, pi[na:]), f"{what} scenario {s} inactive order"
        ps0, pc0 = p.cell_table()
        first = int(np.flatnonzero(mine)[0]) if na else 0
        bc = ccount[cb:cb + nc]
        bs = np.where(bc > 0, cstart[cb:cb + nc] - first, 0)
        assert np.array_equal(bc, pc0) and np.array_equal(bs.astype(np.uint32), ps0), f"{what} scenario {s} cell table"
        assert np.array_equal(assign[lo:hi], p.assign()), f"{what} scenario {s} exit choices"
        assert np.array_equal(counters[s], p.choice_counters()), f"{what} scenario {s} choice counters"
        pst, pex, pper = p.evac()
        assert np.array_equal(ev_step[lo:hi], pst) and np.array_equal(ev_exit[lo:hi], pex), f"{what} scenario {s} evacuation"
        assert np.array_equal(ev_cnt[s][: pper.size], pper), f"{what} scenario {s} evacuation counters"


def run_and_compare(b, plains, checkpoints):
    done = 0
    for k in checkpoints:
        b.step(k - done)
        for p in plains:
            p.step(k - done)
        done = k
        compare(b, plains, f"step {done}")


@pytest.mark.parametrize("layout", list(sh.BATCH_LAYOUTS))
def test_unit_scenario_equals_plain(layout):
    cap = sh.layout_capacity(layout)
    b, plains = make_batch([layout], [0], [2048], [min(2048, cap // 2)])
    run_and_compare(b, plains, [1, 3, 19])


def test_mixed_batch_equals_plain_runs():
    # two scenarios of each layout, uneven sizes (some padded with parked agents, one nearly full), past one evaluation
    layouts = list(sh.BATCH_LAYOUTS)
    scen_layout = [0, 1, 2, 3, 3, 2, 1, 0]
    totals = [1024, 3000, 2500, 1300, 4096, 700, 1500, 5000]
    actives = [900, 2000, 2500, 1000, 3400, 700, 1100, 4800]
    b, plains = make_batch(layouts, scen_layout, totals, actives)
    run_and_compare(b, plains, [1, 2, 18, 24])


def test_candidate_budget_per_scenario():
    """A budget of half a move breaks the lists of the scenarios whose bodies move farther (their own flags: the other
    scenarios keep their lists); every scenario still equals its plain run without lists (the ring every iteration)."""
    layouts = list(sh.BATCH_LAYOUTS)
    scen_layout = [0, 1, 2, 3, 2, 0]
    totals = [2048, 1500, 3000, 1600, 2200, 4096]
    actives = [1800, 1500, 2900, 1200, 1000, 4000]
    b, plains = make_batch(layouts, scen_layout, totals, actives)
    b.set_candidate_budget(0.5)
    for p in plains:
        p.set_candidate_budget(0.0)
    run_and_compare(b, plains, [1, 5, 20, 40])


def test_launch_shapes_and_autotune_keep_results():
    layouts = list(sh.BATCH_LAYOUTS)
    scen_layout = [0, 1, 2, 3]
    totals = [2000, 1800, 3000, 1500]
    actives = [1500, 1800, 2600, 1200]
    b, plains = make_batch(layouts, scen_layout, totals, actives)
    for shape in [(64, 0), (256, 0), (128, 64), (256, 16)]:
        b.set_physics_shape(*shape)
        b.step(2)
        for p in plains:
            p.step(2)
        compare(b, plains, f"shape {shape}")
    before = b.state()
    ms = b.autotune(3)
    assert ms > 0.0
    after = b.state()
    for f in FIELDS:
        assert_bits(after[f], before[f], f"autotune restored {f}")
    blk, grd = b.physics_shape()
    assert blk in (64, 128, 192, 256) and grd >= 0
    b.step(19)
    for p in plains:
        p.step(19)
    compare(b, plains, "after autotune")


def test_ensemble_workload_sampled_scenarios():
    # the 256-scenario ensemble (about 1.4 M agents, 0.65 M active) stepped together; a sample of scenarios
    # re-run on their own
    scen_layout, totals, actives, agents = sh.batch_workload()
    layouts = list(sh.BATCH_LAYOUTS)
    scenes = [scene(l) for l in layouts]
    b = sh.BatchSim(scenes, scen_layout, totals)
    b.upload(agents)
    b.step(20)
    st = b.state()
    _, n_active, _ = b.control()
    assert n_active <= int(actives.sum())
    assert np.all(np.isfinite(st["x"])) and np.all(np.isfinite(st["y"]))
    key, _ = b.grid()
    act = key[key != sh.INACTIVE_CELL]
    assert np.all(np.diff(act.astype(np.int64)) >= 0), "global keys sorted"
    for s in (0, 1, 2, 3, 101, 255):
        lo = int(b.agent_base[s])
        n = int(totals[s])
        part = {k: v[lo:lo + n] for k, v in agents.items()}
        p = sh.Sim(scenes[scen_layout[s]], n)
        p.upload(part)
        p.step(20)
        ps = p.state()
        for f in FIELDS:
            assert_bits(st[f][lo:lo + n], ps[f], f"ensemble scenario {s} {f}")


def test_batch_scenarios_match_the_oracle_directly():
    """Each scenario of a batch (room, corridor, bowl eighth, pillar room; uneven counts, parked agents) against the
    oracle run on that scenario's own inputs: state every step for 20 steps, then exit choices and evacuation events,
    bit for bit. This pins the batched path to the CPU restatement without going through a plain GPU run."""
    from common import oracle_for

    layouts = [sh.LAYOUT_ROOM, sh.LAYOUT_CORRIDOR, sh.LAYOUT_BOWL_EIGHTH, sh.LAYOUT_PILLAR]
    scen_layout = np.array([0, 1, 2, 3, 2], np.int32)
    totals = np.array([1500, 1024, 2300, 1800, 1100], np.int32)
    actives = np.array([1400, 1024, 2000, 1700, 900], np.int32)
    scenes = [scene(l) for l in layouts]
    parts = [sh.place_agents(layouts[scen_layout[s]], int(actives[s]), int(totals[s]), 4242 + s) for s in range(len(totals))]
    agents = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    b = sh.BatchSim(scenes, scen_layout, totals)
    b.upload(agents)
    oracles = []
    for s, p in enumerate(parts):
        o = oracle_for(scenes[scen_layout[s]], int(totals[s]))
        o.set_state(p)
        oracles.append(o)
    for k in range(1, 21):
        b.step(1)
        st = b.state()
        assign = b.assign()
        for s, o in enumerate(oracles):
            o.step(1)
            lo, hi = int(b.agent_base[s]), int(b.agent_base[s]) + int(totals[s])
            os_ = o.state()
            for f in FIELDS:
                assert_bits(st[f][lo:hi], os_[f], f"step {k} scenario {s} {f}")
            assert np.array_equal(assign[lo:hi], o.assign()), f"step {k} scenario {s} exit choices"
    ev_step, ev_exit, _ = b.evac()
    for s, o in enumerate(oracles):
        lo, hi = int(b.agent_base[s]), int(b.agent_base[s]) + int(totals[s])
        cs, ce = o.evac()
        assert np.array_equal(ev_step[lo:hi], cs) and np.array_equal(ev_exit[lo:hi], ce), f"scenario {s} evacuation"
