"""GPU parity: the HIP path through the C ABI against the reference's golden vectors and the oracle.

Bars: bit-exact for everything (hash keys, sorted permutation, cell table, exit choices and counters,
positions, velocities, speeds, flags): the kernels evaluate the oracle's float operations in the same
order, so no tolerance is needed. The per-step bar of the task (1e-4 m, 1e-4 m/s) is strictly weaker.
"""
import numpy as np
import pytest

import swarm_hip as sh
from common import agents_from_golden, assert_bits, golden, kat_rect_scene, oracle_for, terrace_scene

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "vx", "vy", "speed", "w")


def compare_step(sim, o, what, cells=True):
    g, c = sim.state(), o.state()
    for f in FIELDS:
        assert_bits(g[f], c[f], f"{what} {f}")
    gk, gi = sim.grid()
    ck, ci = o.grid()
    assert np.array_equal(gk, ck), f"{what} keys"
    assert np.array_equal(gi, ci), f"{what} permutation"
    if cells:
        gs, gc = sim.cell_table()
        cs, cc = o.cell_table()
        assert np.array_equal(gs, cs) and np.array_equal(gc, cc), f"{what} cell table"


# ---------------- against the reference's own golden vectors ----------------
@pytest.mark.parametrize("tag,steps", [("t_sparse", 6), ("t_dense", 8), ("t_bsparse", 6), ("t_bdense", 8)])
@pytest.mark.parametrize("graph", [False, True])
def test_gpu_matches_reference_terrace(tag, steps, graph):
    """t_b*: the reference sorted by its GPU's bitonic network (SortMode::GpuBitonic), this path in sort mode 1."""
    kat = golden("terrace_kat.npz")
    a = agents_from_golden(kat, tag)
    n = a["x"].size
    sim = sh.Sim(terrace_scene(), n, use_graph=graph)
    sim.set_sort_mode(tag.startswith("t_b"))
    sim.upload(a)
    sim.set_drive(kat[tag + "__in__dx"], kat[tag + "__in__dy"])
    for k in range(1, steps + 1):
        sim.step(1)
        p = f"{tag}__step{k}__"
        s = sim.state()
        for f in FIELDS:
            assert_bits(s[f], kat[p + f], f"{tag} step {k} {f}")
        key, idx = sim.grid()
        assert np.array_equal(key, kat[p + "key"]) and np.array_equal(idx, kat[p + "id"])
        st, cn = sim.cell_table()
        assert np.array_equal(st, kat[p + "cell_start"]) and np.array_equal(cn, kat[p + "cell_count"])


def test_gpu_choice_matches_reference():
    import test_choice_golden as tcg

    ck = golden("choice_kat.npz")
    room = kat_rect_scene(golden("nav_kat.npz"), "room2")
    tcg.check_all(lambda n: sh.Sim(room, n), ck)


# ---------------- against the oracle on the benchmark floors ----------------
def run_pair(layout, active, total, seed, steps, check_every=1, proj_iters=32, cells=True, **place):
    scene = sh.Scene.layout(layout)
    a = sh.place_agents(layout, active, total, seed, **place)
    sim = sh.Sim(scene, total, proj_iters=proj_iters)
    sim.upload(a)
    o = oracle_for(scene, total, proj_iters=proj_iters)
    o.set_state(a)
    done = 0
    while done < steps:
        k = min(check_every, steps - done)
        sim.step(k)
        o.step(k)
        done += k
        compare_step(sim, o, f"layout {layout} step {done}", cells=cells)
        assert np.array_equal(sim.assign(), o.assign()), f"layout {layout} step {done} exit choices"
        assert np.array_equal(sim.choice_counters(), o.choice_counters()), f"layout {layout} step {done} choice counters"
    return sim, o


def test_room_4096_steps_bit_exact():
    """The small plain config: 4,096 agents, 40 x 30 m room; 60 steps, every quantity every step (crosses
    three exit-choice evaluations)."""
    run_pair(sh.LAYOUT_ROOM, 4096, 4096, 1337, 60)


def test_room_evacuation_curve():
    """Evacuation events (agent, step, exit) identical over 600 steps."""
    sim, o = run_pair(sh.LAYOUT_ROOM, 4096, 4096, 99, 600, check_every=100, cells=False)
    gs, ge, gper = sim.evac()
    cs, ce = o.evac()
    assert np.array_equal(gs, cs) and np.array_equal(ge, ce)
    assert (gs >= 0).sum() > 100   # people did leave
    assert gper.sum() == (gs >= 0).sum()


def test_room_full_evacuation():
    """The small plain config run to its end (every agent out, within 5,000 steps; graph replays): positions,
    velocities and exit choices every 500 steps, then the evacuated-count curve step by step, the final
    evacuation step and every agent's (step, exit), identical to the oracle's."""
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    a = sh.place_agents(sh.LAYOUT_ROOM, 4096, 4096, 99)
    sim = sh.Sim(scene, 4096)
    sim.upload(a)
    o = oracle_for(scene, 4096)
    o.set_state(a)
    done = 0
    while done < 5000:
        sim.step(500)
        o.step(500)
        done += 500
        compare_step(sim, o, f"room evacuation step {done}", cells=False)
        assert np.array_equal(sim.assign(), o.assign()), f"room evacuation step {done} exit choices"
        if (sim.evac()[0] >= 0).all():
            break
    gs, ge, gper = sim.evac()
    cs, ce = o.evac()
    assert (gs >= 0).all(), f"agents left after {done} steps: {(gs < 0).sum()}"
    assert np.array_equal(gs, cs) and np.array_equal(ge, ce)
    assert gs.max() == cs.max() and gs.max() < done
    curve = lambda s: np.cumsum(np.bincount(s, minlength=done))
    assert np.array_equal(curve(gs), curve(cs))
    assert np.array_equal(gper[: scene.info.E + 1], np.bincount(ce, minlength=scene.info.E + 1))


def test_bowl_63424_agents_bit_exact():
    """The large plain config: 63,424 agents on the 200 x 150 m bowl, sort size 65,536 (2,112 parked)."""
    run_pair(sh.LAYOUT_BOWL, 63424, 65536, 1337, 3)


def test_bowl_63424_agents_150_steps():
    """The large plain config over 150 steps (the crowd leaves its seats; an exit-choice evaluation every 17
    steps), every quantity at steps 50, 100 and 150."""
    run_pair(sh.LAYOUT_BOWL, 63424, 65536, 1337, 150, check_every=50)


@pytest.mark.parametrize("layout", [sh.LAYOUT_CORRIDOR, sh.LAYOUT_PILLAR, sh.LAYOUT_BOWL_EIGHTH])
def test_other_floors_bit_exact(layout):
    cap = sh.layout_capacity(layout)
    run_pair(layout, min(3000, cap), 3001, 5, 20, check_every=5)


def test_edge_cases_sizes_and_flags():
    """Agent counts that are not multiples of the block size, a single agent, all-inactive, no projection."""
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    for active, total, it in [(1, 1, 32), (200, 777, 32), (0, 300, 32), (500, 513, 0), (500, 513, 1)]:
        a = sh.place_agents(sh.LAYOUT_ROOM, active, total, 3)
        sim = sh.Sim(scene, total, proj_iters=it)
        sim.upload(a)
        o = oracle_for(scene, total, proj_iters=it)
        o.set_state(a)
        for k in range(3):
            sim.step(1)
            o.step(1)
            compare_step(sim, o, f"edge {active}/{total}/{it} step {k}")


def test_nan_position_reset():
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    a = sh.place_agents(sh.LAYOUT_ROOM, 300, 300, 11)
    a["x"][17] = np.nan
    sim = sh.Sim(scene, 300)
    sim.upload(a)
    o = oracle_for(scene, 300)
    o.set_state(a)
    sim.step(1)
    o.step(1)
    compare_step(sim, o, "nan step 1")
    assert not np.isnan(sim.state()["x"]).any()


def test_graph_and_eager_identical():
    """Determinism: a hipGraph replay and eager launches give the same bits (two independent runs)."""
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    a = sh.place_agents(sh.LAYOUT_ROOM, 4096, 4096, 21)
    s1 = sh.Sim(scene, 4096, use_graph=True)
    s2 = sh.Sim(scene, 4096, use_graph=False)
    s1.upload(a)
    s2.upload(a)
    s1.step(41)
    s2.step(41)
    x1, x2 = s1.state(), s2.state()
    for f in FIELDS:
        assert_bits(x1[f], x2[f], f"graph vs eager {f}")
    assert np.array_equal(s1.assign(), s2.assign())


@pytest.mark.parametrize("layout,active,total,steps", [(sh.LAYOUT_ROOM, 4096, 4096, 120), (sh.LAYOUT_BOWL, 63424, 65536, 12)])
def test_candidate_pruning_is_exact(layout, active, total, steps):
    """Pruning the projection's contact candidates (every 1, 4 or 8 iterations), 2, 4 or 8 lanes per slot in the
    list iterations, in the first (collecting) iteration and in predict (its ring and its wall contacts) give the
    single-lane, unpruned run's bits (evacuation records included)."""
    scene = sh.Scene.layout(layout)
    a = sh.place_agents(layout, active, total, 8)
    runs = []
    for cfg in ((0, 1, 1, 1), (1, 1, 1, 1), (4, 1, 1, 1), (8, 1, 1, 1), (0, 2, 1, 1), (0, 4, 1, 1), (8, 4, 1, 1), (0, 8, 1, 1),
                (0, 1, 2, 1), (0, 1, 4, 1), (0, 1, 8, 1), (8, 1, 4, 1), (0, 1, 1, 2), (0, 1, 1, 4), (0, 1, 1, 8), (8, 1, 4, 4)):
        every, lanes, collect, pred = cfg
        s = sh.Sim(scene, total)
        s.set_projection_prune(every)
        s.set_projection_lanes(lanes)
        s.set_projection_collect_lanes(collect)
        s.set_predict_lanes(pred)
        s.upload(a)
        s.step(steps)
        st, ev = s.state(), s.evac()
        runs.append((cfg, st, s.assign(), ev))
    for cfg, st, asg, ev in runs[1:]:
        for f in FIELDS:
            assert_bits(st[f], runs[0][1][f], f"prune every {cfg[0]}, lanes: {cfg[1]} list, {cfg[2]} collecting, {cfg[3]} predict: {f}")
        assert np.array_equal(asg, runs[0][2])
        for u, v in zip(ev, runs[0][3]):
            assert np.array_equal(u, v), f"evacuations, {cfg}"


@pytest.mark.parametrize("layout,active,total,steps", [(sh.LAYOUT_ROOM, 4096, 4096, 150), (sh.LAYOUT_BOWL, 63424, 65536, 40)])
def test_candidate_budget_is_exact(layout, active, total, steps):
    """The candidate lists' displacement budget (swarm_set_candidate_budget): no budget (the lists reach over all P
    moves, no checks), the default 12 moves, 1 and 0.5 moves (the lists break in most steps and those steps' later
    iterations scan the ring), with pruning and 4 lanes per slot in the list and collecting iterations, and no lists at
    all (the ring every iteration), give the same bits every 10 steps (evacuation records included); the default lists
    are shorter than the unbudgeted ones."""
    scene = sh.Scene.layout(layout)
    a = sh.place_agents(layout, active, total, 21)
    runs = []
    counts = {}
    for cfg in ((float("inf"), 0, 1), (float("inf"), 8, 1), (12.0, 8, 1), (1.0, 8, 4), (0.5, 0, 1), (0.0, 8, 1), (12.0, 0, 4)):
        moves, every, lanes = cfg
        s = sh.Sim(scene, total)
        s.set_candidate_budget(moves)
        s.set_projection_prune(every)
        s.set_projection_lanes(lanes)
        s.set_projection_collect_lanes(lanes)
        s.upload(a)
        print(f"candidate budget {cfg}", flush=True)
        snaps = []
        for k in range(steps // 10):
            s.step(10)
            snaps.append(s.state())
        counts[cfg] = s.candidate_lists()[0]
        runs.append((cfg, snaps, s.assign(), s.evac()))
        s.close()
    cfg0, snaps0, asg0, ev0 = runs[0]
    for cfg, snaps, asg, ev in runs[1:]:
        for k, (st, st0) in enumerate(zip(snaps, snaps0)):
            for f in FIELDS:
                assert_bits(st[f], st0[f], f"budget {cfg}, step {10 * (k + 1)}: {f}")
        assert np.array_equal(asg, asg0)
        for u, v in zip(ev, ev0):
            assert np.array_equal(u, v), f"evacuations, {cfg}"
    # the lists of the last step's collecting iteration (unpruned runs): shorter with the budget
    c_inf, c_12 = counts[(float("inf"), 0, 1)], counts[(12.0, 0, 4)]
    live = (c_inf != 0xFFFFFFFF) & (c_12 != 0xFFFFFFFF)
    assert live.sum() > active // 2 and c_12[live].sum() < c_inf[live].sum()


def test_collecting_lanes_write_the_same_lists():
    """After the first iteration the candidate lists (count, indices, contact distances) of 2, 4 and 8 collecting
    lanes are the single lane's, entry for entry, on the bowl and in the dense room."""
    for layout, active, total in ((sh.LAYOUT_BOWL, 63424, 65536), (sh.LAYOUT_ROOM, 4096, 4096)):
        scene = sh.Scene.layout(layout)
        a = sh.place_agents(layout, active, total, 3)
        lists = []
        for collect in (1, 2, 4, 8):
            s = sh.Sim(scene, total, proj_iters=2)   # the first iteration, then the final pass (which reads the lists)
            s.set_projection_prune(0)
            s.set_projection_collect_lanes(collect)
            s.upload(a)
            s.step(3)
            key, ids = s.grid()
            moving = (key != sh.INACTIVE_CELL) & (s.state()["w"][ids] != 0)   # the slots that collected this step
            lists.append((s.candidate_lists(), s.state(), moving))
        (c0, k0, d0), st0, mv = lists[0]
        c0 = np.where(mv, c0, 0)
        for (c, k, d), st, _ in lists[1:]:
            assert np.array_equal(np.where(mv, c, 0), c0)
            live = np.arange(k.shape[0])[:, None] < np.where(c0 == 0xFFFFFFFF, 0, c0)[None, :]
            assert np.array_equal(np.where(live, k, 0), np.where(live, k0, 0))
            assert np.array_equal(np.where(live, d.view(np.uint32), 0), np.where(live, d0.view(np.uint32), 0))
            for f in FIELDS:
                assert_bits(st[f], st0[f], f"collecting lanes {f}")
        assert ((c0 > 0) & (c0 != 0xFFFFFFFF)).sum() > 100   # lists were written and compared


def test_sort_grid_on_given_keys():
    """CrowdSim::sortGrid: the sort stage alone on keys of the caller's own, against a stable sort by (key, index)."""
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    n = 5000
    sim = sh.Sim(scene, n)
    sim.upload(sh.place_agents(sh.LAYOUT_ROOM, 100, n, 1))
    rng = np.random.default_rng(7)
    for dense in (3, 200, sim.cells):
        keys = rng.integers(0, dense, n).astype(np.uint32)
        keys[rng.random(n) < 0.1] = sh.INACTIVE_CELL
        sk, ids = sim.sort_grid(keys)
        order = np.argsort(keys.astype(np.uint64), kind="stable")
        assert np.array_equal(ids, order.astype(np.uint32)), f"permutation ({dense} distinct keys)"
        assert np.array_equal(sk, keys[order]), "sorted keys"
        st, cn = sim.cell_table()
        act = keys[keys != sh.INACTIVE_CELL]
        assert np.array_equal(cn, np.bincount(act, minlength=sim.cells).astype(np.uint32)), "cell counts"
    bad = np.full(n, sim.cells, np.uint32)
    with pytest.raises(RuntimeError):
        sim.sort_grid(bad)


def test_previous_state_is_last_input():
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    a = sh.place_agents(sh.LAYOUT_ROOM, 1000, 1024, 4)
    sim = sh.Sim(scene, 1024)
    sim.upload(a)
    p0 = sim.prev_state()
    assert_bits(p0["x"], a["x"], "before the first step: the upload")
    for k in range(3):
        before = sim.state()
        sim.step(1)
        prev = sim.prev_state()
        for f in FIELDS:
            assert_bits(prev[f], before[f], f"previous state after step {k + 1} {f}")


# ---------------- the bitonic sort mode ----------------
@pytest.mark.parametrize("layout,active,total,steps,every", [(sh.LAYOUT_ROOM, 4096, 4096, 60, 1), (sh.LAYOUT_BOWL, 63424, 65536, 3, 1)])
def test_bitonic_mode_bit_exact(layout, active, total, steps, every):
    """The reference GPU's bitonic network as the sort (its unstable within-cell order changes every neighbour sum):
    the whole step against the oracle in the same mode, every quantity every step."""
    scene = sh.Scene.layout(layout)
    a = sh.place_agents(layout, active, total, 1337)
    sim = sh.Sim(scene, total)
    sim.set_sort_mode(True)
    sim.upload(a)
    o = oracle_for(scene, total)
    o.set_sort_mode(True)
    o.set_state(a)
    first = None
    for k in range(steps):
        sim.step(1)
        o.step(1)
        compare_step(sim, o, f"bitonic layout {layout} step {k + 1}")
        assert np.array_equal(sim.assign(), o.assign())
        assert np.array_equal(sim.choice_counters(), o.choice_counters())
        if first is None:
            first = sim.grid()[1]
    counting = sh.Sim(scene, total)
    counting.upload(a)
    counting.step(1)
    assert not np.array_equal(counting.grid()[1], first)   # the two sorts do differ within cells


def test_bitonic_sort_grid_on_given_keys():
    """CrowdSim::sortGrid in the bitonic mode: the network's permutation on keys of the caller's own (few distinct
    keys, all equal, descending, inactive), sizes 2 .. 65,536, against the oracle's network."""
    import oracle_py as op

    rng = np.random.default_rng(3)
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    for n in (2, 256, 1024, 4096, 8192, 65536):
        sim = sh.Sim(scene, n)
        sim.upload(sh.place_agents(sh.LAYOUT_ROOM, min(n, 100), n, 1))
        sim.set_sort_mode(True)
        cases = [rng.integers(0, 7, n), np.zeros(n, np.int64), np.arange(n)[::-1] % sim.cells, rng.integers(0, sim.cells, n)]
        for keys in cases:
            keys = keys.astype(np.uint32)
            keys[rng.random(n) < 0.1] = sh.INACTIVE_CELL
            sk, ids = sim.sort_grid(keys)
            ok, oi = op.bitonic_sort(keys)
            assert np.array_equal(sk, ok) and np.array_equal(ids, oi), f"n {n}"
            st, cn = sim.cell_table()
            act = keys[keys != sh.INACTIVE_CELL]
            assert np.array_equal(cn, np.bincount(act, minlength=sim.cells).astype(np.uint32))
    with pytest.raises(RuntimeError):
        sh.Sim(scene, 300).set_sort_mode(True)   # not a power of two
