"""GPU: the seated venues' own 3D step (a context on swarm_scene_build_seated) against the reference and the oracle.

The reference's own CpuCrowdSim on venue S (seated_kat.npz: its 100 % seat spawn, 3,728 of 4,096 slots) is reproduced
bit for bit by the HIP step: positions with heights, velocities with vertical components and speeds, after a forced
random drive of 40 steps and after the default model's first 200 steps (with its sorted grid, assigned exits and choice
counters). Then venue S runs to its last evacuee against the oracle (whose seated step is the reference's, pinned by the
same fixtures on the CPU): the state every 250 steps and every agent's evacuation step and exit, bit for bit; the last
seated agent leaves at step index 1,360, the reference's published CPU run (results/RESULTS.md:90: step 1361). Venue M
(32 exits: the exit-type rule of its candidate table) after 10 default steps, bit for bit the reference's.
"""
import os
import tempfile

import numpy as np
import pytest

import oracle_py as op
import swarm_hip as sh
from common import golden

pytestmark = pytest.mark.gpu

N = 4096
CACHE = os.path.join(tempfile.gettempdir(), "swarm_scene_cache")


@pytest.fixture(scope="module")
def kat():
    return golden("seated_kat.npz")


@pytest.fixture(scope="module")
def scene_s():
    return sh.Scene.seated("S", cache_dir=CACHE)


def assert_same(sim, kat, p, tag):
    pos, vel = sim.state3()
    rp, rv = kat[p + "pos"], kat[p + "vel"]
    bad = ~((pos.view(np.uint32) == rp.view(np.uint32)).all(1) & (vel.view(np.uint32) == rv.view(np.uint32)).all(1))
    if bad.any():
        i = int(np.nonzero(bad)[0][0])
        raise AssertionError(f"{tag}: {int(bad.sum())} agents differ; agent {i}: {pos[i]} {vel[i]} vs {rp[i]} {rv[i]}")


def test_seated_drive_matches_reference(kat, scene_s):
    sim = sh.Sim(scene_s, N)
    sim.upload3(kat["seated__S__init__pos"], kat["seated__S__init__vel"])
    d = kat["seated__S__drive__dir"]
    sim.set_drive(d[:, 0], d[:, 1])
    print("venue S seated step, forced drive", flush=True)
    done = 0
    for s in (1, 2, 3, 5, 10, 20, 40):
        sim.step(s - done)
        done = s
        assert_same(sim, kat, f"seated__S__drive__step{s}__", f"drive step {s}")
    sim.close()


def test_seated_default_matches_reference(kat, scene_s):
    sim = sh.Sim(scene_s, N)
    sim.upload3(kat["seated__S__init__pos"], kat["seated__S__init__vel"])
    print("venue S seated step, default model", flush=True)
    done = 0
    for s in (1, 2, 3, 5, 10, 25, 50, 100, 200):
        sim.step(s - done)
        done = s
        p = f"seated__S__default__step{s}__"
        assert_same(sim, kat, p, f"default step {s}")
        assert np.array_equal(sim.assign(), kat[p + "assign"]), f"default step {s} exits"
        assert np.array_equal(sim.choice_counters(), kat[p + "counters"]), f"default step {s} counters"
        if s == 1:
            key, sid = sim.grid()
            assert np.array_equal(key, kat["seated__S__default__grid_key"]) and np.array_equal(sid, kat["seated__S__default__grid_id"])
    sim.close()


def test_seated_s_evacuation_matches_oracle(kat, scene_s):
    kind, lay = scene_s.venue_kind()
    assert kind == 2
    o = op.SeatedOracle(op.venue_layout_arrays_of(lay), (lay.domain_min, lay.domain_span, lay.grid_res),
                        (lay.exit_depth, lay.exit_max_z, lay.domain_min + lay.domain_span), N, nav=(scene_s.info, scene_s.tables()))
    o.set_state3(kat["seated__S__init__pos"], kat["seated__S__init__vel"])
    sim = sh.Sim(scene_s, N)
    sim.upload3(kat["seated__S__init__pos"], kat["seated__S__init__vel"])
    active = kat["seated__S__init__pos"][:, 3] > 0
    done = 0
    while done < 2000:
        sim.step(250)
        o.step(250)
        done += 250
        pos, vel = sim.state3()
        opos, ovel = o.state3()
        assert np.array_equal(pos.view(np.uint32), opos.view(np.uint32)), f"step {done} positions"
        assert np.array_equal(vel.view(np.uint32), ovel.view(np.uint32)), f"step {done} velocities"
        print(f"venue S seated step {done}: {int((pos[:, 3] > 0).sum())} active", flush=True)
        if (pos[active, 3] == 0).all():
            break
    gs, ge, per = sim.evac()
    cs, ce = o.evac()
    assert np.array_equal(gs, cs) and np.array_equal(ge, ce)
    assert gs[active].max() == 1360   # the reference's CPU run: its last evacuee at step 1361 (1-based)
    assert per[:12].sum() == active.sum() and (per[:12] > 0).all()
    sim.close()


def test_seated_m_matches_reference(kat):
    """Venue M (32 exits: the candidate table's exit-type rule), 15,832 seated in 16,384 slots: 10 default steps."""
    print("venue M seated scene: building", flush=True)
    scene = sh.Scene.seated("M", cache_dir=CACHE)
    sim = sh.Sim(scene, 16384)
    sim.upload3(kat["seated__M__init__pos"], kat["seated__M__init__vel"])
    sim.step(10)
    p = "seated__M__default__step10__"
    assert_same(sim, kat, p, "venue M step 10")
    assert np.array_equal(sim.assign(), kat[p + "assign"]) and np.array_equal(sim.choice_counters(), kat[p + "counters"])
    sim.close()


def test_seated_candidate_lists_match_ring_scan(kat, scene_s):
    """The projection's candidate lists (swarm_set_candidate_budget) against the ring scan every iteration (budget 0):
    venue S from its seat spawn through the congested gangways, 400 steps, bit for bit with no budget beyond the
    projection's P - 1 maximal moves (infinity), at the default 12 and at 0.62 moves (which breaks its lists early in most
    steps and falls back to the ring); the lists are collected (counts within capacity for nearly every active slot)
    and shorter at the smaller budget."""
    runs = {}
    for budget in (0.0, float("inf"), 12.0, 0.62):
        sim = sh.Sim(scene_s, N)
        sim.set_candidate_budget(budget)
        sim.upload3(kat["seated__S__init__pos"], kat["seated__S__init__vel"])
        print(f"venue S seated step, candidate budget {budget}", flush=True)
        states = []
        for k in range(4):
            sim.step(100)
            states.append(sim.state3())
        cnt = sim.candidate_lists()[0]
        runs[budget] = (states, cnt)
        sim.close()
    for budget in (float("inf"), 12.0, 0.62):
        for k, ((p, v), (rp, rv)) in enumerate(zip(runs[budget][0], runs[0.0][0])):
            assert np.array_equal(p.view(np.uint32), rp.view(np.uint32)), f"budget {budget}, step {100 * (k + 1)}: positions"
            assert np.array_equal(v.view(np.uint32), rv.view(np.uint32)), f"budget {budget}, step {100 * (k + 1)}: velocities"
    active = int((runs[0.0][0][-1][0][:, 3] > 0).sum())
    c1, c2 = runs[float("inf")][1][:active], runs[0.62][1][:active]
    assert active > 1000 and (c1 != 0xFFFFFFFF).mean() > 0.99
    assert c2[c2 != 0xFFFFFFFF].sum() < c1[c1 != 0xFFFFFFFF].sum()


def test_seated_scene_refuses_plain_uploads_and_batches(scene_s):
    sim = sh.Sim(scene_s, 64)
    with pytest.raises(RuntimeError, match="metric"):
        sim.metrics_enable(True)
    sim.close()
    room = sh.Scene.layout(sh.LAYOUT_ROOM)
    plain = sh.Sim(room, 64)
    with pytest.raises(RuntimeError, match="seated"):
        plain.upload3(np.zeros((64, 4), np.float32), np.zeros((64, 4), np.float32))
    plain.close()
