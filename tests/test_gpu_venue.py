"""GPU: the analytic seated venue's queries and the step on the venues' plans.

The probe kernel (swarm_venue_probe) evaluates seated::sdfGrad / floorHeight / cellAt / wall / floorSlope / exitByLat /
gangwayAt on the GPU: bit for bit the reference's answers at its probe points (venue_kat.npz, written by the reference's
own src/venue.cpp) and the oracle's at fresh random points. The step on the plans of venues S, M and L runs the
reference's seated crowd (its seat spawn, global radius and free speed) under the reference's default venue model (the
incline speed and the stair spacing on the gangways) and is bit for bit the oracle's: state, sorted grid, cell table and
exit choices every step, then the evacuation curve and every agent's (step, exit) to the end.
"""
import os
import tempfile

import numpy as np
import pytest

import oracle_py as op
import swarm_hip as sh
from common import golden, oracle_for
from test_gpu_parity import compare_step
from test_venue_golden import assert_same_answers, reference_answers

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("venue", ["S", "M", "L"])
def test_probe_matches_reference(venue):
    kat = golden("venue_kat.npz")
    p, e = kat[f"venue__{venue}__probe__p"], kat[f"venue__{venue}__probe__e"]
    print(f"venue {venue}: {len(p)} reference probe points", flush=True)
    o = sh.venue_probe(venue, p, e, op.VENUE_REACHES)
    assert_same_answers(o, reference_answers(venue), f"venue {venue} GPU")


@pytest.mark.parametrize("venue,n", [("S", 200000), ("L", 400000)])
def test_probe_matches_oracle_at_random_points(venue, n):
    kat = golden("venue_kat.npz")
    lay = op.venue_layout_arrays(kat, venue)
    L = sh.venue_layout(venue)
    rng = np.random.default_rng(2024)
    H = L.landing_end + 0.02
    xy = rng.uniform(-H, H, (n, 2)).astype(np.float32)
    xy[: n // 8, 1] = xy[: n // 8, 0]   # on the diagonal
    fl = op.venue_eval(lay, np.c_[xy, np.zeros(n, np.float32)], np.zeros((n, 2), np.float32))["floor"]
    z = (fl + rng.choice(np.array([0.0, 0.0064, 0.0128], np.float32), n) + rng.uniform(-0.01, 0.02, n) * (rng.random(n) < 0.5)).astype(np.float32)
    p = np.c_[xy, z].astype(np.float32)
    ang = rng.uniform(0, 2 * np.pi, n)
    e = np.c_[np.cos(ang), np.sin(ang)].astype(np.float32)
    print(f"venue {venue}: {n} random probe points", flush=True)
    got = sh.venue_probe(venue, p, e, op.VENUE_REACHES)
    ref = op.venue_eval(lay, p, e)
    assert_same_answers(got, ref, f"venue {venue} GPU vs oracle")


def venue_pair(venue, n, load=1.0, seed=1337):
    scene = sh.Scene.venue(venue)
    a = sh.venue_agents(venue, n, load, seed)
    sim = sh.Sim(scene, n)
    sim.upload(a)
    o = oracle_for(scene, n)
    o.set_state(a)
    return sim, o, a


def test_venue_s_steps_bit_exact():
    """Venue S's plan, its 3,728 seats filled (4,096 agents): 60 steps, every quantity every step (three exit-choice
    evaluations)."""
    sim, o, _ = venue_pair("S", 4096)
    for k in range(1, 61):
        sim.step(1)
        o.step(1)
        compare_step(sim, o, f"venue S step {k}")
        assert np.array_equal(sim.assign(), o.assign()), f"venue S step {k} exit choices"
        assert np.array_equal(sim.choice_counters(), o.choice_counters()), f"venue S step {k} choice counters"


def test_venue_s_full_evacuation():
    """Venue S's plan run to its end: state every 250 steps, then every agent's evacuation step and exit and the
    per-exit counts identical to the oracle's; every seat empties through the avenues and the doors."""
    sim, o, a = venue_pair("S", 4096)
    done = 0
    while done < 3000:
        sim.step(250)
        o.step(250)
        done += 250
        compare_step(sim, o, f"venue S evacuation step {done}", cells=False)
        if (sim.evac()[0][a["w"] > 0] >= 0).all():
            break
    gs, ge, gper = sim.evac()
    cs, ce = o.evac()
    assert np.array_equal(gs, cs) and np.array_equal(ge, ce)
    active = a["w"] > 0
    assert (gs[active] >= 0).all(), "every seated agent leaves within 3,000 steps"
    assert gper[:-1].sum() == active.sum() and (gper[:-1] > 0).all()


def test_venue_s_incline_walkers_bit_exact():
    """Single agents on venue S's gangways, driven up and down the inclines and along the level concourse in all four
    quadrants, two of them behind a standing agent: 60 steps, every quantity every step, against the oracle (whose
    incline rules are pinned to the reference's inclineFreeSpeed / inclineHeadwayTime); the free walkers settle at the
    stair speeds of the rake and at the level speed."""
    scene = sh.Scene.venue("S")
    # per quadrant, one walk in each of its two wings (x <-> y): up the gangway, down it, up and down off-centre, and
    # along the level concourse; no two walkers share a gangway
    walks = {(1, 1): [(0.3498, 0.50, 0.0, 1.0)] * 2, (-1, 1): [(0.3498, 0.80, 0.0, -1.0)] * 2,
             (-1, -1): [(0.3402, 0.62, 0.0, 1.0), (0.3594, 0.70, 0.0, -1.0)], (1, -1): [(0.10, 0.912, 1.0, 0.0)] * 2}
    starts = []
    for (sx, sy), pair in walks.items():
        for swap, (lat, dep, dl, dd) in zip((False, True), pair):
            x, y, dx, dy = (dep, lat, dd, dl) if swap else (lat, dep, dl, dd)
            starts.append((sx * x, sy * y, sx * dx, sy * dy))
    # a standing agent just ahead of two of them: the stair time gap in the headway rule
    n = 64
    a = dict(x=np.full(n, 1000.0, np.float32), y=np.full(n, 1000.0, np.float32), vx=np.zeros(n, np.float32),
             vy=np.zeros(n, np.float32), w=np.zeros(n, np.float32), r=np.full(n, 0.0064, np.float32),
             v0=np.full(n, 0.75, np.float32))
    dx = np.zeros(n, np.float32)
    dy = np.zeros(n, np.float32)
    for i, (x, y, ex, ey) in enumerate(starts):
        a["x"][i], a["y"][i], a["w"][i], dx[i], dy[i] = x, y, 1.0, ex, ey
    for j, i in enumerate((0, 2)):
        a["x"][40 + j] = a["x"][i] + 0.02 * dx[i]
        a["y"][40 + j] = a["y"][i] + 0.02 * dy[i]
        a["w"][40 + j] = 1.0
    sim = sh.Sim(scene, n)
    sim.upload(a)
    sim.set_drive(dx, dy)
    o = oracle_for(scene, n)
    o.set_state(a)
    o.set_drive(dx, dy)
    print(f"venue S: {len(starts)} incline walkers", flush=True)
    for k in range(1, 61):
        sim.step(1)
        o.step(1)
        compare_step(sim, o, f"venue S incline walkers step {k}")
    (v_up, v_down), _ = op.incline(np.array([0.5, -0.5], np.float32))
    st = sim.state()
    speed = np.hypot(st["vx"], st["vy"])
    assert abs(speed[1] - v_up) < 1e-3 * v_up   # up the gangway, nobody ahead
    assert abs(speed[3] - v_down) < 1e-3 * v_down   # down it
    assert abs(speed[6] - 0.75) < 1e-3 * 0.75   # along the concourse


def test_venue_adjoint_is_refused():
    """The adjoint differentiates a rectangle world's step; a venue plan's incline rules have none, so it refuses."""
    scene = sh.Scene.venue("S")
    a = sh.venue_agents("S", 4096, 1.0, 1337)
    sim = sh.Sim(scene, 4096)
    sim.upload(a)
    with pytest.raises(RuntimeError, match="incline"):
        sim.adjoint_forward(4)


def test_venue_m_steps_bit_exact():
    """Venue M's plan, 15,832 of 16,384 agents seated: 20 steps with every quantity every step, then on into the
    congested regime (the gangways and doors queueing) to step 400, checked every 40 steps, with the evacuation record."""
    sim, o, _ = venue_pair("M", 16384)
    for k in range(1, 21):
        sim.step(1)
        o.step(1)
        compare_step(sim, o, f"venue M step {k}")
        assert np.array_equal(sim.assign(), o.assign()), f"venue M step {k} exit choices"
    for k in range(60, 401, 40):
        sim.step(40)
        o.step(40)
        compare_step(sim, o, f"venue M step {k}", cells=False)
        assert np.array_equal(sim.assign(), o.assign()), f"venue M step {k} exit choices"
    gs, ge, _ = sim.evac()
    cs, ce = o.evac()
    assert np.array_equal(gs, cs) and np.array_equal(ge, ce)
    assert (gs >= 0).sum() > 1000   # people did leave


def test_venue_l_spawn_steps_bit_exact():
    """Venue L's plan at spawn, 63,424 of 65,536 agents seated (the workload of the reference's venue L timings): 3 steps,
    every quantity every step."""
    scene = sh.Scene.venue("L", cache_dir=os.path.join(tempfile.gettempdir(), "swarm_scene_cache"))
    a = sh.venue_agents("L", 65536, 1.0, 1337)
    assert int(a["w"].sum()) == 63424
    sim = sh.Sim(scene, 65536)
    sim.upload(a)
    o = oracle_for(scene, 65536)
    o.set_state(a)
    for k in range(1, 4):
        sim.step(1)
        o.step(1)
        compare_step(sim, o, f"venue L step {k}")
        assert np.array_equal(sim.assign(), o.assign()), f"venue L step {k} exit choices"
