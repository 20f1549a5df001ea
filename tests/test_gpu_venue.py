"""GPU: the analytic seated venue's queries and the step on the venues' plans.

The probe kernel (swarm_venue_probe) evaluates seated::sdfGrad / floorHeight / cellAt / wall / floorSlope / exitByLat /
gangwayAt on the GPU: bit for bit the reference's answers at its probe points (venue_kat.npz, written by the reference's
own src/venue.cpp) and the oracle's at fresh random points. The step on the plan of venue S (and M at spawn) runs the
reference's seated crowd (its seat spawn, global radius and free speed) and is bit for bit the oracle's: state, sorted
grid, cell table and exit choices every step, then the evacuation curve and every agent's (step, exit) to the end.
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
