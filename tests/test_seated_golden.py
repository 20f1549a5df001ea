"""The seated venues' own 3D step (the reference's CpuCrowdSim on an analytic venue) against the reference itself.

seated_kat.npz holds what the reference's own code (src/cpu_crowd_sim.cpp, src/nav_field.cpp, src/exit_choice.cpp, built
by oracle/ref_build.sh and run by oracle/ref_driver.cpp) makes on venue S from its 100 % seat spawn (3,728 of 4,096):
a forced random drive for 40 steps (every seated-step branch: floors, gravity, the exact walls, the wall plane, the
projection's height test, the step-up limit with its slides and face margin, the fence) and the default model for 200
steps (its navigation field and candidate table, the queue-aware choice with the surface rule, the incline rules),
with the navigation field's and table's FNV-1a checksums.

Checked here, on the CPU:
  - the product's seated scene builder (swarm_scene_build_seated) makes the reference's navigation classes and candidate
    table for venues S and M (M: 32 exits, more than a cell's 24 candidates: the exit-type rule) bit for bit: the FNV-1a of the cell classes, of the index (class bit, entry number) and of every
    entry's 24 slots and 12 distance words, rebuilt in the reference's layout from the product's dense table, equal the
    reference's; its capacities and scalars equal the reference's; every 997th cell's entry equals the reference's;
  - the oracle's seated step equals the reference's bit for bit: positions, velocities (with speed) and contact bits at
    the recorded steps of both runs, and the sorted grid, the assigned exits and the choice counters of the default run;
    on venue M (15,832 seated in 16,384 slots) after 10 default steps.
"""
import os
import tempfile

import numpy as np
import pytest

import oracle_py as op
import swarm_hip as sh
from common import golden

CHOICE_FIT = 0x80000000
EMPTY_SLOT = CHOICE_FIT | (63 << 25)


@pytest.fixture(scope="module")
def kat():
    return golden("seated_kat.npz")


CACHE = os.path.join(tempfile.gettempdir(), "swarm_scene_cache")
_scenes = {}


def seated_scene(name):
    if name not in _scenes:
        _scenes[name] = sh.Scene.seated(name, cache_dir=CACHE)
    return _scenes[name]


@pytest.fixture(scope="module")
def scene_s():
    return seated_scene("S")


def ref_fnv(kat, venue):
    w = kat[f"seated__{venue}__nav__fnv"].astype(np.uint64)
    return [int(w[2 * k]) | (int(w[2 * k + 1]) << 32) for k in range(7)]


@pytest.mark.parametrize("venue", ["S", "M"])
def test_seated_nav_table_matches_reference(kat, venue):
    scene = seated_scene(venue)
    i = scene.info
    t = scene.tables()
    pre = f"seated__{venue}__nav__"
    res, origin, cell, inv = kat[pre + "grid"]
    assert (i.nav_res_x, i.nav_res_y) == (int(res), int(res)) and np.float32(i.nav_origin_x) == np.float32(origin)
    assert np.float32(i.nav_cell) == np.float32(cell) and np.float32(i.nav_inv_cell) == np.float32(inv)
    h = ref_fnv(kat, venue)
    K = i.K
    info = t["nav_info"]
    fit = (info >> 31) & 1
    assert op.fnv64(np.where(fit == 1, 0, 1).astype(np.uint8)) == h[0]   # the classes (FIT 0, NOFIT 1)
    has = (info >> 30) & 1
    cells = np.nonzero(has)[0]
    n_cells, n_entries = kat[pre + "sizes"]
    assert info.size == n_cells and cells.size == n_entries
    entry = np.zeros(info.size, np.uint32)
    entry[cells] = np.arange(1, cells.size + 1, dtype=np.uint32)
    assert op.fnv64((fit.astype(np.uint32) << np.uint32(31)) | entry) == h[4]
    slots = np.full((cells.size, 24), EMPTY_SLOT, np.uint32)
    slots[:, :K] = t["nav_slot"].reshape(-1, K)[cells] | np.uint32(CHOICE_FIT)
    assert op.fnv64(slots) == h[5]
    q = np.full((cells.size, 24), 0xFFFF, np.uint32)
    q[:, :K] = t["nav_dist"].reshape(-1, K)[cells]
    assert op.fnv64((q[:, 0::2] | (q[:, 1::2] << 16)).astype(np.uint32)) == h[6]
    sp, dq = kat[pre + "scalars"]
    assert np.float32(i.seconds_per_unit) == sp and np.float32(i.dist_quantum) == dq
    assert np.array_equal(t["inv_cap_s"], kat[pre + "inv_cap_s"])
    # the sampled cells, entry by entry
    sc = kat[pre + "sample_cell"]
    si = kat[pre + "sample_index"]
    assert np.array_equal((si & CHOICE_FIT) != 0, fit[sc] == 1) and np.array_equal((si & 0x7FFFFFFF) != 0, has[sc] == 1)
    live = (si & 0x7FFFFFFF) != 0
    ss = kat[pre + "sample_slots"][live]
    assert np.array_equal(t["nav_slot"].reshape(-1, K)[sc[live]] | np.uint32(CHOICE_FIT), ss[:, :K])
    assert (ss[:, K:] == EMPTY_SLOT).all()


def seated_oracle(kat_venue, n, nav=None, venue="S"):
    vf, vi = kat_venue[f"venue__{venue}__venue_f"], kat_venue[f"venue__{venue}__venue_i"]
    return op.SeatedOracle(op.venue_layout_arrays(kat_venue, venue), (float(vf[0]), float(vf[1]), int(vi[0])),
                           (float(vf[2]), float(vf[3]), float(vf[0] + vf[1])), n, nav=nav)


def assert_state(o, kat, p, tag):
    pos, vel = o.state3()
    rp, rv = kat[p + "pos"], kat[p + "vel"]
    assert np.array_equal(pos.view(np.uint32), rp.view(np.uint32)), tag + " positions"
    assert np.array_equal(vel.view(np.uint32), rv.view(np.uint32)), tag + " velocities"
    assert np.array_equal(o.contacts(), kat[p + "contacts"]), tag + " contacts"


def test_oracle_seated_drive_matches_reference(kat):
    o = seated_oracle(golden("venue_kat.npz"), 4096)
    o.set_state3(kat["seated__S__init__pos"], kat["seated__S__init__vel"])
    d = kat["seated__S__drive__dir"]
    o.set_drive(d[:, 0], d[:, 1])
    done = 0
    for s in (1, 2, 3, 5, 10, 20, 40):
        o.step(s - done)
        done = s
        assert_state(o, kat, f"seated__S__drive__step{s}__", f"drive step {s}")
    c = kat["seated__S__drive__step40__contacts"]
    assert ((c & 16) != 0).sum() > 1000 and ((c & 128) != 0).sum() > 10 and ((c & 4) != 0).sum() > 0   # walls, step-ups, fence


def test_oracle_seated_default_matches_reference(kat, scene_s):
    o = seated_oracle(golden("venue_kat.npz"), 4096, nav=(scene_s.info, scene_s.tables()))
    o.set_state3(kat["seated__S__init__pos"], kat["seated__S__init__vel"])
    done = 0
    for s in (1, 2, 3, 5, 10, 25, 50, 100, 200):
        o.step(s - done)
        done = s
        p = f"seated__S__default__step{s}__"
        assert_state(o, kat, p, f"default step {s}")
        assert np.array_equal(o.assign(), kat[p + "assign"]), f"default step {s} exits"
        assert np.array_equal(o.choice_counters(), kat[p + "counters"]), f"default step {s} counters"
        if s == 1:
            key, sid = o.grid()
            assert np.array_equal(key, kat["seated__S__default__grid_key"]) and np.array_equal(sid, kat["seated__S__default__grid_id"])
    assert (o.state3()[0][:, 3] == 0).sum() - 368 > 0   # agents left through the exits (368 slots are parked from the start)


def test_oracle_seated_m_matches_reference(kat):
    """Venue M (32 exits: the candidate table's exit-type rule), its 100 % seat spawn in 16,384 slots: 10 default steps."""
    scene = seated_scene("M")
    o = seated_oracle(golden("venue_kat.npz"), 16384, nav=(scene.info, scene.tables()), venue="M")
    o.set_state3(kat["seated__M__init__pos"], kat["seated__M__init__vel"])
    o.step(10)
    p = "seated__M__default__step10__"
    assert_state(o, kat, p, "venue M step 10")
    assert np.array_equal(o.assign(), kat[p + "assign"]) and np.array_equal(o.choice_counters(), kat[p + "counters"])
