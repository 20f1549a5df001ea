"""The seated venues' default model on the plan's gangway inclines (defaultModelOptions, src/exit_choice.cpp:51-58:
the incline speed and the EMPIRICAL stair spacing, applied in the step at cpu_crowd_sim.cpp:727-733).

Checked here, on the CPU:
  - the oracle's seated::inclineFreeSpeed / inclineHeadwayTime equal the reference's bit for bit at 1,040 slopes
    (venue_kat.npz, written by the reference's own src/venue.cpp), with the reference's own known answers (tools/
    test_gpu.cpp:4319-4331: the level values at g = 0, the stair values from the reference slope on);
  - the oracle's floor slope, the one its step reads, equals the reference's seated::floorSlope at every probe point;
  - a venue plan carries its layout (swarm_scene_read_venue), equal to the reference's; other scenes carry none;
  - the oracle's step on venue S: a lone agent walking up or down a gangway settles at the stair speed of the rake
    (0.5: ascending 0.3251, descending 0.3670 model units, against the level 0.75), across the concourse at the level
    speed; an agent behind a standing one on an incline takes the stair time gap, not the level one, in the headway rule
    (the first step's velocity from rest is v_des (1 - exp(-DT / tau)) along e_goal).
"""
import numpy as np
import pytest

import oracle_py as op
import swarm_hip as sh
from common import golden, oracle_for

f32 = np.float32
RADIUS = f32(0.0064)
V_MAX = f32(0.75)


def kat(key):
    return golden("venue_kat.npz")["venue__" + key]


def test_incline_rules_match_the_reference():
    gs = kat("incline__g")
    v, t = op.incline(gs)
    assert np.array_equal(v.view(np.uint32), kat("incline__v_free").view(np.uint32))
    assert np.array_equal(t.view(np.uint32), kat("incline__t_headway").view(np.uint32))
    v_max, v_up, v_down, t_level, t_up, t_down, inv_ref = kat("incline__consts")
    c = op.constants()
    assert (c[26], c[27], c[28], c[29], c[30], c[31], c[32]) == (v_max, t_level, v_up, v_down, inv_ref, t_up, t_down)
    # the reference's own known answers
    v, t = op.incline(np.array([0.0, 0.6, -0.6, 0.5, -0.5], np.float32))
    assert v[0] == v_max and v[1] == v_up and v[2] == v_down
    assert t[0] == t_level and t[1] == t_up and t[2] == t_down
    # the rake of every venue (0.5) lies just under the reference slope: within 0.2 % of the stair values
    assert v_up < v[3] < v_up * 1.002 and v_down < v[4] < v_down * 1.002
    assert t_up * 0.998 < t[3] < t_up and t_down * 0.998 < t[4] < t_down


@pytest.mark.parametrize("venue", ["S", "M", "L"])
def test_oracle_slope_matches_the_reference(venue):
    lay = op.venue_layout_arrays(golden("venue_kat.npz"), venue)
    L = op.lib()
    h = L.oracle_venue_new(*[a.ctypes.data for a in lay])
    try:
        p, e, want = kat(venue + "__probe__p"), kat(venue + "__probe__e"), kat(venue + "__probe__slope")
        got = np.array([L.oracle_venue_slope(h, float(p[i, 0]), float(p[i, 1]), float(e[i, 0]), float(e[i, 1])) for i in range(len(p))],
                       np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert np.count_nonzero(want) > 0
    finally:
        L.oracle_venue_free(h)


@pytest.fixture(scope="module")
def venue_s():
    return sh.Scene.venue("S")


def test_venue_plan_carries_the_layout(venue_s):
    lay = venue_s.venue_layout()
    assert lay is not None
    want = op.venue_layout_arrays(golden("venue_kat.npz"), "S")
    got = op.venue_layout_arrays_of(lay)
    for k in (0, 2, 3, 4, 5, 6, 7):
        assert np.array_equal(np.asarray(got[k]).view(np.uint32), np.asarray(want[k]).view(np.uint32)), k
    assert np.array_equal(got[1][:4], want[1][:4])
    room = sh.Scene.layout(sh.LAYOUT_ROOM)
    assert room.venue_layout() is None


def _lone(scene, x, y, dx, dy, steps, extra=None):
    """Agent 0 at (x, y) driven along (dx, dy) from rest; extra: one more agent at rest with no drive."""
    n = 2
    a = dict(x=np.array([x, 1000.0], np.float32), y=np.array([y, 1000.0], np.float32), vx=np.zeros(n, np.float32),
             vy=np.zeros(n, np.float32), w=np.array([1.0, 0.0], np.float32), r=np.full(n, RADIUS, np.float32),
             v0=np.full(n, V_MAX, np.float32))
    if extra is not None:
        a["x"][1], a["y"][1], a["w"][1] = extra[0], extra[1], 1.0
    o = oracle_for(scene, n, threads=1)
    o.set_state(a)
    o.set_drive(np.array([dx, 0.0], np.float32), np.array([dy, 0.0], np.float32))
    o.step(steps)
    s = o.state()
    o.close()
    return s


GANGWAY_LAT = 0.3498   # venue S: the gangway spans lat 0.3306 .. 0.3690; rows from dep 0.4292 to 0.8644


def test_oracle_walks_the_gangway_at_the_stair_speed(venue_s):
    v_up, v_down = op.incline(np.array([0.5, -0.5], np.float32))[0]
    # tolerance 1e-3: the seat blocks' obstacle avoidance adds a small term where the gangway's distance samples are not
    # symmetric; the alternatives (the level speed, the other direction's stair speed) lie 10 % and more away
    up = _lone(venue_s, GANGWAY_LAT, 0.5, 0.0, 1.0, 40)
    down = _lone(venue_s, GANGWAY_LAT, 0.8, 0.0, -1.0, 40)
    assert abs(up["vy"][0] - v_up) < 1e-3 * v_up
    assert abs(down["vy"][0] + v_down) < 1e-3 * v_down
    # the same walk in another octant (x <-> y, mirrored)
    side = _lone(venue_s, -0.5, -GANGWAY_LAT, -1.0, 0.0, 40)
    assert abs(side["vx"][0] + v_up) < 1e-3 * v_up
    # the concourse is level: the level free speed
    level = _lone(venue_s, 0.1, 0.912, 1.0, 0.0, 40)
    assert abs(level["vx"][0] - V_MAX) < 1e-3 * V_MAX


def test_oracle_keeps_the_stair_time_gap(venue_s):
    tau, relax = sh.model_params(0.5)
    t_up, t_down = op.incline(np.array([0.5, -0.5], np.float32))[1]
    v_up, v_down = op.incline(np.array([0.5, -0.5], np.float32))[0]
    gap_centres = 0.02   # no separation force (it starts below R_SEP + the size margin, 0.016)
    for y0, dy, t_stair, v_free in ((0.6, 1.0, t_up, v_up), (0.7, -1.0, t_down, v_down)):
        lead = np.float32(y0 + dy * gap_centres)
        s = _lone(venue_s, GANGWAY_LAT, y0, 0.0, dy, 1, extra=(GANGWAY_LAT, lead))
        gap = abs(float(lead) - float(np.float32(y0))) - 2 * float(RADIUS)
        v_des = float(gap) / float(t_stair)
        assert v_des < v_free   # the headway rule binds; the level rule (gap / 0.3 s) would not
        want = v_des * (1.0 - float(relax))
        assert abs(abs(s["vy"][0]) - want) < 2e-5 * want, (abs(s["vy"][0]), want)
        assert np.sign(s["vy"][0]) == dy
