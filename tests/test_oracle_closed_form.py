"""Closed-form pins of the oracle where the reference's own vectors are silent.

The reference's known answers (terrace_kat.npz) use a uniform radius and speed and keep agents away from
walls. The configs' generalisations (per-agent radius r_i and desired speed v0_i, stair bands) and the
wall contacts are pinned here by what the model fixes in closed form, each chosen so that a plausible
mistake fails it:
- a free walker relaxes as v_n = vDes (1 - RELAX_STEP^n) with vDes = stair(x) v0_i (pins the relaxation,
  the per-agent free speed and the stair factor; a dropped factor or the global V_MAX fails);
- the along-goal cap and the speed clamp |v| <= v0_i from an over-fast start;
- headway: vDes = (d - (r_i + r_j)) HEADWAY_RATE behind a standing agent (pins D_ij = r_i + r_j);
- the projection separates an overlapping pair to exactly r_i + r_j and leaves a pair at r_i + r_j + eps alone;
- a walker driven into a flat wall ends at clearance r_i with no gain, and at the avoidance equilibrium
  d = OBSTACLE_RANGE - v0 / (W_OBS gain tau) with a large gain (pins the sign, gain and tau of a_obs);
- the baked SDF grid is the analytic distance to the wall rectangles.
Units are the reference's model units (1 L = 31.25 m, 1 T = 17.490671 s).
"""
import numpy as np
import pytest

import oracle_py as op
import swarm_hip as sh
from common import oracle_for

C = op.constants()
DT, HEADWAY_RATE, PROJ_MAX_SHIFT, W_OBS, OBSTACLE_RANGE = (float(C[0]), float(C[5]), float(C[7]), float(C[8]), float(C[10]))

WALL = (1.0, 0.6, 1.2, 1.4)   # a wall block whose left face is the plane x = 1.0
STAIR = (0.5, 0.8, 1.5, 1.2)


def world(walls=(), stairs=(), factors=()):
    return sh.Scene.rects(domain=(0.0, 0.0, 2.0, 2.0), hash_cell=0.016, sdf_spacing=0.0032, sdf_clamp=0.1,
                          nav=(200, 200, 0.0, 0.0, np.float32(0.01), np.float32(100.0)), nav_body=0.0,
                          walls=np.asarray(walls, np.float32).reshape(-1, 4), exits=np.zeros((0, 4), np.float32),
                          stairs=np.asarray(stairs, np.float32).reshape(-1, 4), stair_factor=np.asarray(factors, np.float32),
                          max_radius=0.008, max_v0=0.75)


def agents(xy, r, v0, v=None):
    n = len(xy)
    a = dict(x=np.array([p[0] for p in xy], np.float32), y=np.array([p[1] for p in xy], np.float32),
             r=np.asarray(r, np.float32), v0=np.asarray(v0, np.float32), w=np.ones(n, np.float32),
             vx=np.zeros(n, np.float32), vy=np.zeros(n, np.float32))
    if v is not None:
        a["vx"] = np.array([q[0] for q in v], np.float32)
        a["vy"] = np.array([q[1] for q in v], np.float32)
    return a


def run(scene, a, drive, steps, **kw):
    o = oracle_for(scene, a["x"].size, threads=1, **kw)
    o.set_state(a)
    o.set_drive(np.array([d[0] for d in drive], np.float32), np.array([d[1] for d in drive], np.float32))
    out = []
    for _ in range(steps):
        o.step(1)
        out.append(o.state())
    return out


def relax_step(tau_s=0.5):
    tau = tau_s / sh.S_PER_UNIT
    return np.exp(-DT / tau), tau


@pytest.fixture(scope="module")
def open_scene():
    return world(stairs=[STAIR], factors=[0.6])


@pytest.mark.parametrize("where,factor", [((0.3, 1.0), 1.0), ((0.6, 1.0), 0.6)])
def test_free_walk_relaxation_and_stair_factor(open_scene, where, factor):
    v0 = 0.6
    st = run(open_scene, agents([where], [0.007], [v0]), [(1.0, 0.0)], 30)
    rho, _ = relax_step()
    vdes = factor * v0
    x = float(where[0])
    for n, s in enumerate(st, start=1):
        assert s["vx"][0] == pytest.approx(vdes * (1.0 - rho ** n), rel=2e-6)
        assert s["vy"][0] == 0.0
        x += float(s["vx"][0]) * DT
        assert float(s["x"][0]) == pytest.approx(x, abs=2e-6)
    assert float(st[-1]["speed"][0]) == pytest.approx(abs(float(st[-1]["vx"][0])), rel=1e-6)


def test_along_goal_cap_and_speed_clamp(open_scene):
    v0 = 0.6
    st = run(open_scene, agents([(0.3, 1.0)], [0.007], [v0], v=[(1.5 * v0, 0.3)]), [(1.0, 0.0)], 1)
    rho, _ = relax_step()
    # along the goal the velocity is capped at vDes = v0; across it relaxes; then |v| is clamped to v0
    vx, vy = v0, 0.3 * rho
    scale = v0 / np.hypot(vx, vy)
    assert float(st[0]["vx"][0]) == pytest.approx(vx * scale, rel=2e-6)
    assert float(st[0]["vy"][0]) == pytest.approx(vy * scale, rel=2e-6)
    assert np.hypot(float(st[0]["vx"][0]), float(st[0]["vy"][0])) <= v0 * (1 + 1e-6)


def test_headway_uses_the_pair_contact_distance(open_scene):
    r1, r2, d = 0.006, 0.008, 0.02      # gap 0.006 beyond contact; outside the separation support r1 + r2 + margin
    st = run(open_scene, agents([(0.3, 1.0), (0.3 + d, 1.0)], [r1, r2], [0.6, 0.6]), [(1.0, 0.0), (0.0, 0.0)], 1)
    rho, _ = relax_step()
    vdes = min(0.6, (d - (r1 + r2)) * HEADWAY_RATE)
    assert vdes < 0.6
    assert float(st[0]["vx"][0]) == pytest.approx(vdes * (1.0 - rho), rel=1e-5)
    assert float(st[0]["vx"][1]) == 0.0 and float(st[0]["x"][1]) == np.float32(0.3 + d)


def test_projection_separates_to_sum_of_radii(open_scene):
    r1, r2 = 0.006, 0.008
    a = agents([(1.0, 1.0), (1.012, 1.0)], [r1, r2], [0.6, 0.6])
    st = run(open_scene, a, [(0.0, 0.0), (0.0, 0.0)], 1)[0]
    d = float(st["x"][1]) - float(st["x"][0])
    assert d == pytest.approx(r1 + r2, abs=2e-7)
    assert st["y"][0] == a["y"][0] and st["y"][1] == a["y"][1]
    m0, m1 = float(a["x"][0]) - float(st["x"][0]), float(st["x"][1]) - float(a["x"][1])
    assert m0 > 0 and m1 > 0 and m1 >= m0          # the larger radius has the larger per-iteration cap
    assert m0 <= 32 * PROJ_MAX_SHIFT * r1 + 1e-7 and m1 <= 32 * PROJ_MAX_SHIFT * r2 + 1e-7
    # just beyond contact: untouched
    b = agents([(1.0, 1.0), (1.0141, 1.0)], [r1, r2], [0.6, 0.6])
    st2 = run(open_scene, b, [(0.0, 0.0), (0.0, 0.0)], 1)[0]
    assert st2["x"][0] == b["x"][0] and st2["x"][1] == b["x"][1]


@pytest.fixture(scope="module")
def wall_scene():
    return world(walls=[WALL])


def test_wall_clearance_without_avoidance(wall_scene):
    r = 0.007
    st = run(wall_scene, agents([(0.9, 1.0)], [r], [0.6]), [(1.0, 0.0)], 120, obstacle_gain=0.0)
    gaps = np.array([WALL[0] - float(s["x"][0]) for s in st])
    assert gaps.min() >= r - 1e-6, "no penetration"
    assert gaps[-1] == pytest.approx(r, abs=2e-6), "held at clearance r"
    assert abs(float(st[-1]["vx"][0])) < 1e-3 * 0.6, "the into-wall velocity is removed"


def test_wall_avoidance_equilibrium(wall_scene):
    r, v0, gain = 0.007, 0.6, 200.0
    _, tau = relax_step()
    d_eq = OBSTACLE_RANGE - v0 / (W_OBS * gain * tau)
    assert d_eq > r
    st = run(wall_scene, agents([(0.9, 1.0)], [r], [v0]), [(1.0, 0.0)], 1500, obstacle_gain=gain)
    gaps = np.array([WALL[0] - float(s["x"][0]) for s in st])
    assert gaps.min() >= r - 1e-6
    assert gaps[-1] == pytest.approx(d_eq, abs=2e-5)


def test_sdf_grid_is_the_analytic_box_distance(wall_scene):
    i = wall_scene.info
    sdf = wall_scene.tables()["sdf"].reshape(i.sdf_ny, i.sdf_nx).astype(np.float64)
    xs = i.sdf_ox + np.arange(i.sdf_nx) * np.float64(i.sdf_h)
    ys = i.sdf_oy + np.arange(i.sdf_ny) * np.float64(i.sdf_h)
    X, Y = np.meshgrid(xs, ys)
    lx, ly, hx, hy = WALL
    dx = np.maximum(np.maximum(lx - X, X - hx), 0.0)
    dy = np.maximum(np.maximum(ly - Y, Y - hy), 0.0)
    outside = np.hypot(dx, dy)
    inside = np.minimum(np.minimum(X - lx, hx - X), np.minimum(Y - ly, hy - Y))
    exact = np.where((dx > 0) | (dy > 0), outside, -inside)
    exact = np.minimum(exact, 0.1)
    assert np.max(np.abs(sdf - exact)) < 2e-6
