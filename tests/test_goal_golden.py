"""e_goal's inputs and its wall rule, pinned without the GPU.

nav_kat.npz holds, at probe points of three rectangle worlds (a seeded uniform sample, points beside every wall face
in the NOFIT band and its 3 x 3 / 5 x 5 fallback rings, anchor centres and points inside and just beyond
GOAL_DEADZONE of them, exit-rectangle corners and edges and their float neighbours), the reference's own
nav::choiceCell, nav::choiceDirection for every exit (src/exit_choice.cpp:512-566) and RectGeometry::exitAt
(src/nav_field.cpp:155-159), written by oracle/ref_driver.cpp. The oracle must reproduce them bit for bit.

The wall rule (choiceGoal, src/exit_choice.cpp:568-578) queries the nearest wall within FEELER_REACH. On a rectangle
world that query is the distance to a union of boxes, which is pinned here against a float64 brute force, and the
rule's result against what it must be: unchanged bit for bit where no wall is within reach or e does not lead into
it; otherwise a unit vector (or 0) orthogonal to the wall direction, along the projection of e, never against e.
The GPU's wall index (per hash cell, the walls it can reach) is checked complete against the brute force.
"""
import numpy as np
import pytest

import swarm_hip as sh
from common import golden, kat_rect_scene, oracle_for

WORLDS = ["room2", "lcorr", "room3"]


@pytest.fixture(scope="module")
def kat():
    return golden("nav_kat.npz")


@pytest.mark.parametrize("world", WORLDS)
def test_choice_cell_direction_exit_at_match_reference(kat, world):
    sc = kat_rect_scene(kat, world)
    E = sc.info.E
    o = oracle_for(sc, 16)
    x, y = kat[world + "__probe__x"], kat[world + "__probe__y"]
    assert np.array_equal(o.choice_cell(x, y), kat[world + "__probe__cell"])
    assert np.array_equal(o.exit_at(x, y), kat[world + "__probe__exit_at"])
    rx, ry = kat[world + "__probe__dir_x"], kat[world + "__probe__dir_y"]
    for e in range(E):
        ex, ey, _ = o.goal(x, y, 0.2, e, wall_rule=False)
        assert np.array_equal(ex.view(np.uint32), rx[:, e].view(np.uint32)), f"{world} exit {e}: direction x"
        assert np.array_equal(ey.view(np.uint32), ry[:, e].view(np.uint32)), f"{world} exit {e}: direction y"
    # the probes reach every branch: NOFIT cells (the fallback rings), the dead zone, points on exit edges
    cls = kat[world + "__cls"]
    gx = np.floor((x - sc.info.nav_origin_x) * np.float32(sc.info.nav_inv_cell)).astype(int)
    gy = np.floor((y - sc.info.nav_origin_y) * np.float32(sc.info.nav_inv_cell)).astype(int)
    inside = (gx >= 0) & (gy >= 0) & (gx < sc.info.nav_res_x) & (gy < sc.info.nav_res_y)
    own = gx[inside] + gy[inside] * sc.info.nav_res_x
    assert np.count_nonzero(cls[own] != 0) > 20, "probes in NOFIT cells"
    assert np.count_nonzero(o.choice_cell(x, y)[inside] != own) > 20, "probes that take a fallback cell"
    zero = (rx == 0) & (ry == 0)
    assert zero.any() and (~zero).any()
    assert np.count_nonzero(kat[world + "__probe__exit_at"] >= 0) > 10


def brute_wall(walls, x, y, reach):
    """Nearest wall box within reach, float64: (index or -1, distance, direction away from it)."""
    x = x.astype(np.float64)[:, None]
    y = y.astype(np.float64)[:, None]
    lo_x, lo_y, hi_x, hi_y = (walls[:, k].astype(np.float64)[None, :] for k in range(4))
    qx, qy = np.clip(x, lo_x, hi_x), np.clip(y, lo_y, hi_y)
    d = np.hypot(x - qx, y - qy)
    k = np.argmin(d, axis=1)
    dk = d[np.arange(len(k)), k]
    ux = (x[:, 0] - qx[np.arange(len(k)), k]) / np.where(dk > 0, dk, 1)
    uy = (y[:, 0] - qy[np.arange(len(k)), k]) / np.where(dk > 0, dk, 1)
    second = np.partition(d, 1, axis=1)[:, 1] if d.shape[1] > 1 else np.full(len(k), np.inf)
    return np.where(dk < reach, k, -1), dk, ux, uy, second


def near_wall_points(walls, rng, n, reach):
    k = rng.integers(0, len(walls), n)
    w = walls[k]
    side = rng.integers(0, 4, n)
    a = rng.random(n).astype(np.float32)
    off = (rng.random(n) * 1.6 * reach).astype(np.float32)
    fx = w[:, 0] + a * (w[:, 2] - w[:, 0])
    fy = w[:, 1] + a * (w[:, 3] - w[:, 1])
    x = np.where(side == 0, w[:, 0] - off, np.where(side == 1, w[:, 2] + off, fx)).astype(np.float32)
    y = np.where(side == 2, w[:, 1] - off, np.where(side == 3, w[:, 3] + off, fy)).astype(np.float32)
    # half of them around the boxes' corners (door jambs, row ends: where a path wraps round a wall)
    c = rng.random(n) < 0.5
    ang = rng.random(n) * 2.0 * np.pi
    rad = rng.random(n) * 1.2 * reach
    cx = np.where(rng.random(n) < 0.5, w[:, 0], w[:, 2]) + rad * np.cos(ang)
    cy = np.where(rng.random(n) < 0.5, w[:, 1], w[:, 3]) + rad * np.sin(ang)
    return np.where(c, cx, x).astype(np.float32), np.where(c, cy, y).astype(np.float32)


@pytest.mark.parametrize("layout", [sh.LAYOUT_ROOM, sh.LAYOUT_PILLAR, sh.LAYOUT_BOWL_EIGHTH])
def test_wall_query_matches_brute_force(layout):
    sc = sh.Scene.layout(layout)
    t = sc.tables()
    walls = t["walls"].reshape(-1, 4)[: sc.info.n_walls]
    o = oracle_for(sc, 16)
    reach = np.float32(1.25 * 0.25 / sh.M_PER_UNIT)
    rng = np.random.default_rng(11)
    x, y = near_wall_points(walls, rng, 6000, reach)
    wx, wy, wd = o.wall(x, y, reach)
    k, d, ux, uy, second = brute_wall(walls, x, y, reach)
    clear = np.abs(d - reach) > 1e-6 * reach
    hit = (k >= 0) & clear & (d > 1e-7)
    miss = (k < 0) & clear
    assert hit.sum() > 1000 and miss.sum() > 500
    assert np.all(wd[miss] == np.float32(1e9))
    np.testing.assert_allclose(wd[hit], d[hit], rtol=2e-6, atol=1e-9)
    # the direction points away from the nearest box (where the nearest is not a tie)
    unique = hit & (second - d > 1e-5 * reach)
    np.testing.assert_allclose(wx[unique], ux[unique], atol=2e-5)
    np.testing.assert_allclose(wy[unique], uy[unique], atol=2e-5)


@pytest.mark.parametrize("layout", [sh.LAYOUT_ROOM, sh.LAYOUT_BOWL_EIGHTH, sh.LAYOUT_BOWL])
def test_wall_index_is_complete(layout):
    """The GPU's per-hash-cell wall lists give every point the brute force's nearest wall (float32, ascending
    candidates, the first of equal distances)."""
    sc = sh.Scene.layout(layout)
    i, t = sc.info, sc.tables()
    walls = t["walls"].reshape(-1, 4)[: i.n_walls]
    start, lst = t["wall_bin_start"], t["wall_bin_list"]
    reach = np.float32(np.float32(1.25) * np.float32(0.25 / sh.M_PER_UNIT))
    rng = np.random.default_rng(5)
    x, y = near_wall_points(walls, rng, 4000, reach)
    inside = (x >= i.xmin) & (x <= i.xmax) & (y >= i.ymin) & (y <= i.ymax)
    x, y = x[inside], y[inside]
    f32 = np.float32
    gx = np.clip(((x - f32(i.xmin)) / f32(i.hash_span_x) * f32(i.hash_nx)).astype(np.int64), 0, i.hash_nx - 1)
    gy = np.clip(((y - f32(i.ymin)) / f32(i.hash_span_y) * f32(i.hash_ny)).astype(np.int64), 0, i.hash_ny - 1)
    cell = gx + gy * i.hash_nx
    r2 = reach * reach

    def nearest(px, py, cand):
        best, bk = r2, -1
        for k in cand:
            lx, ly, hx, hy = walls[k]
            qx, qy = min(max(px, lx), hx), min(max(py, ly), hy)
            dx, dy = f32(px - qx), f32(py - qy)
            d2 = f32(f32(dx * dx) + f32(dy * dy))
            if d2 < best:
                best, bk = d2, k
        return bk

    every = np.arange(len(walls))
    found = 0
    for j in range(len(x)):
        c = cell[j]
        mine = nearest(x[j], y[j], lst[start[c]:start[c + 1]])
        # the brute force over the walls that can be within reach at all (a cheap prefilter in float64)
        near = every[(np.maximum.reduce([walls[:, 0] - x[j], x[j] - walls[:, 2], np.zeros(len(walls))]) < 2 * reach) &
                     (np.maximum.reduce([walls[:, 1] - y[j], y[j] - walls[:, 3], np.zeros(len(walls))]) < 2 * reach)]
        assert mine == nearest(x[j], y[j], near), f"point {j} ({x[j]}, {y[j]}) cell {c}"
        found += mine >= 0
    assert found > len(x) // 3


@pytest.mark.parametrize("layout", [sh.LAYOUT_ROOM, sh.LAYOUT_BOWL_EIGHTH])
def test_wall_rule_removes_the_into_wall_component(layout):
    sc = sh.Scene.layout(layout)
    t = sc.tables()
    walls = t["walls"].reshape(-1, 4)[: sc.info.n_walls]
    o = oracle_for(sc, 16)
    rng = np.random.default_rng(23)
    r = np.float32(0.2 / sh.M_PER_UNIT)
    reach = np.float32(1.25) * r
    x, y = near_wall_points(walls, rng, 6000, reach)
    # keep points on the floor (outside every wall box, inside the domain)
    i = sc.info
    inside_wall = np.zeros(len(x), bool)
    for w in walls:
        inside_wall |= (x >= w[0]) & (x <= w[2]) & (y >= w[1]) & (y <= w[3])
    keep = ~inside_wall & (x > i.xmin) & (x < i.xmax) & (y > i.ymin) & (y < i.ymax)
    x, y = x[keep], y[keep]
    applied_any = 0
    for e in range(sc.info.E):
        e0x, e0y, _ = o.goal(x, y, r, e, wall_rule=False)
        ex, ey, ap = o.goal(x, y, r, e)
        wx, wy, wd = o.wall(x, y, reach)
        into = e0x * wx + e0y * wy
        # applied exactly where a wall lies within reach and e leads into it
        assert np.array_equal(ap, (wd < reach) & (into < 0))
        # elsewhere e_goal is the table's direction bit for bit
        assert np.array_equal(ex[~ap].view(np.uint32), e0x[~ap].view(np.uint32))
        assert np.array_equal(ey[~ap].view(np.uint32), e0y[~ap].view(np.uint32))
        a = ap & ((ex != 0) | (ey != 0))
        applied_any += a.sum()
        n = np.hypot(ex[a].astype(np.float64), ey[a].astype(np.float64))
        np.testing.assert_allclose(n, 1.0, atol=1e-6)
        # no component into the wall, the tangential part of e kept (same side, parallel), never against e
        dot_w = ex[a] * wx[a].astype(np.float64) + ey[a] * wy[a].astype(np.float64)
        tx = e0x[a] - wx[a] * into[a].astype(np.float64)
        ty = e0y[a] - wy[a] * into[a].astype(np.float64)
        # FP32 rounding of the projection, amplified by the renormalisation of a short vector (1 / |t|)
        assert np.all(np.abs(dot_w) < 3e-7 / np.hypot(tx, ty) + 1e-7)
        cross = ex[a].astype(np.float64) * ty - ey[a].astype(np.float64) * tx
        assert np.all(np.abs(cross) < 3e-7 + 1e-7 * np.hypot(tx, ty))
        assert np.all(ex[a] * tx + ey[a] * ty > 0)
        assert np.all(ex[a] * e0x[a] + ey[a] * e0y[a] >= 0)
    assert applied_any > 20


def test_wall_rule_at_a_straight_wall_in_closed_form():
    """A 12 x 9 m room: at a point 0.1 m from the long south wall, every direction with a southward component becomes
    due east or due west (the sign of its x component), every other direction stays as it is."""
    walls = np.array([[-0.3, -0.3, 12.3, 0.0], [-0.3, 9.0, 12.3, 9.3], [-0.3, 0.0, 0.0, 9.0], [12.0, 0.0, 12.3, 9.0]], np.float32)
    exits = np.array([[12.0, 4.0, 12.3, 5.0]], np.float32)
    walls[3] = [12.0, 5.0, 12.3, 9.0]
    walls = np.vstack([walls, [[12.0, 0.0, 12.3, 4.0]]]).astype(np.float32)
    sc = sh.Scene.rects(domain=(-1.0, -1.0, 13.0, 10.0), hash_cell=0.5, sdf_spacing=0.1, sdf_clamp=1.0,
                        nav=(140, 110, -1.0, -1.0, np.float32(0.1), np.float32(10.0)), nav_body=0.2, walls=walls, exits=exits,
                        metres_per_unit=1.0, max_radius=0.25, max_v0=1.34)
    o = oracle_for(sc, 4)
    wx, wy, wd = o.wall(np.array([6.0], np.float32), np.array([0.1], np.float32), np.float32(0.3125))
    assert (wx[0], wy[0]) == (0.0, 1.0) and abs(wd[0] - 0.1) < 1e-7
    _, _, wd = o.wall(np.array([6.0], np.float32), np.array([0.4], np.float32), np.float32(0.3125))
    assert wd[0] == np.float32(1e9)
