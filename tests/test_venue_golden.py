"""The analytic seated venues S, M and L: the layout, the seats and the heightfield queries against the reference.

venue_kat.npz holds what the reference's own code (src/venue.cpp, built by oracle/ref_build.sh and run by
oracle/ref_driver.cpp) makes for its venues: buildSeatedLayout's layout (every float bit for bit), the cell table of every
band, the seats (venue L: every 17th), and seated::sdfGrad / floorHeight / cellAt / wall (at three reaches) / floorSlope /
exitByLat / gangwayAt at seeded probe points (uniform at the step's wall plane, above and below the surface, exactly on it,
on cell and band boundaries in every octant, on the diagonals and on the axes).

Checked here, on the CPU:
  - the product's layout generator and seats (swarm_venue_layout_get, swarm_venue_seats) equal the reference's bit for bit;
  - the oracle's cell table and its answers to every query equal the reference's bit for bit;
  - the oracle's distance is the exact distance to the heightfield: brute force in float64 over the unfolded cells (each
    a box or an incline prism in all eight octants), for points in air within 1.25 m of the surface; its wall query is
    the exact 2D distance to the footprints of the cells above the plane;
  - the plan the step runs on venue S: every probe point inside a plan wall lies on a seat block, barrier, side or outer
    wall cell or on the fenced pitch, and every other point does not; every exit rectangle lies beyond the outer face and
    maps to the reference's exit of that opening.
"""
import numpy as np
import pytest

import oracle_py as op
import swarm_hip as sh
from common import golden

VENUES = ["S", "M", "L"]
LAYOUT_F = ["stand_front", "row0", "row_depth", "clearway", "conc_front", "conc_back", "outer_face", "landing_end", "avenue_half",
            "lat0", "gangway_lo", "gangway_hi", "exit_lo", "exit_hi", "corner_inset", "seat_inset", "face_margin", "wall_above",
            "wall_clear", "first_row", "row_rise", "seat_top", "barrier_top", "concourse_z", "wall_top", "rake", "seat_width",
            "domain_half", "corner_door_lo"]
GROUND, TREAD, SEAT, INCLINE, BARRIER, WALL, CONC, SIDE = range(8)
WALL_KINDS = (SEAT, BARRIER, WALL, SIDE)
FENCE_HALF_OUTER = np.float32(0.335432)


def g(name, key):
    return golden("venue_kat.npz")["venue__%s__%s" % (name, key)]


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("venue", VENUES)
def test_layout_matches_reference(venue):
    L = sh.venue_layout(venue)
    got = np.array([getattr(L, n) for n in LAYOUT_F], np.float32)
    ref = g(venue, "layout_f")
    bad = [LAYOUT_F[i] for i in range(len(LAYOUT_F)) if bits(got)[i] != bits(ref)[i]]
    assert not bad, f"venue {venue}: layout fields differ: {bad}"
    assert [L.rows, L.max_band_steps, L.gangways, L.doors, L.seat_total] == list(g(venue, "layout_i"))
    for field, key in (("gw_lo", "layout_gw"), ("gw_hi", "layout_gw_hi"), ("door_lo", "layout_door_lo"), ("door_hi", "layout_door_hi")):
        assert np.array_equal(bits(np.array(getattr(L, field)[:], np.float32)), bits(g(venue, key))), field
    assert list(L.gw_row) == list(g(venue, "layout_gw_row")) and list(L.gw_seated) == list(g(venue, "layout_gw_seated"))
    assert [L.door_width_m, L.corner_door_width_m, L.total_exit_width_m, L.required_exit_width_m] == list(g(venue, "layout_d"))
    assert np.array_equal(bits(np.array([L.domain_min, L.domain_span, L.exit_depth, L.exit_max_z], np.float32)), bits(g(venue, "venue_f")))
    assert [L.grid_res, L.metric_grid_res, L.exit_count, L.gangway_count] == list(g(venue, "venue_i"))


@pytest.mark.parametrize("venue", VENUES)
def test_seats_match_reference(venue):
    x, y, z, where = sh.venue_seats(venue)
    assert len(x) == int(g(venue, "seat_count")[0]) == sh.venue_layout(venue).seat_total
    sf, si = g(venue, "seats_f"), g(venue, "seats_i")
    idx = si[:, 0]
    assert np.array_equal(bits(x[idx]), bits(sf[:, 0])) and np.array_equal(bits(y[idx]), bits(sf[:, 1]))
    assert np.array_equal(bits(z[idx]), bits(sf[:, 2]))
    assert np.array_equal(where[idx], si[:, 1:])


def test_venue_refuses_unknown_name():
    with pytest.raises(RuntimeError):
        sh.venue_layout("XL")


@pytest.mark.parametrize("venue", VENUES)
def test_oracle_cells_match_reference(venue):
    ci, cf = op.venue_cells(op.venue_layout_arrays(golden("venue_kat.npz"), venue))
    assert np.array_equal(ci, g(venue, "cells_i"))
    assert np.array_equal(bits(cf), bits(g(venue, "cells_f")))


def reference_answers(venue):
    return dict(sdf_grad=g(venue, "probe__sdf_grad"), floor=g(venue, "probe__floor"), cell=g(venue, "probe__cell"),
                wall=np.stack([g(venue, "probe__wall_feeler"), g(venue, "probe__wall_plane"), g(venue, "probe__wall_wide")], 1),
                slope=g(venue, "probe__slope"), exit=g(venue, "probe__exit"), gangway=g(venue, "probe__gangway"))


def assert_same_answers(got, ref, what):
    for k in ("sdf_grad", "floor", "wall", "slope"):
        a, b = bits(got[k]).reshape(len(got[k]), -1), bits(ref[k]).reshape(len(ref[k]), -1)
        bad = np.nonzero((a != b).any(axis=1))[0]
        assert bad.size == 0, f"{what}: {k} differs at {bad.size} points, first {bad[:5]}"
    for k in ("cell", "exit", "gangway"):
        a, b = np.asarray(got[k]).reshape(len(got[k]), -1), np.asarray(ref[k]).reshape(len(ref[k]), -1)
        bad = np.nonzero((a != b).any(axis=1))[0]
        assert bad.size == 0, f"{what}: {k} differs at {bad.size} points, first {bad[:5]}"


@pytest.mark.parametrize("venue", VENUES)
def test_oracle_queries_match_reference(venue):
    o = op.venue_eval(op.venue_layout_arrays(golden("venue_kat.npz"), venue), g(venue, "probe__p"), g(venue, "probe__e"))
    assert_same_answers(o, reference_answers(venue), f"venue {venue} oracle")


# --- brute force over the unfolded cells ---

def primitives(venue, far=1.0e3):
    """Every cell unfolded to plan rectangles in the eight octants: (x0, x1, y0, y1, top0, slope, axis, kind); an incline's
    top rises by slope per unit of |coordinate| along axis from top0 at |coordinate| = row0. A cell that reaches the
    diagonal covers, with its mirror, [l0, hi] x [lo, hi] in each wing."""
    L = sh.venue_layout(venue)
    ci, cf = g(venue, "cells_i"), g(venue, "cells_f").astype(np.float64)
    out = []
    for (j, c, kind, row), (l0, l1, lo, hi) in zip(ci, cf):
        hi = min(hi, far)
        b = min(l1, hi)
        if b <= l0:
            continue
        if kind == INCLINE:
            top0, slope = 0.0, float(L.rake)
        else:   # the tops in float32, as the heightfield computes them
            f = np.float32
            tread = f(f(L.first_row) + f(f(row) * f(L.row_rise)))
            top = {GROUND: f(0), TREAD: tread, SEAT: f(tread + f(L.seat_top)), BARRIER: f(L.barrier_top), WALL: f(L.wall_top),
                   CONC: f(L.concourse_z), SIDE: f(L.barrier_top) if row < 0 else f(tread + f(L.seat_top))}[int(kind)]
            top0, slope = float(top), 0.0
        for sx in (1, -1):
            for sy in (1, -1):
                xs = sorted((sx * l0, sx * b)); ys = sorted((sy * lo, sy * hi))
                out.append((xs[0], xs[1], ys[0], ys[1], top0, slope, 1, kind))   # y wing: depth along y
                xs2 = sorted((sx * lo, sx * hi)); ys2 = sorted((sy * l0, sy * b))
                out.append((xs2[0], xs2[1], ys2[0], ys2[1], top0, slope, 0, kind))
    return np.array(out, np.float64), float(L.row0)


def brute_distance(prims, row0, p):
    """Distance from points p in air to the union of the primitives' solids, float64."""
    x0, x1, y0, y1, top0, slope, axis = (prims[:, k][None, :] for k in range(7))
    px, py, pz = (p[:, k][:, None] for k in range(3))
    # lateral axis clamp, then the nearest point of {u in [u0, u1], z <= s (|u| - row0)} in the rise axis' plane
    lat_p = np.where(axis == 1, px, py)
    lat0, lat1 = np.where(axis == 1, x0, y0), np.where(axis == 1, x1, y1)
    u_p = np.where(axis == 1, py, px)
    u0, u1 = np.where(axis == 1, y0, x0), np.where(axis == 1, y1, x1)
    dl = lat_p - np.clip(lat_p, lat0, lat1)
    # the rise runs along |u|: work in |u| (a rectangle never straddles u = 0 for an incline)
    sgn = np.where(u0 >= 0, 1.0, -1.0)
    a, b = np.minimum(sgn * u0, sgn * u1), np.maximum(sgn * u0, sgn * u1)
    u = sgn * u_p
    best = np.full(u.shape, np.inf)
    flat = slope == 0
    # boxes: clamp
    du = u - np.clip(u, a, b)
    dz = np.maximum(pz - top0, 0.0)
    best = np.where(flat, dl * dl + du * du + dz * dz, best)
    # incline prisms: the top segment or the two vertical edges
    t0, t1 = slope * (a - row0), slope * (b - row0)
    uc = np.clip(u, a, b)
    below = pz <= slope * (uc - row0)
    L2 = (b - a) ** 2 + (t1 - t0) ** 2
    tt = np.clip(((u - a) * (b - a) + (pz - t0) * (t1 - t0)) / np.where(L2 > 0, L2, 1), 0, 1)
    seg = (u - (a + tt * (b - a))) ** 2 + (pz - (t0 + tt * (t1 - t0))) ** 2
    e0 = (u - a) ** 2 + np.maximum(pz - t0, 0) ** 2
    e1 = (u - b) ** 2 + np.maximum(pz - t1, 0) ** 2
    inc = dl * dl + np.where(below, du * du, np.minimum(np.minimum(seg, e0), e1))
    best = np.where(flat, best, inc)
    return np.sqrt(best.min(axis=1))


@pytest.mark.parametrize("venue", ["S", "M"])
def test_oracle_distance_is_exact(venue):
    kat = golden("venue_kat.npz")
    lay = op.venue_layout_arrays(kat, venue)
    rng = np.random.default_rng(11)
    L = sh.venue_layout(venue)
    H = L.landing_end + 0.02
    xy = rng.uniform(-H, H, (3000, 2)).astype(np.float32)
    fl = op.venue_eval(lay, np.c_[xy, np.zeros(len(xy), np.float32)], np.zeros((len(xy), 2), np.float32))["floor"]
    z = (fl + rng.uniform(0.0005, 0.03, len(xy))).astype(np.float32)
    p = np.c_[xy, z].astype(np.float32)
    o = op.venue_eval(lay, p, np.zeros((len(p), 2), np.float32))
    d = o["sdf_grad"][:, 3].astype(np.float64)
    assert np.all(d > 0)
    prims, row0 = primitives(venue)
    ref = np.concatenate([brute_distance(prims, row0, p[i:i + 200].astype(np.float64)) for i in range(0, len(p), 200)])
    near = ref < 0.04
    assert near.sum() > 2000
    err = np.abs(d - ref)[near]
    assert err.max() < 2e-6, f"venue {venue}: sdf vs brute force max error {err.max():.3g}"
    # the gradient is the unit direction away from the nearest surface point
    gl = np.linalg.norm(o["sdf_grad"][:, :3].astype(np.float64), axis=1)
    assert np.all(np.abs(gl - 1) < 1e-5)


@pytest.mark.parametrize("venue", ["S", "M"])
def test_oracle_wall_is_exact(venue):
    kat = golden("venue_kat.npz")
    lay = op.venue_layout_arrays(kat, venue)
    p = g(venue, "probe__p").astype(np.float32)
    o = op.venue_eval(lay, p, g(venue, "probe__e"))
    prims, _ = primitives(venue)
    rng = 0.02
    for i in np.nonzero(g(venue, "probe__kind") == 0)[0][:400]:
        x, y, z = p[i].astype(np.float64)
        walls = prims[(prims[:, 5] == 0) & (prims[:, 4] > z)]
        dx = np.maximum(np.maximum(walls[:, 0] - x, x - walls[:, 1]), 0)
        dy = np.maximum(np.maximum(walls[:, 2] - y, y - walls[:, 3]), 0)
        dist = np.sqrt(dx * dx + dy * dy).min() if len(walls) else np.inf
        c = o["wall"][i, 2, 2]
        if dist < rng and dist > 0:
            assert abs(c - dist) < 1e-6, (i, c, dist)
        elif dist >= rng:
            assert c >= 1e9 or abs(c - dist) < 1e-6, (i, c, dist)


# --- the plan the step runs on ---

def inside(rects, x, y):
    r = rects.reshape(-1, 4)
    return ((x[:, None] >= r[None, :, 0]) & (x[:, None] <= r[None, :, 2]) & (y[:, None] >= r[None, :, 1]) & (y[:, None] <= r[None, :, 3]))


def test_plan_walls_and_exits_follow_the_cells():
    sc = sh.Scene.venue("S")
    t = sc.tables()
    L = sh.venue_layout("S")
    walls = t["walls"][:4 * sc.info.n_walls].reshape(-1, 4)
    p = g("S", "probe__p")
    cell = g("S", "probe__cell")
    x, y = p[:, 0], p[:, 1]
    lat = np.minimum(np.abs(x), np.abs(y))
    dep = np.maximum(np.abs(x), np.abs(y))
    # points strictly inside a cell (not on a face): the plan agrees with the cell's kind
    cf, ci = g("S", "cells_f"), g("S", "cells_i")
    interior = np.ones(len(p), bool)
    for (l0, l1, lo, hi) in cf:
        for v, e in ((lat, l0), (lat, l1), (dep, lo), (dep, hi)):
            interior &= np.abs(v - e) > 1e-6
    interior &= np.abs(lat - dep) > 1e-6
    in_wall = inside(walls, x, y).any(axis=1)
    pitch = dep < FENCE_HALF_OUTER
    want = np.isin(cell[:, 0], WALL_KINDS) | pitch
    bad = np.nonzero(interior & (in_wall != want))[0]
    assert bad.size == 0, f"plan walls disagree with the cells at {bad[:10]}"
    # exits: every rectangle beyond the outer face; the plan exit of a point past the face is the reference's exit there
    ex = t["exit_rects"][:4 * sc.info.E].reshape(-1, 4)
    ref_exit = sh.venue_exit_map("S")
    assert len(ref_exit) == sc.info.E == L.exit_count
    for r in ex:
        lo = np.minimum(np.abs(r[[0, 1]]), np.abs(r[[2, 3]]))
        assert max(abs(r[0]), abs(r[2])) > L.outer_face or max(abs(r[1]), abs(r[3])) > L.outer_face
    past = (dep > L.outer_face) & (dep < L.landing_end)
    hit = inside(ex, x, y)
    first = np.where(hit.any(axis=1), hit.argmax(axis=1), -1)
    e_ref = g("S", "probe__exit")
    for i in np.nonzero(past & (e_ref >= 0))[0]:
        assert first[i] >= 0 and ref_exit[first[i]] == e_ref[i], (i, p[i], first[i], e_ref[i])
