"""Shared helpers of the tests: the known-answer worlds and the oracle's setup over a built scene."""
import os

import numpy as np

import oracle_py as op
import swarm_hip as sh

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

BODY_RADIUS = np.float32(0.0064)
V_MAX = np.float32(0.75)


def golden(name):
    return np.load(os.path.join(GOLDEN, name))


def terrace_scene():
    """The reference's terrace neighbour grid (GRID_RES 40 over [-1, 1]^2, include/swarm/sim_params.hpp:57-60),
    no walls and no exits: the scene of its pitch known answers with forced drive directions."""
    return sh.Scene.rects(domain=(-1.0, -1.0, 1.0, 1.0), hash_cell=0.05, sdf_spacing=1.0, sdf_clamp=1.0,
                          nav=(1, 1, -1.0, -1.0, 2.0, 0.5), nav_body=0.0, walls=np.zeros((0, 4)), exits=np.zeros((0, 4)),
                          max_radius=float(BODY_RADIUS), max_v0=float(V_MAX))


def kat_rect_scene(g, prefix, body=None):
    """A rectangle world of nav_kat.npz (metres; the reference's square test grid)."""
    res, origin, cell, inv, b = [float(v) for v in g[prefix + "__grid"]]
    res = int(res)
    walls = g[prefix + "__walls"]
    exits = g[prefix + "__exits"]
    widths = g[prefix + "__widths"] if g[prefix + "__widths"].size else None
    span = res * cell
    return sh.Scene.rects(domain=(origin, origin, origin + span, origin + span), hash_cell=0.5, sdf_spacing=0.1, sdf_clamp=1.0,
                          nav=(res, res, origin, origin, np.float32(cell), np.float32(inv)), nav_body=b if body is None else body,
                          walls=walls, exits=exits, exit_width_m=widths, metres_per_unit=1.0, max_radius=0.25, max_v0=1.34)


def oracle_for(scene, n, proj_iters=32, threads=0, tau_s=0.5, obstacle_gain=20.0):
    """The oracle over the scene's tables (and, for a venue plan, the layout its incline rules read)."""
    tau, relax = sh.model_params(tau_s)
    t = scene.tables()
    lay = scene.venue_layout()
    venue = None if lay is None else op.venue_layout_arrays_of(lay)
    return op.Oracle(scene.info, t, n, relax, tau, obstacle_gain=obstacle_gain, proj_iters=proj_iters, threads=threads, venue=venue)


def agents_from_golden(g, tag):
    n = g[tag + "__in__x"].size
    return dict(x=g[tag + "__in__x"], y=g[tag + "__in__y"], vx=g[tag + "__in__vx"], vy=g[tag + "__in__vy"],
                w=g[tag + "__in__w"], r=np.full(n, BODY_RADIUS, np.float32), v0=np.full(n, V_MAX, np.float32))


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float32).view(np.uint32)
    b = np.ascontiguousarray(b, np.float32).view(np.uint32)
    return np.array_equal(a, b)


def assert_bits(a, b, what):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    bad = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
    assert bad.size == 0, f"{what}: {bad.size} values differ, first at {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"
