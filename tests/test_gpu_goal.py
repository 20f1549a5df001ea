"""The GPU's navigation lookups and e_goal against the reference's own vectors and the oracle.

At the probe points of nav_kat.npz (written by the reference's nav::choiceCell / nav::choiceDirection on its own
tables, src/exit_choice.cpp:512-566) the kernels' lookups, run through swarm_nav_probe (the reference's probe dispatch
of physics.comp's nav section, tools/test_gpu.cpp:4155-4191), must give the reference's cells and directions bit for
bit, and e_goal with the wall rule (choiceGoal, src/exit_choice.cpp:568-578) the oracle's, whose rule is pinned in
test_goal_golden.py. The benchmark floors' near-wall points exercise the GPU's wall index.
"""
import numpy as np
import pytest

import swarm_hip as sh
from common import golden, kat_rect_scene, oracle_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("world", ["room2", "lcorr", "room3"])
def test_gpu_nav_lookups_match_reference(world):
    kat = golden("nav_kat.npz")
    sc = kat_rect_scene(kat, world)
    sim = sh.Sim(sc, 64)
    o = oracle_for(sc, 16)
    x, y = kat[world + "__probe__x"], kat[world + "__probe__y"]
    rx, ry = kat[world + "__probe__dir_x"], kat[world + "__probe__dir_y"]
    for r in (0.2, 0.25):
        for e in range(sc.info.E):
            cell, d, g = sim.nav_probe(x, y, r, e)
            assert np.array_equal(cell, kat[world + "__probe__cell"])
            assert np.array_equal(d[:, 0].view(np.uint32), rx[:, e].view(np.uint32))
            assert np.array_equal(d[:, 1].view(np.uint32), ry[:, e].view(np.uint32))
            gx, gy, _ = o.goal(x, y, r, e)
            assert np.array_equal(g[:, 0].view(np.uint32), gx.view(np.uint32))
            assert np.array_equal(g[:, 1].view(np.uint32), gy.view(np.uint32))
    sim.close()


@pytest.mark.parametrize("layout", [sh.LAYOUT_ROOM, sh.LAYOUT_BOWL_EIGHTH, sh.LAYOUT_BOWL])
def test_gpu_goal_wall_rule_matches_oracle(layout):
    from test_goal_golden import near_wall_points

    sc = sh.Scene.layout(layout)
    t = sc.tables()
    walls = t["walls"].reshape(-1, 4)[: sc.info.n_walls]
    rng = np.random.default_rng(31)
    reach = np.float32(1.25 * 0.25 / sh.M_PER_UNIT)
    x, y = near_wall_points(walls, rng, 20000, reach)
    r = (rng.uniform(0.20, 0.25, x.size) / sh.M_PER_UNIT).astype(np.float32)
    sim = sh.Sim(sc, 64)
    o = oracle_for(sc, 16)
    applied = 0
    for e in range(sc.info.E):
        _, _, g = sim.nav_probe(x, y, r, e)
        gx, gy, ap = o.goal(x, y, r, e)
        assert np.array_equal(g[:, 0].view(np.uint32), gx.view(np.uint32))
        assert np.array_equal(g[:, 1].view(np.uint32), gy.view(np.uint32))
        applied += int(ap.sum())
    assert applied > 50
    sim.close()
