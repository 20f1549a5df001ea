"""A floor's safety-metric grid (swarm_scene_metric_grid), on the CPU: 1 m cells from whole metres below the floor's
lower corner, and each cell's walkable area from 40 x 40 sample points at the centres of a regular subgrid
(src/metrics.cpp:143-164 samples the terrace's SDF the same way; here the test is the floor's wall rectangles).

The walkable areas are checked against a direct evaluation of every sample point (inside the floor, outside every
wall, bounds inclusive), which pins the scene builder's per-wall index ranges; the grid's offsets against the cell
a point lands in.
"""
import numpy as np

import swarm_hip as sh

MPU = 31.25
SUB = 40


def _scene(domain_m, walls_m):
    u = lambda v: np.asarray(v, np.float64) / MPU
    return sh.Scene.rects(domain=tuple(np.float32(u(domain_m))), hash_cell=float(np.float32(1.0 / MPU)), sdf_spacing=float(np.float32(0.5 / MPU)),
                          sdf_clamp=float(np.float32(2.0 / MPU)), nav=(1, 1, -4.0, -4.0, 8.0, 0.125), nav_body=0.0,
                          walls=np.float32(u(walls_m)).reshape(-1, 4), exits=np.zeros((0, 4)))


def _direct(domain_m, walls_m, rx, ry):
    """Every sample point, tested directly: inside the floor (in model units, as float) and in no wall."""
    x0, y0 = np.floor(domain_m[0]), np.floor(domain_m[1])
    xs = x0 + (np.arange(rx * SUB) + 0.5) / SUB
    ys = y0 + (np.arange(ry * SUB) + 0.5) / SUB
    X, Y = np.meshgrid(xs, ys)
    xu, yu = (X / MPU).astype(np.float32), (Y / MPU).astype(np.float32)
    d = np.float32(np.asarray(domain_m, np.float64) / MPU)
    free = (xu >= d[0]) & (xu <= d[2]) & (yu >= d[1]) & (yu <= d[3])
    for w in np.asarray(walls_m, np.float64).reshape(-1, 4):
        f = lambda v: np.float64(np.float32(v / MPU)) * MPU
        free &= ~((X >= f(w[0])) & (X <= f(w[2]))
                  & (Y >= f(w[1])) & (Y <= f(w[3])))
    counts = free.reshape(ry, SUB, rx, SUB).sum(axis=(1, 3))
    return (counts * (1.0 / SUB) ** 2).astype(np.float32)


def test_walkable_area_matches_direct_sampling():
    domain = (-3.3, -1.7, 9.45, 6.2)                     # metres: not on whole metres, negative corner
    walls = [(-3.3, -1.7, 9.45, -1.2),                   # a wall along the bottom edge
             (2.0, 0.0, 2.3, 6.2),                       # a thin partition
             (5.51, 2.49, 7.013, 3.77),                  # a block off the sample lattice
             (8.0, 4.0, 20.0, 20.0)]                     # a wall reaching past the floor
    s = _scene(domain, walls)
    rx, ry, cpu, ox, oy, w = s.metric_grid()
    assert (rx, ry) == (int(np.ceil(9.45 - np.floor(-3.3))), int(np.ceil(6.2 - np.floor(-1.7))))
    assert cpu == np.float32(MPU) and ox == np.float32(4 - 0.5) and oy == np.float32(2 - 0.5)
    want = _direct(domain, walls, rx, ry)
    assert np.array_equal(w, want), np.argwhere(w != want)[:5]
    assert ((w > 0) & (w < 1)).any() and (w == 0).any() and w.max() == 1.0


def test_grid_offsets_put_points_in_their_metre_cell():
    domain = (-3.3, -1.7, 9.45, 6.2)
    rx, ry, cpu, ox, oy, _ = _scene(domain, []).metric_grid()
    rng = np.random.default_rng(3)
    pm = np.stack([rng.uniform(-3.2, 9.4, 500), rng.uniform(-1.6, 6.1, 500)], 1)
    pu = (pm / MPU).astype(np.float32)
    # the cloud-in-cell deposit's floor(g) is the cell whose centre lies below-left: g = xy * cpu + off, cell ix
    # covers [ix - 0.5, ix + 0.5) of g, so the metre cell of a point is floor(g + 0.5)
    g = pu * np.float32(cpu) + np.array([ox, oy], np.float32)
    cell = np.floor(g + 0.5).astype(int)
    want = np.floor(pm - np.floor(np.array(domain[:2]))).astype(int)
    near = np.abs(pm - np.round(pm)) < 1e-4   # points within float rounding of a cell edge
    ok = (cell == want) | near
    assert ok.all()
    assert (cell >= 0).all() and (cell[:, 0] < rx).all() and (cell[:, 1] < ry).all()


def test_benchmark_floors_grids():
    for lay in (sh.LAYOUT_ROOM, sh.LAYOUT_BOWL, sh.LAYOUT_CORRIDOR, sh.LAYOUT_PILLAR, sh.LAYOUT_BOWL_EIGHTH):
        s = sh.Scene.layout(lay)
        rx, ry, cpu, ox, oy, w = s.metric_grid()
        i = s.info
        # the grid covers the floor: its corners land in the first and last cells
        for x, y in ((i.xmin, i.ymin), (i.xmax, i.ymax)):
            cx, cy = np.floor(np.float32(x) * np.float32(cpu) + np.float32(ox) + 0.5), np.floor(np.float32(y) * np.float32(cpu) + np.float32(oy) + 0.5)
            assert 0 <= cx <= rx and 0 <= cy <= ry
        area = (i.xmax - i.xmin) * (i.ymax - i.ymin) * MPU * MPU
        assert 0.5 * area < w.sum() <= area + 1e-3 * area
        assert w.min() >= 0.0 and w.max() == 1.0
