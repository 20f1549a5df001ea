"""Spawners: the seat spawn against the reference's own seatedSpawnSeats; the open floors' uniform spawn by what it must
give.

spawn_kat.npz holds the reference's seatedSpawnSeats (src/crowd_sim.cpp:652-674; mt19937's raw output, partial
Fisher-Yates) on its venues S, M and L (3,728, 15,832 and 63,424 seats) for full, partial and tiny loads and several
seeds, written by oracle/ref_driver.cpp: swarm_spawn_seats must choose the same seats. The uniform spawn (rejection,
then the Gauss-Seidel relaxation of relaxSpawn, src/crowd_sim.cpp:544-597, with per-agent radii) must leave no pair
overlapping by more than its tolerance (checked by brute force), every agent on its floor clear of the walls by its
radius, radii and speeds in their ranges, and the same placement for the same seed.
"""
import numpy as np
import pytest

import swarm_hip as sh
from common import golden


@pytest.mark.parametrize("venue", ["S", "M", "L"])
def test_seat_spawn_matches_reference(venue):
    g = golden("spawn_kat.npz")
    seats = int(g[f"spawn__{venue}__seats"][0])
    cases = g[f"spawn__{venue}__cases"].reshape(-1, 3)
    ref = g[f"spawn__{venue}__chosen"]
    off = 0
    for n, load_milli, seed in cases:
        got = sh.spawn_seats(seats, int(n), load_milli / 1000.0, int(seed))
        assert np.array_equal(got, ref[off:off + n]), f"venue {venue}: n {n}, load {load_milli / 1000}, seed {seed}"
        off += int(n)
    assert off == ref.size


def test_seat_spawn_refuses_more_seats_than_agents():
    with pytest.raises(RuntimeError):
        sh.spawn_seats(1000, 100, 0.5, 1)


FLOORS = {sh.LAYOUT_ROOM: ((0.5, 0.5, 40.5, 30.5), []), sh.LAYOUT_CORRIDOR: ((0.5, 0.5, 60.5, 20.5), []),
          sh.LAYOUT_PILLAR: ((0.5, 0.5, 30.5, 30.5), [(13.5, 13.5, 17.5, 17.5)])}


@pytest.mark.parametrize("layout,n", [(sh.LAYOUT_ROOM, 4096), (sh.LAYOUT_PILLAR, 2500), (sh.LAYOUT_CORRIDOR, 3000)])
def test_uniform_spawn_is_overlap_free_and_on_the_floor(layout, n):
    a, sweeps, max_ov = sh.spawn_uniform(layout, n, n + 100, 1337)
    b, _, _ = sh.spawn_uniform(layout, n, n + 100, 1337)
    for k in a:
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    assert 0 < sweeps < 2000 and max_ov <= 1e-6
    m = sh.M_PER_UNIT
    x, y, r = a["x"][:n].astype(np.float64) * m, a["y"][:n].astype(np.float64) * m, a["r"][:n].astype(np.float64) * m
    assert np.all(a["w"][:n] == 1) and np.all(a["w"][n:] == 0)
    assert np.all((r >= 0.20 - 1e-6) & (r <= 0.25 + 1e-6))
    v = a["v0"][:n].astype(np.float64) * m / sh.S_PER_UNIT
    assert np.all((v >= 0.5 - 1e-6) & (v <= 2.0 + 1e-6)) and abs(np.mean(v) - 1.34) < 0.03
    (x0, y0, x1, y1), holes = FLOORS[layout]
    assert np.all(x - r >= x0 - 1e-5) and np.all(x + r <= x1 + 1e-5) and np.all(y - r >= y0 - 1e-5) and np.all(y + r <= y1 + 1e-5)
    for hx0, hy0, hx1, hy1 in holes:
        gap = np.hypot(np.maximum(np.maximum(hx0 - x, x - hx1), 0), np.maximum(np.maximum(hy0 - y, y - hy1), 0))
        assert np.all(gap >= r - 1e-5)
    # brute force over every pair within a cell grid's reach
    order = np.argsort(x)
    worst = 0.0
    for k, i in enumerate(order):
        j = order[k + 1:k + 400]
        j = j[x[j] - x[i] < 0.5]
        if j.size:
            worst = max(worst, float(np.max(r[i] + r[j] - np.hypot(x[i] - x[j], y[i] - y[j]))))
    assert worst <= 1e-6 * m, f"largest overlap {worst} m"
    # uniform over the floor: the four quadrants of the box hold similar counts (the pillar is central)
    qx, qy = x < 0.5 * (x0 + x1), y < 0.5 * (y0 + y1)
    counts = [np.sum(qx & qy), np.sum(qx & ~qy), np.sum(~qx & qy), np.sum(~qx & ~qy)]
    assert max(counts) - min(counts) < 0.1 * n
