"""The safety-metrics oracle (oracle/metrics_oracle.cpp) against the reference's own CPU reference of the metric
passes (src/metrics.cpp: cicDeposit, speedMass, MetricVenue::cellInfo, MetricsReference::record), whose outputs
tests/golden/metrics_kat.npz holds (tests/golden/make_golden.py, oracle/ref_driver.cpp metricsCases): bit for bit,
on the reference's 64 x 64 grid of 1 m cells. Plus the invariants the fixed point guarantees (metrics.hpp:26-33).
The reference's exit slots are the terrace's avenue gaps; the oracle's are a floor's exit rectangles, so exits are
compared as their total (the evacuated count).
"""
import numpy as np

import oracle_py as op
from common import golden

RES = 64
CPU = np.float32(31.25)          # cells per model unit at 1 m cells (metrics.hpp:66)
OFF = np.float32(0.5 * RES - 0.5)   # the centred grid's offset (metrics.hpp:59)


def test_cic_weights_match_reference():
    g = golden("metrics_kat.npz")
    cells, w = op.metrics_cic(g["metrics__cic__xy"], RES, RES, CPU, OFF, OFF)
    assert np.array_equal(cells, g["metrics__cic__cell"])
    assert np.array_equal(w, g["metrics__cic__weight"])
    assert np.all(w.sum(axis=1) == 65536)   # one person per agent, exactly


def test_speed_mass_matches_reference():
    g = golden("metrics_kat.npz")
    got = op.metrics_speed_mass(g["metrics__speed__weight"], g["metrics__speed__mps"])
    assert np.array_equal(got, g["metrics__speed__mass"])


def test_thresholds_match_reference():
    g = golden("metrics_kat.npz")
    walk = g["metrics__info__walkable_m2"]
    thr = op.metrics_thresholds(walk, walk >= np.float32(0.10))
    assert np.array_equal(thr, g["metrics__info__threshold_mass"])


def test_recorded_steps_match_reference():
    g = golden("metrics_kat.npz")
    states = g["metrics__run__states"]   # steps + 1, n, (x, y, w, vx, vy)
    steps, n = states.shape[0] - 1, states.shape[1]
    o = op.MetricsOracle(n, RES, RES, CPU, OFF, OFF, g["metrics__info__threshold_mass"])
    mass_one = 65536
    for k in range(steps):
        prev = dict(x=states[k, :, 0], y=states[k, :, 1], w=states[k, :, 2])
        cur = dict(x=states[k + 1, :, 0], y=states[k + 1, :, 1], w=states[k + 1, :, 2], vx=states[k + 1, :, 3], vy=states[k + 1, :, 4])
        o.record(prev, cur)
        ref = g["metrics__run__counters"][k]   # step, active, evacuated, activeAccum, sum of exits
        assert list(o.counters[:4]) == list(ref[:4]), f"step {k + 1} counters"
        assert int(o.counters[4:].sum()) == int(ref[4])
        assert np.array_equal(o.cells[:, 0], g["metrics__run__last_mass"][k]), f"step {k + 1} mass"
        assert np.array_equal(o.cells[:, 1], g["metrics__run__last_speed_mass"][k]), f"step {k + 1} speed mass"
        assert int(o.cells[:, 0].astype(np.int64).sum()) == int(o.counters[1]) * mass_one   # mass is conserved exactly
    assert np.array_equal(o.evac_step, g["metrics__run__evac_step"])
    assert np.array_equal(o.cells, g["metrics__run__cells"])
    assert o.cells[:, 6].max() > 0   # the dense cluster passes the crush threshold


def test_exit_rectangles_and_grid_offsets():
    # this build's generalisations: an evacuated agent counts at the floor exit holding its last position (E when
    # none), and a grid off the origin only shifts g by its offset
    rects = np.array([[0.0, 0.0, 0.1, 0.1], [0.5, 0.5, 0.6, 0.6]], np.float32)
    thr = np.full((32 * 16, 4), 0xFFFFFFFF, np.uint32)
    o = op.MetricsOracle(3, 32, 16, np.float32(31.25), np.float32(-0.5), np.float32(-0.5), thr, rects)
    prev = dict(x=np.array([0.05, 0.55, 0.3], np.float32), y=np.array([0.05, 0.55, 0.3], np.float32), w=np.ones(3, np.float32))
    cur = dict(x=prev["x"], y=prev["y"], w=np.zeros(3, np.float32), vx=np.zeros(3, np.float32), vy=np.zeros(3, np.float32))
    o.record(prev, cur)
    assert list(o.counters[4:7]) == [1, 1, 1] and o.counters[2] == 3 and o.counters[1] == 0
    # an agent at a cell centre of the offset grid puts its whole mass in that cell
    cells, w = op.metrics_cic(np.array([[(3 + 0.5) / 31.25, (2 + 0.5) / 31.25]], np.float32), 32, 16, np.float32(31.25),
                              np.float32(-0.5), np.float32(-0.5))
    assert cells[0, 0] == 3 + 2 * 32 and w[0, 0] == 65536 and w[0, 1:].sum() == 0
