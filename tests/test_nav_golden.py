"""The scene builder's navigation tables against the reference's own builder.

nav_kat.npz holds, for rectangle worlds (the two-door room of tools/choice_check.hpp, the L corridor of
tools/nav_check.hpp, a three-exit room with a body radius), the reference's nav::buildGrid classes and
nav::buildChoiceTable candidates (src/nav_field.cpp:184-342, src/exit_choice.cpp:143-317): per cell the
class, and per candidate its exit, anchor cell and 16-bit distance, ascending. The builder restated here
(Lazy Theta* per exit, the merge, the quantisation) must give the same table bit for bit.
"""
import numpy as np
import pytest

from common import golden, kat_rect_scene

WORLDS = ["room2", "lcorr", "room3"]
CHOICE_INDEX_FIT = 0x80000000
CHOICE_INDEX_MASK = 0x7FFFFFFF


@pytest.fixture(scope="module")
def kat():
    return golden("nav_kat.npz")


@pytest.mark.parametrize("world", WORLDS)
def test_candidate_table_matches_reference(kat, world):
    sc = kat_rect_scene(kat, world)
    t = sc.tables()
    info = sc.info
    E, K = info.E, info.K
    ref_index = kat[world + "__table__index"]
    ref_slots = kat[world + "__table__slots"]
    ref_dists = kat[world + "__table__dists"]
    assert info.nav_res_x * info.nav_res_y == ref_index.size
    # the class of every cell
    fit = (t["nav_info"] >> 31) & 1
    assert np.array_equal(fit, ((ref_index & CHOICE_INDEX_FIT) != 0).astype(np.uint32))
    assert np.array_equal(fit, (kat[world + "__cls"] == 0).astype(np.uint32))
    # the cells with candidates and their candidates
    has = (t["nav_info"] >> 30) & 1
    entry = ref_index & CHOICE_INDEX_MASK
    assert np.array_equal(has, (entry != 0).astype(np.uint32))
    cells = np.nonzero(entry)[0]
    mine_slots = t["nav_slot"].reshape(-1, K)[cells]
    ref_s = ref_slots[entry[cells] - 1][:, :K] & 0x7FFFFFFF   # without the class bit every reference slot carries
    assert np.array_equal(mine_slots, ref_s)
    words = ref_dists[entry[cells] - 1]
    ref_q = np.stack([(words[:, k >> 1] >> (16 * (k & 1))) & 0xFFFF for k in range(K)], axis=1)
    mine_q = t["nav_dist"].reshape(-1, K)[cells].astype(np.uint32)
    assert np.array_equal(mine_q, ref_q)
    sp, dq = kat[world + "__table__scalars"]
    assert np.float32(info.seconds_per_unit) == sp and np.float32(info.dist_quantum) == dq
    assert np.array_equal(t["inv_cap_s"][:E], kat[world + "__table__inv_cap_s"])


@pytest.mark.parametrize("world", WORLDS)
def test_exit_fields_anchor_and_reachability(kat, world):
    """Every per-exit field of the reference: a cell reachable there holds that exit among its candidates,
    with the same anchor."""
    sc = kat_rect_scene(kat, world)
    t = sc.tables()
    K = sc.info.K
    slots = t["nav_slot"].reshape(-1, K)
    ex = (slots >> 25) & 63
    an = slots & ((1 << 25) - 1)
    fit = kat[world + "__cls"] == 0
    for e in range(sc.info.E):
        d = kat[f"{world}__field{e}__dist"]
        a = kat[f"{world}__field{e}__anchor"]
        reach = np.isfinite(d) & fit
        hit = ex == e
        assert np.array_equal(hit.any(axis=1), reach)
        k = np.argmax(hit, axis=1)
        rows = np.nonzero(reach)[0]
        assert np.array_equal(an[rows, k[rows]], a[rows].astype(np.uint32))


def test_known_answer_corridor_distance(kat):
    """The L corridor's field against the geodesic: in the straight arm the distance to the exit is the exact
    shortest path (within a cell), an independent check of the search itself."""
    sc = kat_rect_scene(kat, "lcorr")
    i = sc.info
    t = sc.tables()
    q = t["nav_dist"].reshape(-1, sc.info.K)[:, 0].astype(np.float64) * i.dist_quantum
    res = i.nav_res_x
    # a cell in the vertical arm, x 19 m, y 10 m: straight up to the exit region's lower edge (15.025 m)
    ix, iy = int((19.0 - i.nav_origin_x) / i.nav_cell), int((10.0 - i.nav_origin_y) / i.nav_cell)
    cy = i.nav_origin_y + (iy + 0.5) * i.nav_cell
    exact = 15.025 - cy
    assert abs(q[ix + iy * res] - exact) < i.nav_cell
