"""The oracle against golden vectors written by the reference's own CpuCrowdSim (src/cpu_crowd_sim.cpp).

terrace_kat.npz: the reference's terrace grid, agents on the fenced pitch (flat floor, no walls in reach)
with forced drive directions (CpuCrowdSim::setDriveTable), uniform BODY_RADIUS and V_MAX, counting sort; the
t_b* cases are the same scenes sorted by the GPU's bitonic network (SortMode::GpuBitonic), whose unstable
within-cell order changes the order of every neighbour sum.
On this case the 2D restatement must reproduce the reference bit for bit: positions, velocities, speed,
flags, the sorted (cell, agent) pairs and the cell table after every step. The cases cover separation,
headway with mutual standoffs, the Jacobi projection (dense overlaps), inactive agents, agents beyond the
domain (parked), and a NaN position (reset by the hash).
"""
import numpy as np
import pytest

from common import agents_from_golden, assert_bits, golden, oracle_for, terrace_scene

CASES = [("t_sparse", 6), ("t_dense", 8), ("t_bsparse", 6), ("t_bdense", 8)]


@pytest.fixture(scope="module")
def kat():
    return golden("terrace_kat.npz")


@pytest.fixture(scope="module")
def scene():
    return terrace_scene()


@pytest.mark.parametrize("tag,steps", CASES)
def test_oracle_matches_reference_steps(kat, scene, tag, steps):
    a = agents_from_golden(kat, tag)
    n = a["x"].size
    o = oracle_for(scene, n, threads=2)
    o.set_sort_mode(tag.startswith("t_b"))
    o.set_state(a)
    o.set_drive(kat[tag + "__in__dx"], kat[tag + "__in__dy"])
    for k in range(1, steps + 1):
        o.step(1)
        p = f"{tag}__step{k}__"
        s = o.state()
        active = kat[p + "w"] != 0
        # the reference keeps z: every active agent rests at one height, the flat floor's
        z = kat[p + "z"][active]
        assert np.all(z == z[0])
        for f in ("x", "y", "vx", "vy", "speed", "w"):
            assert_bits(s[f], kat[p + f], f"{tag} step {k} {f}")
        key, idx = o.grid()
        assert np.array_equal(key, kat[p + "key"]), f"{tag} step {k} keys"
        assert np.array_equal(idx, kat[p + "id"]), f"{tag} step {k} permutation"
        st, cn = o.cell_table()
        assert np.array_equal(st, kat[p + "cell_start"]) and np.array_equal(cn, kat[p + "cell_count"]), f"{tag} step {k} cells"


def test_kat_exercises_every_branch(kat):
    """The golden cases are not trivial: contacts, parks, inactive agents and a NaN reset happen."""
    a = agents_from_golden(kat, "t_dense")
    assert np.isnan(a["x"]).sum() == 1 and (a["w"] == 0).sum() == 37
    s1 = {f: kat["t_dense__step1__" + f] for f in ("x", "w")}
    assert (s1["w"] == 0).sum() == 37 + 5                    # the 5 agents beyond the domain were parked
    assert not np.isnan(s1["x"]).any()                       # the NaN position was reset
    moved = np.abs(kat["t_dense__step1__x"] - a["x"])
    assert np.nanmax(moved[a["w"] != 0]) > 0


def test_bitonic_network_closed_forms():
    # the network on tiny inputs against the rules it must satisfy: sorted keys, a permutation, and on distinct
    # keys the unique sorted order; with equal keys its deterministic within-key order is the reference's (the
    # t_b* cases above), and it is not the stable one
    kat = golden("terrace_kat.npz")
    for tag in ("t_bsparse", "t_bdense"):
        key, idx = kat[tag + "__step1__key"], kat[tag + "__step1__id"]
        assert np.all(np.diff(key.astype(np.int64)) >= 0)
        assert np.array_equal(np.sort(idx), np.arange(idx.size))
        stable = np.lexsort((idx, key))
        assert not np.array_equal(idx, idx[stable])   # some cell holds a non-ascending run


def test_bitonic_network_entry_matches_reference():
    # the network alone (oracle_bitonic_sort) on the hash keys of each recorded step, in agent order, gives the
    # reference's sorted (key, agent) sequence of that step
    kat = golden("terrace_kat.npz")
    import oracle_py as op

    for tag, steps in (("t_bsparse", 6), ("t_bdense", 8)):
        for k in range(1, steps + 1):
            key, idx = kat[f"{tag}__step{k}__key"], kat[f"{tag}__step{k}__id"]
            by_agent = np.empty_like(key)
            by_agent[idx] = key
            sk, ids = op.bitonic_sort(by_agent)
            assert np.array_equal(sk, key) and np.array_equal(ids, idx), f"{tag} step {k}"
