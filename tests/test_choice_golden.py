"""The exit-choice pass against the reference's nav::choiceStep (src/exit_choice.cpp:683-791).

choice_kat.npz: test_gpu's nav.exit_choice known answers (tools/test_gpu.cpp:3968-4100) on the two-door
room of tools/choice_check.hpp, run through the reference's CPU pass: (a) newcomers choose the free door
over a queued one (and the mirror, and the nearer door without a queue), (b) the hysteresis margins, (c) a
closed loop of 190 agents over 40 evaluations (the switch cap, no oscillation). Assignments and every
counter must match exactly.
"""
import numpy as np
import pytest

from common import golden, kat_rect_scene, oracle_for

CS = 17


@pytest.fixture(scope="module")
def ck():
    return golden("choice_kat.npz")


@pytest.fixture(scope="module")
def room():
    return kat_rect_scene(golden("nav_kat.npz"), "room2")


def agents(ck, prefix):
    x = ck[prefix + "x"]
    n = x.size
    return dict(x=x, y=ck[prefix + "y"], w=ck[prefix + "w"], vx=np.zeros(n, np.float32), vy=np.zeros(n, np.float32),
                r=np.full(n, 0.2, np.float32), v0=np.full(n, 1.34, np.float32))


class OracleChoice:
    def __init__(self, scene, n):
        self.o = oracle_for(scene, n, threads=1)

    def __getattr__(self, k):
        return getattr(self.o, k)


def run_a(make, ck, variant):
    p = f"choice_a{variant}__"
    s0, s1 = agents(ck, p + "s0__"), agents(ck, p + "s1__")
    sim = make(s0["x"].size)
    sim.set_state(s0) if hasattr(sim, "set_state") else sim.upload(s0)
    sim.choice_step()
    out = [(sim.assign(), sim.choice_counters())]
    sim.restore(s1)
    sim.choice_step()
    out.append((sim.assign(), sim.choice_counters()))
    return out


def run_b(make, ck, variant):
    p = f"choice_b{variant}__"
    st = agents(ck, p)
    step, ql, qr = [int(v) for v in ck[p + "set"]]
    sim = make(2)
    sim.set_state(st) if hasattr(sim, "set_state") else sim.upload(st)
    queue = np.zeros(128, np.uint32)
    par = step & 1
    queue[par * 64 + 0], queue[par * 64 + 1] = ql, qr
    ahead = np.zeros(2 * 64 * 1024, np.uint32)
    ahead[par * 64 * 1024 + 0 * 1024] = ql
    ahead[par * 64 * 1024 + 1 * 1024] = qr
    sim.set_choice_state(step, assign=np.zeros(2, np.uint32), queue=queue, ahead=ahead)
    sim.choice_step()
    return sim.assign(), sim.choice_counters()


def run_c(make, ck):
    st = agents(ck, "choice_c__")
    n = st["x"].size
    sim = make(n)
    sim.set_state(st) if hasattr(sim, "set_state") else sim.upload(st)
    queue = np.zeros(128, np.uint32)
    queue[0] = n
    sim.set_choice_state(0, assign=np.zeros(n, np.uint32), queue=queue)
    hist = []
    for s in range(40 * CS + 1):
        sim.choice_step()
        if s % CS == 0:
            hist.append(sim.assign())
    return np.stack(hist), sim.assign(), sim.choice_counters()


def check_all(make, ck):
    for v in range(3):
        out = run_a(make, ck, v)
        for k, (a, c) in enumerate(out):
            assert np.array_equal(a, ck[f"choice_a{v}__s{k}__assign"]), f"(a{v}) step {k} assignments"
            assert np.array_equal(c, ck[f"choice_a{v}__s{k}__counters"]), f"(a{v}) step {k} counters"
    for v in range(2):
        a, c = run_b(make, ck, v)
        assert np.array_equal(a, ck[f"choice_b{v}__assign"]), f"(b{v}) assignments"
        assert np.array_equal(c, ck[f"choice_b{v}__counters"]), f"(b{v}) counters"
    hist, a, c = run_c(make, ck)
    assert np.array_equal(hist, ck["choice_c__assign_at_evals"]), "(c) assignments at the evaluations"
    assert np.array_equal(c, ck["choice_c__final__counters"]), "(c) counters"


def test_oracle_choice_matches_reference(ck, room):
    check_all(lambda n: OracleChoice(room, n), ck)


def test_known_answers_hold(ck):
    """The fixtures carry the reference's known answers themselves (test_gpu nav.exit_choice's checks)."""
    a = ck["choice_a0__s1__assign"]
    assert (a[:50] == 0).all() and (a[50:] == 1).all()          # queued left, newcomers right
    a = ck["choice_a1__s1__assign"]
    assert (a[50:] == 0).all()                                  # mirrored
    a = ck["choice_a2__s1__assign"]
    assert np.array_equal(a, np.arange(8) % 2)                  # no queue: the nearer door
    assert list(ck["choice_b0__assign"]) == [0, 0] and list(ck["choice_b1__assign"]) == [1, 1]
    h = ck["choice_c__assign_at_evals"]
    # one boundary: every right-door agent lies right of every left-door agent, and nobody switches back
    for row in h:
        assert np.all(np.diff(row.astype(int)) >= 0)
    assert np.all(np.diff(h.astype(int), axis=0) >= 0)
