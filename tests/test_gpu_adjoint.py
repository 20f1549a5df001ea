"""The GPU adjoint of K steps against its FP64 reference and FP64 central finite differences.

The adjoint differentiates the step with every discrete decision frozen at the forward's values (include/
swarm_adjoint.h). Its checker is oracle/adjoint_ref.cpp: the same frozen-decision model evaluated in double
precision from the forward's definitions, giving directional derivatives by forward-mode dual numbers and the loss
itself for central differences (eps 1e-6 along unit directions in parameter units). On 64 random directions:
- the GPU gradient (FP32, reverse mode, checkpoints every 20 steps) equals the FP64 product of the step Jacobians
  along the same recorded trajectory (the implementation's exactness);
- the FP64 model's forward mode equals its own central differences (the model's differentiability);
- the GPU gradient, projected on the directions, equals FP64 central differences to 1e-3 relative (the vector
  of directional derivatives, norm-wise) over a horizon where the FP32 and FP64 trajectories stay within rounding
  of each other.
Beyond tens of steps this crowd is chaotic (the last test measures it on the plain forward), so the pathwise
derivative of a long-horizon loss grows exponentially and no finite-difference step resolves it.

Workload: the 40 x 30 m room with two 2 m exits (the small plain configuration's floor), seeded placement; loss
= sum of squared nav-field distances to the exits at step K + 0.1 x the summed squared overlaps of every step.
"""
import numpy as np
import pytest

import oracle_py as op
import swarm_hip as sh
from common import assert_bits

pytestmark = pytest.mark.gpu

P = 32
GOAL_WALL = 0x4000   # SWARM_REC_GOAL_WALL
# parameter units of a direction: v0 (model speed), tau_s (s), radius scale, wall gain
UNITS = (0.05, 0.05, 0.02, 1.0)


def make(n, K, every=20, seed=1337, scale=1.0, ovw=0.1):
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    a = sh.place_agents(sh.LAYOUT_ROOM, n, n, seed)
    sim = sh.Sim(scene, n)
    sim.upload(a)
    loss = sim.adjoint_forward(K, every, scale, ovw)
    return scene, a, sim, loss


def reference(scene, a, sim, K, ovw=0.1, scale=1.0, tau_s=0.5, gain=20.0):
    ref = op.AdjRef(scene.info, scene.tables(), sim.n, P, K, ovw, a, tau_s, scale, gain)
    tape = sh.new_tape(sim.n, P)
    ex = (sh.ExitRec * sim.n)()
    for s in range(K):
        sim.adjoint_tape(s, tape, ex if s == K - 1 else None)
        ref.add_step(s, tape["view"])
    ref.set_exit(ex)
    return ref


def directions(n, m, seed=7):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((m, n + 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:, :n] *= UNITS[0]
    d[:, n] *= UNITS[1]
    d[:, n + 1] *= UNITS[2]
    d[:, n + 2] *= UNITS[3]
    return d


def rel_err(a, b):
    """per direction: |a - b| / max(|b|, 1e-2 max |b|)"""
    floor = 1e-2 * np.max(np.abs(b))
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def vec_err(a, b):
    """the relative error of the projected gradient as a vector: ||a - b|| / ||b|| over the directions"""
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_adjoint_forward_is_the_plain_forward():
    n, K = 1024, 45
    scene, a, sim, loss = make(n, K)
    plain = sh.Sim(scene, n)
    plain.upload(a)
    tape = sh.new_tape(n, P)
    for s in range(K):
        st = plain.state()
        sim.adjoint_tape(s, tape)
        for f in ("x", "y", "vx", "vy", "w"):
            assert_bits(tape[f], st[f], f"taped step {s} input {f}")
        plain.step(1)
    assert np.isfinite(loss) and loss > 0.0


def check_directions(n, K, m=64, seed=1337, free=True):
    scene, a, sim, loss = make(n, K, seed=seed)
    gv0, g = sim.adjoint_backward()
    ref = reference(scene, a, sim, K)
    d = directions(n, m)
    gpu = d @ np.concatenate([gv0.astype(np.float64), g[:3]])
    along, _ = ref.directional(d, anchored=True)
    out = dict(loss=loss, gv0=gv0, g=g, gpu=gpu, along=along, ref=ref, sim=sim, a=a, K=K)
    if free:
        out["free"], _ = ref.directional(d, anchored=False)
        eps = 1e-6
        out["fd"] = np.array([(ref.loss(u / np.linalg.norm(u), eps)[0] - ref.loss(-u / np.linalg.norm(u), eps)[0]) / (2 * eps)
                              * np.linalg.norm(u) for u in d])
    return out


@pytest.fixture(scope="module")
def room_run():
    # the differentiable configuration's floor and agents (4,096 in the 40 x 30 m room), K = 20 steps (within the
    # crowd's predictability horizon, see test_the_crowd_is_chaotic_beyond_tens_of_steps), checkpoints every 20
    return check_directions(4096, 20)


def test_fp64_model_follows_the_forward(room_run):
    loss64, traj = room_run["ref"].loss()
    assert traj < 1e-4, f"FP64 model drifts from the FP32 forward by {traj}"
    assert loss64 == pytest.approx(room_run["loss"], rel=1e-5)


def test_gradient_is_the_tangent_along_the_trajectory(room_run):
    # the GPU adjoint (FP32, reverse mode, 2 checkpoint segments) against the FP64 forward-mode product of the step
    # Jacobians along the same recorded trajectory: the implementation's exactness
    e = rel_err(room_run["gpu"], room_run["along"])
    print(f"gpu vs FP64 tangent along the trajectory: max rel err {e.max():.2e}")
    assert e.max() < 1e-4
    # the run exercises e_goal's wall rule (agents beside the walls whose goal direction leads into them)
    sim, K = room_run["sim"], room_run["K"]
    tape = sh.new_tape(sim.n, P)
    ruled = 0
    for s in range(K):
        sim.adjoint_tape(s, tape)
        f = np.array([p.flags for p in tape["pred"]], np.uint32)
        ruled += int(np.count_nonzero(f & GOAL_WALL))
    print(f"wall-rule applications in the {K} steps: {ruled}")
    assert ruled > 0


def test_fp64_forward_mode_matches_central_differences(room_run):
    # the FP64 model's forward mode against its own central differences (eps 1e-6, unit directions in parameter units)
    e = rel_err(room_run["free"], room_run["fd"])
    print(f"FP64 forward mode vs central differences: max rel err {e.max():.2e}")
    assert e.max() < 1e-4


def test_gradient_matches_central_differences():
    # the GPU gradient against FP64 central differences of the model, over a horizon where the FP32 and FP64
    # trajectories stay within rounding of each other (10 steps: they differ by ~1e-6 of the room)
    r = check_directions(4096, 10)
    e = rel_err(r["gpu"], r["fd"])
    v = vec_err(r["gpu"], r["fd"])
    print(f"gpu vs FP64 central differences (K = 10): relative error {v:.2e}; per direction max {e.max():.2e}, median {np.median(e):.2e}")
    assert v < 1e-3 and e.max() < 5e-3


def test_gradient_near_central_differences_at_20_steps(room_run):
    # at 20 steps the FP32 and FP64 trajectories have drifted apart by the crowd's amplification of rounding (a few
    # 1e-5 of the room), and single directions differ from the FP64 differences by a few 1e-3
    e = rel_err(room_run["gpu"], room_run["fd"])
    v = vec_err(room_run["gpu"], room_run["fd"])
    print(f"gpu vs FP64 central differences (K = 20): relative error {v:.2e}; per direction max {e.max():.2e}, median {np.median(e):.2e}")
    assert v < 2e-3 and np.median(e) < 1e-3


def test_gradient_is_the_tangent_over_longer_horizons():
    # 60 steps, 3 checkpoint segments: along the recorded trajectory the FP32 adjoint stays the exact tangent up to
    # FP32 rounding of tangents that have grown by orders of magnitude
    r = check_directions(4096, 60, m=16, free=False)
    e = rel_err(r["gpu"], r["along"])
    v = vec_err(r["gpu"], r["along"])
    print(f"gpu vs FP64 tangent along the trajectory (K = 60): relative error {v:.2e}; per direction max {e.max():.2e}, "
          f"median {np.median(e):.2e}")
    assert v < 5e-3 and np.median(e) < 1e-3


def test_the_crowd_is_chaotic_beyond_tens_of_steps():
    # Why the checks above stop at tens of steps: the plain forward itself (no adjoint involved) amplifies a
    # 1e-6 perturbation of the initial positions by orders of magnitude within 40 steps in this crowd (4,096
    # agents in the room: hard-disc contacts), so the pathwise derivative of a K-step loss grows like exp(lambda K)
    # and central differences at eps 1e-6 leave the linear regime.
    n = 4096
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    a = sh.place_agents(sh.LAYOUT_ROOM, n, n, 1337)
    b = {k: v.copy() for k, v in a.items()}
    b["x"] = (b["x"] + 1e-6 * np.random.default_rng(3).standard_normal(n)).astype(np.float32)
    s1, s2 = sh.Sim(scene, n), sh.Sim(scene, n)
    s1.upload(a)
    s2.upload(b)
    s1.step(40)
    s2.step(40)
    x1, x2 = s1.state(), s2.state()
    act = (x1["w"] > 0) & (x2["w"] > 0)
    growth = np.max(np.hypot(x1["x"][act] - x2["x"][act], x1["y"][act] - x2["y"][act])) / 1e-6
    print(f"largest displacement after 40 steps / initial perturbation: {growth:.3e}")
    assert growth > 1e2


def test_torch_autograd_wrapper(room_run):
    torch = pytest.importorskip("torch")
    sim, a, K = room_run["sim"], room_run["a"], room_run["K"]

    class CrowdLoss(torch.autograd.Function):
        """The K-step loss as an autograd function of (v0, [tau_s, radius scale, wall gain]) over the C ABI pair."""

        @staticmethod
        def forward(ctx, v0, glob):
            st = dict(a)
            st["v0"] = v0.detach().cpu().numpy().astype(np.float32)
            sim.set_model(float(glob[0]), float(glob[2]))
            sim.upload(st)
            return torch.tensor(sim.adjoint_forward(K, 20, float(glob[1]), 0.1), dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            gv0, g = sim.adjoint_backward()
            return gout * torch.from_numpy(gv0.astype(np.float64)), gout * torch.from_numpy(g[:3])

    v0 = torch.tensor(a["v0"].astype(np.float64), requires_grad=True)
    glob = torch.tensor([0.5, 1.0, 20.0], dtype=torch.float64, requires_grad=True)
    loss = CrowdLoss.apply(v0, glob)
    loss.backward()
    assert float(loss.detach()) == pytest.approx(room_run["loss"], rel=1e-6)
    np.testing.assert_allclose(v0.grad.numpy(), room_run["gv0"], rtol=1e-5, atol=1e-6 * np.abs(room_run["gv0"]).max())
    np.testing.assert_allclose(glob.grad.numpy(), room_run["g"][:3], rtol=1e-5)


def test_captured_segments_match_eager():
    # full segments replay and run backwards as captured graphs (one per parity), the partial last one eagerly;
    # dL/dv0 equals the all-eager one bit for bit, also on a second backward (cached graphs) and with crossing seeds
    # and a truncation window; the per-scenario globals (relaxation time, wall gain) sum wave partials with float
    # atomics, whose order differs from run to run
    n, K = 1024, 45
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    a = sh.place_agents(sh.LAYOUT_ROOM, n, n, 5)
    rng = np.random.default_rng(2)
    steps = rng.integers(-1, K + 1, n).astype(np.int32)
    gin = rng.standard_normal((n, 2)).astype(np.float32)
    gout = rng.standard_normal((n, 2)).astype(np.float32)
    out = {}
    for graph in (True, False):
        sim = sh.Sim(scene, n, use_graph=graph)
        sim.set_tape_budget(0)   # checkpoints and recompute, whatever fits
        sim.upload(a)
        loss = sim.adjoint_forward(K, 20, 1.0, 0.1)
        assert sim.adjoint_segments() == 3
        r1 = sim.adjoint_backward()
        r2 = sim.adjoint_backward()
        sim.set_adjoint_terms(steps, gin, gout, exit_term=False, window=10)
        r3 = sim.adjoint_backward()
        out[graph] = (loss, r1, r2, r3)
    assert out[True][0] == out[False][0]
    for i in (1, 2, 3):
        assert_bits(out[True][i][0], out[False][i][0], f"dL/dv0, backward {i}")
        np.testing.assert_allclose(out[True][i][1], out[False][i][1], rtol=1e-5)
    assert_bits(out[True][1][0], out[True][2][0], "second backward")


def test_crossing_terms_match_differences():
    # The evacuation-crossing terms (the calibration's gradient path): L = sum_i c_i tau_i over the agents' refined
    # exit-crossing times, its adjoint through swarm_adjoint_set_terms (position adjoints at each crossing's two
    # endpoints, injected at their states), against central differences of the same L over the agents' desired
    # speeds. A sparse crowd (300 agents in the room, no contacts over 40 steps) keeps the forward smooth; an agent
    # counts when its perturbed runs cross at the same steps and the differences at h and h / 2 agree to 0.5 % (no
    # discrete event, such as a nav-cell or wall-contact change, inside the interval).
    import calibrate as cb

    n, K = 300, 40
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    a = sh.place_agents(sh.LAYOUT_ROOM, n, n, 3)
    rects = scene.tables()["exit_rects"].reshape(-1, 4)[: scene.info.E].astype(np.float64)
    c = np.random.default_rng(4).standard_normal(n)
    sim = sh.Sim(scene, n)

    def run(v0):
        st = dict(a)
        st["v0"] = v0.astype(np.float32)
        sim.upload(st)
        sim.adjoint_forward(K, 20, 1.0, 0.0)
        step, ex, ins, bef = sim.evac_points()
        tau, ok, d_in, d_bf = cb.crossing_times(step, ex, ins, bef, rects, scene.info.E)
        return step, ok, tau, d_in, d_bf

    v0 = a["v0"].astype(np.float64)
    step, ok, tau, d_in, d_bf = run(v0)
    crossed = np.nonzero(ok)[0]
    assert crossed.size >= 5, "too few crossings in the window"
    w = np.where(ok, c, 0.0)
    sim.set_adjoint_terms(np.where(ok, step, -1).astype(np.int32), (w[:, None] * d_in).astype(np.float32),
                          (w[:, None] * d_bf).astype(np.float32), exit_term=False, window=0)
    gv0, _ = sim.adjoint_backward()
    sim.set_adjoint_terms(None, None, None, exit_term=True, window=0)
    def central(j, h):
        vp, vm = v0.copy(), v0.copy()
        vp[j] += h
        vm[j] -= h
        sp, okp, tp, _, _ = run(vp)
        sm, okm, tm, _, _ = run(vm)
        if not (np.array_equal(sp, step) and np.array_equal(sm, step)):
            return None   # a crossing moved to another step: the loss is not differentiable there
        return (np.sum(c[okp] * tp[okp]) - np.sum(c[okm] * tm[okm])) / (2 * h)

    errs = []
    for j in crossed[:24]:
        f1, f2 = central(j, 2e-3 * v0[j]), central(j, 1e-3 * v0[j])
        if f1 is None or f2 is None or abs(f1 - f2) > 5e-3 * max(abs(f1), abs(f2)):
            print(f"agent {j}: skipped (not smooth at this scale)")
            continue
        errs.append(abs(gv0[j] - f2) / abs(f2))
        print(f"agent {j}: adjoint {gv0[j]:.6e}, central differences {f1:.6e} (h), {f2:.6e} (h / 2); rel err {errs[-1]:.1e}")
    # a crossing seeded at the wrong state would be off by a whole step's motion (order 1); the frozen-decision
    # model's first-order wall clearances (exact for flat walls, not at the exit gap's wall ends) leave a few
    # percent on some agents
    assert len(errs) >= 5
    assert np.median(errs) < 1e-2 and max(errs) < 6e-2


def test_taped_forward_segments_match_recompute():
    # the forward tapes its last segment (the backward starts there without recomputing it); a checkpoint interval
    # >= K, or a run whose whole tape fits the tape budget (the default one here), tapes the whole run and
    # recomputes nothing: every split gives the same loss and dL/dv0 bit for bit, also with a truncation window
    n, K = 1024, 45
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    a = sh.place_agents(sh.LAYOUT_ROOM, n, n, 6)
    out = {}
    for every, budget, segs in ((20, 0, 3), (7, 0, 7), (45, 0, 1), (100, 0, 1), (20, None, 1)):
        for graph in (True, False):
            sim = sh.Sim(scene, n, use_graph=graph)
            if budget is not None:
                sim.set_tape_budget(budget)
            sim.upload(a)
            loss = sim.adjoint_forward(K, every, 1.0, 0.1)
            assert sim.adjoint_segments() == segs, (every, budget)
            r1 = sim.adjoint_backward()
            sim.set_adjoint_terms(window=10)
            r2 = sim.adjoint_backward()
            out[(every, budget, graph)] = (loss, r1[0], r2[0], sim.tape_bytes())
    ref = out[(20, 0, False)]
    for key, (loss, g1, g2, nbytes) in out.items():
        assert loss == ref[0], key
        assert_bits(g1, ref[1], f"dL/dv0, checkpoint every {key[0]}, budget {key[1]}, graph {key[2]}")
        assert_bits(g2, ref[2], f"dL/dv0 with a window, checkpoint every {key[0]}, budget {key[1]}, graph {key[2]}")
    assert out[(45, 0, True)][3] > out[(7, 0, True)][3]   # the whole run's tape is the larger


def test_candidate_budget_in_the_tape():
    """The taped steps use the candidate lists' budget too: the default (12 moves), and 0.3 moves, which breaks in the
    dense room (a broken taped step makes the call run again with whole lists), give the loss and dL/dv0 of whole lists
    bit for bit, through the forward and through a backward that recomputes its segments, and keep giving them after."""
    n, K = 2048, 40
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    a = sh.place_agents(sh.LAYOUT_ROOM, n, n, 11)
    out = {}
    for moves in (float("inf"), 12.0, 0.3):
        for tape_budget in (None, 0):   # the whole run taped by the forward; every segment recomputed
            sim = sh.Sim(scene, n)
            sim.set_candidate_budget(moves)
            if tape_budget is not None:
                sim.set_tape_budget(tape_budget)
            sim.upload(a)
            print(f"adjoint with candidate budget {moves}, tape budget {tape_budget}", flush=True)
            loss = sim.adjoint_forward(K, 10, 1.0, 0.1)
            g1 = sim.adjoint_backward()[0]
            sim.upload(a)
            loss2 = sim.adjoint_forward(K, 10, 1.0, 0.1)   # a second run on the same context
            g2 = sim.adjoint_backward()[0]
            out[(moves, tape_budget)] = (loss, g1, loss2, g2)
            sim.close()
    ref = out[(float("inf"), None)]
    for key, (loss, g1, loss2, g2) in out.items():
        assert loss == ref[0] and loss2 == ref[0], key
        assert_bits(g1, ref[1], f"dL/dv0, budget {key}")
        assert_bits(g2, ref[1], f"dL/dv0 of the second run, budget {key}")


def run_k200(n, budget=0, seed=1337):
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    a = sh.place_agents(sh.LAYOUT_ROOM, n, n, seed)
    sim = sh.Sim(scene, n)
    if budget is not None:
        sim.set_tape_budget(budget)
    sim.upload(a)
    loss = sim.adjoint_forward(200, 20, 1.0, 0.1)
    return scene, a, sim, loss


def test_k200_gradient_matches_central_differences():
    # The differentiable configuration's horizon, K = 200 steps with checkpoints every 20 and recompute, in a crowd
    # sparse enough that the FP32 and FP64 trajectories stay within rounding of each other over the whole evacuation
    # (128 agents in the room; the configuration's 4,096 amplify rounding beyond any difference step, see the last
    # tests). 64 random directions: GPU gradient vs the FP64 tangent along the recorded trajectory, the FP64 model's
    # forward mode vs its own central differences (eps 1e-6), and the GPU gradient vs those differences, 1e-3 relative.
    K, n = 200, 128
    scene, a, sim, loss = run_k200(n)
    assert sim.adjoint_segments() == 10
    gv0, g = sim.adjoint_backward()
    assert np.all(np.isfinite(gv0)) and np.all(np.isfinite(g))
    ref = reference(scene, a, sim, K)
    loss64, traj = ref.loss()
    assert traj < 1e-4 and loss64 == pytest.approx(loss, rel=1e-5)
    d = directions(n, 64)
    gpu = d @ np.concatenate([gv0.astype(np.float64), g[:3]])
    along, _ = ref.directional(d, anchored=True)
    free, _ = ref.directional(d, anchored=False)
    eps = 1e-6
    fd = np.array([(ref.loss(u / np.linalg.norm(u), eps)[0] - ref.loss(-u / np.linalg.norm(u), eps)[0]) / (2 * eps)
                   * np.linalg.norm(u) for u in d])
    print(f"K = 200, {n} agents: gpu vs FP64 tangent {vec_err(gpu, along):.2e}; FP64 forward mode vs FD {vec_err(free, fd):.2e}; "
          f"gpu vs FD {vec_err(gpu, fd):.2e} (per direction median {np.median(rel_err(gpu, fd)):.2e})")
    assert vec_err(gpu, along) < 1e-4
    assert vec_err(free, fd) < 1e-4
    assert vec_err(gpu, fd) < 1e-3


def test_k200_recompute_is_the_tape_and_the_tangent():
    # 2,048 agents, K = 200: checkpoints every 20 with recompute give the whole-run tape's loss and dL/dv0 bit for bit;
    # along the recorded trajectory the gradient is the FP64 tangent (grown here by twenty orders of magnitude, so the
    # comparison is norm-wise over 16 directions); the state after the backward is the forward's final state
    K, n = 200, 2048
    scene, a, sim, loss = run_k200(n, budget=0)
    after_fwd = sim.state()
    gv0, g = sim.adjoint_backward()
    after_bwd = sim.state()
    for f in after_fwd:
        assert_bits(after_bwd[f], after_fwd[f], f"state after the backward: {f}")
    _, _, whole, loss_w = run_k200(n, budget=None)
    assert whole.adjoint_segments() == 1 and loss_w == loss
    gw, gwg = whole.adjoint_backward()
    assert_bits(gw, gv0, "dL/dv0, whole tape vs recompute")
    ref = reference(scene, a, sim, K)
    d = directions(n, 16)
    gpu = d @ np.concatenate([gv0.astype(np.float64), g[:3]])
    along, _ = ref.directional(d, anchored=True)
    v = vec_err(gpu, along)
    print(f"K = 200, {n} agents: |dL/dv0| {np.linalg.norm(gv0.astype(np.float64)):.3e}; gpu vs FP64 tangent {v:.2e}")
    assert v < 2e-2


def test_k200_overflow_reports_its_status():
    # the configuration's own crowd (4,096 agents in the room): the pathwise derivative of the 200-step loss grows
    # past the float range; the backward writes what it has and returns SWARM_ENONFINITE
    scene, a, sim, loss = run_k200(4096)
    with pytest.raises(sh.NonFiniteGradient) as e:
        sim.adjoint_backward()
    assert not (np.all(np.isfinite(e.value.grad_v0)) and np.all(np.isfinite(e.value.grad_globals)))
    # a truncation window keeps it finite
    sim.set_adjoint_terms(window=25)
    gv0, g = sim.adjoint_backward()
    assert np.all(np.isfinite(gv0)) and np.all(np.isfinite(g))


def test_tape_follows_the_iteration_count():
    # a forward at 8 projection iterations, then 32: the tape is made again for 32 (its mask and position arrays hold
    # an entry per iteration) and the gradient equals a fresh context's; a backward after the iteration count changed
    # is refused (the forward's run no longer applies)
    n, K = 1024, 30
    scene = sh.Scene.layout(sh.LAYOUT_ROOM)
    a = sh.place_agents(sh.LAYOUT_ROOM, n, n, 9)
    sim = sh.Sim(scene, n, proj_iters=8)
    sim.upload(a)
    sim.adjoint_forward(K, 10, 1.0, 0.1)
    sim.adjoint_backward()
    sim.set_proj_iters(32)
    with pytest.raises(RuntimeError):
        sim.adjoint_backward()
    sim.upload(a)
    loss = sim.adjoint_forward(K, 10, 1.0, 0.1)
    gv0, _ = sim.adjoint_backward()
    fresh = sh.Sim(scene, n)
    fresh.upload(a)
    assert fresh.adjoint_forward(K, 10, 1.0, 0.1) == loss
    assert_bits(fresh.adjoint_backward()[0], gv0, "dL/dv0 after the iteration count changed")
