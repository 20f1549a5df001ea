"""The batched calibration's gradient (reduced batched calibration configuration) against unbatched adjoints.

calibrate.py fits (speed mean, speed sd, relaxation time) to exit-flow curves with the batched forward + adjoint of
scenarios on the 1/4 bowl. Here a reduced form of it (3 scenarios of 2,000 seated agents each, K = 150 steps,
checkpoints every 25, the headway estimator's crossing seeds and truncation window): the batched dL/dv0 of every
scenario must equal, bit for bit, that scenario's own plain context run through the same forward, the same crossing
seeds and the same backward; the per-scenario relaxation-time adjoints sum wave partials with float atomics (their
order varies), so they agree to rounding.
"""
import numpy as np
import pytest

import calibrate as cb
import swarm_hip as sh
from common import assert_bits

pytestmark = pytest.mark.gpu


def test_batched_calibration_gradient_is_the_unbatched_ones():
    S, N, K, every = 3, 2000, 150, 25
    cal = cb.Calibration(scenarios=S, agents=N, K=K, every=every, estimator="headway", window=25)
    cal.set_target((1.20, 0.30, 0.5))
    cap = {}
    set_terms, backward = cal.sim.set_adjoint_terms, cal.sim.adjoint_backward

    def capture_terms(step=None, gin=None, gout=None, **kw):
        if step is not None:
            cap.update(step=step.copy(), gin=gin.copy(), gout=gout.copy(), kw=dict(kw))
        return set_terms(step, gin, gout, **kw)

    def capture_backward():
        cap["grad"] = backward()
        return cap["grad"]

    cal.sim.set_adjoint_terms = capture_terms
    cal.sim.adjoint_backward = capture_backward
    theta = (1.34, 0.26, 0.7)
    loss, grad = cal.run(theta)
    assert np.isfinite(loss) and np.all(np.isfinite(grad))
    gv0_b, glob_b = cap["grad"]
    assert np.count_nonzero(cap["step"] >= 0) > 20, "crossings seeded in the window"
    v0, _ = cal.speeds(theta[0], theta[1])
    for s in range(S):
        sl = slice(s * N, (s + 1) * N)
        plain = sh.Sim(cal.scene, N)
        plain.set_model(theta[2], 20.0)
        a = {k: v[sl] for k, v in cal.base.items()}
        a["v0"] = v0[sl]
        plain.upload(a)
        plain.adjoint_forward(K, every, 1.0, 0.0)
        st_p, ex_p, _, _ = plain.evac_points()
        st_b, ex_b, _, _ = cal.sim.evac_points()
        assert np.array_equal(st_p, st_b[sl]) and np.array_equal(ex_p, ex_b[sl])
        plain.set_adjoint_terms(cap["step"][sl], cap["gin"][sl], cap["gout"][sl], **cap["kw"])
        gv0_p, glob_p = plain.adjoint_backward()
        assert np.count_nonzero(gv0_p) > 0
        assert_bits(gv0_b[sl], gv0_p, f"scenario {s}: dL/dv0")
        np.testing.assert_allclose(glob_b.reshape(-1, 3)[s], glob_p[:3], rtol=1e-4, atol=1e-6 * np.abs(glob_p[:3]).max())
        plain.close()
