"""The HLL fallback at extreme compression on the GPU (no counterpart in luna-cpp) against the oracle's: static and
moving 3D Sedov steps with flagged cells, dt, Q and fallbacks to the bit; an unreachable threshold gives the plain HLLC
step's bits; the combinations it needs double and the face limiter for are refused, and so are the derivatives.
"""
import numpy as np
import pytest

import cases
import oracle
from luna_hydro import hydro, mesh, problems

pytestmark = pytest.mark.gpu
KAPPA = 0.2


def say(msg):
    print(msg, flush=True)


def sedov(n=10):
    m = cases.voronoi(n, 3)
    q, _ = problems.sedov_state(m)
    return m, q


def test_static_steps_match_oracle():
    m, q = sedov()
    gpu = hydro.FiniteVolume(m, 5.0 / 3.0, hll_compression=KAPPA)
    orc = oracle.FiniteVolume(m, 5.0 / 3.0, hll_compression=KAPPA)
    gpu.set_state(q)
    orc.set_state(q)
    flagged = 0
    try:
        for s in range(5):
            dt = orc.time_step(0.3)
            assert gpu.time_step(0.3) == dt, s
            flagged += int(orc.hll.sum())
            gpu.step(dt)
            orc.step(dt)
            assert np.array_equal(gpu.state(), orc.q), s
            assert gpu.fallbacks() == orc.fallbacks()
    finally:
        gpu.close()
    say("static: %d flagged cell-steps" % flagged)
    assert flagged > 0


def test_moving_steps_match_oracle():
    m, q = sedov()
    gpu = hydro.FiniteVolume(m, 5.0 / 3.0, motion="moving", hll_compression=KAPPA)
    orc = oracle.FiniteVolume(m, 5.0 / 3.0, moving=True, hll_compression=KAPPA)
    gpu.set_state(q)
    orc.set_state(q)
    flagged = 0
    try:
        for s in range(4):
            dt = orc.time_step(0.3)
            assert gpu.time_step(0.3) == dt, s
            flagged += int(orc.hll.sum())
            xg = gpu.moving_begin(dt)
            assert np.array_equal(xg, orc.moving_begin(dt))
            m = mesh.build(xg, 1.0)
            gpu.moving_finish(m, dt)
            orc.moving_finish(m, dt)
            assert np.array_equal(gpu.state(), orc.q), s
    finally:
        gpu.close()
    say("moving: %d flagged cell-steps" % flagged)
    assert flagged > 0


def test_unreachable_threshold_is_plain_hllc():
    m, q = sedov()
    a = hydro.FiniteVolume(m, 5.0 / 3.0, hll_compression=1e300)
    b = hydro.FiniteVolume(m, 5.0 / 3.0)
    try:
        for fv in (a, b):
            fv.set_state(q)
            for _ in range(4):
                fv.step(fv.time_step(0.3))
        assert np.array_equal(a.state(), b.state())
    finally:
        a.close()
        b.close()


def test_refusals():
    m, q = sedov(6)
    for kw in ({"precision": "grad16"}, {"limiter": "learned"}, {"order": 0}):
        with pytest.raises(ValueError):
            hydro.FiniteVolume(m, 5.0 / 3.0, hll_compression=KAPPA, **kw)
    with pytest.raises(ValueError):
        hydro.FiniteVolume(m, 5.0 / 3.0, hll_compression=-1.0)
    fv = hydro.FiniteVolume(m, 5.0 / 3.0, hll_compression=KAPPA)
    try:
        fv.set_state(q)
        with pytest.raises(RuntimeError):
            fv.step_tangent(1e-4, np.zeros((m["cells"], 5)))
    finally:
        fv.close()
