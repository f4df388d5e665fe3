"""The reduced-precision policies on the GPU (grad32, grad16: luna-cpp src/hydro/precision.hpp:26-54) against the
oracle's float restatement on the same seeded meshes and states: dt equal, the stored gradient rows (decoded) equal to
the bit, the conserved state after every step within 1e-12 relative (to the bit in practice), fallbacks equal; static
and moving, 2D and 3D, HLLC and HLL. And the refusals: the learned limiter and the derivatives need double.
"""
import numpy as np
import pytest

import cases
import oracle
from luna_hydro import hydro, mesh

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("precision", ["grad32", "grad16"])
@pytest.mark.parametrize("dim,n,kw", [
    (3, 10, {}),
    (3, 10, {"riemann": "hll"}),
    (3, 10, {"limiter": "none"}),
    (2, 24, {}),
    (2, 24, {"step_rule": "luna"}),
])
def test_static_steps_match_oracle(precision, dim, n, kw):
    say("precision %s static dim=%d n=%d %s" % (precision, dim, n, kw))
    m = cases.voronoi(n, dim)
    q = cases.rough_state(m)
    gpu = hydro.FiniteVolume(m, 5.0 / 3.0, precision=precision, **kw)
    orc = oracle.FiniteVolume(m, 5.0 / 3.0, precision=precision, **kw)
    gpu.set_state(q)
    orc.set_state(q)
    try:
        for s in range(4):
            dt = orc.time_step(0.3)
            assert gpu.time_step(0.3) == dt, s
            assert np.array_equal(gpu.gradients(), orc.grad), s
            gpu.step(dt)
            orc.step(dt)
            qg, qo = gpu.state(), orc.q
            assert np.max(np.abs(qg - qo) / np.maximum(np.abs(qo), 1e-300)) <= TOL, s
            assert gpu.fallbacks() == orc.fallbacks()
        say("  bits equal: %s" % np.array_equal(gpu.state(), orc.q))
    finally:
        gpu.close()


@pytest.mark.parametrize("precision", ["grad32", "grad16"])
@pytest.mark.parametrize("dim,n", [(3, 10), (2, 24)])
def test_moving_steps_match_oracle(precision, dim, n):
    say("precision %s moving dim=%d n=%d" % (precision, dim, n))
    m = cases.voronoi(n, dim)
    q = cases.rough_state(m)
    gpu = hydro.FiniteVolume(m, 5.0 / 3.0, motion="moving", precision=precision)
    orc = oracle.FiniteVolume(m, 5.0 / 3.0, moving=True, precision=precision)
    gpu.set_state(q)
    orc.set_state(q)
    try:
        for s in range(3):
            dt = orc.time_step(0.3)
            assert gpu.time_step(0.3) == dt
            assert np.array_equal(gpu.gradients(), orc.grad), s
            assert np.array_equal(gpu.point_velocities(), orc.w)
            xg = gpu.moving_begin(dt)
            assert np.array_equal(xg, orc.moving_begin(dt)), s
            m = mesh.build(xg, 1.0)
            gpu.moving_finish(m, dt)
            orc.moving_finish(m, dt)
            qg, qo = gpu.state(), orc.q
            assert np.max(np.abs(qg - qo) / np.maximum(np.abs(qo), 1e-300)) <= TOL, s
            assert gpu.fallbacks() == orc.fallbacks()
        say("  bits equal: %s" % np.array_equal(gpu.state(), orc.q))
    finally:
        gpu.close()


def test_sedov_64_grad16_matches_oracle():
    """A 64^3 Sedov blast (the headline problem at an eighth of its cells), two steps under grad16."""
    from luna_hydro import problems

    m = cases.voronoi(64, 3, 0.25, 1)
    q, _ = problems.sedov_state(m)
    gpu = hydro.FiniteVolume(m, 5.0 / 3.0, precision="grad16")
    orc = oracle.FiniteVolume(m, 5.0 / 3.0, precision="grad16")
    gpu.set_state(q)
    orc.set_state(q)
    try:
        for s in range(2):
            dt = orc.time_step(0.3)
            assert gpu.time_step(0.3) == dt
            gpu.step(dt)
            orc.step(dt)
            assert np.max(np.abs(gpu.state() - orc.q) / np.maximum(np.abs(orc.q), 1e-300)) <= TOL, s
        say("sedov 64 grad16 bits equal: %s" % np.array_equal(gpu.state(), orc.q))
    finally:
        gpu.close()


def test_reduced_policies_refuse_what_needs_double():
    m = cases.voronoi(6, 3)
    with pytest.raises(ValueError):  # LUNA_ERR_ARGUMENT
        hydro.FiniteVolume(m, 5.0 / 3.0, limiter="learned", precision="grad16")
    fv = hydro.FiniteVolume(m, 5.0 / 3.0, precision="grad32")
    try:
        fv.set_state(cases.rough_state(m))
        with pytest.raises(RuntimeError):  # LUNA_ERR_LOGIC
            fv.step_tangent(1e-3, np.zeros((m["cells"], 5)))
    finally:
        fv.close()
