"""LaneSweep (finite_volume.hpp:880-1037; luna_fv_options.sweep_lanes = 2, 4, 8) on the GPU: dt, gradients, point velocities,
positions and Q equal to the oracle's restatement to the bit over three steps, static and moving, 2D (both dt rules) and
3D, HLLC, HLL and the HLL fallback, the face-centroid limiter and none; within rounding of the one-lane sweep; and the
derivatives refuse it."""
import numpy as np
import pytest

import cases
import oracle
from luna_hydro import hydro, mesh

pytestmark = pytest.mark.gpu
GAMMA = 5.0 / 3.0


@pytest.mark.parametrize("dim,n,kw", [(3, 10, {}), (3, 10, {"riemann": "hll"}), (3, 10, {"hll_compression": 0.5}),
                                      (2, 24, {}), (2, 24, {"step_rule": "luna", "limiter": "none"})])
@pytest.mark.parametrize("motion", ["static", "moving"])
@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_lane_sweep_matches_oracle(dim, n, kw, motion, lanes):
    print("LaneSweep %d lanes dim=%d n=%d %s %s" % (lanes, dim, n, kw, motion), flush=True)
    m = cases.voronoi(n, dim)
    q = cases.rough_state(m)
    gpu = hydro.FiniteVolume(m, GAMMA, motion=motion, sweep_lanes=lanes, **kw)
    one = hydro.FiniteVolume(m, GAMMA, motion=motion, **kw)
    orc = oracle.FiniteVolume(m, GAMMA, moving=motion == "moving", sweep_lanes=lanes, **kw)
    try:
        for fv in (gpu, one, orc):
            fv.set_state(q)
        for s in range(3):
            dt = orc.time_step(0.3)
            assert gpu.time_step(0.3) == dt, s
            one.time_step(0.3)
            g = gpu.gradients()
            assert np.array_equal(g, orc.grad.reshape(g.shape)), s
            g1 = one.gradients()
            scale = np.max(np.abs(g1)) + 1e-300
            assert np.max(np.abs(g - g1)) <= 1e-12 * scale, s
            if motion == "static":
                gpu.step(dt)
                orc.step(dt)
                one.set_state(gpu.state())
            else:
                x = gpu.moving_begin(dt)
                assert np.array_equal(gpu.point_velocities(), orc.w), s
                assert np.array_equal(orc.moving_begin(dt), x), s
                m1 = mesh.build(x, 1.0)
                gpu.moving_finish(m1, dt)
                orc.moving_finish(m1, dt)
                one.close()
                one = hydro.FiniteVolume(m1, GAMMA, motion=motion, **kw)
                one.set_state(gpu.state())
            assert np.array_equal(gpu.state(), orc.q), s
            assert gpu.fallbacks() == orc.fallbacks()
    finally:
        gpu.close()
        one.close()


def test_lane_sweep_refusals():
    m = cases.voronoi(6, 3)
    for kw in ({"order": 0}, {"precision": "grad16"}, {"limiter": "learned"}):
        with pytest.raises(ValueError):
            hydro.FiniteVolume(m, GAMMA, sweep_lanes=4, **kw)
    with pytest.raises(ValueError):
        hydro.FiniteVolume(m, GAMMA, sweep_lanes=3)
    fv = hydro.FiniteVolume(m, GAMMA, sweep_lanes=4)
    try:
        fv.set_state(cases.rough_state(m))
        dt = fv.time_step(0.3)
        with pytest.raises(RuntimeError):
            fv.step_adjoint(dt, np.ones((m["cells"], 5)))
    finally:
        fv.close()
