"""Update::per_cell (finite_volume.hpp:1544-1566, fluxes_and_update: each cell computes its faces' fluxes itself) on the GPU:
the same bits as the per-face loop, which equals the oracle's; static and moving, 2D and 3D, HLLC, HLL and the HLL fallback,
the reduced-precision policies; the fallbacks counted alike."""
import numpy as np
import pytest

import cases
import oracle
from luna_hydro import hydro, mesh

pytestmark = pytest.mark.gpu
GAMMA = 5.0 / 3.0


@pytest.mark.parametrize("dim,n,kw", [(3, 10, {}), (3, 10, {"riemann": "hll"}), (3, 10, {"hll_compression": 0.5}), (3, 10, {"precision": "grad16"}),
                                      (2, 24, {}), (2, 24, {"step_rule": "luna", "limiter": "none"})])
@pytest.mark.parametrize("motion", ["static", "moving"])
def test_per_cell_equals_per_face(dim, n, kw, motion):
    print("per-cell update dim=%d n=%d %s %s" % (dim, n, kw, motion), flush=True)
    m = cases.voronoi(n, dim)
    q = cases.rough_state(m)
    a = hydro.FiniteVolume(m, GAMMA, motion=motion, update="per_cell", **kw)
    b = hydro.FiniteVolume(m, GAMMA, motion=motion, **kw)
    okw = {k: v for k, v in kw.items()}
    orc = oracle.FiniteVolume(m, GAMMA, moving=motion == "moving", **okw)
    try:
        for fv in (a, b, orc):
            fv.set_state(q)
        for s in range(3):
            dt = b.time_step(0.3)
            assert a.time_step(0.3) == dt and orc.time_step(0.3) == dt, s
            if motion == "static":
                a.step(dt)
                b.step(dt)
                orc.step(dt)
            else:
                x = b.moving_begin(dt)
                assert np.array_equal(a.moving_begin(dt), x)
                assert np.array_equal(orc.moving_begin(dt), x)
                m1 = mesh.build(x, 1.0)
                a.moving_finish(m1, dt)
                b.moving_finish(m1, dt)
                orc.moving_finish(m1, dt)
            assert np.array_equal(a.state(), b.state()), s
            assert np.array_equal(b.state(), orc.q), s
            assert a.fallbacks() == b.fallbacks() == orc.fallbacks()
    finally:
        a.close()
        b.close()


def test_per_cell_refuses_order_0():
    with pytest.raises(ValueError):
        hydro.FiniteVolume(cases.voronoi(6, 3), GAMMA, order=0, update="per_cell")
