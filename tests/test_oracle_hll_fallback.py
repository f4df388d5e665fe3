"""The oracle's HLL fallback at extreme compression (no counterpart in luna-cpp; the 256^3 moving Sedov config asks for
it): the sweep's flags against the indicator recomputed in numpy from the oracle's own gradients and primitives
(r (div v) / c_s < -kappa), and each face's flux against the plain HLLC and HLL runs of the same step (HLL where a side
is flagged, HLLC elsewhere, to the bit); conservation; the derivatives refuse it.
"""
import math

import numpy as np
import pytest

import cases
import oracle
from luna_hydro import problems

KAPPA = 0.2


def sedov_after(steps, kappa=KAPPA, **kw):
    m = cases.voronoi(10, 3)
    q, _ = problems.sedov_state(m)
    fv = oracle.FiniteVolume(m, 5.0 / 3.0, hll_compression=kappa, **kw)
    fv.set_state(q)
    for _ in range(steps):
        fv.step(fv.time_step(0.3))
    return m, fv


def test_flags_are_the_compression_indicator():
    m, fv = sedov_after(1)
    fv.time_step(0.3)
    D, V = 3, 5
    g = fv.grad.reshape(-1, V, D)
    div = g[:, 1, 0] + g[:, 2, 1] + g[:, 3, 2]
    want = (div * fv.prims[:, V + 1] / fv.prims[:, V] < -KAPPA).astype(np.uint8)
    assert np.array_equal(fv.hll, want)
    assert 0 < int(fv.hll.sum()) < m["cells"]


def test_flagged_faces_take_hll_the_rest_hllc():
    m, fv = sedov_after(1)
    q = fv.q.copy()
    runs = {}
    for name, kw in (("fallback", {"hll_compression": KAPPA}), ("hllc", {}), ("hll", {"riemann": "hll"})):
        o = oracle.FiniteVolume(m, 5.0 / 3.0, **kw)
        o.set_state(q)
        dt = o.time_step(0.3)
        o.step(dt)
        runs[name] = (o, dt)
    assert runs["fallback"][1] == runs["hllc"][1] == runs["hll"][1]
    fb = runs["fallback"][0]
    flags = fb.hll.astype(bool)
    owner, neighbour = np.asarray(m["owner"]), np.asarray(m["neighbour"])
    hll_face = flags[owner] | flags[neighbour]
    F = m["faces"]
    assert hll_face.any() and not hll_face.all()
    assert np.array_equal(fb.flux[:F][hll_face], runs["hll"][0].flux[:F][hll_face])
    assert np.array_equal(fb.flux[:F][~hll_face], runs["hllc"][0].flux[:F][~hll_face])


def test_no_flags_is_hllc_and_conserves():
    m, fv = sedov_after(3, kappa=1e300)
    m2 = cases.voronoi(10, 3)
    q, _ = problems.sedov_state(m2)
    ref = oracle.FiniteVolume(m2, 5.0 / 3.0)
    ref.set_state(q)
    for _ in range(3):
        ref.step(ref.time_step(0.3))
    assert np.array_equal(fv.q, ref.q)
    m, fv = sedov_after(4)
    for v in range(5):
        assert abs(math.fsum(fv.q[:, v]) - math.fsum(q[:, v])) <= 1e-14 * max(1.0, math.fsum(np.abs(q[:, v])))


def test_derivatives_refuse_the_fallback():
    m, fv = sedov_after(0)
    with pytest.raises(ValueError):
        fv.step_tangent(1e-4, np.zeros((m["cells"], 5)))
