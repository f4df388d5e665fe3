"""The oracle's reduced-precision policies (luna-cpp src/hydro/precision.hpp:26-130, finite_volume.hpp:463-641 and
:1288-1329 in float): the FP16 codec against numpy's float16 (an independent IEEE binary16 implementation), the group
exponent against its defining interval, grad16 rows as the FP16 round trip of grad32's (numpy's encoder), and both
policies against double on the same mesh and state (the gradients to float precision, the conserved totals exactly).
"""
import math

import numpy as np
import pytest

import cases
import oracle


def test_half_decode_matches_numpy_for_every_pattern():
    h = np.arange(65536, dtype=np.uint16)
    ref = h.view(np.float16).astype(np.float32)
    got = np.array([oracle.half_to_float(int(x)) for x in h], dtype=np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32))


def test_half_encode_matches_numpy():
    """Random magnitudes across the normal and subnormal ranges, every midpoint between adjacent halves (ties to even),
    and the overflow boundary."""
    rng = np.random.default_rng(0)
    f = (rng.standard_normal(20000) * 2.0 ** rng.integers(-30, 18, 20000)).astype(np.float32)
    halves = np.arange(0, 0x7c00, dtype=np.uint16).view(np.float16).astype(np.float64)
    mid = ((halves[:-1] + halves[1:]) / 2).astype(np.float32)  # exact in float
    edge = np.float32([65504, 65519.99, 65520, 65536, 1e30, 2.0 ** -24, 2.0 ** -25, 2.0 ** -26, 0.0, -0.0, -2.0 ** -25, 6.1e-5])
    f = np.concatenate([f, mid, -mid, edge])
    with np.errstate(over="ignore"):
        ref = f.astype(np.float16).view(np.uint16)
    got = np.array([oracle.half_from_float(float(x)) for x in f], dtype=np.uint16)
    assert np.array_equal(got, ref)


def test_group_exponent_interval():
    rng = np.random.default_rng(1)
    for m in (rng.random(2000) * 2.0 ** rng.integers(-100, 100, 2000)).astype(np.float32):
        k = oracle.group_exponent(float(m))
        if m == 0:
            assert k == 0
            continue
        scaled = float(m) * 2.0 ** k
        if -114 < k < 126:
            assert 2.0 ** 14 <= scaled < 2.0 ** 15, (m, k)
    assert oracle.group_exponent(0.0) == 0 and oracle.group_exponent(-1.0) == 0
    assert oracle.group_exponent(3e38) == 14 - 127  # the largest binade: 3e38 2^-113 in [2^14, 2^15)
    assert oracle.group_exponent(float("inf")) == -114  # the clamp below
    assert oracle.group_exponent(1e-40) == 126  # a subnormal: its exponent bits read as 2^-127


def _swept(m, q, precision, limiter="face", moving=False):
    fv = oracle.FiniteVolume(m, 5.0 / 3.0, limiter=limiter, moving=moving, precision=precision)
    fv.set_state(q)
    dt = fv.time_step(0.3)
    return fv, dt


def _groups(D):
    return np.array([0 if c < D else 1 if c < D + D * D else 2 for c in range((D + 2) * D)])


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 20)])
def test_grad16_rows_are_fp16_round_trip_of_grad32(dim, n):
    """The same float fit and limiter, then each group scaled by 2^k (k from its largest |g|), rounded to FP16 by numpy,
    scaled back."""
    m = cases.voronoi(n, dim)
    q = cases.rough_state(m)
    f32, dt32 = _swept(m, q, "grad32")
    f16, dt16 = _swept(m, q, "grad16")
    assert dt32 == dt16
    g32 = f32.grad.astype(np.float32)
    assert np.array_equal(g32.astype(np.float64), f32.grad)  # the rows are floats
    groups = _groups(dim)
    expect = np.empty_like(f16.grad)
    for i in range(m["cells"]):
        for grp in range(3):
            sel = groups == grp
            largest = np.max(np.abs(g32[i, sel]))
            k = oracle.group_exponent(float(largest))
            scaled = (g32[i, sel] * np.float32(2.0 ** k)).astype(np.float32)
            expect[i, sel] = (scaled.astype(np.float16).astype(np.float32) * np.float32(2.0 ** -k)).astype(np.float64)
    assert np.array_equal(f16.grad, expect)


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 20)])
def test_float_fit_near_double(dim, n):
    """Unlimited gradients of a rough state: grad32 within float rounding of double (relative to each variable's largest
    gradient), grad16 within FP16's (2^-11 of the group's largest, plus grad32's)."""
    m = cases.voronoi(n, dim)
    q = cases.rough_state(m)
    d64, dt = _swept(m, q, "double", limiter="none")
    d32, dt32 = _swept(m, q, "grad32", limiter="none")
    d16, _ = _swept(m, q, "grad16", limiter="none")
    assert dt32 == dt  # the time step reads no gradient
    V = dim + 2
    for v in range(V):
        cols = slice(v * dim, (v + 1) * dim)
        scale = np.max(np.abs(d64.grad[:, cols]), axis=1, keepdims=True)
        scale = np.maximum(scale, 1e-300)
        e32 = np.max(np.abs(d32.grad[:, cols] - d64.grad[:, cols]) / scale)
        assert e32 < 2e-4, (v, e32)  # float sums of a condition-limited fit
        groups = _groups(dim)
        gscale = np.max(np.abs(d64.grad[:, groups == groups[v * dim]]), axis=1, keepdims=True)
        e16 = np.max(np.abs(d16.grad[:, cols] - d32.grad[:, cols]) / np.maximum(gscale, 1e-300))
        assert e16 <= 2.0 ** -11 * 1.01, (v, e16)


@pytest.mark.parametrize("precision", ["grad32", "grad16"])
@pytest.mark.parametrize("moving", [False, True])
def test_reduced_steps_conserve_and_stay_near_double(precision, moving):
    """Three steps of a 3D rough state: totals of mass, momentum and energy as double's to rounding (the fluxes are
    summed in double and cancel pairwise), and the states within a small relative L1 of double's."""
    from luna_hydro import mesh as meshlib

    m0 = cases.voronoi(8, 3)
    q = cases.rough_state(m0)
    runs = {}
    for p in ("double", precision):
        m = m0
        fv = oracle.FiniteVolume(m, 5.0 / 3.0, moving=moving, precision=p)
        fv.set_state(q)
        for _ in range(3):
            dt = fv.time_step(0.3)
            if moving:
                x = fv.moving_begin(dt)
                m = meshlib.build(x, 1.0)
                fv.moving_finish(m, dt)
            else:
                fv.step(dt)
        runs[p] = fv.q.copy()
    a, b = runs["double"], runs[precision]
    for v in range(5):
        ta, tb = math.fsum(a[:, v]), math.fsum(b[:, v])
        assert abs(ta - tb) <= 1e-13 * max(1.0, math.fsum(np.abs(a[:, v]))), (v, ta, tb)
    rel = np.sum(np.abs(a - b), axis=0) / np.sum(np.abs(a), axis=0)
    assert np.max(rel) < {"grad32": 1e-7, "grad16": 1e-4}[precision], rel  # measured 5e-9 and 6e-6


def test_reduced_policies_refuse_derivatives_and_learned():
    m = cases.voronoi(6, 3)
    fv = oracle.FiniteVolume(m, 5.0 / 3.0, precision="grad32")
    fv.set_state(cases.rough_state(m))
    with pytest.raises(Exception):
        fv.step_tangent(1e-3, np.zeros((m["cells"], 5)))
    with pytest.raises(ValueError):
        oracle.FiniteVolume(m, 5.0 / 3.0, limiter="learned", precision="grad16")
