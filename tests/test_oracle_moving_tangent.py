"""The oracle's tangent of the moving step (forward mode, fixed topology; finite_volume.hpp:2428-2486 order 0, :2764-2870 order 1
of the reference), with the mesh library's geometry tangent for both meshes. What pins it:
  - its pieces against central differences: the face velocity's tangent on both sides of the 0.9 |d| limit, and the moving
    face's (|A| F(L, R, A / |A|, w) in L, R, A and w, HLLC in each of its wave cases and HLL);
  - the dot test against the reverse (<w, J v> = <J^T w, v>, the reverse pinned by central differences in
    test_oracle_moving_adjoint.py): orders 0 and 1, 2D and 3D, HLLC and HLL, steps held and free;
  - central differences of the oracle's forward with the whole grid steps held, along Q, whole-grid-step positions and both.
"""
import numpy as np
import pytest

import oracle
from luna_hydro import mesh, problems

GAMMA = 5.0 / 3.0


def rng_of(seed):
    return np.random.Generator(np.random.MT19937(seed))


@pytest.mark.parametrize("D", [2, 3])
def test_face_velocity_tangent_against_central_differences(D):
    rng = rng_of(7 + D)
    worst = 0.0
    for frac in (0.2, 0.7, 1.3, 3.0):
        for _ in range(20):
            d = rng.uniform(-1, 1, D)
            c = rng.normal(size=D)
            c -= d * np.dot(c, d) / np.dot(d, d)
            c *= frac * np.linalg.norm(d) / np.linalg.norm(c)
            xf = 0.5 * d + c + 0.01 * rng.normal(size=D)  # off the bisector too: the formula's own arguments
            wL, wR = rng.uniform(-2, 2, D), rng.uniform(-2, 2, D)
            t = [rng.normal(size=D) for _ in range(4)]
            got = oracle.face_velocity_tangent(d, xf, wL, wR, *t)
            h = 1e-6
            fp = oracle.face_velocity(d + h * t[0], xf + h * t[1], wL + h * t[2], wR + h * t[3])
            fm = oracle.face_velocity(d - h * t[0], xf - h * t[1], wL - h * t[2], wR - h * t[3])
            fd = (fp - fm) / (2 * h)
            worst = max(worst, float(np.max(np.abs(got - fd)) / (1 + np.max(np.abs(fd)))))
    assert worst < 1e-7, worst


def moving_face(L, R, A, w, riemann):
    s = np.linalg.norm(A)
    return s * oracle.flux(L, R, A / s, w, GAMMA, riemann)


@pytest.mark.parametrize("riemann", ["hllc", "hll"])
@pytest.mark.parametrize("D", [2, 3])
def test_moving_face_tangent_against_central_differences(D, riemann):
    rng = rng_of(17 + D)
    cases = 0
    for flow in (-3.0, -0.5, 0.0, 0.5, 3.0):  # supersonic left / right, subsonic both ways: each wave case
        for _ in range(10):
            n = rng.normal(size=D)
            n /= np.linalg.norm(n)
            L = np.concatenate([[rng.uniform(0.5, 2)], flow * n + 0.3 * rng.normal(size=D), [rng.uniform(0.5, 2)]])
            R = np.concatenate([[rng.uniform(0.5, 2)], flow * n + 0.3 * rng.normal(size=D), [rng.uniform(0.5, 2)]])
            A = n * rng.uniform(0.5, 2)
            w = 0.3 * rng.normal(size=D)
            dL, dR = rng.normal(size=D + 2), rng.normal(size=D + 2)
            dA, dw = rng.normal(size=D), rng.normal(size=D)
            got = oracle.moving_face_tangent(L, R, A, w, dL, dR, dA, dw, GAMMA, riemann)
            h = 1e-6
            fd = (moving_face(L + h * dL, R + h * dR, A + h * dA, w + h * dw, riemann) -
                  moving_face(L - h * dL, R - h * dR, A - h * dA, w - h * dw, riemann)) / (2 * h)
            assert np.max(np.abs(got - fd)) <= 1e-6 * (1 + np.max(np.abs(fd))), (flow, got, fd)
            cases += 1
    assert cases == 50


def _state(m, rng, amp=0.2, drift=0.0):
    D, N = m["dim"], m["cells"]
    V = D + 2
    rho = 1 + amp * rng.uniform(-1, 1, N)
    P = 1 + amp * rng.uniform(-1, 1, N)
    vel = drift + amp * rng.uniform(-1, 1, (N, D))
    vol = m["volume"]
    q = np.zeros((N, V))
    q[:, 0] = rho * vol
    q[:, 1:1 + D] = (rho[:, None] * vel) * vol[:, None]
    q[:, V - 1] = (P / (GAMMA - 1) + 0.5 * rho * np.sum(vel ** 2, axis=1)) * vol
    return q


def _forward(q, x, dt, steps=None, riemann="hllc", order=1):
    m0 = mesh.build(x, 1.0)
    fv = oracle.FiniteVolume(m0, GAMMA, moving=True, riemann=riemann, order=order)
    fv.set_state(q)
    fv.time_step(0.3)
    if steps is not None:
        fv.c.hold_steps = 1
        fv.steps[...] = steps
    xn = fv.moving_begin(dt)
    m1 = mesh.build(xn, 1.0)
    fv.moving_finish(m1, dt)
    return fv.q.copy(), fv.x.copy(), fv.steps.copy(), m0, m1


def _setup(D, n, seed, riemann="hllc", order=1, drift=0.0):
    rng = rng_of(seed)
    m = mesh.build(problems.jittered_lattice(n, D, 0.25, seed), 1.0)
    q = _state(m, rng, drift=drift)
    fv0 = oracle.FiniteVolume(m, GAMMA, moving=True, riemann=riemann, order=order)
    fv0.set_state(q)
    dt = fv0.time_step(0.3)
    Q1, X1, steps, m0, m1 = _forward(q, m["positions"], dt, riemann=riemann, order=order)
    return rng, m, q, dt, steps, m0, m1


def _solver(m0, q, steps, riemann, order):
    fv = oracle.FiniteVolume(m0, GAMMA, moving=True, riemann=riemann, order=order)
    fv.set_state(q)
    fv.c.hold_steps = 1
    fv.steps[...] = steps
    return fv


def _tangent(m0, m1, q, dt, dq, dx, steps, fixed, riemann="hllc", order=1):
    g0, g1 = mesh.GeometryDerivative(m0), mesh.GeometryDerivative(m1)
    return _solver(m0, q, steps, riemann, order).moving_tangent(dt, dq, dx, m1, g0.tangent, g1.tangent, steps_fixed=fixed)


def _adjoint(m0, m1, q, dt, qbar, xbar, steps, fixed, riemann="hllc", order=1):
    g0, g1 = mesh.GeometryDerivative(m0), mesh.GeometryDerivative(m1)
    return _solver(m0, q, steps, riemann, order).moving_adjoint(dt, qbar, xbar, m1, lambda *a: g0.adjoint(*a), lambda *a: g1.adjoint(*a),
                                                                steps_fixed=fixed)


@pytest.mark.parametrize("D,n,riemann,order", [(2, 10, "hllc", 1), (2, 10, "hll", 1), (3, 5, "hllc", 1), (3, 5, "hll", 1),
                                               (2, 10, "hllc", 0), (3, 5, "hll", 0)])
def test_moving_tangent_dot_test_against_the_reverse(D, n, riemann, order):
    rng, m, q, dt, steps, m0, m1 = _setup(D, n, 3, riemann, order)
    N, V = m["cells"], D + 2
    for fixed in (True, False):
        dq, dx = rng.normal(size=(N, V)) * np.abs(q), rng.normal(size=(N, D))
        qbar, xbar = rng.normal(size=(N, V)), rng.normal(size=(N, D))
        tq, tx = _tangent(m0, m1, q, dt, dq, dx, steps, fixed, riemann, order)
        aq, ax = _adjoint(m0, m1, q, dt, qbar, xbar, steps, fixed, riemann, order)
        lhs = np.sum(qbar * tq) + np.sum(xbar * tx)
        rhs = np.sum(aq * dq) + np.sum(ax * dx)
        scale = np.sum(np.abs(qbar * tq)) + np.sum(np.abs(xbar * tx))
        assert abs(lhs - rhs) <= 1e-12 * scale, (fixed, lhs, rhs)


@pytest.mark.parametrize("D,n,riemann", [(2, 10, "hllc"), (3, 5, "hll")])
def test_moving_tangent_against_central_differences(D, n, riemann):
    rng, m, q, dt, steps, m0, m1 = _setup(D, n, 3, riemann)
    N, V, x = m["cells"], D + 2, m["positions"]
    for uq, ux in ((1, 0), (0, 1), (1, 1)):
        dq = rng.normal(size=(N, V)) * np.abs(q) * 1e-7 * uq
        dx = rng.integers(-200, 201, size=(N, D)).astype(float) * m["step"] * ux
        Qp, Xp, _, m0p, m1p = _forward(q + dq, x + dx, dt, steps, riemann)
        Qm, Xm, _, m0m, m1m = _forward(q - dq, x - dx, dt, steps, riemann)
        for a, b in ((m0p, m0), (m0m, m0), (m1p, m1), (m1m, m1)):
            assert np.array_equal(a["owner"], b["owner"]) and np.array_equal(a["neighbour"], b["neighbour"])
        dX = Xp - Xm
        dX -= np.round(dX)
        tq, tx = _tangent(m0, m1, q, dt, dq, dx, steps, True, riemann)
        fq = (Qp - Qm) / 2
        assert np.max(np.abs(tq - fq)) <= 1e-6 * np.max(np.abs(fq)), (uq, ux)
        assert np.max(np.abs(tx - dX / 2)) <= 1e-12 * (1 + np.max(np.abs(dX)))
