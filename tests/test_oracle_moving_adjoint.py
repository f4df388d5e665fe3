"""The oracle's reverse of the moving step (order 1, fixed topology; finite_volume.hpp:2233-3155 of the reference), with the
mesh library's geometry derivative for both meshes. What pins it (the reference's own checks, hydro_derivatives_test.cpp:
1470-1706, with the project's own meshes):
  - central differences of the oracle's forward with the whole grid steps held, along directions of Q alone, of the
    positions alone (whole grid steps, so no snapping error; the topology unchanged) and of both: 2D and 3D, HLLC and HLL;
  - with the steps free (the quantisation taken as the identity, x^n+1 = x^n + dt w), central differences of the
    quantised forward at a larger step, where the held-steps adjoint is far off (so the point velocity's path counts);
  - conservation: the sums of Q are conserved for any state and mesh, so J^T of a uniform Q cotangent is itself, with no
    position part; a zero cotangent gives zero."""
import numpy as np
import pytest

import oracle
from luna_hydro import mesh, problems

GAMMA = 5.0 / 3.0


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


def _forward(q, x, dt, steps=None, riemann="hllc"):
    m0 = mesh.build(x, 1.0)
    fv = oracle.FiniteVolume(m0, GAMMA, moving=True, riemann=riemann)
    fv.set_state(q)
    fv.time_step(0.3)
    if steps is not None:
        fv.c.hold_steps = 1
        fv.steps[...] = steps
    xn = fv.moving_begin(dt)
    m1 = mesh.build(xn, 1.0)
    fv.moving_finish(m1, dt)
    return fv.q.copy(), fv.x.copy(), fv.steps.copy(), m0, m1


def _setup(D, n, seed, riemann="hllc", drift=0.0):
    rng = np.random.default_rng(seed)
    m = mesh.build(problems.jittered_lattice(n, D, 0.25, seed), 1.0)
    q = _state(m, rng, drift=drift)
    fv0 = oracle.FiniteVolume(m, GAMMA, moving=True, riemann=riemann)
    fv0.set_state(q)
    dt = fv0.time_step(0.3)
    Q1, X1, steps, m0, m1 = _forward(q, m["positions"], dt, riemann=riemann)
    return rng, m, q, dt, steps, m0, m1


def _adjoint(m0, m1, q, dt, qbar, xbar, steps, fixed, riemann="hllc"):
    fv = oracle.FiniteVolume(m0, GAMMA, moving=True, riemann=riemann)
    fv.set_state(q)
    fv.c.hold_steps = 1
    fv.steps[...] = steps
    g0, g1 = mesh.GeometryDerivative(m0), mesh.GeometryDerivative(m1)
    return fv.moving_adjoint(dt, qbar, xbar, m1, lambda *a: g0.adjoint(*a), lambda *a: g1.adjoint(*a), steps_fixed=fixed)


@pytest.mark.parametrize("D,n,riemann", [(2, 10, "hllc"), (2, 10, "hll"), (3, 5, "hllc"), (3, 5, "hll")])
def test_moving_adjoint_against_central_differences(D, n, riemann):
    rng, m, q, dt, steps, m0, m1 = _setup(D, n, 3, riemann)
    N, V, x = m["cells"], D + 2, m["positions"]
    qbar, xbar = rng.normal(size=(N, V)), rng.normal(size=(N, D))
    qb, xb = _adjoint(m0, m1, q, dt, qbar, xbar, steps, True, riemann)
    for uq, ux in ((1, 0), (0, 1), (1, 1)):
        dq = rng.normal(size=(N, V)) * np.abs(q) * 1e-7 * uq
        dx = rng.integers(-200, 201, size=(N, D)).astype(float) * m["step"] * ux
        Qp, Xp, _, m0p, m1p = _forward(q + dq, x + dx, dt, steps, riemann)
        Qm, Xm, _, m0m, m1m = _forward(q - dq, x - dx, dt, steps, riemann)
        for a, b in ((m0p, m0), (m0m, m0), (m1p, m1), (m1m, m1)):
            assert np.array_equal(a["owner"], b["owner"]) and np.array_equal(a["neighbour"], b["neighbour"])
        dX = Xp - Xm
        dX -= np.round(dX)
        # the Q part alone (the x part passes through the held steps unchanged)
        fd_q = np.sum(qbar * (Qp - Qm)) / 2
        ad_q = np.sum(_adjoint(m0, m1, q, dt, qbar, np.zeros((N, D)), steps, True, riemann)[0] * dq) + \
            np.sum(_adjoint(m0, m1, q, dt, qbar, np.zeros((N, D)), steps, True, riemann)[1] * dx)
        assert abs(fd_q - ad_q) < 1e-6 * abs(fd_q), (uq, ux, fd_q, ad_q)
        fd = fd_q + np.sum(xbar * dX) / 2
        assert abs(fd - (np.sum(qb * dq) + np.sum(xb * dx))) < 1e-6 * abs(fd)


@pytest.mark.parametrize("D,n", [(2, 10), (3, 5)])
def test_moving_adjoint_with_free_steps(D, n):
    rng, m, q, dt, steps, m0, m1 = _setup(D, n, 3, drift=0.5)
    N, V, x = m["cells"], D + 2, m["positions"]
    qbar, xbar = rng.normal(size=(N, V)), rng.normal(size=(N, D))
    free = _adjoint(m0, m1, q, dt, qbar, xbar, steps, False)
    held = _adjoint(m0, m1, q, dt, qbar, xbar, steps, True)
    dq = rng.normal(size=(N, V)) * np.abs(q) * 3e-4
    Qp, Xp, sp, *_ = _forward(q + dq, x, dt)
    Qm, Xm, sm, *_ = _forward(q - dq, x, dt)
    assert np.max(np.abs(sp - sm)) > 1000  # the steps move by thousands of grid steps: the quantisation is a small part
    dX = Xp - Xm
    dX -= np.round(dX)
    fd = (np.sum(qbar * (Qp - Qm)) + np.sum(xbar * dX)) / 2
    assert abs(np.sum(free[0] * dq) - fd) < 1e-2 * abs(fd)
    assert abs(np.sum(held[0] * dq) - fd) > 0.3 * abs(fd)


@pytest.mark.parametrize("D,n", [(2, 10), (3, 5)])
def test_moving_adjoint_conservation_and_zero(D, n):
    rng, m, q, dt, steps, m0, m1 = _setup(D, n, 5)
    N, V = m["cells"], D + 2
    for v in range(V):
        qbar = np.zeros((N, V))
        qbar[:, v] = 1.0
        for fixed in (True, False):
            qb, xb = _adjoint(m0, m1, q, dt, qbar, np.zeros((N, D)), steps, fixed)
            assert np.array_equal(qb, qbar) and not np.any(xb)  # each face's G bar is -dt/2 (qbar_o - qbar_j) = 0 exactly
    qb, xb = _adjoint(m0, m1, q, dt, np.zeros((N, V)), np.zeros((N, D)), steps, False)
    assert not np.any(qb) and not np.any(xb)
