"""The oracle's order-0 step on a moving mesh (luna-cpp finite_volume.hpp:216-225: advance<0, true>() with the sweep's
point velocities, then move() off the grid, :1211-1250): a uniform flow on a lattice moving with it stays uniform to the
bit (Galilean invariance of the moving-face flux; the translated lattice has the same cells); conservation; the points
move by dt w, wrapped into the box."""
import math

import numpy as np
import pytest

import cases
import oracle
from luna_hydro import mesh, problems


@pytest.mark.parametrize("dim,n", [(3, 8), (2, 16)])
def test_uniform_flow_on_a_moving_lattice_stays_uniform(dim, n):
    m = cases.voronoi(n, dim, jitter=0.0)
    W = np.zeros((m["cells"], dim + 2))
    W[:, 0], W[:, dim + 1] = 1.0, 1.0
    W[:, 1:dim + 1] = [0.3, -0.2, 0.1][:dim]
    U0 = problems.conserved_rows(W, 5.0 / 3.0)
    fv = oracle.FiniteVolume(m, 5.0 / 3.0, order=0, moving=True)
    fv.set_state(U0 * np.asarray(m["volume"])[:, None])
    for _ in range(3):
        dt = fv.time_step(0.3)
        x0 = fv.x.copy()
        x = fv.moving_begin(dt)
        moved = x0 + dt * fv.w
        assert np.array_equal(x, moved - np.floor(moved))  # unit box
        m = mesh.build(x, 1.0)
        fv.moving_finish(m, dt)
    assert np.max(np.abs(fv.q / np.asarray(m["volume"])[:, None] - U0)) == 0.0


def test_conserves():
    m = cases.voronoi(8, 3)
    q = cases.rough_state(m)
    fv = oracle.FiniteVolume(m, 5.0 / 3.0, order=0, moving=True)
    fv.set_state(q)
    for _ in range(3):
        dt = fv.time_step(0.3)
        m = mesh.build(fv.moving_begin(dt), 1.0)
        fv.moving_finish(m, dt)
    for v in range(5):
        assert abs(math.fsum(fv.q[:, v]) - math.fsum(q[:, v])) <= 1e-14 * math.fsum(np.abs(q[:, v]))


def _forward0(q, x, dt, riemann):
    m0 = mesh.build(x, 1.0)
    fv = oracle.FiniteVolume(m0, 5.0 / 3.0, order=0, moving=True, riemann=riemann)
    fv.set_state(q)
    fv.time_step(0.3)
    xn = fv.moving_begin(dt)
    m1 = mesh.build(xn, 1.0)
    fv.moving_finish(m1, dt)
    return fv.q.copy(), xn.copy(), m0


@pytest.mark.parametrize("D,n,riemann", [(2, 10, "hllc"), (3, 5, "hllc"), (3, 5, "hll")])
def test_order0_reverse_against_central_differences(D, n, riemann):
    """The order-0 moving reverse (finite_volume.hpp:2490-2598; no quantisation: x^n+1 = x^n + dt w) against central
    differences of the order-0 forward along Q, along the positions (whole grid steps) and along both, the topology
    unchanged."""
    from test_oracle_moving_adjoint import _state

    rng = np.random.default_rng(11)
    m = mesh.build(problems.jittered_lattice(n, D, 0.25, 11), 1.0)
    q = _state(m, rng)
    N, V, x = m["cells"], D + 2, np.asarray(m["positions"]).reshape(-1, D)
    fv = oracle.FiniteVolume(m, 5.0 / 3.0, order=0, moving=True, riemann=riemann)
    fv.set_state(q)
    dt = fv.time_step(0.3)
    qbar, xbar = rng.normal(size=(N, V)), rng.normal(size=(N, D))
    g0 = mesh.GeometryDerivative(m)
    qb, xb = fv.moving_adjoint(dt, qbar, xbar, m, lambda *a: g0.adjoint(*a), lambda *a: None)
    for uq, ux in ((1, 0), (0, 1), (1, 1)):
        dq = rng.normal(size=(N, V)) * np.abs(q) * 1e-6 * uq
        dx = rng.integers(-200, 201, size=(N, D)).astype(float) * m["step"] * ux  # whole grid steps: the producer snaps
        Qp, Xp, mp = _forward0(q + dq, x + dx, dt, riemann)
        Qm, Xm, mm = _forward0(q - dq, x - dx, dt, riemann)
        for a in (mp, mm):
            assert np.array_equal(a["owner"], m["owner"]) and np.array_equal(a["neighbour"], m["neighbour"])
        dX = Xp - Xm
        dX -= np.round(dX)
        fd = (np.sum(qbar * (Qp - Qm)) + np.sum(xbar * dX)) / 2
        ad = np.sum(qb * dq) + np.sum(xb * dx)
        assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1e-300), (uq, ux, fd, ad)
