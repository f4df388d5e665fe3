"""The GPU's tangent of the moving step (luna_fv_moving_tangent, orders 0 and 1, fixed topology) against the oracle's (pinned by
central differences and by the dot test against the pinned reverse, tests/test_oracle_moving_tangent.py), on the same meshes
with the mesh library's geometry tangent: 2D and 3D, HLLC and HLL, the grid steps held or free; and the dot test of the GPU's
tangent against the GPU's reverse."""
import numpy as np
import pytest

import oracle
from luna_hydro import hydro, mesh, problems

pytestmark = pytest.mark.gpu
GAMMA = 5.0 / 3.0


def _case(D, n, seed, riemann, order=1):
    rng = np.random.default_rng(seed)
    m0 = mesh.build(problems.jittered_lattice(n, D, 0.25, seed), 1.0)
    N, V = m0["cells"], D + 2
    rho = 1 + 0.2 * rng.uniform(-1, 1, N)
    P = 1 + 0.2 * rng.uniform(-1, 1, N)
    vel = 0.3 + 0.2 * rng.uniform(-1, 1, (N, D))
    vol = m0["volume"]
    q = np.zeros((N, V))
    q[:, 0] = rho * vol
    q[:, 1:1 + D] = (rho[:, None] * vel) * vol[:, None]
    q[:, V - 1] = (P / (GAMMA - 1) + 0.5 * rho * np.sum(vel ** 2, axis=1)) * vol
    fv = hydro.FiniteVolume(m0, GAMMA, riemann=riemann, motion="moving", order=order)
    fv.set_state(q)
    dt = fv.time_step(0.3)
    m1 = mesh.build(fv.moving_begin(dt), 1.0)
    fv.moving_finish(m1, dt)
    return rng, fv, m0, m1, q, dt


def _oracle_tangent(m0, m1, q, dt, dq, dx, fixed, riemann, order):
    fvo = oracle.FiniteVolume(m0, GAMMA, moving=True, riemann=riemann, order=order)
    fvo.set_state(q)
    g0, g1 = mesh.GeometryDerivative(m0), mesh.GeometryDerivative(m1)
    return fvo.moving_tangent(dt, dq, dx, m1, g0.tangent, g1.tangent, steps_fixed=fixed)


@pytest.mark.parametrize("D,n,riemann,order", [(2, 12, "hllc", 1), (2, 12, "hll", 1), (3, 6, "hllc", 1), (3, 8, "hll", 1),
                                               (2, 12, "hllc", 0), (3, 6, "hll", 0)])
@pytest.mark.parametrize("fixed", [True, False])
def test_moving_tangent_matches_oracle(D, n, riemann, order, fixed):
    print("moving tangent D=%d n=%d %s order=%d steps_fixed=%s" % (D, n, riemann, order, fixed), flush=True)
    rng, fv, m0, m1, q, dt = _case(D, n, 11, riemann, order)
    N, V = m0["cells"], D + 2
    dq, dx = rng.normal(size=(N, V)) * np.abs(q), rng.normal(size=(N, D)) * 1e-3
    fv.set_state(q)
    tg, xg = fv.moving_tangent(dt, m0, m1, dq, dx, steps_fixed=fixed)
    to, xo = _oracle_tangent(m0, m1, q, dt, dq, dx, fixed, riemann, order)
    for got, want, name in ((tg, to, "dq"), (xg, xo, "dx")):
        scale = np.max(np.abs(want))
        err = np.max(np.abs(got - want)) / scale
        print("  %s: max |GPU - oracle| / max |oracle| = %.2e (bits equal: %s)" % (name, err, np.array_equal(got, want)), flush=True)
        assert err < 1e-10, (name, err)
    assert np.array_equal(fv.state(), q)  # the handle is back on t^n with Q as set
    # the dot test against the GPU's own reverse
    qbar, xbar = rng.normal(size=(N, V)), rng.normal(size=(N, D))
    fv.set_state(q)
    aq, ax = fv.moving_adjoint(dt, m0, m1, qbar, xbar, steps_fixed=fixed)
    lhs = np.sum(qbar * tg) + np.sum(xbar * xg)
    rhs = np.sum(aq * dq) + np.sum(ax * dx)
    assert abs(lhs - rhs) <= 1e-11 * (np.sum(np.abs(qbar * tg)) + np.sum(np.abs(xbar * xg))), (lhs, rhs)
    fv.close()
